"""Pins of the oracle's error estimates and drag that do not go through the
oracle's own formulas.

* RKF7(8) error estimate (int_rungekutta8.cu:44-46 weights, 250-260 error
  terms, 693-696 norm): the estimate lambda |dt| max|k0 + k10 - k11 - k12|
  must approach the TRUE local error of the 7th-order step, which here comes
  from an exact two-body propagation (Kepler's equation solved in long
  double).  A wrong lambda shifts the ratio away from 1; a dropped or
  sign-flipped term leaves an O(dt) remainder instead of O(dt^8).
* RK4's embedded estimate (int_rungekutta4.cu:216-229) must scale as dt^4
  (the local error of the third-order embedded solution).
* Quadratic drag (pp_disk.cu:87-155, 486-521) against
  a = 1/2 C_D rho_g pi R^2 |u| u / (4/3 pi R^3 rho_p), with rho_g and the gas
  velocity written from the disk model in long double, at points outside the
  0.1 AU taper, inside it and off the midplane.
* The gas-disk parameters used for the N = 131072 gas workload: rho_g
  integrated over z must give Sigma = 1700 r^-1.5 g/cm^2.
"""
import ctypes as C
import os
import sys

import numpy as np
import pytest

import oracle_lib as O
from redamd_types import VAR4, PARAM, BODY_MD, GasParams
from test_oracle_physics import kepler_binary

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "red.cuda_amd"))

L = np.longdouble
PI = 4 * np.arctan(L(1))   # pi in long double (np.pi is a double)


def kepler_exact(r, v, p, dt, period=False):
    """Barycentric two-body state after dt from Kepler's equation (f and g
    functions in the eccentric anomaly), in long double; with period=True the
    orbit's period instead."""
    m0, m1 = L(p["mass"][0]), L(p["mass"][1])
    M = m0 + m1
    x0 = np.array([r[c][1] - r[c][0] for c in "xyz"], L)
    u0 = np.array([v[c][1] - v[c][0] for c in "xyz"], L)
    r0 = np.sqrt((x0 * x0).sum())
    a = 1 / (2 / r0 - (u0 * u0).sum() / M)
    n = np.sqrt(M / a ** 3)
    if period:
        return 2 * PI / n
    ecosE, esinE = 1 - r0 / a, (x0 * u0).sum() / np.sqrt(M * a)
    e = np.sqrt(ecosE ** 2 + esinE ** 2)
    E0 = np.arctan2(esinE, ecosE)
    Mt = E0 - esinE + n * L(dt)
    E = Mt
    for _ in range(60):
        E = E - (E - e * np.sin(E) - Mt) / (1 - e * np.cos(E))
    dE = E - E0
    f = 1 - a / r0 * (1 - np.cos(dE))
    g = L(dt) - (dE - np.sin(dE)) / n
    x1 = f * x0 + g * u0
    r1 = np.sqrt((x1 * x1).sum())
    u1 = (-np.sqrt(M * a) / (r1 * r0) * np.sin(dE)) * x0 + (1 - a / r1 * (1 - np.cos(dE))) * u0
    pos = np.zeros((2, 4), L)
    vel = np.zeros((2, 4), L)
    pos[0, :3], pos[1, :3] = -m1 / M * x1, m0 / M * x1
    vel[0, :3], vel[1, :3] = -m1 / M * u1, m0 / M * u1
    return np.concatenate([pos.ravel(), vel.ravel()])


def flat_state(d):
    _, rr, vv, _, _ = d.state()
    return np.concatenate([np.stack([rr[c] for c in "xyzw"], 1).ravel(),
                           np.stack([vv[c] for c in "xyzw"], 1).ravel()])


def one_attempt(method, e, div):
    r, v, p, b, nbt, P = kepler_binary(e)
    d = O.ODisk(r, v, p, b, nbt)
    dt = P / div
    it = O.OInteg(d, method, dt, adaptive=True, tol=1.0)   # accepted at once
    assert it.step() == dt
    err, terms = it.last_error()
    true = np.abs(flat_state(d).astype(L) - kepler_exact(r, v, p, dt)).astype(np.float64)
    return dt, err, terms, true


def test_kepler_exact_is_exact():
    # the propagator itself: a full period returns to the start, a half
    # period of the circular orbit is the antipode (to 2 a e: the double
    # initial conditions leave e ~ 1e-16)
    r, v, p, b, nbt, _ = kepler_binary(0.6)
    P = kepler_exact(r, v, p, 0.0, period=True)
    start = kepler_exact(r, v, p, 0.0)
    assert np.abs(kepler_exact(r, v, p, P) - start).max() < 1e-16
    r, v, p, b, nbt, _ = kepler_binary(0.0)
    P = kepler_exact(r, v, p, 0.0, period=True)
    start = kepler_exact(r, v, p, 0.0)
    half = kepler_exact(r, v, p, P / 2)
    assert np.abs(half + start).max() < 1e-14


def test_rkf78_estimate_is_the_true_local_error():
    ratios = []
    errs = []
    for div in (24, 48):
        dt, err, terms, true = one_attempt(4, 0.0, div)
        # the controller's error is lambda |dt| times the largest term
        assert err == abs(41.0 / 840.0 * dt * terms.max())
        ratios.append(err / true.max())
        errs.append(err)
        # component by component where the error is not tiny
        m = true > 0.1 * true.max()
        cw = 41.0 / 840.0 * dt * terms[m] / true[m]
        assert np.all(np.abs(cw - 1.0) < 0.1), cw
    # y8 - y7 = (y_exact - y7) + O(dt^9): the ratio tends to 1 as dt halves
    assert abs(ratios[0] - 1.0) < 0.08 and abs(ratios[1] - 1.0) < 0.02, ratios
    assert abs(ratios[1] - 1.0) < 0.5 * abs(ratios[0] - 1.0), ratios
    # and the estimate is O(dt^8)
    slope = np.log2(errs[0] / errs[1])
    assert 7.6 < slope < 8.4, slope


def test_rkf78_estimate_on_an_eccentric_orbit():
    # the same on e = 0.3 in the max norm, more loosely (the O(dt^9) part is
    # larger near periapsis); a wrong lambda (e.g. 1/10) doubles the ratio
    for div in (32, 48):
        dt, err, terms, true = one_attempt(4, 0.3, div)
        assert 0.9 < err / true.max() < 1.2, (div, err / true.max())


def test_rk4_embedded_estimate_order():
    # err = |dt/10 (k3 - k4)| is the local error of the embedded third-order
    # solution: O(dt^4), while RK4's own local error is O(dt^5)
    out = [one_attempt(2, 0.0, div) for div in (64, 128)]
    slope = np.log2(out[0][1] / out[1][1])
    assert 3.7 < slope < 4.3, slope
    true_slope = np.log2(out[0][3].max() / out[1][3].max())
    assert 4.6 < true_slope < 5.4, true_slope
    for dt, err, terms, true in out:
        assert err == abs(0.1 * dt * terms.max())
        assert err > 10 * true.max()


# ---- drag --------------------------------------------------------------------

def gas_rho(g, x, y, z):
    """rho_g of the analytic disk (pp_disk.cu:87-105) in long double: a power
    law times exp(-(z/h)^2), h = sch0 r^sch1, with the power law replaced by
    rho(0.1) (r/0.1)^4 inside 0.1 AU."""
    r = np.sqrt(L(x) ** 2 + L(y) ** 2)
    h = L(g.sch[0]) * r ** L(g.sch[1])
    vert = np.exp(-(L(z) / h) ** 2)
    if r > L("0.1"):
        return L(g.rho[0]) * r ** L(g.rho[1]) * vert
    return L(g.rho[0]) * L("0.1") ** L(g.rho[1]) * (r / L("0.1")) ** 4 * vert


def drag_accel(g, m_star, pos, vel, cd, radius, dens_p):
    """a = F / m with F = 1/2 C_D rho_g (pi R^2) |u| u and m = 4/3 pi R^3 rho_p,
    u = v_gas - v; the gas moves on circles at v_K sqrt(1 - 2 eta r^eta_p)
    counter-clockwise about the z axis."""
    x, y, z = (L(c) for c in pos)
    r = np.sqrt(x * x + y * y)
    vg = np.sqrt(L(m_star) / r) * np.sqrt(1 - 2 * L(g.eta[0]) * r ** L(g.eta[1]))
    u = np.array([-vg * y / r, vg * x / r, L(0)]) - np.array([L(c) for c in vel])
    un = np.sqrt((u * u).sum())
    area = PI * L(radius) ** 2
    mass = L(4) / 3 * PI * L(radius) ** 3 * L(dens_p)
    return L("0.5") * L(cd) * gas_rho(g, x, y, z) * area * un * u / mass


def test_drag_against_quadratic_drag_law():
    import redamd as R
    g = R.gas_params()
    # star + planetesimals: outside the taper on the midplane, off the
    # midplane, inside the taper on and off the midplane, all four quadrants
    pts = [((1.3, 0.4, 0.0), 1.0), ((-2.2, 1.1, 0.02), 0.5), ((0.7, -3.1, -0.09), 2.0),
           ((-0.04, -0.05, 0.0), 1.0), ((0.06, 0.03, 0.004), 0.7), ((4.5, 0.0, 0.3), 1.0)]
    n = len(pts) + 1
    rng = np.random.default_rng(11)
    r = np.zeros(n, VAR4)
    v = np.zeros(n, VAR4)
    p = np.zeros(n, PARAM)
    b = np.zeros(n, BODY_MD)
    p[0] = (1.0, 4.6e-3, 1.0, 0.0)
    b[0] = (1, 0, 0, 0.0)
    for k, (pos, cd) in enumerate(pts, 1):
        r[k] = (*pos, 0.0)
        rr = np.hypot(pos[0], pos[1])
        vk = np.sqrt(1.0 / rr)
        # near-Keplerian velocity with a random 1% offset (|u| ~ the headwind)
        v[k] = (-vk * pos[1] / rr * (1 + 0.01 * rng.standard_normal()),
                vk * pos[0] / rr * (1 + 0.01 * rng.standard_normal()), 0.01 * vk * rng.standard_normal(), 0.0)
        rad = 10.0 ** rng.uniform(np.log10(1.0), np.log10(100.0)) / 1.495978707e8   # 1-100 km in AU
        p[k] = (1e-12, rad, 2.0 * 1.495978707e13 ** 3 / 1.98911e33, cd)            # 2 g/cm^3 in Msun/AU^3
        b[k] = (k + 1, 5, 0, 0.0)
    nbt = np.array([1, 0, 0, 0, 0, n - 1, 0], np.uint32)
    a = np.zeros(n, VAR4)
    O.lib.or_drag(0.0, C.byref(g), 0, n, O.ptr(nbt), O.ptr(p), O.ptr(r), O.ptr(v), O.ptr(a))
    for k in range(1, n):
        ref = drag_accel(g, p["mass"][0], (r["x"][k], r["y"][k], r["z"][k]), (v["x"][k], v["y"][k], v["z"][k]),
                         p["cd"][k], p["radius"][k], p["density"][k])
        got = np.array([a["x"][k], a["y"][k], a["z"][k]], L)
        # u = v_gas - v cancels ~2 digits (|u| ~ 1% of |v|), so double
        # rounding shows at ~1e-14; a wrong factor or exponent is >= 1e-2 off
        assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max(), (k, got, ref)


def test_drag_decrease_laws():
    # gas_decrease (pp_disk.cu:60-84): linear between t0 and t1, exponential
    # with the e-folding time; the drag scales with the factor
    import redamd as R
    base = R.gas_params()
    n = 2
    r = np.zeros(n, VAR4)
    v = np.zeros(n, VAR4)
    p = np.zeros(n, PARAM)
    b = np.zeros(n, BODY_MD)
    p[0] = (1.0, 4.6e-3, 1.0, 0.0)
    p[1] = (1e-12, 1e-7, 2.0e6, 1.0)
    b[0], b[1] = (1, 0, 0, 0.0), (2, 5, 0, 0.0)
    r[1] = (1.0, 0.2, 0.01, 0.0)
    v[1] = (0.0, 0.9, 0.0, 0.0)
    nbt = np.array([1, 0, 0, 0, 0, 1, 0], np.uint32)

    def ax(g, t):
        a = np.zeros(n, VAR4)
        O.lib.or_drag(t, C.byref(g), 0, n, O.ptr(nbt), O.ptr(p), O.ptr(r), O.ptr(v), O.ptr(a))
        return a["x"][1]

    full = ax(base, 0.0)
    lin = R.gas_params()
    lin.gas_decrease, lin.t0, lin.t1 = 1, 10.0, 30.0
    assert ax(lin, 5.0) == full
    assert abs(ax(lin, 25.0) / full - 0.25) < 1e-14
    assert ax(lin, 31.0) == 0.0
    ex = R.gas_params()
    ex.gas_decrease, ex.t0, ex.e_folding_time = 2, 10.0, 7.0
    assert abs(ax(ex, 24.0) / full - np.exp(-2.0)) < 1e-14


def test_config3_gas_disk_integrates_to_sigma():
    """Sigma = 1700 r^-1.5 g/cm^2, H/r = 0.05 r^0.25: the engine's gas
    parameters give back Sigma when rho_g is integrated over z."""
    import redamd as R
    from scipy.integrate import quad
    g = R.gas_params()
    assert (g.sch[0], g.sch[1]) == (0.05, 1.25)
    au_cm, msun_g = 1.495978707e13, 1.98911e33   # red_constants.h
    for rr in (0.3, 1.0, 2.5, 5.0):
        h = 0.05 * rr ** 1.25
        col, _ = quad(lambda z: float(gas_rho(g, rr, 0.0, z)), -12 * h, 12 * h, epsabs=0, epsrel=1e-12, limit=200)
        sigma = col * msun_g / au_cm ** 2
        assert abs(sigma / (1700.0 * rr ** -1.5) - 1.0) < 1e-9, (rr, sigma)
    # and the midplane density at 1 AU in g/cm^3 (Sigma / (sqrt(pi) H))
    rho1 = float(gas_rho(g, 1.0, 0.0, 0.0)) * msun_g / au_cm ** 3
    assert abs(rho1 / (1700.0 / (np.sqrt(np.pi) * 0.05 * au_cm)) - 1.0) < 1e-12


# ---- step controller -----------------------------------------------------------

@pytest.mark.parametrize("tol", [1e-12, 1e-10])
def test_rkf78_controller_fixed_point(tol):
    """dt *= 0.9 (tol/err)^(1/q) (int_rungekutta8.cu:697-704) on an estimate
    err ~ C dt^8 settles where err/tol = 0.9^q: with q = 7, 0.478.  On a
    circular orbit C is constant up to the max-norm of a rotating vector, so
    the geometric mean over whole orbits sits at the fixed point.  A safety
    factor of 0.8 would give 0.21, an exponent of 1/8 0.43."""
    r, v, p, b, nbt, P = kepler_binary(0.0)
    d = O.ODisk(r, v, p, b, nbt)
    it = O.OInteg(d, 4, P / 40, adaptive=True, tol=tol)
    logs = []
    for _ in range(2000):   # ~250 steps at tol 1e-12; a broken estimate may need far more
        if d.t >= 4 * P:
            break
        it.step()
        if d.t > 2 * P:
            logs.append(np.log(it.last_error()[0] / tol))
    assert d.t >= 4 * P and logs
    assert abs(np.exp(np.mean(logs)) / 0.9 ** 7 - 1.0) < 0.005, np.exp(np.mean(logs))
    tried, passed, failed = it.counters()
    assert tried == passed + failed and failed <= 1


def test_rk4_controller_fixed_point():
    # the adaptive RK4 estimate is O(dt^4) and its exponent is 1/4
    # (int_rungekutta4.cu:229-237): one step lands on err/tol = 0.9^4
    r, v, p, b, nbt, P = kepler_binary(0.0)
    d = O.ODisk(r, v, p, b, nbt)
    it = O.OInteg(d, 2, P / 40, adaptive=True, tol=1e-10)
    for _ in range(40):
        it.step()
    assert abs(it.last_error()[0] / 1e-10 / 0.9 ** 4 - 1.0) < 0.005


@pytest.mark.parametrize("method", [2, 4])
def test_zero_error_doubles_the_step(method):
    # a lone star in uniform motion: every stage derivative is the same, the
    # error terms cancel exactly and the step doubles (int_rungekutta8.cu:705-708)
    r, v, p, b = np.zeros(1, VAR4), np.zeros(1, VAR4), np.zeros(1, PARAM), np.zeros(1, BODY_MD)
    r[0] = (0.1, 0.2, 0.3, 0.0)
    v[0] = (0.01, -0.02, 0.005, 0.0)
    p[0] = (1.0, 4.6e-3, 1.0, 0.0)
    b[0] = (1, 0, 0, 0.0)
    d = O.ODisk(r, v, p, b, np.array([1, 0, 0, 0, 0, 0, 0], np.uint32))
    it = O.OInteg(d, method, 0.1, adaptive=True, tol=1e-12)
    for k in range(4):
        assert it.step() == 0.1 * 2 ** k
        assert it.last_error()[0] == 0.0 and it.dt_next() == 0.1 * 2 ** (k + 1)
