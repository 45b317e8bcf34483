"""The oracle's force, drag, integrators and events against what the
mathematics fixes: brute force in long double, closed forms, invariants,
convergence order.  A wrong sign, index or dropped term fails one of these."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
from redamd_types import VAR4, PARAM, BODY_MD, GasParams
from systems import small_system, as3


def brute(r, p, b, nbt):
    """Long-double all-pairs over the reference classes (red_type.h:670-700)."""
    x = np.stack([r["x"], r["y"], r["z"]], 1).astype(np.longdouble)
    m = p["mass"].astype(np.longdouble)
    n_si = int(nbt[:4].sum())
    n_mas = n_si + int(nbt[4:6].sum())
    a = np.zeros((len(r), 3), np.longdouble)
    for i in range(len(r)):
        if b["id"][i] <= 0:
            continue
        hi = n_mas if (i < n_si or i >= n_mas) else n_si
        for j in range(hi):
            if j == i or b["id"][j] < 0:
                continue
            d = x[j] - x[i]
            a[i] += m[j] * d / np.sqrt((d * d).sum()) ** 3
    return a


@pytest.mark.parametrize("types,inactive", [
    (None, ()),
    (None, (3, 7)),
    ([0, 1, 3, 3, 3, 5, 5, 5, 6, 6], (4,)),
])
def test_gravity_gpu_formula_vs_long_double(types, inactive):
    n = 10 if types else 24
    r, v, p, b, nbt = small_system(n, 1, types, inactive)
    a = np.zeros(n, VAR4)
    O.lib.or_grav_gpu(n, O.ptr(nbt), O.ptr(b), O.ptr(p), O.ptr(r), O.ptr(a))
    ref = brute(r, p, b, nbt)
    got = as3(a)
    scale = np.abs(ref).max(axis=1, keepdims=True) + 1e-300
    assert np.all(np.abs(got - ref.astype(np.float64)) <= 1e-14 * scale)
    for i in inactive:
        assert np.all(got[i] == 0.0)


def test_gravity_cpu_symmetric_matches_gpu_formula_to_rounding():
    r, v, p, b, nbt = small_system(64, 2)
    a1, a2 = np.zeros(64, VAR4), np.zeros(64, VAR4)
    O.lib.or_grav_gpu(64, O.ptr(nbt), O.ptr(b), O.ptr(p), O.ptr(r), O.ptr(a1))
    assert O.lib.or_grav_cpu(64, O.ptr(nbt), O.ptr(b), O.ptr(p), O.ptr(r), O.ptr(a2)) == 0
    g1, g2 = as3(a1), as3(a2)
    assert np.all(np.abs(g1 - g2) <= 1e-13 * np.abs(g1).max(axis=1, keepdims=True))


def test_gravity_cpu_refuses_nsi():
    r, v, p, b, nbt = small_system(10, 3, [0, 3, 3, 3, 5, 5, 5, 5, 5, 5])
    a = np.zeros(10, VAR4)
    assert O.lib.or_grav_cpu(10, O.ptr(nbt), O.ptr(b), O.ptr(p), O.ptr(r), O.ptr(a)) == 3


def test_newton3_and_two_body():
    r, v, p, b, nbt = small_system(32, 4)
    a = np.zeros(32, VAR4)
    O.lib.or_grav_gpu(32, O.ptr(nbt), O.ptr(b), O.ptr(p), O.ptr(r), O.ptr(a))
    f = as3(a) * p["mass"][:, None]
    assert np.all(np.abs(f.sum(0)) <= 1e-14 * np.abs(f).max())
    # two bodies: |a_1| = m_2 / d^2 along the separation
    r2 = r[:2].copy()
    r2[1] = (1.5, 0.0, 0.0, 0.0)
    r2[0] = (0.0, 0.0, 0.0, 0.0)
    a2 = np.zeros(2, VAR4)
    nb = np.array([1, 0, 0, 1, 0, 0, 0], np.uint32)
    O.lib.or_grav_gpu(2, O.ptr(nb), O.ptr(b[:2].copy()), O.ptr(p[:2].copy()), O.ptr(r2), O.ptr(a2))
    assert abs(a2["x"][0] - p["mass"][1] / 2.25) <= 1e-16 * abs(a2["x"][0])
    assert abs(a2["x"][1] + p["mass"][0] / 2.25) <= 1e-16


def test_symmetric_ring_zero_net_force_on_centre():
    n = 9
    r, v, p, b, nbt = small_system(n, 5)
    ang = np.arange(8) * np.pi / 4
    r["x"][0] = r["y"][0] = r["z"][0] = 0.0
    r["x"][1:], r["y"][1:], r["z"][1:] = np.cos(ang), np.sin(ang), 0.0
    p["mass"][1:] = 1e-6
    a = np.zeros(n, VAR4)
    O.lib.or_grav_gpu(n, O.ptr(nbt), O.ptr(b), O.ptr(p), O.ptr(r), O.ptr(a))
    assert abs(a["x"][0]) < 1e-21 and abs(a["y"][0]) < 1e-21


def gas():
    g = GasParams()
    g.rho[0], g.rho[1] = 1e-9, -2.75
    g.sch[0], g.sch[1] = 0.05, 1.25
    g.eta[0], g.eta[1] = 0.002, 0.0
    return g


def test_drag_closed_forms():
    n = 6
    r, v, p, b, nbt = small_system(n, 6, [0, 3, 5, 5, 5, 5])
    g = gas()
    a = np.zeros(n, VAR4)
    O.lib.or_drag(0.0, C.byref(g), 0, n, O.ptr(nbt), O.ptr(p), O.ptr(r), O.ptr(v), O.ptr(a))
    # only NSI bodies feel the gas (get_bound_GD)
    assert np.all(as3(a)[:2] == 0.0) and np.all(np.abs(as3(a)[2:]).sum(1) > 0)
    # a body moving with the gas feels nothing; doubling |u| quadruples |a|
    k = 3
    R = np.hypot(r["x"][k], r["y"][k])
    vc = np.sqrt(p["mass"][0] / R) * np.sqrt(1 - 2 * 0.002)
    v2 = v.copy()
    v2["x"][k], v2["y"][k], v2["z"][k] = -vc * r["y"][k] / R, vc * r["x"][k] / R, 0.0
    a2 = np.zeros(n, VAR4)
    O.lib.or_drag(0.0, C.byref(g), 0, n, O.ptr(nbt), O.ptr(p), O.ptr(r), O.ptr(v2), O.ptr(a2))
    assert np.abs(as3(a2)[k]).max() < 1e-12 * np.abs(as3(a)[k]).max()
    v3 = v2.copy()
    v3["z"][k] = 1e-3
    v4 = v2.copy()
    v4["z"][k] = 2e-3
    a3, a4 = np.zeros(n, VAR4), np.zeros(n, VAR4)
    O.lib.or_drag(0.0, C.byref(g), 0, n, O.ptr(nbt), O.ptr(p), O.ptr(r), O.ptr(v3), O.ptr(a3))
    O.lib.or_drag(0.0, C.byref(g), 0, n, O.ptr(nbt), O.ptr(p), O.ptr(r), O.ptr(v4), O.ptr(a4))
    assert abs(a4["z"][k] / a3["z"][k] - 4.0) < 1e-6
    assert a3["z"][k] < 0.0   # opposes the relative motion
    # gas switched off below the 1e-6 reduction threshold (pp_disk.cu:21, 502-505)
    g.gas_decrease = 2
    g.t0, g.e_folding_time = 0.0, 1.0
    a5 = np.zeros(n, VAR4)
    O.lib.or_drag(20.0, C.byref(g), 0, n, O.ptr(nbt), O.ptr(p), O.ptr(r), O.ptr(v), O.ptr(a5))
    assert np.all(as3(a5) == 0.0)


def kepler_binary(e=0.3):
    """Star + one protoplanet on an eccentric orbit, barycentric, G = 1."""
    m0, m1, a = 1.0, 1e-3, 1.0
    M = m0 + m1
    rp = a * (1 - e)
    vp = np.sqrt(M * (1 + e) / rp)
    r = np.zeros(2, VAR4)
    v = np.zeros(2, VAR4)
    r["x"] = [-m1 / M * rp, m0 / M * rp]
    v["y"] = [-m1 / M * vp, m0 / M * vp]
    p = np.zeros(2, PARAM)
    p["mass"] = [m0, m1]
    b = np.zeros(2, BODY_MD)
    b["id"] = [1, 2]
    b["body_type"] = [0, 3]
    nbt = np.array([1, 0, 0, 1, 0, 0, 0], np.uint32)
    P = 2 * np.pi * np.sqrt(a ** 3 / M)
    return r, v, p, b, nbt, P


@pytest.mark.parametrize("method,order,steps", [(0, 1, (4096, 8192)), (1, 2, (512, 1024)), (2, 4, (64, 128)),
                                               (4, 7, (64, 128))])
def test_fixed_step_convergence_order(method, order, steps):
    # Euler (int_euler.cu), midpoint RK2 (int_rungekutta2.cu), RK4, RKF7(8): the
    # error after one period halves 2^order times when the step halves
    errs = []
    for nsteps in steps:
        r, v, p, b, nbt, P = kepler_binary()
        d = O.ODisk(r, v, p, b, nbt)
        it = O.OInteg(d, method, P / nsteps, adaptive=False)
        for _ in range(nsteps):
            it.step()
        _, rr, vv, _, _ = d.state()
        errs.append(np.hypot(rr["x"][1] - r["x"][1], rr["y"][1] - r["y"][1]))
    rate = np.log2(errs[0] / errs[1])
    assert rate > order - 0.6, (errs, rate)
    assert rate < order + 0.6, (errs, rate)


@pytest.mark.parametrize("method", [0, 1])
def test_euler_rk2_keep_their_step(method):
    # euler::step and rungekutta2::step use dt_try, the constructor's dt, and
    # set dt_next back to it (int_euler.cu:78-105, int_rungekutta2.cu:99-131):
    # the driver's clamp to the output interval does not shorten their steps
    r, v, p, b, nbt, P = kepler_binary()
    d = O.ODisk(r, v, p, b, nbt)
    h = P / 500
    it = O.OInteg(d, method, h, adaptive=True, tol=1e-12)
    assert it.step() == h
    O.lib.or_set_dt_next(it.h, h / 7)
    assert it.step() == h and it.dt_next() == h
    assert it.counters() == (2, 2, 0)
    assert d.t == h + h


def test_rkf78_adaptive_closed_orbit_and_controller():
    r, v, p, b, nbt, P = kepler_binary(0.6)
    d = O.ODisk(r, v, p, b, nbt)
    it = O.OInteg(d, 4, P / 50, adaptive=True, tol=1e-12)
    while d.t < P:
        it.step()
    tried, passed, failed = it.counters()
    assert tried == passed + failed and failed > 0
    E0 = d.integral()["E"]
    assert np.isfinite(E0)


@pytest.mark.parametrize("method", [2, 4])
def test_backward_step_mirrors_forward(method):
    """dt < 0 (main.cu sets it for a negative simulation length): without drag
    the right-hand side is odd in v, so a backward run from (r, v) is the
    forward run from (r, -v) with r, t and v mirrored.  Every operation on the
    path is symmetric under a sign flip, so the mirror is exact; the error
    norm takes |lambda dt err| (int_rungekutta8.cu:696) and the controller's
    next step keeps the sign."""
    r, v, p, b, nbt = small_system(48, 17)
    vm = v.copy()
    for c in "xyz":
        vm[c] = -v[c]
    fwd = O.ODisk(r, vm, p, b, nbt)
    bwd = O.ODisk(r, v, p, b, nbt)
    h = 0.05
    fi = O.OInteg(fwd, method, h, adaptive=True, tol=1e-12)
    bi = O.OInteg(bwd, method, -h, adaptive=True, tol=1e-12)
    for _ in range(6):
        hf = fi.step()
        assert bi.step() == -hf and hf > 0.0
        assert bi.dt_next() == -fi.dt_next()
    assert fi.counters() == bi.counters()
    tf, rf, vf, _, _ = fwd.state()
    tb, rb, vb, _, _ = bwd.state()
    assert tb == -tf
    for c in "xyz":
        assert np.array_equal(rf[c], rb[c])
        assert np.array_equal(vf[c], -vb[c])


def test_backward_rkf78_retraces_the_orbit():
    # one period forward then one period back along the same eccentric orbit
    r, v, p, b, nbt, P = kepler_binary(0.3)
    d = O.ODisk(r, v, p, b, nbt)
    it = O.OInteg(d, 4, P / 128, adaptive=False)
    for _ in range(128):
        it.step()
    back = O.OInteg(d, 4, -P / 128, adaptive=False)
    for _ in range(128):
        back.step()
    _, rr, vv, _, _ = d.state()
    assert abs(d.t) < 1e-12 * P
    assert np.hypot(rr["x"][1] - r["x"][1], rr["y"][1] - r["y"][1]) < 1e-10
    assert np.hypot(vv["x"][1] - v["x"][1], vv["y"][1] - v["y"][1]) < 1e-10


def test_rkf78_tableau_row_sums():
    """The oracle's own Fehlberg 7(8) table (int_rungekutta8.cu:26-48) against
    the published fractions (Fehlberg, NASA TR R-287), plus what the method
    fixes on it: sum_j a_rj = c_r, and the quadrature conditions
    sum_i b_i c_i^k = 1/(k+1) for k = 0..7 of a 7th-order formula."""
    from fractions import Fraction as F
    rows = {1: [F(2, 27)], 2: [F(1, 36), F(1, 12)], 3: [F(1, 24), 0, F(1, 8)],
            4: [F(5, 12), 0, F(-25, 16), F(25, 16)], 5: [F(1, 20), 0, 0, F(1, 4), F(1, 5)],
            6: [F(-25, 108), 0, 0, F(125, 108), F(-65, 27), F(125, 54)],
            7: [F(31, 300), 0, 0, 0, F(61, 225), F(-2, 9), F(13, 900)],
            8: [2, 0, 0, F(-53, 6), F(704, 45), F(-107, 9), F(67, 90), 3],
            9: [F(-91, 108), 0, 0, F(23, 108), F(-976, 135), F(311, 54), F(-19, 60), F(17, 6), F(-1, 12)],
            10: [F(2383, 4100), 0, 0, F(-341, 164), F(4496, 1025), F(-301, 82), F(2133, 4100), F(45, 82),
                 F(45, 164), F(18, 41)],
            11: [F(3, 205), 0, 0, 0, 0, F(-6, 41), F(-3, 205), F(-3, 41), F(3, 41), F(6, 41), 0],
            12: [F(-1777, 4100), 0, 0, F(-341, 164), F(4496, 1025), F(-289, 82), F(2193, 4100), F(51, 82),
                 F(33, 164), F(12, 41), 0, 1]}
    cp = [0, F(2, 27), F(1, 9), F(1, 6), F(5, 12), F(1, 2), F(5, 6), F(1, 6), F(2, 3), F(1, 3), 1, 0, 1]
    bp = [F(41, 840), 0, 0, 0, 0, F(34, 105), F(9, 35), F(9, 35), F(9, 280), F(9, 280), F(41, 840)]
    c, a, b = np.zeros(13), np.zeros(79), np.zeros(11)
    O.lib.or_rk8_tableau(O.ptr(c), O.ptr(a), O.ptr(b))
    assert list(c) == [float(x) for x in cp]
    assert list(b) == [float(x) for x in bp]
    for r in range(1, 13):
        row = a[1 + r * (r - 1) // 2: 1 + r * (r - 1) // 2 + r]
        assert list(row) == [float(x) for x in rows[r]], r
        assert sum(rows[r]) == cp[r]
        assert abs(row.sum() - c[r]) <= 4e-16 * max(1.0, np.abs(row).sum()), r
    for k in range(8):
        assert sum(x * y ** k for x, y in zip(bp, cp)) == F(1, k + 1), k
        assert abs((b * c[:11] ** k).sum() - 1.0 / (k + 1)) < 1e-15, k
    # ... and not k = 8: the 7th-order weights are not 8th order
    assert sum(x * y ** 8 for x, y in zip(bp, cp)) != F(1, 9)
