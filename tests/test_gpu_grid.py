"""FARGO grid drag on the GPU (rd_disk_set_gas_grid; the reference's CPU loop
pp_disk.cu:778-814 with calc_linear_index pp_disk.cu:158-193) against the
oracle, through the C ABI.  The drag uses device atan2 / sin / cos, so forces
are compared to a relative 1e-13 and trajectories to 1e-11 AU, as for the
analytic disk."""
import numpy as np
import pytest

import oracle_lib as O
import redamd as R
from systems import small_system, mixed_types, as3, bits_equal, fargo_grid

pytestmark = pytest.mark.gpu

KDAY = R.K_GAUSS


def ring_system(n, seed, types=None):
    """Bodies inside the grid's span (0.4 - 2.5 AU) on near-circular orbits."""
    r, v, p, b, nbt = small_system(n, seed, types)
    rng = np.random.default_rng(seed + 1)
    rad, ang = rng.uniform(0.5, 2.3, n), rng.uniform(0, 2 * np.pi, n)
    r["x"], r["y"], r["z"] = rad * np.cos(ang), rad * np.sin(ang), rng.uniform(-0.02, 0.02, n)
    vc = np.sqrt(1.0 / rad) * rng.uniform(0.98, 1.02, n)
    v["x"], v["y"], v["z"] = -vc * np.sin(ang), vc * np.cos(ang), rng.uniform(-0.01, 0.01, n)
    r[0], v[0] = 0, 0
    p["radius"] = 10 ** rng.uniform(-8, -6, n)           # 1.5 - 150 km
    p["density"] = 3.4e6                                  # 2 g/cm^3 in Msun/AU^3
    p["mass"][1:] = 10 ** rng.uniform(-12, -9, n - 1)
    return r, v, p, b, nbt


# gas of 30 Msun/AU^3 puts the drag at 0.1 - 0.001 of gravity (the force tests
# then check the drag itself); 2e-3 (a minimum-mass nebula's midplane at 1 AU,
# 1.4e-9 g/cm^3) keeps trajectories non-stiff while moving velocities by ~1e-6
# in 12 days
STRONG, NEBULA = 30.0, 2e-3


@pytest.mark.parametrize("mode", [R.RD_DRAG_NSI, R.RD_DRAG_ALL_MASSIVE])
def test_grid_drag_vs_oracle(mode):
    r, v, p, b, nbt = ring_system(400, 31, mixed_types(400))
    grid = fargo_grid(20, 64, 5, rho1=STRONG)
    d = R.Disk(r, v, p, b, nbt, t=3.0)
    o = O.ODisk(r, v, p, b, nbt, t=3.0)
    d.set_drag_mode(mode)
    O.lib.or_disk_set_drag_mode(o.h, mode)
    d.set_gas_grid(*grid)
    o.set_gas_grid(*grid)
    dy = as3(d.calc_dydx(1, 2, 3.0, r, v))
    _, ody = o.calc_dydx(1, 2, 3.0, r, v)
    ref = as3(ody)
    assert np.all(np.abs(dy - ref) <= 1e-13 * np.abs(ref).max(axis=1, keepdims=True))
    # the drag is there, at the size of gravity: the same disk without gas differs in the dragged range
    plain = as3(R.Disk(r, v, p, b, nbt, t=3.0).calc_dydx(1, 2, 3.0, r, v))
    lo = int(nbt[:4].sum()) if mode == R.RD_DRAG_NSI else 1
    hi = int(nbt[:6].sum())
    assert np.all(np.any(dy[lo:hi] != plain[lo:hi], axis=1))
    assert np.median(np.abs(dy[lo:hi] - plain[lo:hi]).max(axis=1) / np.abs(plain[lo:hi]).max(axis=1)) > 1e-3
    assert bits_equal(dy[:lo], plain[:lo]) and bits_equal(dy[hi:], plain[hi:])


def test_grid_replaces_and_restores_the_analytic_disk():
    r, v, p, b, nbt = ring_system(300, 41, mixed_types(300))
    g = R.gas_params()
    d = R.Disk(r, v, p, b, nbt, gas=g)
    anal = as3(d.calc_dydx(1, 0, 0.0, r, v))
    d.set_gas_grid(*fargo_grid(10, 24, 2, rho1=STRONG))
    grid = as3(d.calc_dydx(1, 0, 0.0, r, v))
    assert not np.array_equal(grid, anal)
    d.set_gas_grid(None, None, None, None)
    assert bits_equal(as3(d.calc_dydx(1, 0, 0.0, r, v)), anal)


@pytest.mark.parametrize("prec", [R.RD_PREC_FP64, R.RD_PREC_MIXED])
def test_grid_rkf78_vs_oracle(prec):
    """The driver loop (rd_run, main.cu:403-640) over 0.2 tau (11.6 days)
    through a grid of piecewise-constant cells: bodies cross cell edges, the
    drag jumps there and the controller shortens the steps, in the oracle as
    on the GPU."""
    r, v, p, b, nbt = ring_system(500, 7, mixed_types(500))
    grid = fargo_grid(32, 96, 11, rho1=NEBULA)
    length = 0.2
    d = R.Disk(r, v, p, b, nbt)
    d.set_precision(prec)
    d.set_drag_mode(R.RD_DRAG_ALL_MASSIVE)
    d.set_gas_grid(*grid)
    it = R.Integrator(d, R.RD_RKF78, 1.0 * KDAY, adaptive=True, tol=1e-10)
    it.run(length, length)
    t, rr, vv, _, _ = d.state()
    ends = []
    for with_grid in (True, False):
        o = O.ODisk(r, v, p, b, nbt, drag_mode=1)
        if with_grid:
            o.set_gas_grid(*grid)
        oi = O.OInteg(o, 4, 1.0 * KDAY, adaptive=True, tol=1e-10)
        rc, _ = oi.run(length, length, 2 ** 62)
        assert rc == 0
        # the loop ends with the first step past the length, so the end time
        # follows the step sequence: ulp-level drag differences (or the FP32
        # pair sum) may move it, and without gas it differs by steps; carry
        # the oracle to the engine's end time with one fixed step
        if with_grid:
            assert abs(o.t - t) <= (1e-9 if prec == R.RD_PREC_FP64 else 0.1) * t
        if o.t != t:
            O.OInteg(o, 4, t - o.t, adaptive=False).step()
        ends.append(o.state())
    ot, orr, ovv, _, _ = ends[0]
    assert ot == t and ends[1][0] == t
    tol = 1e-11 if prec == R.RD_PREC_FP64 else 1e-8
    assert np.abs(as3(rr) - as3(orr)).max() < tol
    assert np.abs(as3(vv) - as3(ovv)).max() < tol
    # the drag moved the bodies: the run without gas ends elsewhere
    assert np.abs(as3(ends[1][2]) - as3(vv)).max() > 1e-7


def test_grid_batch_shares_the_grid():
    grid = fargo_grid(16, 48, 4, rho1=NEBULA)
    systems = [ring_system(n, s) for n, s in ((200, 1), (333, 2))]
    cat = [np.concatenate([s[q] for s in systems]) for q in range(4)]
    nbt = np.concatenate([s[4] for s in systems])
    d = R.Disk(*cat, nbt, batch=True)
    d.set_gas_grid(*grid)
    it = R.Integrator(d, R.RD_RKF78, 1.0 * KDAY, adaptive=True, tol=1e-10)
    it.steps(6)
    for k, (r, v, p, b, nb) in enumerate(systems):
        o = O.ODisk(r, v, p, b, nb)
        o.set_gas_grid(*grid)
        oi = O.OInteg(o, 4, 1.0 * KDAY, adaptive=True, tol=1e-10)
        for _ in range(6):
            oi.step()
        _, rr, vv, _, _ = d.state(k)
        _, orr, ovv, _, _ = o.state()
        assert np.abs(as3(rr) - as3(orr)).max() < 1e-11
        assert np.abs(as3(vv) - as3(ovv)).max() < 1e-11


def test_grid_set_after_a_captured_attempt():
    """An integrator whose attempt is already captured in a graph picks the
    grid up on its next step (the capture is redone): the same trajectory as a
    disk that had the grid from the start."""
    r, v, p, b, nbt = ring_system(400, 9)
    grid = fargo_grid(12, 36, 8, rho1=NEBULA)
    runs = []
    for late in (False, True):
        d = R.Disk(r, v, p, b, nbt)
        d.set_tuning(use_graph=1)
        if not late:
            d.set_gas_grid(*grid)
        it = R.Integrator(d, R.RD_RKF78, 1.0 * KDAY, adaptive=True, tol=1e-10)
        if late:
            it.steps(1)
            d.set_state(0.0, r, v)
            it.set_dt_next(1.0 * KDAY)
            d.set_gas_grid(*grid)
        it.steps(5)
        runs.append(d.state())
    assert runs[0][0] == runs[1][0]
    assert bits_equal(runs[0][1], runs[1][1]) and bits_equal(runs[0][2], runs[1][2])


def test_grid_arguments_are_checked():
    r, v, p, b, nbt = ring_system(50, 3)
    d = R.Disk(r, v, p, b, nbt)
    ur, de, vr, vt = fargo_grid(4, 8, 1)
    bad = ur.copy()
    bad[2], bad[3] = bad[3], bad[2]
    with pytest.raises(R.RedError, match="increasing"):
        d.set_gas_grid(bad, de, vr, vt)
    bad = ur.copy()
    bad[1] = np.nan
    with pytest.raises(R.RedError, match="finite"):
        d.set_gas_grid(bad, de, vr, vt)
    with pytest.raises(ValueError):
        d.set_gas_grid(ur[:-1], de, vr, vt)
