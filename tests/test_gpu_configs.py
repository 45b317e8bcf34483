"""The benchmark configurations through the C ABI on the GPU.

config 1: star + 256, RK4 h = 0.1 d, 10^4 steps, FP64 -- bit-identical to the
          golden fixture written by the oracle (tests/golden/make_golden.py).
config 2: N = 16384, RKF7(8) tol 1e-10 -- mixed precision against the FP64
          kernel on the same GPU (tolerances written below) and energy.
config 3: gas disk + collisions + hit-centrum, mixed precision, at reduced N
          against the oracle (drag uses device pow/exp: tolerance, not bits).
The C++ driver over the host classes runs once as a smoke test."""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as O
import redamd as R
from systems import bits_equal, as3

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
KDAY = R.K_GAUSS


def test_config1_rk4_golden_bitexact():
    g = np.load(os.path.join(ROOT, "tests", "golden", "config1_rk4.npz"))
    r, v, p, b, nbt = R.make_disk(256, 1)
    assert np.array_equal(as3(r), g["r0"]) and np.array_equal(as3(v), g["v0"])
    d = R.Disk(r, v, p, b, nbt)
    d.set_tuning(use_graph=1)
    it = R.Integrator(d, R.RD_RK4, 0.1 * KDAY, adaptive=False)
    it.steps(10000)
    t, rr, vv, _, _ = d.state()
    assert t == g["t"]
    assert np.array_equal(as3(rr), g["r"]) and np.array_equal(as3(vv), g["v"])


def test_config2_mixed_vs_fp64_short_run():
    """Equal-time comparison (fixed-step RKF7(8)) of the two pair kernels, then
    the adaptive controller's energy check.  Adaptive runs are not compared
    position by position: the Fehlberg estimate |k0 + k10 - k11 - k12| is a
    difference of nearly equal stage values, so FP32 rounding moves the chosen
    step sizes by a few percent and the two runs sample different times."""
    r, v, p, b, nbt = R.make_disk(16384, 2)
    h = 0.02 * KDAY
    st = {}
    for prec in (R.RD_PREC_FP64, R.RD_PREC_MIXED):
        d = R.Disk(r, v, p, b, nbt)
        d.set_precision(prec)
        it = R.Integrator(d, R.RD_RKF78, h, adaptive=False)
        it.steps(5)
        st[prec] = d.state()
    assert st[0][0] == st[1][0]
    # tolerance: force difference (<= 2e-9 relative, |a| <= 5 in tau units) over
    # 5 steps of h = 3.4e-4 tau is < 1e-13 AU in position, 1e-11 in velocity
    assert np.abs(as3(st[0][1]) - as3(st[1][1])).max() < 1e-13
    assert np.abs(as3(st[0][2]) - as3(st[1][2])).max() < 1e-11
    for prec in (R.RD_PREC_FP64, R.RD_PREC_MIXED):
        d = R.Disk(r, v, p, b, nbt)
        d.set_precision(prec)
        it = R.Integrator(d, R.RD_RKF78, 0.05 * KDAY, adaptive=True, tol=1e-10)
        E0 = d.integral()["E"]
        it.steps(20)
        assert abs((d.integral()["E"] - E0) / E0) < 1e-11


def test_config3_gas_collisions_vs_oracle():
    # config 3 physics at N = 600: gas drag on all massive bodies, perfect merging
    # with radius factor 1 (inflated to 3000 to produce events quickly), hit-centrum 0.1 AU
    r, v, p, b, nbt = R.make_disk(600, 3, size_mode=1, lo=1.0, hi=100.0, cd=1.0, a_min=0.9, a_max=1.1,
                                  r_star=1e-5)
    g = R.gas_params()
    thr = (0.1, 100.0, 3000.0)
    d = R.Disk(r, v, p, b, nbt, threshold=thr, gas=g)
    d.set_drag_mode(R.RD_DRAG_ALL_MASSIVE)
    o = O.ODisk(r, v, p, b, nbt, thr=thr, gas=g, drag_mode=1)
    it = R.Integrator(d, R.RD_RKF78, 1.0 * KDAY, adaptive=True, tol=1e-10)
    oi = O.OInteg(o, 4, 1.0 * KDAY, adaptive=True, tol=1e-10)
    it.run(1e9, 1e9, 8)
    rc, steps = oi.run(1e9, 1e9, 8)
    assert rc == 0 and steps == 8
    ev, oev = d.event_log(), o.event_log()
    assert len(ev) == len(oev) and len(ev) > 0
    for f in ("event_name", "id1", "id2", "idx1", "idx2"):
        assert np.array_equal(ev[f], oev[f])
    t, rr, vv, pp, bb = d.state()
    ot, orr, ovv, opp, obb = o.state()
    assert np.array_equal(bb["id"], obb["id"])
    assert abs(t - ot) <= 1e-12 * abs(ot)
    assert np.abs(as3(rr) - as3(orr)).max() < 1e-11
    assert np.abs(as3(vv) - as3(ovv)).max() < 1e-11


def test_cpp_driver_smoke():
    exe = os.path.join(ROOT, "red.cuda_amd", "red_run")
    out = subprocess.run([exe, "2000", "4", "rkf78", "1e-10", "20", "mixed", "0.1", "100", "1"],
                         capture_output=True, text=True, timeout=120, check=True).stdout
    rec = json.loads(out.strip().splitlines()[-1])
    assert rec["steps"] == 20 and rec["tried"] >= 20 and rec["rel_energy_error"] < 1e-9


def test_collision_grid_full_size():
    """The collision check at config-3 size (N = 131072, radii 1-100 km) with
    the radius factor raised to 5000 so that hundreds of pairs touch: the
    engine's pairs (hashed grid) equal an independent search -- a scipy k-d
    tree over the same FP64 positions, then the reference's test
    d^2 < f^2 (R_i + R_j)^2 evaluated in the same operation order."""
    from scipy.spatial import cKDTree
    r, v, p, b, nbt = R.make_disk(131072, 3, size_mode=1, lo=1.0, hi=100.0, cd=1.0, r_star=1e-5)
    f = 5000.0
    d = R.Disk(r, v, p, b, nbt, threshold=(0.0, 0.0, f))
    n = d.check_collision()
    ev = d.events()
    got = {(int(min(a, c)), int(max(a, c))) for a, c in zip(ev["idx1"], ev["idx2"])}
    assert len(got) == n
    x = as3(r)
    rad = p["radius"].astype(np.float64)
    act = b["id"] > 0
    reach = 2.0 * f * rad[act].max() * (1.0 + 1e-6)
    cand = cKDTree(x).query_pairs(reach, output_type="ndarray")
    i, j = cand.min(1), cand.max(1)
    keep = act[i] & act[j]
    i, j = i[keep], j[keep]
    dx, dy, dz = (x[j, 0] - x[i, 0]), (x[j, 1] - x[i, 1]), (x[j, 2] - x[i, 2])
    r2 = dx * dx + dy * dy + dz * dz
    s = rad[i] + rad[j]
    hit = r2 < (f * f) * (s * s)
    want = {(int(a), int(c)) for a, c in zip(i[hit], j[hit])}
    assert len(want) > 100
    assert got == want
