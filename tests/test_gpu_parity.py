"""HIP engine vs the CPU oracle, through the C ABI (include/red_abi.h).

FP64 mode restates the reference GPU kernel (pp_disk.cu:374-416) in the same
operation order, so forces and whole RK4 / RKF78 trajectories must be
bit-identical to the oracle.  Mixed precision is checked against the FP64
oracle with the tolerance written in each test.  Drag uses device pow/exp,
so it is compared to a relative 1e-13."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
import redamd as R
from systems import small_system, mixed_types, bits_equal, as3

pytestmark = pytest.mark.gpu

KDAY = R.K_GAUSS     # 1 day in tau units


def pair(r, v, p, b, nbt, thr=(0.0, 0.0, 0.0), gas=None, t=0.0):
    d = R.Disk(r, v, p, b, nbt, t=t, threshold=thr, gas=gas)
    o = O.ODisk(r, v, p, b, nbt, t=t, thr=thr, gas=gas)
    return d, o


# below ~4096 bodies the FP64 kernel takes 4-sink items, above 16-sink ones;
# "extremes5000": sources of mass 1e-95 and 0 and a pair 1e-31 AU apart, outside
# the range of the 16-sink items' short root and division sequences (the
# library path takes those pairs)
@pytest.mark.parametrize("case", ["disk256", "inactive1000", "classes700", "ragged1", "disk5000", "extremes5000"])
def test_force_fp64_bitexact(case):
    if case == "disk256":
        r, v, p, b, nbt = R.make_disk(255, 11)
    elif case == "disk5000":
        r, v, p, b, nbt = R.make_disk(4999, 15)
    elif case == "extremes5000":
        r, v, p, b, nbt = R.make_disk(4999, 16)
        p["mass"][[10, 2000, 4500]] = 1e-95
        p["mass"][3333] = 0.0
        # coordinates this small resolve a 1e-31 AU offset (r2 = 1e-62 < 2^-200)
        r["x"][3000], r["y"][3000], r["z"][3000] = 1e-20, 2e-20, 0.0
        r["x"][3001], r["y"][3001], r["z"][3001] = 1e-20 + 1e-31, 2e-20, 0.0
    elif case == "inactive1000":
        r, v, p, b, nbt = small_system(1000, 12, inactive=(1, 2, 17, 256, 511, 999))
    elif case == "classes700":
        r, v, p, b, nbt = small_system(700, 13, mixed_types(700), inactive=(5, 400))
    else:
        r, v, p, b, nbt = small_system(2, 14)
    d, o = pair(r, v, p, b, nbt)
    dy = d.calc_dydx(1, 0, 0.0, r, v)
    rc, ody = o.calc_dydx(1, 0, 0.0, r, v)
    assert rc == 0
    assert bits_equal(dy, ody)
    dv = d.calc_dydx(0, 0, 0.0, r, v)
    assert bits_equal(dv, v)


def test_fp64_lane_kernel_bitexact():
    """A batch of at least 1024 256-sink groups (4 per CU) runs the FP64 pair
    kernel with one sink per lane; its accelerations are the oracle's bit for
    bit, checked on three of its 260 disks of 1025 bodies."""
    parts = [R.make_disk(1024, 900 + k) for k in range(260)]
    cat = [np.concatenate([q[i] for q in parts]) for i in range(4)]
    bd = R.Disk(*cat, np.stack([q[4] for q in parts]), batch=True)
    for k in (0, 129, 259):
        r, v, p, b, nbt = parts[k]
        o = O.ODisk(r, v, p, b, nbt)
        dy = bd.calc_dydx(1, 0, 0.0, r, v, k=k)
        rc, ody = o.calc_dydx(1, 0, 0.0, r, v)
        assert rc == 0 and bits_equal(dy, ody), k


@pytest.mark.parametrize("mode,spl", [(R.RD_PAIRS_ONE_SIDED, 1), (R.RD_PAIRS_ONE_SIDED, 2),
                                      (R.RD_PAIRS_ONE_SIDED, 4), (R.RD_PAIRS_SYMMETRIC, 2)])
def test_force_mixed_vs_fp64(mode, spl):
    r, v, p, b, nbt = R.make_disk(3000, 21)
    d, o = pair(r, v, p, b, nbt)
    d.set_precision(R.RD_PREC_MIXED)
    d.set_tuning(sinks_per_lane=spl, chunk=512, pair_mode=mode)
    dy = as3(d.calc_dydx(1, 0, 0.0, r, v))
    rc, ody = o.calc_dydx(1, 0, 0.0, r, v)
    ref = as3(ody)
    # mutual part alone (star term removed) for a relative check of the FP32 path
    r1, v1, p1, b1, _ = r.copy(), v, p.copy(), b.copy(), nbt
    p1["mass"][0] = 0.0
    o2 = O.ODisk(r1, v1, p1, b1, nbt)
    _, mut = o2.calc_dydx(1, 0, 0.0, r1, v1)
    mut = as3(mut)
    err = np.linalg.norm(dy - ref, axis=1)
    # tolerance: FP32 pair terms (rel. 2^-22 each incl. rsqrt) summed over <= 3000 sources
    assert np.all(err[1:] <= 2e-5 * np.linalg.norm(mut[1:], axis=1) + 1e-18), err.max()
    # whole acceleration of the disk bodies (the star's is the FP32 mutual sum alone)
    assert np.all(err[1:] <= 1e-9 * np.linalg.norm(ref[1:], axis=1))


def test_force_mixed_classes_and_inactive():
    r, v, p, b, nbt = small_system(900, 22, mixed_types(900), inactive=(7, 300, 650))
    d, o = pair(r, v, p, b, nbt)
    d.set_precision(R.RD_PREC_MIXED)
    dy = as3(d.calc_dydx(1, 0, 0.0, r, v))
    _, ody = o.calc_dydx(1, 0, 0.0, r, v)
    ref = as3(ody)
    err = np.linalg.norm(dy - ref, axis=1)
    assert np.all(err <= 1e-6 * np.linalg.norm(ref, axis=1) + 1e-20)
    for i in (7, 300, 650):
        assert np.all(dy[i] == 0.0)


@pytest.mark.parametrize("mode", [R.RD_DRAG_NSI, R.RD_DRAG_ALL_MASSIVE])
def test_drag_vs_oracle(mode):
    r, v, p, b, nbt = small_system(400, 31, mixed_types(400))
    g = R.gas_params()
    d, o = pair(r, v, p, b, nbt, gas=g, t=3.0)
    d.set_drag_mode(mode)
    O.lib.or_disk_set_drag_mode(o.h, mode)
    dy = as3(d.calc_dydx(1, 2, 3.0, r, v))
    _, ody = o.calc_dydx(1, 2, 3.0, r, v)
    ref = as3(ody)
    assert np.all(np.abs(dy - ref) <= 1e-13 * np.abs(ref).max(axis=1, keepdims=True))


def test_rk4_fixed_bitexact():
    r, v, p, b, nbt = R.make_disk(256, 1)
    d, o = pair(r, v, p, b, nbt)
    h = 0.1 * KDAY
    it = R.Integrator(d, R.RD_RK4, h, adaptive=False)
    oi = O.OInteg(o, 2, h, adaptive=False)
    it.steps(25)
    for _ in range(25):
        oi.step()
    t, rr, vv, _, _ = d.state()
    ot, orr, ovv, _, _ = o.state()
    assert t == ot
    assert bits_equal(rr, orr) and bits_equal(vv, ovv)
    for _ in range(5):
        assert it.step() == oi.step()
    t, rr, vv, _, _ = d.state()
    ot, orr, ovv, _, _ = o.state()
    assert t == ot and bits_equal(rr, orr) and bits_equal(vv, ovv)


@pytest.mark.parametrize("case", ["classes", "gas", "graph", "disk5000"])
def test_rk4_fixed_paths_bitexact(case):
    """Fixed-step RK4 in FP64 takes one launch per stage on a small gas-free
    disk (the FP64 pair kernel's summing lanes do the stage) and pair pass +
    stage kernel otherwise (gas, 16-sink items above ~4096 bodies): both are
    the oracle's steps bit for bit, with every body class and inactive bodies,
    direct launches and graph replays."""
    if case == "disk5000":
        r, v, p, b, nbt = R.make_disk(4999, 15)
    else:
        r, v, p, b, nbt = small_system(600, 61, mixed_types(600), inactive=(4, 77, 599))
    gas = R.gas_params() if case == "gas" else None
    d, o = pair(r, v, p, b, nbt, gas=gas)
    if case == "graph":
        d.set_tuning(use_graph=1)
    h = 0.05 * KDAY
    it = R.Integrator(d, R.RD_RK4, h, adaptive=False)
    oi = O.OInteg(o, 2, h, adaptive=False)
    n = 3 if case == "disk5000" else 12
    it.steps(n)
    for _ in range(n):
        oi.step()
    t, rr, vv, _, _ = d.state()
    ot, orr, ovv, _, _ = o.state()
    assert t == ot and bits_equal(rr, orr) and bits_equal(vv, ovv)
    assert it.step() == oi.step()
    t, rr, vv, _, _ = d.state()
    ot, orr, ovv, _, _ = o.state()
    assert t == ot and bits_equal(rr, orr) and bits_equal(vv, ovv)


@pytest.mark.parametrize("method", [R.RD_RKF78, R.RD_RK4])
def test_adaptive_bitexact_with_rejections(method):
    r, v, p, b, nbt = R.make_disk(64, 2, a_min=0.8, a_max=1.2)
    d, o = pair(r, v, p, b, nbt)
    h0 = 60.0 * KDAY                    # too large: the first attempts are rejected
    tol = 1e-12
    it = R.Integrator(d, method, h0, adaptive=True, tol=tol)
    oi = O.OInteg(o, method, h0, adaptive=True, tol=tol)
    for _ in range(12):
        assert it.step() == oi.step()
        assert it.dt_next() == oi.dt_next()
    assert it.counters() == oi.counters()
    assert it.counters()[2] > 0
    t, rr, vv, _, _ = d.state()
    ot, orr, ovv, _, _ = o.state()
    assert t == ot and bits_equal(rr, orr) and bits_equal(vv, ovv)


@pytest.mark.parametrize("method", [R.RD_RKF78, R.RD_RK4])
def test_adaptive_bitexact_many_waves(method):
    """The error norm of a single disk is reduced over ~140 waves (one slot per
    wave, maximum on the host) and stage 0 is skipped on k0 reuse: the step
    and dt sequence still match the oracle bit for bit."""
    r, v, p, b, nbt = R.make_disk(3000, 5)
    d, o = pair(r, v, p, b, nbt)
    h0 = 40.0 * KDAY                    # too large: rejections first
    it = R.Integrator(d, method, h0, adaptive=True, tol=1e-11)
    oi = O.OInteg(o, method, h0, adaptive=True, tol=1e-11)
    for _ in range(3):
        assert it.step() == oi.step()
        assert it.dt_next() == oi.dt_next()
    assert it.counters() == oi.counters()
    assert it.counters()[2] > 0
    t, rr, vv, _, _ = d.state()
    ot, orr, ovv, _, _ = o.state()
    assert t == ot and bits_equal(rr, orr) and bits_equal(vv, ovv)


@pytest.mark.parametrize("method", [R.RD_RKF78, R.RD_RK4])
def test_backward_integration_bitexact(method):
    """A negative step (options.cpp:244-248 negates dt for a negative
    simulation length): the error norm is |lambda dt err|
    (int_rungekutta8.cu:696), so rejections shrink the step and it stays
    negative; bit for bit with the oracle."""
    r, v, p, b, nbt = R.make_disk(300, 7)
    d, o = pair(r, v, p, b, nbt)
    h0 = -30.0 * KDAY                   # too large: rejections first
    it = R.Integrator(d, method, h0, adaptive=True, tol=1e-11)
    oi = O.OInteg(o, method, h0, adaptive=True, tol=1e-11)
    for _ in range(6):
        h = it.step()
        assert h == oi.step() and h < 0.0
        assert it.dt_next() == oi.dt_next() < 0.0
    assert it.counters() == oi.counters()
    assert it.counters()[2] > 0
    t, rr, vv, _, _ = d.state()
    ot, orr, ovv, _, _ = o.state()
    assert t == ot < 0.0 and bits_equal(rr, orr) and bits_equal(vv, ovv)


def mirrored(v):
    m = v.copy()
    for c in "xyz":
        m[c] = -v[c]
    return m


@pytest.mark.parametrize("prec", [R.RD_PREC_FP64, R.RD_PREC_MIXED])
def test_backward_mirrors_forward_full_size(prec):
    """At N = 16384, where no oracle run is quick: without drag the right-hand
    side is odd in v, so the backward run from (r, v) is the forward run from
    (r, -v) with t and v mirrored.  Every operation on the path (the pair
    kernels, the stage sums, the error norm, the controller) is symmetric
    under a sign flip, so the mirror holds bit for bit in both precisions."""
    r, v, p, b, nbt = R.make_disk(16384, 2)
    fwd = R.Disk(r, mirrored(v), p, b, nbt)
    bwd = R.Disk(r, v, p, b, nbt)
    fwd.set_precision(prec)
    bwd.set_precision(prec)
    h0 = 20.0 * KDAY
    fi = R.Integrator(fwd, R.RD_RKF78, h0, adaptive=True, tol=1e-10)
    bi = R.Integrator(bwd, R.RD_RKF78, -h0, adaptive=True, tol=1e-10)
    for _ in range(4):
        hf = fi.step()
        assert hf > 0.0 and bi.step() == -hf
        assert bi.dt_next() == -fi.dt_next()
    assert fi.counters() == bi.counters()
    tf, rf, vf, _, _ = fwd.state()
    tb, rb, vb, _, _ = bwd.state()
    assert tb == -tf
    assert bits_equal(rf, rb) and bits_equal(vf, mirrored(vb))


@pytest.mark.parametrize("method", [R.RD_EULER, R.RD_RK2])
def test_euler_rk2_bitexact(method):
    """Euler (int_euler.cu) and midpoint RK2 (int_rungekutta2.cu) against the
    oracle, bit for bit: single steps, a dt_next the methods must ignore, and
    the chained fixed-step path of rd_steps."""
    r, v, p, b, nbt = small_system(300, 61, mixed_types(300))
    d, o = pair(r, v, p, b, nbt)
    h = 0.5 * KDAY
    it = R.Integrator(d, method, h, adaptive=True, tol=1e-10)   # fixed step whatever adaptive says
    oi = O.OInteg(o, method, h, adaptive=True, tol=1e-10)
    for _ in range(5):
        assert it.step() == oi.step() == h
    it.set_dt_next(h / 3)
    O.lib.or_set_dt_next(oi.h, h / 3)
    for _ in range(5):
        assert it.step() == oi.step() == h
    assert it.dt_next() == oi.dt_next() == h
    it.steps(20)
    for _ in range(20):
        oi.step()
    t, rr, vv, _, _ = d.state()
    ot, orr, ovv, _, _ = o.state()
    assert t == ot
    assert bits_equal(rr, orr) and bits_equal(vv, ovv)
    assert it.counters() == oi.counters() == (30, 30, 0)


def test_graph_matches_direct_launch():
    out = []
    for g in (0, 1):
        r, v, p, b, nbt = R.make_disk(500, 3)
        d = R.Disk(r, v, p, b, nbt)
        d.set_precision(R.RD_PREC_MIXED)
        d.set_tuning(use_graph=g)
        it = R.Integrator(d, R.RD_RKF78, 5 * KDAY, adaptive=True, tol=1e-10)
        for _ in range(4):
            it.step()
        it.steps(3)
        out.append(d.state())
    assert out[0][0] == out[1][0]
    assert bits_equal(out[0][1], out[1][1]) and bits_equal(out[0][2], out[1][2])


def test_integral_vs_oracle():
    r, v, p, b, nbt = small_system(3000, 41, inactive=(3, 9))
    d, o = pair(r, v, p, b, nbt)
    gi, oi = d.integral(), o.integral()
    assert abs(gi["E"] - oi["E"]) <= 1e-13 * abs(oi["E"])
    for f in ("R", "V", "C"):
        for c in "xyz":
            assert abs(gi[f][c] - oi[f][c]) <= 1e-14 * max(1e-300, np.abs([oi[f][q] for q in "xyz"]).max())


def crowded(n, seed):
    """A thin ring of inflated bodies that collides within a few steps."""
    r, v, p, b, nbt = R.make_disk(n, seed, a_min=0.99, a_max=1.01, sig_e=0.02, sig_i=0.001, r_star=1e-5)
    return r, v, p, b, nbt


@pytest.mark.parametrize("method", [R.RD_RKF78, R.RD_RK4])
def test_driver_loop_events_match_oracle(method):
    """Driver loop with collisions, hit-centrum and a rebuild after 16 events:
    adaptive RKF7(8), and fixed-step RK4, whose stages run inside the FP64
    pair kernel on a disk this small."""
    n, length, max_steps = (96, 40.0, 60) if method == R.RD_RKF78 else (128, 400.0, 120)
    r, v, p, b, nbt = crowded(n, 5)
    thr = (0.2, 50.0, 1000.0)       # hit-centrum 0.2 AU, ejection 50 AU, radius factor 1000
    d, o = pair(r, v, p, b, nbt, thr=thr)
    h = 1.0 * KDAY
    adaptive = method == R.RD_RKF78
    it = R.Integrator(d, method, h, adaptive=adaptive, tol=1e-10)
    oi = O.OInteg(o, method, h, adaptive=adaptive, tol=1e-10)
    it.run(length, 2.0, max_steps)
    rc, steps = oi.run(length, 2.0, max_steps)
    assert rc == 0
    ev, oev = d.event_log(), o.event_log()
    assert len(ev) == len(oev) and len(ev) > 16          # enough to trigger a rebuild
    assert bits_equal(ev, oev)
    t, rr, vv, pp, bb = d.state()
    ot, orr, ovv, opp, obb = o.state()
    assert t == ot
    assert bits_equal(rr, orr) and bits_equal(vv, ovv) and bits_equal(pp, opp) and bits_equal(bb, obb)
    assert all(np.array_equal(a, b_) for a, b_ in zip(d.n_objects(), o.counts()))
    # event counters (pp_disk.h:164-167) agree with the log: totals per kind, and
    # the since-last-clear count was reset by the rebuild after 16 events
    c = d.event_counters()
    kinds = ev["event_name"]
    assert c["event"][0] == len(ev)
    assert c["collision"][0] == int(np.sum(kinds == R.EVENT_COLLISION))
    assert c["ejection"][0] == int(np.sum(kinds == R.EVENT_EJECTION))
    assert c["hit_centrum"][0] == int(np.sum(kinds == R.EVENT_HIT_CENTRUM))
    assert c["event"][1] < 16 and c["collision"][2] == 0
    d.set_event_counter(R.RD_EVENT_COUNTER_TOTAL, 7)
    assert all(v[0] == 7 for v in d.event_counters().values())


def test_ejection_and_hit_centrum():
    r, v, p, b, nbt = small_system(50, 51)
    r["x"][10], r["y"][10], r["z"][10] = 0.05, 0.0, 0.0     # inside 0.1 AU
    r["x"][20] = 80.0                                        # beyond 60 AU
    thr = (0.1, 60.0, 0.0)
    d, o = pair(r, v, p, b, nbt, thr=thr)
    n = d.check_ejection_hit_centrum()
    on = O.lib.or_check_ejection_hit_centrum(o.h)
    assert n == on and n >= 2
    d.handle_ejection_hit_centrum()
    assert O.lib.or_handle_ejection_hit_centrum(o.h) == 0
    assert bits_equal(d.event_log(), o.event_log())
    _, rr, vv, pp, bb = d.state()
    _, orr, ovv, opp, obb = o.state()
    assert bits_equal(rr, orr) and bits_equal(vv, ovv) and bits_equal(pp, opp) and bits_equal(bb, obb)


@pytest.mark.parametrize("method", [R.RD_RKF78, R.RD_RK4, R.RD_RK2, R.RD_EULER])
@pytest.mark.parametrize("prec", [R.RD_PREC_FP64, R.RD_PREC_MIXED])
@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_batch_matches_single_disks(prec, method, sign):
    """Batch (general stage kernels) against standalone disks (single-disk
    stage kernels), every integrator, both precisions, forward and backward."""
    sizes = [64, 300, 1000]
    parts = [R.make_disk(n - 1, 100 + k) for k, n in enumerate(sizes)]
    cat = [np.concatenate([q[i] for q in parts]) for i in range(4)]
    nbt = np.stack([q[4] for q in parts])
    bd = R.Disk(*cat, nbt, batch=True)
    bd.set_precision(prec)
    adaptive = method in (R.RD_RKF78, R.RD_RK4)
    h = sign * (3 * KDAY if adaptive else 0.2 * KDAY)
    bit = R.Integrator(bd, method, h, adaptive=adaptive, tol=1e-11)
    bit.steps(4)
    for k, q in enumerate(parts):
        sd = R.Disk(*q)
        sd.set_precision(prec)
        sit = R.Integrator(sd, method, h, adaptive=adaptive, tol=1e-11)
        sit.steps(4)
        a, b_ = bd.state(k), sd.state()
        assert a[0] == b_[0] and a[0] * sign > 0.0
        assert bits_equal(a[1], b_[1]) and bits_equal(a[2], b_[2])
        assert bit.counters(k) == sit.counters()


def test_config4_batch_matches_single_disks():
    """Config 4 at its full size: the 512-disk batch (N ~ U[64, 4096], disk k
    seeded 4000 + k, 1.06 M bodies, mixed precision, general stage kernels)
    against standalone runs (single-disk stage kernels) of eight of its disks
    after 3 accepted RKF7(8) steps: times, states and counters bit for bit."""
    from proposer import batch_sizes
    sizes = batch_sizes()
    parts = [R.make_disk(n, 4000 + k) for k, n in enumerate(sizes)]
    cat = [np.concatenate([q[i] for q in parts]) for i in range(4)]
    nbt = np.stack([q[4] for q in parts])
    bd = R.Disk(*cat, nbt, batch=True)
    bd.set_precision(R.RD_PREC_MIXED)
    h0, tol = 0.1 * KDAY, 1e-10
    bit = R.Integrator(bd, R.RD_RKF78, h0, adaptive=True, tol=tol)
    bit.steps(3)
    order = np.argsort(sizes)
    picks = sorted({int(order[0]), int(order[1]), int(order[-1]), int(order[-2]), 0, 137, 256, 511})
    for k in picks:
        sd = R.Disk(*parts[k])
        sd.set_precision(R.RD_PREC_MIXED)
        sit = R.Integrator(sd, R.RD_RKF78, h0, adaptive=True, tol=tol)
        sit.steps(3)
        a, b_ = bd.state(k), sd.state()
        assert a[0] == b_[0], k
        assert bits_equal(a[1], b_[1]) and bits_equal(a[2], b_[2]), k
        assert bit.counters(k) == sit.counters(), k
        del sit, sd


@pytest.mark.parametrize("prec", [R.RD_PREC_FP64, R.RD_PREC_MIXED])
def test_batch_events_match_single_disks(prec):
    """Batched driver loop (one launch per stage over all disks, batch-wide
    event detection) gives every disk the event list and state of its own
    per-disk run, bit for bit."""
    # (seed offset, N) of crowded rings whose runs stay valid in the reference's
    # semantics (some seeds merge one body twice and the reference's rebuild
    # check refuses them, which the engine reproduces)
    disks = [(0, 40), (1, 96), (3, 70), (4, 120), (6, 140)]
    parts = [R.make_disk(n, 50 + k, a_min=0.99, a_max=1.01, sig_e=0.02, sig_i=0.001, r_star=1e-5)
             for k, n in disks]
    cat = [np.concatenate([q[i] for q in parts]) for i in range(4)]
    nbt = np.stack([q[4] for q in parts])
    thr = (0.2, 50.0, 1000.0)
    bd = R.Disk(*cat, nbt, threshold=thr, batch=True)
    bd.set_precision(prec)
    bit = R.Integrator(bd, R.RD_RKF78, KDAY, adaptive=True, tol=1e-10)
    bit.run(40.0, 2.0, 40)
    total = 0
    for k, q in enumerate(parts):
        sd = R.Disk(*q, threshold=thr)
        sd.set_precision(prec)
        sit = R.Integrator(sd, R.RD_RKF78, KDAY, adaptive=True, tol=1e-10)
        sit.run(40.0, 2.0, 40)
        ev_b, ev_s = bd.event_log(k), sd.event_log()
        assert bits_equal(ev_b, ev_s)
        total += len(ev_s)
        a, b_ = bd.state(k), sd.state()
        assert a[0] == b_[0]
        for x, y in zip(a[1:], b_[1:]):
            assert bits_equal(x, y)
        assert bit.counters(k) == sit.counters()
    assert total > 16


def test_double_merge_refused_like_the_reference():
    """A ring where one body is merged twice in a step: the reference's rebuild
    throws 'Number of copied bodies does not equal to the number of active
    bodies.' (pp_disk.cu:1403-1406); the engine returns RD_EINVAL with that text
    and the oracle refuses the same run."""
    r, v, p, b, nbt = R.make_disk(150, 52, a_min=0.99, a_max=1.01, sig_e=0.02, sig_i=0.001, r_star=1e-5)
    thr = (0.2, 50.0, 1000.0)
    d = R.Disk(r, v, p, b, nbt, threshold=thr)
    it = R.Integrator(d, R.RD_RKF78, KDAY, adaptive=True, tol=1e-10)
    with pytest.raises(R.RedError) as ei:
        it.run(40.0, 2.0, 40)
    assert ei.value.code == R.RD_EINVAL and "Number of copied bodies" in str(ei.value)
    o = O.ODisk(r, v, p, b, nbt, thr=thr)
    oi = O.OInteg(o, 4, KDAY, adaptive=True, tol=1e-10)
    rc, _ = oi.run(40.0, 2.0, 40)
    assert rc == 1
    assert bits_equal(d.event_log(), o.event_log())
