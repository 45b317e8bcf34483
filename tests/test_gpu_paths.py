"""Engine paths beyond the configs' main loop, against the oracle through the
C ABI: the opt-in retry guard and the reference's NaN semantics, the
autotuner (pp_disk::benchmark, pp_disk.cu:553-632), error_check_for_tp
(integrator.h:32, int_rungekutta8.cu:693), the super-planetesimal merge
refusal (pp_disk.cu:1333-1337) and event-buffer overflow (pp_disk.cu:357,
which the reference leaves unguarded)."""
import numpy as np
import pytest

import oracle_lib as O
import redamd as R
from systems import small_system, bits_equal

pytestmark = pytest.mark.gpu

KDAY = R.K_GAUSS


def same_state(d, o, k=0):
    """Engine disk k and oracle disk hold the same state: bit for bit, with
    NaN matched as NaN (CPU and GPU write different NaN payloads)."""
    t, rr, vv, pp, bb = d.state(k)
    ot, orr, ovv, opp, obb = o.state()
    if not (t == ot and bits_equal(pp, opp) and bits_equal(bb, obb)):
        return False
    for a, b in ((rr, orr), (vv, ovv)):
        for c in "xyz":
            na, nb = np.isnan(a[c]), np.isnan(b[c])
            if not np.array_equal(na, nb) or not bits_equal(a[c][~na], b[c][~nb]):
                return False
    return True


def coincident(n=40, seed=3):
    r, v, p, b, nbt = small_system(n, seed)
    for c in "xyz":
        r[c][2] = r[c][1]
    return r, v, p, b, nbt


GUARD_CASES = {
    # name: (system, method, h0, tol, guard)
    "nan_accepted_rkf78": (coincident, R.RD_RKF78, 0.01, 1e-10, {}),
    "nan_accepted_rk4": (coincident, R.RD_RK4, 0.01, 1e-10, {}),
    "nan_refused": (coincident, R.RD_RKF78, 0.01, 1e-10, dict(nonfinite=True)),
    "attempts_rkf78": (lambda: small_system(8, 4), R.RD_RKF78, 5.0, 1e-12, dict(max_attempts=2)),
    "attempts_rk4": (lambda: small_system(8, 4), R.RD_RK4, 5.0, 1e-12, dict(max_attempts=2)),
    "attempts_enough": (lambda: small_system(8, 4), R.RD_RKF78, 5.0, 1e-12, dict(max_attempts=3)),
    "zero_step": (lambda: small_system(8, 4), R.RD_RKF78, 0.0, 1e-12, dict(min_dt=1e-30)),
    "zero_step_stalls": (lambda: small_system(8, 4), R.RD_RKF78, 0.0, 1e-12, {}),
}


@pytest.mark.parametrize("case", sorted(GUARD_CASES))
def test_guard_and_nan_semantics_match_oracle(case):
    make, method, h0, tol, guard = GUARD_CASES[case]
    print(f"# {case}", flush=True)
    r, v, p, b, nbt = make()
    d = R.Disk(r, v, p, b, nbt)
    o = O.ODisk(r, v, p, b, nbt)
    it = R.Integrator(d, method, h0, adaptive=True, tol=tol)
    oi = O.OInteg(o, method, h0, adaptive=True, tol=tol)
    if guard:
        it.set_guard(**guard)
        oi.set_guard(**guard)
    for _ in range(3):
        rc, dt = it.step_rc()
        orc, odt = oi.step_rc()
        assert rc == orc, (rc, orc)
        if rc:
            assert rc == R.RD_ESTATE
            break
        assert dt == odt
    assert it.counters() == oi.counters()
    assert it.dt_next() == oi.dt_next() or (np.isnan(it.dt_next()) and np.isnan(oi.dt_next()))
    assert same_state(d, o)
    if case.startswith("nan_accepted"):
        assert rc == 0 and np.isnan(d.state()[2]["x"][1])
    if case in ("nan_refused", "attempts_rkf78", "attempts_rk4", "zero_step"):
        assert rc == R.RD_ESTATE and d.state()[0] == 0.0


def test_guard_in_a_batch_fails_one_disk_only():
    """Disk 1 of three goes NaN; with the non-finite guard it fails and keeps
    y_n, while disks 0 and 2 take their step exactly as they do alone."""
    parts = [R.make_disk(99, 61), coincident(), R.make_disk(299, 62)]
    cat = [np.concatenate([q[i] for q in parts]) for i in range(4)]
    nbt = np.stack([q[4] for q in parts])
    bd = R.Disk(*cat, nbt, batch=True)
    bit = R.Integrator(bd, R.RD_RKF78, 0.01, adaptive=True, tol=1e-11)
    bit.set_guard(nonfinite=True)
    rc, _ = bit.step_rc()
    assert rc == R.RD_ESTATE and "disk 1" in R.last_error()
    t1, r1, v1, _, _ = bd.state(1)
    assert t1 == 0.0 and bits_equal(r1, parts[1][0]) and bits_equal(v1, parts[1][1])
    assert bit.counters(1) == (1, 0, 1)
    for k in (0, 2):
        sd = R.Disk(*parts[k])
        sit = R.Integrator(sd, R.RD_RKF78, 0.01, adaptive=True, tol=1e-11)
        sit.step()
        a, s = bd.state(k), sd.state()
        assert a[0] == s[0] and bits_equal(a[1], s[1]) and bits_equal(a[2], s[2]), k
        assert bit.counters(k) == sit.counters()


def test_autotune_picks_a_candidate_and_replays_bit_for_bit():
    r, v, p, b, nbt = R.make_disk(3000, 31)
    d = R.Disk(r, v, p, b, nbt)
    d.set_precision(R.RD_PREC_MIXED)
    best = d.autotune()
    assert best.pair_mode in (R.RD_PAIRS_ONE_SIDED, R.RD_PAIRS_SYMMETRIC)
    assert best.sinks_per_lane in (1, 2, 4)
    assert best.chunk in (512, 1024, 2048, 4096, 8192)
    assert best.blocks_per_cu in (1, 2, 4, 8)
    assert best.sym_chunk_min in (1, 2)
    cur = d.tuning()
    assert all(getattr(cur, f) == getattr(best, f) for f, _ in best._fields_)
    it = R.Integrator(d, R.RD_RKF78, 2 * KDAY, adaptive=True, tol=1e-11)
    it.steps(5)
    d2 = R.Disk(r, v, p, b, nbt)
    d2.set_precision(R.RD_PREC_MIXED)
    d2.set_tuning(**{f: getattr(best, f) for f, _ in best._fields_})
    it2 = R.Integrator(d2, R.RD_RKF78, 2 * KDAY, adaptive=True, tol=1e-11)
    it2.steps(5)
    a, s = d.state(), d2.state()
    assert a[0] == s[0] and bits_equal(a[1], s[1]) and bits_equal(a[2], s[2])
    assert it.counters() == it2.counters()


def tp_system():
    """Star, 30 protoplanets at 1-3 AU and 10 test particles on tight
    circular orbits at 0.05-0.1 AU: the test particles set the error."""
    rng = np.random.default_rng(71)
    n_pp, n_tp = 30, 10
    types = [0] + [3] * n_pp + [6] * n_tp
    r, v, p, b, nbt = small_system(len(types), 71, types)
    rad = np.concatenate([rng.uniform(1.0, 3.0, n_pp), rng.uniform(0.05, 0.1, n_tp)])
    ang = rng.uniform(0, 2 * np.pi, n_pp + n_tp)
    r["x"][1:], r["y"][1:], r["z"][1:] = rad * np.cos(ang), rad * np.sin(ang), 0.0
    vc = np.sqrt(1.0 / rad)
    v["x"][1:], v["y"][1:], v["z"][1:] = -vc * np.sin(ang), vc * np.cos(ang), 0.0
    r["x"][0] = r["y"][0] = v["x"][0] = v["y"][0] = 0.0
    p["mass"][1 + n_pp:] = 0.0
    return r, v, p, b, nbt


def test_error_check_for_tp_matches_oracle():
    r, v, p, b, nbt = tp_system()
    seqs = {}
    for on in (False, True):
        d = R.Disk(r, v, p, b, nbt)
        o = O.ODisk(r, v, p, b, nbt)
        it = R.Integrator(d, R.RD_RKF78, 0.05, adaptive=True, tol=1e-11)
        oi = O.OInteg(o, 4, 0.05, adaptive=True, tol=1e-11)
        it.set_error_check_for_tp(on)
        O.lib.or_set_error_check_for_tp(oi.h, 1 if on else 0)
        hs = []
        for _ in range(6):
            h = it.step()
            assert h == oi.step()
            hs.append(h)
        assert it.counters() == oi.counters() and it.dt_next() == oi.dt_next()
        assert same_state(d, o)
        seqs[on] = hs
    # the test particles' error shortens the steps when it counts
    assert sum(seqs[True]) < 0.5 * sum(seqs[False])


def superpl_system(kind):
    """Star, a protoplanet and a super-planetesimal (types sorted); the
    super-planetesimal either touches the protoplanet or sits inside the
    hit-centrum radius."""
    r, v, p, b, nbt = small_system(6, 81, [0, 3, 3, 4, 5, 5])
    p["mass"][1], p["mass"][3] = 1e-6, 1e-9
    p["radius"][1] = p["radius"][3] = 1e-3
    if kind == "collision":
        for c in "xyz":
            r[c][3] = r[c][1] + 1e-4
    else:
        r["x"][3], r["y"][3], r["z"][3] = 0.05, 0.0, 0.0
    return r, v, p, b, nbt


@pytest.mark.parametrize("kind", ["collision", "hit_centrum"])
def test_superplanetesimal_merge_refused_without_side_effects(kind):
    r, v, p, b, nbt = superpl_system(kind)
    thr = (0.1, 0.0, 1.0)
    d = R.Disk(r, v, p, b, nbt, threshold=thr)
    o = O.ODisk(r, v, p, b, nbt, thr=thr)
    if kind == "collision":
        assert d.check_collision() == 1
        import ctypes as C
        n = C.c_uint32()
        assert O.lib.or_check_collision(o.h, C.byref(n)) == 0 and n.value == 1
        ev = d.events()
        assert ev["idx1"][0] == 1 and ev["idx2"][0] == 3
        with pytest.raises(R.RedError) as ei:
            d.handle_collision()
        assert O.lib.or_handle_collision(o.h) == 3
    else:
        assert d.check_ejection_hit_centrum() == 1
        assert O.lib.or_check_ejection_hit_centrum(o.h) == 1
        with pytest.raises(R.RedError) as ei:
            d.handle_ejection_hit_centrum()
        assert O.lib.or_handle_ejection_hit_centrum(o.h) == 3
    assert ei.value.code == R.RD_ENOTIMPL and "super-planetesimal" in str(ei.value)
    # nothing was merged: state, ids, counters and log as before
    t, rr, vv, pp, bb = d.state()
    assert bits_equal(rr, r) and bits_equal(vv, v) and bits_equal(pp, p) and bits_equal(bb, b)
    assert len(d.event_log()) == 0
    assert int(np.sum(d.n_objects()[1])) == 0


def test_event_buffer_overflow_is_reported():
    """128 overlapping bodies make 8128 colliding pairs, more than the event
    buffer holds (4096 records; the reference's buffer has N_total records
    and its kernel writes past it, pp_disk.cu:357).  The kernel stops writing
    at the capacity and the check returns RD_EOVERFLOW; the disk stays usable."""
    n = 128
    r, v, p, b, nbt = small_system(n, 91, spread=1e-6)
    r["z"] = np.linspace(-1e-6, 1e-6, n)
    r["x"][0] = -10.0                       # the star apart
    p["radius"] = 1e-3
    thr = (0.0, 0.0, 1.0)
    d = R.Disk(r, v, p, b, nbt, threshold=thr)
    with pytest.raises(R.RedError) as ei:
        d.check_collision()
    assert ei.value.code == R.RD_EOVERFLOW
    import ctypes as C
    o = O.ODisk(r, v, p, b, nbt, thr=thr)
    cnt = C.c_uint32()
    assert O.lib.or_check_collision(o.h, C.byref(cnt)) == 4    # OR_EOVERFLOW
    d.clear_event_counter()
    # spread the bodies apart again: the next check finds nothing
    r2 = r.copy()
    r2["x"][1:] = np.arange(1, n) * 0.1
    d.set_state(0.0, r2, v)
    assert d.check_collision() == 0


def _steps_with_overlap(make, method, h0, tol, n, overlap, precision, guard=None):
    """rd_steps(n) with the device-side decision, commit and k0 pass on or off:
    (status, per-disk states, counters, dt_next, work counters)."""
    d = make()
    d.set_precision(precision)
    it = R.Integrator(d, method, h0, adaptive=True, tol=tol)
    it.set_overlap(overlap)
    if guard:
        it.set_guard(**guard)
    rc = 0
    try:
        it.steps(n)
    except R.RedError as e:
        rc = e.code
    nd = d.n_disk()
    return (rc, [d.state(k) for k in range(nd)], [it.counters(k) for k in range(nd)],
            [it.dt_next(k) for k in range(nd)], it.work())


def _batch(parts):
    cat = [np.concatenate([q[i] for q in parts]) for i in range(4)]
    return R.Disk(*cat, np.stack([q[4] for q in parts]), batch=True)


OVERLAP_CASES = {
    # name: (disk factory, method, h0, tol, steps, precision, guard)
    "one_disk_mixed_rkf78": (lambda: R.Disk(*R.make_disk(2000, 71)), R.RD_RKF78, 0.1 * KDAY, 1e-10, 40,
                             R.RD_PREC_MIXED, None),
    "one_disk_fp64_rkf78": (lambda: R.Disk(*R.make_disk(300, 72)), R.RD_RKF78, 0.1 * KDAY, 1e-10, 40,
                            R.RD_PREC_FP64, None),
    "one_disk_mixed_rk4": (lambda: R.Disk(*R.make_disk(1000, 73)), R.RD_RK4, 0.1 * KDAY, 1e-9, 40,
                           R.RD_PREC_MIXED, None),
    "batch_mixed_rkf78": (lambda: _batch([R.make_disk(n, 80 + n) for n in (64, 300, 999, 257, 1500)]), R.RD_RKF78,
                          0.1 * KDAY, 1e-10, 25, R.RD_PREC_MIXED, None),
    "batch_fp64_nan_guard": (lambda: _batch([R.make_disk(99, 61), coincident(), R.make_disk(299, 62)]), R.RD_RKF78,
                             0.01, 1e-11, 3, R.RD_PREC_FP64, dict(nonfinite=True)),
    "one_disk_nan_accepted": (lambda: R.Disk(*coincident()), R.RD_RKF78, 0.01, 1e-10, 3, R.RD_PREC_FP64, None),
}


@pytest.mark.parametrize("case", sorted(OVERLAP_CASES))
def test_overlap_changes_no_result(case):
    """rd_steps with the accept decision, commit and next k0 pass made on the
    device while the host computes the step size (rd_integrator_set_overlap)
    gives the results of the host-only sequence bit for bit: states,
    counters, next step sizes, force passes and pair counts, also where a
    disk of a batch fails the non-finite guard or a NaN state is accepted."""
    print(f"# {case}", flush=True)
    make, method, h0, tol, n, prec, guard = OVERLAP_CASES[case]
    on = _steps_with_overlap(make, method, h0, tol, n, True, prec, guard)
    off = _steps_with_overlap(make, method, h0, tol, n, False, prec, guard)
    assert on[0] == off[0]
    if case == "batch_fp64_nan_guard":
        assert on[0] == R.RD_ESTATE
    for (ta, ra, va, pa, ba), (tb, rb, vb, pb, bb) in zip(on[1], off[1]):
        assert ta == tb and bits_equal(pa, pb) and bits_equal(ba, bb)
        for c in "xyz":
            assert np.array_equal(ra[c], rb[c], equal_nan=True) and bits_equal(ra[c][~np.isnan(ra[c])],
                                                                                 rb[c][~np.isnan(rb[c])])
            assert np.array_equal(va[c], vb[c], equal_nan=True)
    assert on[2] == off[2]
    assert np.array_equal(np.array(on[3]), np.array(off[3]), equal_nan=True)
    assert on[4] == off[4]


def test_overlapped_steps_match_the_oracle():
    """FP64 rd_steps with the overlap on against the oracle's step loop, bit for
    bit, over a run with rejections (the device's decision must be the host's)."""
    r, v, p, b, nbt = R.make_disk(200, 74)
    d = R.Disk(r, v, p, b, nbt)
    o = O.ODisk(r, v, p, b, nbt)
    it = R.Integrator(d, R.RD_RKF78, 5.0 * KDAY, adaptive=True, tol=1e-12)
    oi = O.OInteg(o, R.RD_RKF78, 5.0 * KDAY, adaptive=True, tol=1e-12)
    it.steps(30)
    for _ in range(30):
        oi.step()
    assert it.counters() == oi.counters() and it.counters()[2] > 0
    assert it.dt_next() == oi.dt_next()
    assert same_state(d, o)
