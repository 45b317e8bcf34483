"""Edge cases of the GPU path against the oracle: a disk that is only a star,
a two-body disk, and a ragged batch mixing one-, two- and few-body disks with
larger ones (work lists with empty source ranges, partial tiles, single-sink
groups)."""
import numpy as np
import pytest

import oracle_lib as O
import redamd as R
from systems import small_system, bits_equal

pytestmark = pytest.mark.gpu

KDAY = R.K_GAUSS


def _kepler2():
    """Star + one planet on an eccentric orbit (e = 0.3, a = 1 AU), tau units."""
    r, v, p, b, nbt = small_system(2, 7, types=[0, 2])
    p["mass"][1] = 3e-6
    r["x"], r["y"], r["z"] = [0.0, 0.7], [0.0, 0.0], [0.0, 0.0]
    vp = np.sqrt((1.0 + 3e-6) * (1 + 0.3) / 0.7)      # periapsis speed, G = 1
    v["x"], v["y"], v["z"] = [0.0, 0.0], [0.0, vp], [0.0, 0.0]
    return r, v, p, b, nbt


@pytest.mark.parametrize("prec", [R.RD_PREC_FP64, R.RD_PREC_MIXED])
def test_star_only_disk(prec):
    r, v, p, b, nbt = small_system(1, 3)
    d = R.Disk(r, v, p, b, nbt)
    d.set_precision(prec)
    it = R.Integrator(d, R.RD_RKF78, 2 * KDAY, adaptive=True, tol=1e-10)
    it.steps(5)
    o = O.ODisk(r, v, p, b, nbt)
    oi = O.OInteg(o, 4, 2 * KDAY, adaptive=True, tol=1e-10)
    for _ in range(5):
        oi.step()
    t, rr, vv, _, _ = d.state()
    ot, orr, ovv, _, _ = o.state()
    assert t == ot and bits_equal(rr, orr) and bits_equal(vv, ovv)
    assert it.counters() == oi.counters()
    assert it.dt_next() == oi.dt_next()


@pytest.mark.parametrize("method", [R.RD_RK4, R.RD_RKF78])
def test_two_body_disk_bitexact(method):
    r, v, p, b, nbt = _kepler2()
    d = R.Disk(r, v, p, b, nbt)
    adaptive = method == R.RD_RKF78
    it = R.Integrator(d, method, 1 * KDAY, adaptive=adaptive, tol=1e-12)
    o = O.ODisk(r, v, p, b, nbt)
    oi = O.OInteg(o, method, 1 * KDAY, adaptive=adaptive, tol=1e-12)
    it.steps(60)
    for _ in range(60):
        oi.step()
    _, rr, vv, _, _ = d.state()
    _, orr, ovv, _, _ = o.state()
    assert bits_equal(rr, orr) and bits_equal(vv, ovv)
    assert it.counters() == oi.counters()


@pytest.mark.parametrize("prec", [R.RD_PREC_FP64, R.RD_PREC_MIXED])
def test_ragged_batch_with_tiny_disks(prec):
    sizes = [1, 2, 3, 17, 1, 65, 130, 2]
    parts = [small_system(n, 200 + k, spread=1.5) for k, n in enumerate(sizes)]
    cat = [np.concatenate([q[i] for q in parts]) for i in range(4)]
    nbt = np.stack([q[4] for q in parts])
    bd = R.Disk(*cat, nbt, batch=True)
    bd.set_precision(prec)
    bit = R.Integrator(bd, R.RD_RKF78, 0.5 * KDAY, adaptive=True, tol=1e-11)
    bit.steps(6)
    for k, q in enumerate(parts):
        sd = R.Disk(*q)
        sd.set_precision(prec)
        sit = R.Integrator(sd, R.RD_RKF78, 0.5 * KDAY, adaptive=True, tol=1e-11)
        sit.steps(6)
        a, s = bd.state(k), sd.state()
        assert a[0] == s[0], k
        assert bits_equal(a[1], s[1]) and bits_equal(a[2], s[2]), k
        assert bit.counters(k) == sit.counters(), k
        if prec == R.RD_PREC_FP64:
            o = O.ODisk(*q)
            oi = O.OInteg(o, 4, 0.5 * KDAY, adaptive=True, tol=1e-11)
            for _ in range(6):
                oi.step()
            _, orr, ovv, _, _ = o.state()
            assert bits_equal(a[1], orr) and bits_equal(a[2], ovv), k
    # integrals of the tiny disks
    for k in (0, 1, 7):
        I = bd.integral(k)
        assert np.isfinite(I["E"])


def test_exact_kernels_agree_at_batch_scale():
    """FP64 pair sums: a batch large enough for one-sink-per-lane work items
    (80 disks x 4097 bodies, >= 4 items of 256 sinks per CU) against the same
    disks run alone (term-parallel kernel): bit-identical trajectories."""
    parts = [R.make_disk(4096, 300 + k) for k in range(80)]
    cat = [np.concatenate([q[i] for q in parts]) for i in range(4)]
    nbt = np.stack([q[4] for q in parts])
    bd = R.Disk(*cat, nbt, batch=True)
    bit = R.Integrator(bd, R.RD_RKF78, 2 * KDAY, adaptive=True, tol=1e-11)
    bit.steps(2)
    for k in (0, 41, 79):
        sd = R.Disk(*parts[k])
        sit = R.Integrator(sd, R.RD_RKF78, 2 * KDAY, adaptive=True, tol=1e-11)
        sit.steps(2)
        a, s = bd.state(k), sd.state()
        assert a[0] == s[0]
        assert bits_equal(a[1], s[1]) and bits_equal(a[2], s[2]), k
        assert bit.counters(k) == sit.counters()


def test_integrator_destroyed_after_its_disk():
    """A garbage collector may finalize a disk before the integrator that holds
    it (objects in a reference cycle): destroying the integrator then must not
    read the disk, and later calls must not see an error from it."""
    r, v, p, b, nbt = R.make_disk(100, 3)
    d = R.Disk(r, v, p, b, nbt)
    it = R.Integrator(d, R.RD_RKF78, 1.0 * R.K_GAUSS, adaptive=True, tol=1e-10)
    it.steps(2)
    d.close()
    it.close()
    d2 = R.Disk(r, v, p, b, nbt)
    d2.set_precision(R.RD_PREC_MIXED)          # rebuilds the lists and launches
    a = d2.calc_dydx(1, 0, 0.0, r, v)
    assert np.all(np.isfinite(a["x"]))
