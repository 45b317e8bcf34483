"""rd_calc_integral_noalloc (the allocation-free rd_calc_integral) through the
C ABI: the same integrals bit for bit as rd_calc_integral, on two calls in a
row with the second on a smaller disk, on a disk of a batch, after steps,
and a clear refusal above its size limit."""
import ctypes as C

import numpy as np
import pytest

import redamd as R
from redamd_types import INTEGRAL
from systems import bits_equal

pytestmark = pytest.mark.gpu


def noalloc(d, k=0):
    out = np.zeros(1, INTEGRAL)
    rc = R.lib.rd_calc_integral_noalloc(d.h, k, out.ctypes.data_as(C.c_void_p))
    return rc, out[0]


def test_noalloc_integral_matches_rd_calc_integral():
    big = R.Disk(*R.make_disk(20000, 5))
    small = R.Disk(*R.make_disk(3000, 6))
    for d in (big, small, big):               # the second call is on a smaller disk
        rc, got = noalloc(d)
        assert rc == R.RD_OK, R.last_error()
        assert bits_equal(got, d.integral())
    # after a few steps, and on a disk of a batch
    it = R.Integrator(small, R.RD_RKF78, 2 * R.K_GAUSS, adaptive=True, tol=1e-11)
    it.steps(3)
    rc, got = noalloc(small)
    assert rc == R.RD_OK and bits_equal(got, small.integral())
    parts = [R.make_disk(n, 40 + n) for n in (500, 64, 2000)]
    cat = [np.concatenate([q[i] for q in parts]) for i in range(4)]
    bd = R.Disk(*cat, np.stack([q[4] for q in parts]), batch=True)
    for k in (2, 1, 0):
        rc, got = noalloc(bd, k)
        assert rc == R.RD_OK and bits_equal(got, bd.integral(k))


def test_noalloc_integral_refuses_above_its_limit():
    ok = R.Disk(*R.make_disk(65535, 7))       # 65536 bodies: the largest it takes
    rc, got = noalloc(ok)
    assert rc == R.RD_OK and bits_equal(got, ok.integral())
    d = R.Disk(*R.make_disk(65536, 8))        # 65537 bodies
    rc, _ = noalloc(d)
    assert rc == R.RD_EINVAL and "at most 65536 bodies" in R.last_error()
