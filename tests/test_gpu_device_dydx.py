"""rd_calc_dydx_device: calc_dydx on device arrays (the reference's operator
works on the GPU copies it owns, pp_disk.cu:1450-1502), through the C ABI with
device memory from hipMalloc.  The same kernels as rd_calc_dydx, so the result
is the host entry point's bit for bit; pointers that are not device memory of
the disk's device, or allocations too short for the disk, are refused before
any kernel runs."""
import ctypes as C
import os

import numpy as np
import pytest

import redamd as R
from systems import small_system, mixed_types, bits_equal

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    """The HIP runtime the engine links (hipMalloc / hipMemcpy / hipFree)."""
    global _hip
    if _hip is None:
        _hip = C.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so"))
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipMemGetAddressRange.argtypes = [C.POINTER(C.c_void_p), C.POINTER(C.c_size_t), C.c_void_p]
    return _hip


def address_range(p):
    """(base, size) of the allocation holding device address p, as the runtime reports it."""
    base, size = C.c_void_p(), C.c_size_t()
    assert hip().hipMemGetAddressRange(C.byref(base), C.byref(size), p) == 0
    return base.value, size.value


H2D, D2H = 1, 2


class DevArray:
    def __init__(self, nbytes):
        self.p = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.p), nbytes) == 0
        self.nbytes = nbytes

    def put(self, a):
        assert a.nbytes <= self.nbytes
        assert hip().hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, a):
        assert hip().hipMemcpy(a.ctypes.data, self.p, a.nbytes, D2H) == 0
        return a

    def __del__(self):
        if self.p:
            hip().hipFree(self.p)


@pytest.mark.parametrize("prec,gas", [(R.RD_PREC_FP64, False), (R.RD_PREC_MIXED, False), (R.RD_PREC_FP64, True)])
def test_device_dydx_equals_host_dydx(prec, gas):
    n = 700
    r, v, p, b, nbt = small_system(n, 51, mixed_types(n), inactive=(5, 300))
    d = R.Disk(r, v, p, b, nbt, t=2.0, gas=R.gas_params() if gas else None)
    d.set_precision(prec)
    # the state the operator is asked about differs from the disk's own
    r2, v2 = r.copy(), v.copy()
    r2["x"] += 1e-3
    v2["y"] *= 0.99
    dr, dv, ddy = DevArray(r2.nbytes), DevArray(v2.nbytes), DevArray(r2.nbytes)
    dr.put(np.ascontiguousarray(r2))
    dv.put(np.ascontiguousarray(v2))
    for i in (1, 0):
        want = d.calc_dydx(i, 3, 2.5, r2, v2)
        d.calc_dydx_device(i, 3, 2.5, dr.p.value, dv.p.value, ddy.p.value)
        got = ddy.get(np.zeros(n, R.VAR4))
        assert bits_equal(got, want)
    # the disk's own state is untouched
    assert bits_equal(d.state()[1], r)


def test_device_dydx_on_a_batch_disk():
    systems = [small_system(n, s) for n, s in ((300, 1), (450, 2))]
    cat = [np.concatenate([s[q] for s in systems]) for q in range(4)]
    d = R.Disk(*cat, np.concatenate([s[4] for s in systems]), batch=True)
    r, v = systems[1][0], systems[1][1]
    dr, dv, ddy = DevArray(r.nbytes), DevArray(v.nbytes), DevArray(r.nbytes)
    dr.put(r)
    dv.put(v)
    d.calc_dydx_device(1, 0, 0.0, dr.p.value, dv.p.value, ddy.p.value, k=1)
    assert bits_equal(ddy.get(np.zeros(len(r), R.VAR4)), d.calc_dydx(1, 0, 0.0, r, v, k=1))


def test_device_dydx_refuses_bad_pointers():
    n = 100
    r, v, p, b, nbt = small_system(n, 3)
    d = R.Disk(r, v, p, b, nbt)
    good = [DevArray(r.nbytes) for _ in range(3)]
    for g, a in zip(good, (r, v)):
        g.put(a)
    host = np.zeros(n, R.VAR4)
    # a host array where device memory is expected
    with pytest.raises(R.RedError, match="not device memory"):
        d.calc_dydx_device(1, 0, 0.0, host.ctypes.data, good[1].p.value, good[2].p.value)
    with pytest.raises(R.RedError, match="not device memory"):
        d.calc_dydx_device(0, 0, 0.0, good[0].p.value, good[1].p.value, host.ctypes.data)
    # an allocation one record short (the check uses the allocation's size as the
    # runtime reports it; a size rounded up past the last record is memory the
    # records may use)
    short = DevArray(r.nbytes - 32)
    if address_range(short.p) == (short.p.value, short.nbytes):
        with pytest.raises(R.RedError, match="fewer records"):
            d.calc_dydx_device(1, 0, 0.0, good[0].p.value, good[1].p.value, short.p.value)
    # records starting 16 bytes before the reported end always are refused
    base, size = address_range(good[2].p)
    with pytest.raises(R.RedError, match="fewer records"):
        d.calc_dydx_device(1, 0, 0.0, good[0].p.value, good[1].p.value, base + size - 16)
    # an offset into an allocation that still holds the disk's records works
    big = DevArray(r.nbytes + 64)
    d.calc_dydx_device(1, 0, 0.0, good[0].p.value, good[1].p.value, big.p.value + 32)
    out = np.zeros(n + 2, R.VAR4)
    big.get(out)
    assert bits_equal(out[1:n + 1], d.calc_dydx(1, 0, 0.0, r, v))
    # and the engine still runs afterwards (no error left behind)
    it = R.Integrator(d, R.RD_RKF78, 0.1 * R.K_GAUSS, adaptive=True, tol=1e-10)
    it.steps(2)
