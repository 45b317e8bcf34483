"""ctypes binding of librdamd.so (include/red_abi.h) for tests and bench.py.

The engine is C++/HIP; this module only converts numpy arrays to the C ABI's
plain pointers.  Array dtypes are byte-compatible with the reference's body
structs (red.type/red_type.h:206-212, 347-362, 375-381, 415-485).

Importing this module loads the HIP library and raises if it is missing: there
is no CPU fallback.
"""
import ctypes as C
import os

import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librdamd.so")
if HERE not in sys.path:
    sys.path.insert(0, HERE)

from redamd_types import (VAR4, PARAM, BODY_MD, INTEGRAL, EVENT, GasParams, GasGrid, Tuning, DiskSpec,  # noqa: F401
                          RD_OK, RD_EINVAL, RD_EDEVICE, RD_ENOTIMPL, RD_EOVERFLOW, RD_ESTATE, RD_ENOMEM,
                          RD_EULER, RD_RK2, RD_RK4, RD_RKF78, RD_PREC_FP64, RD_PREC_MIXED, RD_DRAG_NSI, RD_DRAG_ALL_MASSIVE,
                          BODY_STAR, BODY_GIANT, BODY_ROCKY, BODY_PROTO, BODY_SUPERPL, BODY_PL, BODY_TP,
                          EVENT_NONE, EVENT_HIT_CENTRUM, EVENT_EJECTION, EVENT_COLLISION, K_GAUSS,
                          RD_EVENT_COUNTER_TOTAL, RD_EVENT_COUNTER_LAST_CLEAR, RD_EVENT_COUNTER_LAST_STEP,
                          RD_PAIRS_ONE_SIDED, RD_PAIRS_SYMMETRIC)


class RedError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"rd error {code}: {msg}")
        self.code = code


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built (run __graft_entry__.build())")
    return C.CDLL(LIB_PATH)


lib = _load()
_P = C.c_void_p
_u32, _u64, _i32, _f64 = C.c_uint32, C.c_uint64, C.c_int32, C.c_double
_SIGS = {
    "rd_last_error": (C.c_char_p, []),
    "rd_version": (C.c_int, []),
    "rd_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "rd_disk_create": (C.c_int, [_u32, _P, _P, _P, _P, _P, _f64, _P, _P, C.c_int, C.POINTER(_P)]),
    "rd_batch_create": (C.c_int, [_u32, _P, _P, _P, _P, _P, _P, _P, _P, C.c_int, C.POINTER(_P)]),
    "rd_disk_destroy": (C.c_int, [_P]),
    "rd_disk_set_precision": (C.c_int, [_P, C.c_int]),
    "rd_disk_set_drag_mode": (C.c_int, [_P, C.c_int]),
    "rd_disk_set_gas_grid": (C.c_int, [_P, _P]),
    "rd_calc_dydx_device": (C.c_int, [_P, _u32, _u32, _u32, _f64, _P, _P, _P]),
    "rd_disk_set_tuning": (C.c_int, [_P, C.POINTER(Tuning)]),
    "rd_disk_get_tuning": (C.c_int, [_P, C.POINTER(Tuning)]),
    "rd_autotune": (C.c_int, [_P, C.POINTER(Tuning)]),
    "rd_disk_count": (C.c_int, [_P, C.POINTER(_u32)]),
    "rd_disk_n_total": (C.c_int, [_P, _u32, C.POINTER(_u32)]),
    "rd_disk_n_objects": (C.c_int, [_P, _u32, _P, _P, _P]),
    "rd_get_state": (C.c_int, [_P, _u32, C.POINTER(_f64), _P, _P, _P, _P]),
    "rd_set_state": (C.c_int, [_P, _u32, _f64, _P, _P]),
    "rd_sync": (C.c_int, [_P]),
    "rd_calc_dydx": (C.c_int, [_P, _u32, _u32, _u32, _f64, _P, _P, _P]),
    "rd_bench_force": (C.c_int, [_P, _u32, C.POINTER(_f64), C.POINTER(_f64)]),
    "rd_bench_copy": (C.c_int, [C.c_int, C.c_size_t, _u32, C.POINTER(_f64)]),
    "rd_bench_force_stage": (C.c_int, [_P, _u32, C.POINTER(_f64)]),
    "rd_integrator_create": (C.c_int, [_P, C.c_int, _f64, C.c_int, _f64, C.POINTER(_P)]),
    "rd_integrator_destroy": (C.c_int, [_P]),
    "rd_step": (C.c_int, [_P, C.POINTER(_f64)]),
    "rd_steps": (C.c_int, [_P, _u32]),
    "rd_get_dt_next": (C.c_int, [_P, _u32, C.POINTER(_f64)]),
    "rd_set_dt_next": (C.c_int, [_P, _u32, _f64]),
    "rd_counters": (C.c_int, [_P, _u32, C.POINTER(_u64), C.POINTER(_u64), C.POINTER(_u64)]),
    "rd_set_error_check_for_tp": (C.c_int, [_P, C.c_int]),
    "rd_integrator_set_guard": (C.c_int, [_P, _u32, _f64, C.c_int]),
    "rd_integrator_set_overlap": (C.c_int, [_P, C.c_int]),
    "rd_bench_stages": (C.c_int, [_P, C.c_int, _u32, C.POINTER(_f64)]),
    "rd_calc_integral_noalloc": (C.c_int, [_P, _u32, _P]),
    "rd_work_counters": (C.c_int, [_P, C.POINTER(_u64), C.POINTER(_f64)]),
    "rd_run": (C.c_int, [_P, _f64, _f64, _u64]),
    "rd_check_ejection_hit_centrum": (C.c_int, [_P, _u32, C.POINTER(_u32)]),
    "rd_check_collision": (C.c_int, [_P, _u32, C.POINTER(_u32)]),
    "rd_get_events": (C.c_int, [_P, _u32, _P, _u32, C.POINTER(_u32)]),
    "rd_handle_collision": (C.c_int, [_P, _u32]),
    "rd_handle_ejection_hit_centrum": (C.c_int, [_P, _u32]),
    "rd_clear_event_counter": (C.c_int, [_P, _u32]),
    "rd_event_log": (C.c_int, [_P, _u32, _P, _u32, C.POINTER(_u32)]),
    "rd_event_counters": (C.c_int, [_P, _u32, _P, _P, _P, _P]),
    "rd_set_event_counter": (C.c_int, [_P, _u32, C.c_int, _u32]),
    "rd_rebuild_vectors": (C.c_int, [_P, _u32]),
    "rd_calc_integral": (C.c_int, [_P, _u32, _P]),
    "rd_make_disk": (C.c_int, [C.POINTER(DiskSpec), _P, _P, _P, _P]),
    "rd_mlp_create": (C.c_int, [C.c_int, _u32, C.POINTER(_P)]),
    "rd_mlp_destroy": (C.c_int, [_P]),
    "rd_mlp_set_norm": (C.c_int, [_P, _P, _P, C.c_float, C.c_float]),
    "rd_mlp_set_target": (C.c_int, [_P, C.c_int, C.c_float]),
    "rd_mlp_set_params": (C.c_int, [_P, _P]),
    "rd_mlp_get_params": (C.c_int, [_P, _P, _P]),
    "rd_mlp_forward": (C.c_int, [_P, _P, _u32, _P]),
    "rd_mlp_loss_grad": (C.c_int, [_P, _P, _P, _u32, C.POINTER(C.c_float), _P]),
    "rd_mlp_train": (C.c_int, [_P, _P, _P, _u32, _u32, _u32, C.c_float, C.POINTER(C.c_float)]),
    "rd_integrator_set_proposer": (C.c_int, [_P, _P, C.c_float]),
    "rd_integrator_set_logging": (C.c_int, [_P, C.c_int]),
    "rd_integrator_log": (C.c_int, [_P, _P, _P, _u32, C.POINTER(_u32)]),
    "rd_integrator_clear_log": (C.c_int, [_P]),
}
for _name, (_res, _args) in _SIGS.items():
    _f = getattr(lib, _name)
    _f.restype = _res
    _f.argtypes = _args

EXPORTED = sorted(_SIGS)
ABI_VERSION = 3   # include/red_abi.h RD_ABI_VERSION: the structs below follow that header
if lib.rd_version() != ABI_VERSION:
    raise ImportError(f"{LIB_PATH} has ABI version {lib.rd_version()}, this binding expects {ABI_VERSION}: rebuild it")


def as_bodies(r, v, p, bmd):
    """Contiguous arrays in the ABI's byte layout (np.concatenate and similar
    drop the 24-byte padded layout of BODY_MD, so every array is recast)."""
    return (np.ascontiguousarray(np.asarray(r).astype(VAR4, copy=False)),
            np.ascontiguousarray(np.asarray(v).astype(VAR4, copy=False)),
            np.ascontiguousarray(np.asarray(p).astype(PARAM, copy=False)),
            np.ascontiguousarray(np.asarray(bmd).astype(BODY_MD, copy=False)))


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def check(rc):
    if rc != RD_OK:
        raise RedError(rc, lib.rd_last_error().decode())
    return rc


def _finalize(obj):
    """Release obj's handle when it is garbage-collected, but not while the
    interpreter shuts down: finalizers then run in no set order, possibly
    after the library or the GPU runtime began its own teardown, and the
    process exit releases every device resource anyway."""
    if sys is None or sys.is_finalizing():
        return
    try:
        obj.close()
    except Exception:
        pass


def last_error():
    """rd_last_error() of the calling thread."""
    return lib.rd_last_error().decode()


def device_count():
    n = C.c_int(0)
    check(lib.rd_device_count(C.byref(n)))
    return n.value


def bench_copy(nbytes, n_iter=10, device=0):
    """Mean time [ms] of a plain device copy of nbytes (calibrates byte counters)."""
    ms = _f64()
    check(lib.rd_bench_copy(device, nbytes, n_iter, C.byref(ms)))
    return ms.value


def make_disk(n, seed, body_type=BODY_PROTO, a_min=0.5, a_max=5.0, a_power=-0.5, sig_e=0.01, sig_i=0.005,
              size_mode=0, lo=1e-10, hi=1e-8, density_cgs=2.0, cd=0.0, m_star=1.0, r_star=1.0 / 215.094):
    """Star + n bodies; returns (r, v, p, bmd, n_by_type)."""
    s = DiskSpec(n, seed, body_type, a_min, a_max, a_power, sig_e, sig_i, size_mode, lo, hi, density_cgs, cd,
                 m_star, r_star)
    r = np.zeros(n + 1, VAR4)
    v = np.zeros(n + 1, VAR4)
    p = np.zeros(n + 1, PARAM)
    b = np.zeros(n + 1, BODY_MD)
    check(lib.rd_make_disk(C.byref(s), _ptr(r), _ptr(v), _ptr(p), _ptr(b)))
    nbt = np.zeros(7, np.uint32)
    nbt[0] = 1
    nbt[body_type] += n
    return r, v, p, b, nbt


def gas_params(sigma_cgs=1700.0, sigma_p=-1.5, h_c=0.05, h_p=1.25, eta=0.002, eta_p=0.0):
    """Analytic gas disk with the reference's exp(-(z/h)^2) profile
    (pp_disk.cu:87-105): rho_mid = Sigma / (sqrt(pi) h), h = h_c r^h_p."""
    AuToMeter = 1.495978707e11
    cm_au = 1.0e-2 / AuToMeter
    gram_solar = 1.0 / (1.0e3 * 1.98911e30)
    sig_solar_au2 = sigma_cgs * gram_solar / (cm_au * cm_au)
    g = GasParams()
    g.rho[0] = sig_solar_au2 / (np.sqrt(np.pi) * h_c)
    g.rho[1] = sigma_p - h_p
    g.sch[0], g.sch[1] = h_c, h_p
    g.eta[0], g.eta[1] = eta, eta_p
    g.gas_decrease = 0
    return g


class Disk:
    """pp_disk counterpart (src/red.cuda/pp_disk.h) over one or more disks."""

    def __init__(self, r, v, p, bmd, n_by_type, t=0.0, threshold=(0.0, 0.0, 0.0), gas=None, device=0,
                 batch=False):
        self.h = C.c_void_p()
        thr = np.ascontiguousarray(threshold, np.float64)
        r, v, p, bmd = as_bodies(r, v, p, bmd)
        self._keep = (r, v, p, bmd)
        gp = C.byref(gas) if gas is not None else None
        if batch:
            nbt = np.ascontiguousarray(n_by_type, np.uint32).reshape(-1)
            nd = nbt.size // 7
            tt = np.ascontiguousarray(np.broadcast_to(np.asarray(t, np.float64), (nd,)))
            check(lib.rd_batch_create(nd, _ptr(nbt), _ptr(r), _ptr(v), _ptr(p), _ptr(bmd), _ptr(tt), _ptr(thr), gp,
                                      device, C.byref(self.h)))
        else:
            nbt = np.ascontiguousarray(n_by_type, np.uint32)
            check(lib.rd_disk_create(len(r), _ptr(nbt), _ptr(r), _ptr(v), _ptr(p), _ptr(bmd), float(t), _ptr(thr),
                                     gp, device, C.byref(self.h)))

    def close(self):
        if self.h:
            lib.rd_disk_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        _finalize(self)

    def set_precision(self, prec):
        check(lib.rd_disk_set_precision(self.h, prec))

    def set_drag_mode(self, mode):
        check(lib.rd_disk_set_drag_mode(self.h, mode))

    def set_gas_grid(self, used_rad, density, vrad, vtheta):
        """Drag through frame 0 of a FARGO grid: used_rad holds the n_rad + 1
        ring interface radii, the three fields are (n_rad, n_sec) arrays in the
        solver's units.  used_rad=None removes the grid."""
        if used_rad is None:
            check(lib.rd_disk_set_gas_grid(self.h, None))
            return
        ur = np.ascontiguousarray(used_rad, np.float64)
        f = [np.ascontiguousarray(x, np.float64) for x in (density, vrad, vtheta)]
        n_rad = ur.size - 1
        if n_rad < 1 or f[0].ndim != 2 or f[0].shape[0] != n_rad or any(x.shape != f[0].shape for x in f):
            raise ValueError("grid fields must be (n_rad, n_sec) with n_rad + 1 radii")
        g = GasGrid(n_rad, f[0].shape[1], ur.ctypes.data, f[0].ctypes.data, f[1].ctypes.data, f[2].ctypes.data)
        check(lib.rd_disk_set_gas_grid(self.h, C.byref(g)))

    def tuning(self):
        t = Tuning()
        check(lib.rd_disk_get_tuning(self.h, C.byref(t)))
        return t

    def set_tuning(self, sinks_per_lane=None, chunk=None, blocks_per_cu=None, use_graph=None, pair_mode=None,
                   sym_chunk_min=None):
        t = self.tuning()
        if sinks_per_lane is not None:
            t.sinks_per_lane = sinks_per_lane
        if chunk is not None:
            t.chunk = chunk
        if blocks_per_cu is not None:
            t.blocks_per_cu = blocks_per_cu
        if use_graph is not None:
            t.use_graph = use_graph
        if pair_mode is not None:
            t.pair_mode = pair_mode
        if sym_chunk_min is not None:
            t.sym_chunk_min = sym_chunk_min
        check(lib.rd_disk_set_tuning(self.h, C.byref(t)))

    def autotune(self):
        t = Tuning()
        check(lib.rd_autotune(self.h, C.byref(t)))
        return t

    def n_disk(self):
        n = _u32()
        check(lib.rd_disk_count(self.h, C.byref(n)))
        return n.value

    def n_total(self, k=0):
        n = _u32()
        check(lib.rd_disk_n_total(self.h, k, C.byref(n)))
        return n.value

    def n_objects(self, k=0):
        a, b, c = (np.zeros(7, np.uint32) for _ in range(3))
        check(lib.rd_disk_n_objects(self.h, k, _ptr(a), _ptr(b), _ptr(c)))
        return a, b, c

    def state(self, k=0):
        n = self.n_total(k)
        t = _f64()
        r, v, p, b = np.zeros(n, VAR4), np.zeros(n, VAR4), np.zeros(n, PARAM), np.zeros(n, BODY_MD)
        check(lib.rd_get_state(self.h, k, C.byref(t), _ptr(r), _ptr(v), _ptr(p), _ptr(b)))
        return t.value, r, v, p, b

    def set_state(self, t, r, v, k=0):
        r, v = np.ascontiguousarray(r, VAR4), np.ascontiguousarray(v, VAR4)
        check(lib.rd_set_state(self.h, k, t, _ptr(r), _ptr(v)))

    def calc_dydx(self, i, rr, t, r, v, k=0):
        r, v = np.ascontiguousarray(r, VAR4), np.ascontiguousarray(v, VAR4)
        dy = np.zeros(len(r), VAR4)
        check(lib.rd_calc_dydx(self.h, k, i, rr, t, _ptr(r), _ptr(v), _ptr(dy)))
        return dy

    def calc_dydx_device(self, i, rr, t, r_dev, v_dev, dy_dev, k=0):
        """rd_calc_dydx_device: r_dev, v_dev and dy_dev are device addresses
        (integers) of rd_var4 record arrays on the disk's device."""
        check(lib.rd_calc_dydx_device(self.h, k, i, rr, t, C.c_void_p(r_dev), C.c_void_p(v_dev),
                                      C.c_void_p(dy_dev)))

    def bench_force(self, n_iter=10):
        ms, pairs = _f64(), _f64()
        check(lib.rd_bench_force(self.h, n_iter, C.byref(ms), C.byref(pairs)))
        return ms.value, pairs.value

    def bench_force_stage(self, n_iter=10):
        ms = _f64()
        check(lib.rd_bench_force_stage(self.h, n_iter, C.byref(ms)))
        return ms.value

    def bench_stages(self, method=None, n_iter=10):
        """Mean ms of one adaptive attempt's stage kernels (rd_bench_stages)."""
        ms = _f64()
        check(lib.rd_bench_stages(self.h, RD_RKF78 if method is None else method, n_iter, C.byref(ms)))
        return ms.value

    def check_ejection_hit_centrum(self, k=0):
        n = _u32()
        check(lib.rd_check_ejection_hit_centrum(self.h, k, C.byref(n)))
        return n.value

    def check_collision(self, k=0):
        n = _u32()
        check(lib.rd_check_collision(self.h, k, C.byref(n)))
        return n.value

    def events(self, k=0):
        n = _u32()
        check(lib.rd_get_events(self.h, k, None, 0, C.byref(n)))
        out = np.zeros(n.value, EVENT)
        check(lib.rd_get_events(self.h, k, _ptr(out), n.value, C.byref(n)))
        return out

    def event_log(self, k=0):
        n = _u32()
        check(lib.rd_event_log(self.h, k, None, 0, C.byref(n)))
        out = np.zeros(n.value, EVENT)
        check(lib.rd_event_log(self.h, k, _ptr(out), n.value, C.byref(n)))
        return out

    def event_counters(self, k=0):
        """{'event'|'collision'|'ejection'|'hit_centrum': [total, since last clear, last step]}"""
        arrs = [np.zeros(3, np.uint32) for _ in range(4)]
        check(lib.rd_event_counters(self.h, k, *[_ptr(a) for a in arrs]))
        return dict(zip(("event", "collision", "ejection", "hit_centrum"), (a.tolist() for a in arrs)))

    def set_event_counter(self, field, value, k=0):
        check(lib.rd_set_event_counter(self.h, k, field, value))

    def handle_collision(self, k=0):
        check(lib.rd_handle_collision(self.h, k))

    def handle_ejection_hit_centrum(self, k=0):
        check(lib.rd_handle_ejection_hit_centrum(self.h, k))

    def clear_event_counter(self, k=0):
        check(lib.rd_clear_event_counter(self.h, k))

    def rebuild_vectors(self, k=0):
        check(lib.rd_rebuild_vectors(self.h, k))

    def integral(self, k=0):
        out = np.zeros(1, INTEGRAL)
        check(lib.rd_calc_integral(self.h, k, _ptr(out)))
        return out[0]

    def sync(self):
        check(lib.rd_sync(self.h))


class Integrator:
    """rungekutta4 / rungekutta8 counterpart (int_rungekutta4.h, int_rungekutta8.h)."""

    def __init__(self, disk, method, dt, adaptive=True, tol=1e-10):
        self.disk = disk
        self.h = C.c_void_p()
        check(lib.rd_integrator_create(disk.h, method, dt, 1 if adaptive else 0, tol, C.byref(self.h)))

    def close(self):
        if self.h:
            lib.rd_integrator_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        _finalize(self)

    def step(self):
        dt = _f64()
        check(lib.rd_step(self.h, C.byref(dt)))
        return dt.value

    def steps(self, n):
        check(lib.rd_steps(self.h, n))

    def run(self, length, output_interval, max_steps=2 ** 62):
        check(lib.rd_run(self.h, length, output_interval, max_steps))

    def dt_next(self, k=0):
        d = _f64()
        check(lib.rd_get_dt_next(self.h, k, C.byref(d)))
        return d.value

    def set_dt_next(self, dt, k=0):
        check(lib.rd_set_dt_next(self.h, k, dt))

    def counters(self, k=0):
        a, b, c = _u64(), _u64(), _u64()
        check(lib.rd_counters(self.h, k, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def set_error_check_for_tp(self, on):
        check(lib.rd_set_error_check_for_tp(self.h, 1 if on else 0))

    def set_guard(self, max_attempts=0, min_dt=0.0, nonfinite=False):
        """Opt-in bound on the retry loop (rd_integrator_set_guard)."""
        check(lib.rd_integrator_set_guard(self.h, max_attempts, min_dt, 1 if nonfinite else 0))

    def set_overlap(self, on=True):
        """rd_steps: accept, commit and the next k0 pass on the device while the
        host reads the error (rd_integrator_set_overlap; on by default)."""
        check(lib.rd_integrator_set_overlap(self.h, 1 if on else 0))

    def step_rc(self):
        """(status, dt_did) of rd_step without raising: the guard's RD_ESTATE."""
        dt = _f64()
        rc = lib.rd_step(self.h, C.byref(dt))
        return rc, dt.value

    def work(self):
        e, p = _u64(), _f64()
        check(lib.rd_work_counters(self.h, C.byref(e), C.byref(p)))
        return e.value, p.value

    def set_proposer(self, mlp, safety=1.0):
        check(lib.rd_integrator_set_proposer(self.h, mlp.h if mlp is not None else None, safety))
        self._mlp = mlp

    def set_logging(self, on):
        check(lib.rd_integrator_set_logging(self.h, 1 if on else 0))

    def log(self):
        n = _u32()
        check(lib.rd_integrator_log(self.h, None, None, 0, C.byref(n)))
        x = np.zeros((n.value, MLP_IN), np.float32)
        t = np.zeros(n.value, np.float32)
        check(lib.rd_integrator_log(self.h, _ptr(x), _ptr(t), n.value, C.byref(n)))
        return x, t

    def clear_log(self):
        check(lib.rd_integrator_clear_log(self.h))


MLP_IN, MLP_HIDDEN = 12, 64
MLP_NPARAM = MLP_IN * MLP_HIDDEN + MLP_HIDDEN + MLP_HIDDEN * MLP_HIDDEN + MLP_HIDDEN + MLP_HIDDEN + 1


class Mlp:
    """The step-size proposer network (12 -> 64 -> 64 -> 1, BF16 weights)."""

    def __init__(self, seed=0, device=0):
        self.h = C.c_void_p()
        check(lib.rd_mlp_create(device, seed, C.byref(self.h)))

    def close(self):
        if self.h:
            lib.rd_mlp_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        _finalize(self)

    def set_norm(self, mean, std, t_mean, t_std):
        m = np.ascontiguousarray(mean, np.float32)
        s = np.ascontiguousarray(std, np.float32)
        check(lib.rd_mlp_set_norm(self.h, _ptr(m), _ptr(s), float(t_mean), float(t_std)))

    def set_target(self, base=-1, quantile=0.0):
        """Train on offsets from feature `base` (-1: none) and fit the mean
        (quantile 0) or the given quantile; the output stays log h."""
        check(lib.rd_mlp_set_target(self.h, int(base), float(quantile)))

    def params(self):
        w = np.zeros(MLP_NPARAM, np.float32)
        wb = np.zeros(MLP_NPARAM, np.uint16)
        check(lib.rd_mlp_get_params(self.h, _ptr(w), _ptr(wb)))
        return w, wb

    def set_params(self, w):
        w = np.ascontiguousarray(w, np.float32)
        check(lib.rd_mlp_set_params(self.h, _ptr(w)))

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros(len(x), np.float32)
        check(lib.rd_mlp_forward(self.h, _ptr(x), len(x), _ptr(y)))
        return y

    def loss_grad(self, x, t):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        g = np.zeros(MLP_NPARAM, np.float32)
        loss = C.c_float()
        check(lib.rd_mlp_loss_grad(self.h, _ptr(x), _ptr(t), len(x), C.byref(loss), _ptr(g)))
        return loss.value, g

    def train(self, x, t, epochs=10, batch=256, lr=1e-3):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        loss = C.c_float()
        check(lib.rd_mlp_train(self.h, _ptr(x), _ptr(t), len(x), epochs, batch, lr, C.byref(loss)))
        return loss.value
