"""ctypes binding of the CPU oracle (oracle/liboracle.so) -- test infrastructure.

Array dtypes are shared with the engine binding (reference byte layout)."""
import ctypes as C
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "red.cuda_amd"))
from redamd_types import VAR4, PARAM, BODY_MD, INTEGRAL, EVENT, GasParams  # noqa: E402

LIB = os.path.join(ROOT, "oracle", "liboracle.so")
lib = C.CDLL(LIB)
_P, _u32, _u64, _i32, _f64 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int32, C.c_double


class VAR4C(C.Structure):
    _fields_ = [(n, C.c_double) for n in "xyzw"]


class OrRng(C.Structure):
    _fields_ = [("I", C.c_uint32 * 624), ("idx", C.c_uint32)]


class OrOrbelem(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("sma", "ecc", "inc", "peri", "node", "mean")]


class OrGrid(C.Structure):
    _fields_ = [("n_rad", C.c_uint32), ("n_sec", C.c_uint32), ("used_rad", C.c_void_p),
                ("density", C.c_void_p), ("vrad", C.c_void_p), ("vtheta", C.c_void_p)]


class OrDiskSpec(C.Structure):
    _fields_ = [("n", C.c_uint32), ("seed", C.c_uint32), ("body_type", C.c_int32),
                ("a_min", C.c_double), ("a_max", C.c_double), ("a_power", C.c_double),
                ("sig_e", C.c_double), ("sig_i", C.c_double), ("size_mode", C.c_int32),
                ("lo", C.c_double), ("hi", C.c_double), ("density_cgs", C.c_double),
                ("cd", C.c_double), ("m_star", C.c_double), ("r_star", C.c_double)]


_SIGS = {
    "or_bound": (None, [_P, C.c_int, _P]),
    "or_calc_radius": (_f64, [_f64, _f64]),
    "or_calc_density": (_f64, [_f64, _f64]),
    "or_calc_mass": (_f64, [_f64, _f64]),
    "or_calc_physical_properties": (None, [_P, _P, _P]),
    "or_calc_position_after_collision": (None, [_f64, _f64, _P, _P, _P]),
    "or_calc_velocity_after_collision": (None, [_f64, _f64, _P, _P, _P]),
    "or_get_total_mass": (_f64, [_u32, _P, _P]),
    "or_calc_bc": (None, [_u32, _P, _P, _P, _P, _f64, _P, _P]),
    "or_transform_to_bc": (None, [_u32, _P, _P, _P, _P]),
    "or_calc_integral": (None, [_u32, _P, _P, _P, _P, _P]),
    "or_kepler_equation_solver": (C.c_int, [_f64, _f64, _f64, C.POINTER(_f64)]),
    "or_calc_phase": (C.c_int, [_f64, C.POINTER(OrOrbelem), _P, _P]),
    "or_rng_seed": (None, [C.POINTER(OrRng), _u32]),
    "or_rng_u32": (_u32, [C.POINTER(OrRng)]),
    "or_rng_uniform": (_f64, [C.POINTER(OrRng), _f64, _f64]),
    "or_rng_rayleigh": (_f64, [C.POINTER(OrRng), _f64]),
    "or_rng_power_law": (_f64, [C.POINTER(OrRng), _f64, _f64, _f64]),
    "or_make_disk": (C.c_int, [C.POINTER(OrDiskSpec), _P, _P, _P, _P]),
    "or_grav_gpu": (None, [_u32, _P, _P, _P, _P, _P]),
    "or_grav_cpu": (C.c_int, [_u32, _P, _P, _P, _P, _P]),
    "or_drag": (None, [_f64, C.POINTER(GasParams), C.c_int, _u32, _P, _P, _P, _P, _P]),
    "or_rotate_2d": (VAR4C, [_f64, VAR4C]),
    "or_grid_cell": (_u32, [C.POINTER(OrGrid), _f64, _f64]),
    "or_grid_drag": (None, [C.POINTER(OrGrid), C.c_int, _P, _P, _P, _P, _P]),
    "or_disk_set_gas_grid": (C.c_int, [_P, _u32, _u32, _P, _P, _P, _P]),
    "or_disk_create": (_P, [_u32, _P, _P, _P, _P, _P, _f64, _P, _P]),
    "or_disk_destroy": (None, [_P]),
    "or_disk_set_formula": (None, [_P, C.c_int]),
    "or_disk_set_drag_mode": (None, [_P, C.c_int]),
    "or_disk_n": (_u32, [_P]),
    "or_disk_t": (_f64, [_P]),
    "or_disk_get": (None, [_P, _P, _P, _P, _P]),
    "or_disk_counts": (None, [_P, _P, _P, _P]),
    "or_calc_dydx": (C.c_int, [_P, _u32, _u32, _f64, _P, _P, _P]),
    "or_disk_integral": (None, [_P, _P]),
    "or_check_ejection_hit_centrum": (_u32, [_P]),
    "or_check_collision": (C.c_int, [_P, C.POINTER(_u32)]),
    "or_handle_ejection_hit_centrum": (C.c_int, [_P]),
    "or_handle_collision": (C.c_int, [_P]),
    "or_clear_event_counter": (None, [_P]),
    "or_rebuild_vectors": (C.c_int, [_P]),
    "or_event_log": (_u32, [_P, _P, _u32]),
    "or_integ_create": (_P, [_P, C.c_int, _f64, C.c_int, _f64]),
    "or_integ_destroy": (None, [_P]),
    "or_step": (C.c_int, [_P, C.POINTER(_f64)]),
    "or_get_dt_next": (_f64, [_P]),
    "or_set_dt_next": (None, [_P, _f64]),
    "or_counters": (None, [_P, _P]),
    "or_set_error_check_for_tp": (None, [_P, C.c_int]),
    "or_last_error": (_f64, [_P]),
    "or_set_guard": (None, [_P, _u32, _f64, C.c_int]),
    "or_rk8_tableau": (None, [_P, _P, _P]),
    "or_last_error_terms": (_u32, [_P, _P, _u32]),
    "or_run": (C.c_int, [_P, _P, _f64, _f64, _u64, C.POINTER(_u64)]),
}
for _n, (_r, _a) in _SIGS.items():
    _f = getattr(lib, _n)
    _f.restype = _r
    _f.argtypes = _a


def ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def make_disk(n, seed, body_type=3, a_min=0.5, a_max=5.0, a_power=-0.5, sig_e=0.01, sig_i=0.005, size_mode=0,
              lo=1e-10, hi=1e-8, density_cgs=2.0, cd=0.0, m_star=1.0, r_star=1.0 / 215.094):
    s = OrDiskSpec(n, seed, body_type, a_min, a_max, a_power, sig_e, sig_i, size_mode, lo, hi, density_cgs, cd,
                   m_star, r_star)
    r, v = np.zeros(n + 1, VAR4), np.zeros(n + 1, VAR4)
    p, b = np.zeros(n + 1, PARAM), np.zeros(n + 1, BODY_MD)
    rc = lib.or_make_disk(C.byref(s), ptr(r), ptr(v), ptr(p), ptr(b))
    assert rc == 0
    nbt = np.zeros(7, np.uint32)
    nbt[0] = 1
    nbt[body_type] += n
    return r, v, p, b, nbt


class OGrid:
    """A FARGO grid for the oracle's free functions (keeps its arrays alive)."""

    def __init__(self, used_rad, density, vrad, vtheta):
        self.arrays = [np.ascontiguousarray(x, np.float64) for x in (used_rad, density, vrad, vtheta)]
        ur, de, vr, vt = self.arrays
        self.g = OrGrid(ur.size - 1, de.shape[1], ur.ctypes.data, de.ctypes.data, vr.ctypes.data, vt.ctypes.data)

    def cell(self, x, y):
        return lib.or_grid_cell(C.byref(self.g), x, y)


class ODisk:
    """Oracle pp_disk."""

    def __init__(self, r, v, p, bmd, nbt, t=0.0, thr=(0.0, 0.0, 0.0), gas=None, formula=0, drag_mode=0):
        self.thr = np.ascontiguousarray(thr, np.float64)
        r, v = np.ascontiguousarray(r, VAR4), np.ascontiguousarray(v, VAR4)
        p, bmd = np.ascontiguousarray(p, PARAM), np.ascontiguousarray(np.asarray(bmd).astype(BODY_MD), BODY_MD)
        self.nbt = np.ascontiguousarray(nbt, np.uint32)
        self.h = lib.or_disk_create(len(r), ptr(self.nbt), ptr(r), ptr(v), ptr(p), ptr(bmd), t, ptr(self.thr),
                                    C.byref(gas) if gas is not None else None)
        lib.or_disk_set_formula(self.h, formula)
        lib.or_disk_set_drag_mode(self.h, drag_mode)

    def __del__(self):
        if getattr(self, "h", None):
            lib.or_disk_destroy(self.h)
            self.h = None

    @property
    def n(self):
        return lib.or_disk_n(self.h)

    @property
    def t(self):
        return lib.or_disk_t(self.h)

    def state(self):
        n = self.n
        r, v, p, b = np.zeros(n, VAR4), np.zeros(n, VAR4), np.zeros(n, PARAM), np.zeros(n, BODY_MD)
        lib.or_disk_get(self.h, ptr(r), ptr(v), ptr(p), ptr(b))
        return self.t, r, v, p, b

    def counts(self):
        a, b, c = (np.zeros(7, np.uint32) for _ in range(3))
        lib.or_disk_counts(self.h, ptr(a), ptr(b), ptr(c))
        return a, b, c

    def set_gas_grid(self, used_rad, density, vrad, vtheta):
        if used_rad is None:
            assert lib.or_disk_set_gas_grid(self.h, 0, 0, None, None, None, None) == 0
            return
        ur, de, vr, vt = (np.ascontiguousarray(x, np.float64) for x in (used_rad, density, vrad, vtheta))
        assert lib.or_disk_set_gas_grid(self.h, ur.size - 1, de.shape[1], ptr(ur), ptr(de), ptr(vr), ptr(vt)) == 0

    def calc_dydx(self, i, rr, t, r, v):
        dy = np.zeros(len(r), VAR4)
        rc = lib.or_calc_dydx(self.h, i, rr, t, ptr(r), ptr(v), ptr(dy))
        return rc, dy

    def integral(self):
        out = np.zeros(1, INTEGRAL)
        lib.or_disk_integral(self.h, ptr(out))
        return out[0]

    def event_log(self):
        n = lib.or_event_log(self.h, None, 0)
        out = np.zeros(n, EVENT)
        lib.or_event_log(self.h, ptr(out), n)
        return out


class OInteg:
    def __init__(self, disk, method, dt, adaptive=True, tol=1e-10):
        self.disk = disk
        self.h = lib.or_integ_create(disk.h, method, dt, 1 if adaptive else 0, tol)

    def __del__(self):
        if getattr(self, "h", None):
            lib.or_integ_destroy(self.h)
            self.h = None

    def step(self):
        dt = _f64()
        rc = lib.or_step(self.h, C.byref(dt))
        assert rc == 0, rc
        return dt.value

    def step_rc(self):
        """(status, dt_did) without asserting success."""
        dt = _f64()
        rc = lib.or_step(self.h, C.byref(dt))
        return rc, dt.value

    def set_guard(self, max_attempts=0, min_dt=0.0, nonfinite=False):
        lib.or_set_guard(self.h, max_attempts, min_dt, 1 if nonfinite else 0)

    def counters(self):
        c = np.zeros(3, np.uint64)
        lib.or_counters(self.h, ptr(c))
        return tuple(int(x) for x in c)

    def dt_next(self):
        return lib.or_get_dt_next(self.h)

    def last_error(self):
        """(lambda |dt| max|E|, |E| per component: positions then velocities)."""
        n = 8 * lib.or_disk_n(self.disk.h)
        terms = np.zeros(n, np.float64)
        got = lib.or_last_error_terms(self.h, ptr(terms), n)
        return lib.or_last_error(self.h), terms[:got]

    def run(self, length, output_interval, max_steps):
        steps = _u64()
        rc = lib.or_run(self.disk.h, self.h, length, output_interval, max_steps, C.byref(steps))
        return rc, steps.value
