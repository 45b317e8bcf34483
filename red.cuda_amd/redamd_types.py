"""Body-array dtypes and small structs of include/red_abi.h (no library load).

Byte-compatible with the reference's structs (red.type/red_type.h:206-212,
347-362, 375-381, 415-485)."""
import ctypes as C

import numpy as np

VAR4 = np.dtype([("x", "<f8"), ("y", "<f8"), ("z", "<f8"), ("w", "<f8")])
PARAM = np.dtype([("mass", "<f8"), ("radius", "<f8"), ("density", "<f8"), ("cd", "<f8")])
BODY_MD = np.dtype({"names": ["id", "body_type", "mig_type", "mig_stop_at"],
                    "formats": ["<i4", "<i4", "<i4", "<f8"],
                    "offsets": [0, 4, 8, 16], "itemsize": 24})
INTEGRAL = np.dtype([("R", VAR4), ("V", VAR4), ("C", VAR4), ("E", "<f8")])
EVENT = np.dtype({"names": ["event_name", "t", "d",
                            "id0", "idx0", "p0", "r0", "v0",
                            "id1", "idx1", "p1", "r1", "v1",
                            "id2", "idx2", "p2", "r2", "v2",
                            "ps", "rs", "vs"],
                  "formats": ["<i4", "<f8", "<f8",
                              "<i4", "<u4", PARAM, VAR4, VAR4,
                              "<i4", "<u4", PARAM, VAR4, VAR4,
                              "<i4", "<u4", PARAM, VAR4, VAR4,
                              PARAM, VAR4, VAR4],
                  "offsets": [0, 8, 16,
                              24, 28, 32, 64, 96,
                              128, 132, 136, 168, 200,
                              232, 236, 240, 272, 304,
                              336, 368, 400],
                  "itemsize": 432})

RD_OK, RD_EINVAL, RD_EDEVICE, RD_ENOTIMPL, RD_EOVERFLOW, RD_ESTATE, RD_ENOMEM = range(7)
RD_EULER, RD_RK2, RD_RK4, RD_RKF78 = 0, 1, 2, 4
RD_EVENT_COUNTER_TOTAL, RD_EVENT_COUNTER_LAST_CLEAR, RD_EVENT_COUNTER_LAST_STEP = 0, 1, 2
RD_PREC_FP64, RD_PREC_MIXED = 0, 1
RD_DRAG_NSI, RD_DRAG_ALL_MASSIVE = 0, 1
RD_PAIRS_ONE_SIDED, RD_PAIRS_SYMMETRIC = 0, 1
BODY_STAR, BODY_GIANT, BODY_ROCKY, BODY_PROTO, BODY_SUPERPL, BODY_PL, BODY_TP = range(7)
EVENT_NONE, EVENT_HIT_CENTRUM, EVENT_EJECTION, EVENT_COLLISION = range(4)

K_GAUSS = 0.01720209895


class GasParams(C.Structure):
    """analytic_gas_disk_params_t (red_type.h:464-485)."""
    _fields_ = [("rho", C.c_double * 2), ("sch", C.c_double * 2), ("eta", C.c_double * 2),
                ("tau", C.c_double * 2), ("mfp", C.c_double * 2), ("temp", C.c_double * 2),
                ("c_vth", C.c_double), ("gas_decrease", C.c_int32),
                ("t0", C.c_double), ("t1", C.c_double), ("e_folding_time", C.c_double),
                ("alpha", C.c_double), ("mean_molecular_weight", C.c_double),
                ("particle_diameter", C.c_double)]


class GasGrid(C.Structure):
    """rd_gas_grid: frame 0 of a FARGO run (fargo_gas_disk.h:8-60)."""
    _fields_ = [("n_rad", C.c_uint32), ("n_sec", C.c_uint32), ("used_rad", C.c_void_p),
                ("density", C.c_void_p), ("vrad", C.c_void_p), ("vtheta", C.c_void_p)]


class Tuning(C.Structure):
    _fields_ = [("sinks_per_lane", C.c_uint32), ("chunk", C.c_uint32),
                ("blocks_per_cu", C.c_uint32), ("use_graph", C.c_uint32), ("pair_mode", C.c_uint32),
                ("sym_chunk_min", C.c_uint32)]


class DiskSpec(C.Structure):
    _fields_ = [("n", C.c_uint32), ("seed", C.c_uint32), ("body_type", C.c_int32),
                ("a_min", C.c_double), ("a_max", C.c_double), ("a_power", C.c_double),
                ("sig_e", C.c_double), ("sig_i", C.c_double), ("size_mode", C.c_int32),
                ("lo", C.c_double), ("hi", C.c_double), ("density_cgs", C.c_double),
                ("cd", C.c_double), ("m_star", C.c_double), ("r_star", C.c_double)]


