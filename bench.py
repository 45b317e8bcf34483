#!/usr/bin/env python3
"""Benchmark of the direct-sum RKF7(8) hot path on one MI355X.

Workload (default): star + N = 16384 planetesimals (seed 2; a from
dN/da ~ a^-0.5 on 0.5-5 AU, e/i Rayleigh 0.01/0.005, log-uniform masses
1e-10..1e-8 Msun, 2 g/cm^3), all-pairs (SI class), RKF7(8) adaptive with the
reference controller at tolerance 1e-10, mixed-precision pair kernel.
A "step" is one accepted RKF7(8) step (13 force evaluations, plus 12 per
rejected attempt).  Inputs are resident in HBM before the timed region.

Prints one JSON line:
  value         ordered pair interactions per second over the timed steps
                (whole job: all stages, rejections, controller included)
  rkf78_steps_per_s, force error vs the FP64 kernel, and the final-state
                difference vs FP64 after 50 fixed-h RKF7(8) steps
  rel_energy_error / energy_ok
                over the workload's whole run: 1000 accepted steps from t = 0
                (warmup, timed steps, then untimed steps up to 1000)
  fp64_twin     the same disk and steps with the reference's FP64 pair
                kernel: steps/s, pairs/s and its FP64 roofline
  roofline      pair kernel: 20 flop per pair over its HIP-event-timed
                launch duration, against the 157.3 TFLOP/s FP32 peak
  roofline_stage
                the RK passes of one attempt (k_comb1 + 13 stage kernels):
                4,512 algorithmic bytes per body over their HIP-event-timed
                duration, against 8 TB/s
  energy_horizon
                relative energy drift over 10^3 orbits at 1 AU (P = 365.25 d):
                bounded runs at the bench tolerance and at a tighter one,
                the drift rate fitted and carried to 10^3 orbits
  cpu_baseline  the oracle's restatement of the reference CPU path
                (symmetric FP64 loop, pp_disk.cu:822-865) on one core, timed
                over the first RKF7(8) steps of the same disk (at least 10 s)
"""
import argparse
import json
import os
import platform
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "red.cuda_amd"))

import numpy as np  # noqa: E402

import redamd as R  # noqa: E402
from proposer import batch_sizes  # noqa: E402

PEAK_FP32 = 157.3e12
PEAK_FP64 = 78.6e12
PEAK_HBM = 8.0e12
FLOP_PER_PAIR = 20.0
# algorithmic HBM bytes of the RK passes per body and adaptive attempt: RKF7(8)
# reads and writes 188 FP64 3-vectors (sum over stages of 1 + nonzero columns +
# 1, y_{n+1}, error), RK4 (adaptive) 36
STAGE_BYTES = {R.RD_RKF78: 188 * 24, R.RD_RK4: 36 * 24}
YEAR_DAYS = 365.25     # orbital period at 1 AU: the energy horizon's orbit
FS_STEPS = 50          # fixed-h steps of the final-state comparison against FP64 (config 2)
GRAPH = 1   # replay each RK attempt from a captured hipGraph (--no-graph: direct launches)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def cpu_baseline(r, v, p, b, nbt, h0, tol, min_seconds=10.0):
    """RKF7(8) steps of the oracle's reference CPU path (1 thread) on the same
    disk, as many as fit in min_seconds (at least one)."""
    O = _oracle()
    o = O.ODisk(r, v, p, b, nbt, formula=1)
    oi = O.OInteg(o, 4, h0, adaptive=True, tol=tol)
    t0 = time.perf_counter()
    while True:
        oi.step()
        el = time.perf_counter() - t0
        if el >= min_seconds:
            break
    tried, passed, failed = oi.counters()
    n = len(r)
    evals = passed + 12 * tried
    pairs = evals * n * (n - 1.0)
    return {"value": pairs / el, "unit": "pair interactions/s", "cores": 1, "kind": "port",
            "steps_per_s": passed / el,
            "sample": f"the first {passed} RKF7(8) steps ({tried} attempts, {evals} force evaluations) of the same "
                      f"N={n - 1} disk, symmetric FP64 loop, {el:.1f} s on {cpu_model()} (nproc {os.cpu_count()})"}


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    return O


def cpu_rk4_baseline(r, v, p, b, nbt, h, steps):
    """Config 1 on the reference CPU path (symmetric FP64 loop, fixed-step RK4, 1 thread): the whole run."""
    O = _oracle()
    o = O.ODisk(r, v, p, b, nbt, formula=1)
    oi = O.OInteg(o, 2, h, adaptive=False)
    t0 = time.perf_counter()
    for _ in range(steps):
        oi.step()
    el = time.perf_counter() - t0
    n = len(r)
    return {"value": 4 * steps * n * (n - 1.0) / el, "unit": "pair interactions/s", "cores": 1, "kind": "port",
            "steps_per_s": steps / el,
            "sample": f"all {steps} RK4 steps of the same N={n - 1} disk, symmetric FP64 loop, {el:.1f} s on "
                      f"{cpu_model()} (nproc {os.cpu_count()})"}


def cpu_force_baseline(r, v, p, b, nbt, gas=None, drag_mode=0, pairs_per_step=None, grid=None):
    """One force evaluation (gravity + drag) of the reference CPU path on the same disk (1 thread)."""
    O = _oracle()
    o = O.ODisk(r, v, p, b, nbt, formula=1, gas=gas, drag_mode=drag_mode)
    if grid is not None:
        o.set_gas_grid(*grid)
    t0 = time.perf_counter()
    o.calc_dydx(1, 0, 0.0, r, v)
    el = time.perf_counter() - t0
    n = len(r)
    out = {"value": n * (n - 1.0) / el, "unit": "pair interactions/s", "cores": 1, "kind": "port",
           "sample": f"one force evaluation (gravity + drag) of the same N={n - 1} disk, symmetric FP64 loop, "
                     f"{el:.1f} s on {cpu_model()} (nproc {os.cpu_count()})"}
    if pairs_per_step:
        out["steps_per_s_extrapolated"] = n * (n - 1.0) / el / pairs_per_step
    return out


def cpu_batch_baseline(parts, h0, tol, n_disks):
    """One RKF7(8) step of each of the first n_disks disks of the batch on the reference CPU
    path, disk after disk (1 thread)."""
    O = _oracle()
    pairs = 0.0
    steps = 0
    el = 0.0
    for r, v, p, b, nbt in parts[:n_disks]:
        o = O.ODisk(r, v, p, b, nbt, formula=1)
        oi = O.OInteg(o, 4, h0, adaptive=True, tol=tol)
        t0 = time.perf_counter()
        oi.step()
        el += time.perf_counter() - t0
        tried, passed, _ = oi.counters()
        n = len(r)
        pairs += (1 + 12 * tried) * n * (n - 1.0)
        steps += passed
    return {"value": pairs / el, "unit": "pair interactions/s", "cores": 1, "kind": "port",
            "disk_steps_per_s": steps / el,
            "sample": f"one RKF7(8) step of each of the first {n_disks} disks of the same batch, symmetric FP64 "
                      f"loop, {el:.1f} s on {cpu_model()} (nproc {os.cpu_count()})"}


def profiled_traffic(kernel, workload_file, tuning):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary of this
    workload (FETCH_SIZE + WRITE_SIZE, copy-calibrated), when it was taken at the same
    launch shape; None otherwise."""
    try:
        with open(os.path.join(ROOT, "profiles", workload_file)) as f:
            text = f.read()
    except OSError:
        return None
    m = re.search(r"--tuning (\d+),(\d+),(\d+)(?:,(\d+))?(?:,(\d+))?", text)
    t = re.search(r"Traffic per launch: ([0-9.e+]+) B", text)
    if not t or kernel not in text:
        return None
    if kernel == "k_force_exact":
        # the FP64 kernel has one launch shape per disk
        return float(t[1])
    if not m:
        return None
    shape = (int(m[1]), int(m[2]), int(m[3]), int(m[4]) if m[4] else R.RD_PAIRS_ONE_SIDED)
    chunk_min = int(m[5]) if m[5] else 1
    if tuning is None:
        return None
    if kernel == "k_force_sym":
        # symmetric items and slots depend on the disk and the minimum chunk
        # alone: S, chunk and the grid size change neither (the workload has
        # no one-sided pairs)
        if shape[3] != tuning.pair_mode or chunk_min != max(1, tuning.sym_chunk_min):
            return None
    elif shape != (tuning.sinks_per_lane, tuning.chunk, tuning.blocks_per_cu, tuning.pair_mode):
        return None
    return float(t[1])


def pair_kernel(precision, tuning):
    """Name of the pair kernel a force evaluation runs."""
    if precision != "mixed":
        return "k_force_exact"
    return "k_force_sym" if tuning.pair_mode == R.RD_PAIRS_SYMMETRIC else "k_force_mixed"


def tuning_json(t, **extra):
    return {"sinks_per_lane": t.sinks_per_lane, "chunk": t.chunk, "blocks_per_cu": t.blocks_per_cu,
            "pairs": "symmetric" if t.pair_mode == R.RD_PAIRS_SYMMETRIC else "one-sided",
            "sym_chunk_min": t.sym_chunk_min, **extra}


def set_tuning_arg(d, text):
    """--tuning S,CHUNK,BLOCKS_PER_CU[,PAIR_MODE[,SYM_CHUNK_MIN]] (pair mode 1 = symmetric, 0 = one-sided)."""
    f = [int(x) for x in text.split(",")]
    d.set_tuning(sinks_per_lane=f[0], chunk=f[1], blocks_per_cu=f[2],
                 pair_mode=f[3] if len(f) > 3 else R.RD_PAIRS_SYMMETRIC,
                 sym_chunk_min=f[4] if len(f) > 4 else 1)
    return d.tuning()


def roofline(d, precision, traffic=None):
    ms, pairs_launch = d.bench_force(20)
    achieved = FLOP_PER_PAIR * pairs_launch / (ms * 1e-3)
    return {"bound": "valu-fp32" if precision == "mixed" else "valu-fp64", "achieved": achieved / 1e12,
            "peak": (PEAK_FP32 if precision == "mixed" else PEAK_FP64) / 1e12, "unit": "TFLOP/s",
            "frac": achieved / (PEAK_FP32 if precision == "mixed" else PEAK_FP64), "traffic": traffic,
            "kernel": pair_kernel(precision, d.tuning()),
            "ms_per_launch": ms, "pairs_per_launch": pairs_launch}


def roofline_stage(d, n_bodies, method=R.RD_RKF78, traffic=None, drag_bytes=0):
    """HBM roofline of one attempt's RK passes (rd_bench_stages, HIP events);
    drag_bytes: what the drag reads per body and attempt on top of the RK
    vectors."""
    ms = d.bench_stages(method, 20)
    alg = (STAGE_BYTES[method] + drag_bytes) * n_bodies
    achieved = alg / (ms * 1e-3)
    return {"bound": "hbm", "achieved": achieved / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
            "frac": achieved / PEAK_HBM, "traffic": traffic, "algorithmic_bytes": alg,
            "kernels": "k_comb1 + k_stage_rk8 x13" if method == R.RD_RKF78 else "k_comb1 + k_stage_rk4 x5",
            "ms_per_attempt": ms}


def profiled_stage_traffic(workload_file):
    """HBM bytes of one attempt's stage kernels from the committed PMC summary."""
    try:
        with open(os.path.join(ROOT, "profiles", workload_file)) as f:
            m = re.search(r"Stage traffic per attempt: ([0-9.e+]+) B", f.read())
    except OSError:
        return None
    return float(m[1]) if m else None


def energy_horizon_leg(parts, tol, tuning, seconds, h0, orbits=1000.0, period_days=YEAR_DAYS):
    """RKF7(8) on a fresh copy of the disk for about `seconds` of wall time,
    |E/E0 - 1| every 500 steps; the drift rate is the least-squares slope of
    the error against time through the origin, carried to `orbits` periods."""
    r, v, p, b, nbt = parts
    d = R.Disk(r, v, p, b, nbt)
    d.set_precision(R.RD_PREC_MIXED)
    d.set_tuning(sinks_per_lane=tuning.sinks_per_lane, chunk=tuning.chunk, blocks_per_cu=tuning.blocks_per_cu,
                 pair_mode=tuning.pair_mode, sym_chunk_min=tuning.sym_chunk_min, use_graph=GRAPH)
    it = R.Integrator(d, R.RD_RKF78, h0, adaptive=True, tol=tol)
    E0 = d.integral()["E"]
    P = period_days * R.K_GAUSS
    ts, es = [], []
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds and (not ts or ts[-1] < orbits * P):
        it.steps(500)
        ts.append(d.state()[0])
        es.append(abs(d.integral()["E"] / E0 - 1.0))
    wall = time.perf_counter() - t0
    ts, es = np.array(ts), np.array(es)
    rate = float((ts * es).sum() / (ts * ts).sum())        # per time unit
    tried, passed, _ = it.counters()
    _, pairs = it.work()
    out = {"tol": tol, "period_days": period_days, "orbits_target": orbits, "orbits_run": float(ts[-1] / P),
           "steps": passed, "attempts": tried, "wall_s": wall, "steps_per_s": passed / wall,
           "pair_interactions_per_s": pairs / wall,
           "rel_energy_error_at_end": float(es[-1]), "max_rel_energy_error": float(es.max()),
           "drift_per_orbit": rate * P, "extrapolated_rel_error": rate * P * orbits,
           "ok": bool(rate * P * orbits <= 1e-9),
           "method": "linear fit |E/E0 - 1| = s t over the run, carried to the target"}
    del it, d
    return out


def emit(out):
    print(json.dumps(out))


def bench_config1(args):
    """Config 1: star + 256, RK4 fixed h = 0.1 d, 10^4 steps, FP64 (bit-exact path)."""
    steps = args.steps if args.steps != 1000 else 10000
    r, v, p, b, nbt = R.make_disk(256, 1)
    d = R.Disk(r, v, p, b, nbt)
    d.set_tuning(use_graph=GRAPH)
    it = R.Integrator(d, R.RD_RK4, 0.1 * R.K_GAUSS, adaptive=False)
    it.steps(args.warmup)
    E0 = d.integral()["E"]
    ev0, p0 = it.work()
    d.sync()
    t0 = time.perf_counter()
    it.steps(steps)
    d.sync()
    el = time.perf_counter() - t0
    ev1, p1 = it.work()
    rel_e = abs(d.integral()["E"] - E0) / abs(E0)
    out = {"metric": "pair interactions/sec + RK4 steps/sec (config 1)", "value": (p1 - p0) / el,
           "unit": "pair interactions/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
           "ms_per_step": el * 1e3 / steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": "synthetic seeded disk",
           "config": {"workload": "star + 256 planetesimals, RK4 fixed h = 0.1 d, FP64, seed 1"},
           "rk4_steps_per_s": steps / el, "rel_energy_error": rel_e,
           "roofline": roofline(d, "fp64", traffic=profiled_traffic("k_force_exact", "k_force_exact_disk256_rk4.txt",
                                                                    None))}
    if not args.no_cpu:
        print("# cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_rk4_baseline(r, v, p, b, nbt, 0.1 * R.K_GAUSS, 10000)
    emit(out)


# a FARGO-shaped frame 0 (512 rings x 1024 sectors, the reference's FARGO test
# size, src/red.code_snippet.cuda.test/main.cu:1142-1143) over 0.4 - 6 AU
C3_GRID = (512, 1024, 0.4, 6.0)


def grid_from_analytic(g, n_rad, n_sec, r_min, r_max, m_star=1.0):
    """Frame 0 sampled from the analytic disk g at the cell centres: midplane
    density rho_c r^rho_p (pp_disk.cu:87-105 with z = 0), no radial velocity,
    and the azimuthal offset sqrt(m/r) (sqrt(1 - 2 eta_c r^eta_p) - 1) of the
    sub-Keplerian gas (pp_disk.cu:145-155)."""
    used_rad = r_min * (r_max / r_min) ** (np.arange(n_rad + 1) / n_rad)
    rc = np.sqrt(used_rad[:-1] * used_rad[1:])
    dens = np.repeat((g.rho[0] * rc ** g.rho[1])[:, None], n_sec, 1)
    vt = np.sqrt(m_star / rc) * (np.sqrt(1.0 - 2.0 * g.eta[0] * rc ** g.eta[1]) - 1.0)
    return used_rad, dens, np.zeros_like(dens), np.repeat(vt[:, None], n_sec, 1)


def bench_config3(args):
    """Config 3: N = 131072 in a gas disk, drag on all massive bodies, perfect merging
    (radius factor 1), hit-centrum 0.1 AU, mixed precision, RKF7(8).  --gas grid
    drags through a FARGO frame sampled from the same analytic disk."""
    n = 131072 if args.n == 16384 else args.n
    steps = args.steps if args.steps != 1000 else 40
    r, v, p, b, nbt = R.make_disk(n, 3, size_mode=1, lo=1.0, hi=100.0, cd=1.0)
    g = R.gas_params()
    grid = grid_from_analytic(g, *C3_GRID) if args.gas == "grid" else None
    d = R.Disk(r, v, p, b, nbt, threshold=(0.1, 0.0, 1.0), gas=g if grid is None else None)
    if grid is not None:
        d.set_gas_grid(*grid)
    d.set_drag_mode(R.RD_DRAG_ALL_MASSIVE)
    d.set_precision(R.RD_PREC_MIXED)
    tuning = set_tuning_arg(d, args.tuning) if args.tuning else (d.autotune() if not args.no_autotune else d.tuning())
    d.set_tuning(use_graph=GRAPH)
    a_min = np.min(np.hypot(r["x"][1:], r["y"][1:]))
    h0 = 2 * np.pi * np.sqrt(a_min ** 3) / 1000.0
    it = R.Integrator(d, R.RD_RKF78, h0, adaptive=True, tol=args.tol)
    it.run(1e30, 1e30, args.warmup)
    E0 = d.integral()["E"]
    ev0, p0 = it.work()
    c0 = it.counters()
    d.sync()
    t0 = time.perf_counter()
    it.run(1e30, 1e30, steps)
    d.sync()
    el = time.perf_counter() - t0
    ev1, p1 = it.work()
    c1 = it.counters()
    rel_e = abs(d.integral()["E"] - E0) / abs(E0)
    out = {"metric": "pair interactions/sec + RKF78 steps/sec (config 3)", "value": (p1 - p0) / el,
           "unit": "pair interactions/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
           "ms_per_step": el * 1e3 / steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64 state / f32 pair force", "data": "synthetic seeded disk",
           "config": {"workload": f"N={n} in a gas disk (Sigma=1700 r^-1.5, H/r=0.05 r^0.25, eta=0.002"
                                  + (", as a FARGO frame 0 of %dx%d cells over %g-%g AU" % C3_GRID
                                     if grid is not None else "")
                                  + "), drag on all massive bodies, R 1-100 km, merging f=1, hit-centrum 0.1 AU, "
                                  f"mixed, RKF7(8) tol {args.tol:g}",
                      "gas": args.gas, "tuning": tuning_json(tuning)},
           "rkf78_steps_per_s": steps / el, "attempts": c1[0] - c0[0], "rejected": c1[2] - c0[2],
           "events": int(len(d.event_log())), "rel_energy_error": rel_e,
           "note": "energy is not conserved with gas drag; the error column is the drag dissipation",
           "roofline": roofline(d, "mixed", traffic=profiled_traffic(
               pair_kernel("mixed", tuning), f"{pair_kernel('mixed', tuning)}_disk131072_gas_rkf78.txt", tuning)
               if n == 131072 and grid is None else None),
           # the drag reads R, rho_p and C_D per body and stage, the grid also its cell's three fields
           "roofline_stage": roofline_stage(d, n + 1, drag_bytes=13 * (24 + (24 if grid is not None else 0)),
                                            traffic=profiled_stage_traffic(
                                                "k_stage_rk8_disk131072_gridgas_rkf78.txt" if grid is not None
                                                else "k_force_sym_disk131072_gas_rkf78.txt") if n == 131072 else None)}
    if not args.no_cpu:
        print("# cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_force_baseline(r, v, p, b, nbt, gas=g if grid is None else None, drag_mode=1,
                                                 pairs_per_step=(p1 - p0) / steps, grid=grid)
    emit(out)


def bench_config4(args):
    """Config 4: 512 independent disks, N_d ~ U[64, 4096], disk d seeded 4000 + d,
    one persistent launch per stage over all disks, attempts replayed from a hipGraph."""
    sizes = batch_sizes()
    parts = [R.make_disk(n, 4000 + k) for k, n in enumerate(sizes)]
    cat = [np.concatenate([q[i] for q in parts]) for i in range(4)]
    nbt = np.stack([q[4] for q in parts])
    d = R.Disk(*cat, nbt, batch=True)
    d.set_precision(R.RD_PREC_MIXED if args.precision == "mixed" else R.RD_PREC_FP64)
    if args.tuning:
        tuning = set_tuning_arg(d, args.tuning)
    else:
        tuning = d.autotune() if (args.precision == "mixed" and not args.no_autotune) else d.tuning()
    d.set_tuning(use_graph=GRAPH)
    steps = args.steps if args.steps != 1000 else 100
    it = R.Integrator(d, R.RD_RKF78, 0.1 * R.K_GAUSS, adaptive=True, tol=args.tol)
    it.set_overlap(not args.no_overlap)
    probe = list(range(0, len(sizes), 32))          # energy checked on every 32nd disk, from t = 0
    E0 = [d.integral(k)["E"] for k in probe]
    it.steps(args.warmup)
    ev0, p0 = it.work()
    d.sync()
    t0 = time.perf_counter()
    it.steps(steps)
    d.sync()
    el = time.perf_counter() - t0
    ev1, p1 = it.work()
    tried = sum(it.counters(k)[0] for k in range(len(sizes)))
    rel_e = max(abs(d.integral(k)["E"] / e - 1.0) for k, e in zip(probe, E0))
    out = {"metric": "pair interactions/sec + RKF78 steps/sec (config 4, batched)", "value": (p1 - p0) / el,
           "unit": "pair interactions/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
           "ms_per_step": el * 1e3 / steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64 state / f32 pair force" if args.precision == "mixed" else "f64",
           "data": "synthetic seeded disks",
           "config": {"workload": f"512 disks, N_d ~ U[64, 4096] (total {sum(sizes) + len(sizes)} bodies), "
                                  f"RKF7(8) tol {args.tol:g}, {args.precision}",
                      "tuning": tuning_json(tuning)},
           "disk_steps_per_s": steps * len(sizes) / el, "attempts_all_disks": tried,
           "rel_energy_error": rel_e, "energy_ok": bool(rel_e <= 1e-9),
           "energy_basis": f"max |E/E0 - 1| over disks {probe[0]}, {probe[1]}, ..., {probe[-1]} after "
                           f"{args.warmup + steps} accepted steps from t = 0",
           "roofline": roofline(d, args.precision, traffic=profiled_traffic(
               pair_kernel(args.precision, tuning), f"{pair_kernel(args.precision, tuning)}_batch512_rkf78.txt", tuning)
               if args.precision == "mixed" else None)}
    if not args.no_cpu:
        print("# cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_batch_baseline(parts, 0.1 * R.K_GAUSS, args.tol, 64)
    emit(out)


def rk_evals(res):
    """Force evaluations summed over disks: RKF7(8) spends 12 per attempt plus
    k0 once per step (int_rungekutta8.cu:648-656)."""
    return 12 * res["tried"] + res["passed"]


def bench_config5(args):
    """Config 5: the learned step proposer on the config-3 batch. Samples are logged
    from seeded runs of a 512-disk batch with the config-3 physics (N_d ~ U[64, 4096],
    gas drag on every massive body, merging f = 1, hit-centrum 0.1 AU; disk k seeded
    7000 + k), the MLP is fitted on the GPU, and the ablation integrates held-out disks
    to the same simulated time with the classical controller (mixed and plain FP64) and
    with the proposer: the config-3 disk (N = 131072, seed 3) and the held-out batch
    (seeds 4000 + k)."""
    import proposer as P
    sizes = batch_sizes()
    tol = args.tol
    dt0 = 0.1 * R.K_GAUSS
    print(f"# logging {args.log_steps} steps x {len(sizes)} disks", file=sys.stderr, flush=True)
    d = P.make_batch(sizes, 7000, gas=True, use_graph=GRAPH)
    t0 = time.perf_counter()
    X, T = P.collect(d, args.log_steps, dt0, tol)
    t_log = time.perf_counter() - t0
    d.close()
    print(f"# {len(T)} samples in {t_log:.1f} s; training", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    m, loss = P.fit(X, T, epochs=args.epochs, base=args.c5_base, quantile=args.c5_quantile)
    t_fit = time.perf_counter() - t0
    rms = float(np.sqrt(np.mean((m.forward(X) - T) ** 2)))
    n_samples = int(len(T))
    del X, T
    runs = (("classical_fp64", R.RD_PREC_FP64, None), ("classical_mixed", R.RD_PREC_MIXED, None),
            ("learned_mixed", R.RD_PREC_MIXED, m))
    disk3 = {}
    for name, prec, prop in runs:
        print(f"# config-3 disk {name}", file=sys.stderr, flush=True)
        d, h0 = P.config3_disk(precision=prec, use_graph=GRAPH)
        disk3[name] = P.run(d, length=args.c5_length, proposer=prop, safety=args.safety, dt0=h0, tol=tol)
        disk3[name]["rk_evals"] = rk_evals(disk3[name])
        d.close()
    batch = {}
    for name, prec, prop in runs[1:]:
        print(f"# held-out batch {name}", file=sys.stderr, flush=True)
        d = P.make_batch(sizes, 4000, precision=prec, gas=True, use_graph=GRAPH)
        batch[name] = P.run(d, length=args.c5_batch_length, proposer=prop, safety=args.safety, dt0=dt0, tol=tol,
                            energy_disks=range(0, len(sizes), 32))
        batch[name]["rk_evals"] = rk_evals(batch[name])
        d.close()
    # what any proposer could save against the classical controller (mixed precision), from one
    # more classical run of each workload with sample logging on
    print("# ceilings", file=sys.stderr, flush=True)
    d, h0 = P.config3_disk(use_graph=GRAPH)
    ceil = {"config3_disk": P.ceilings(d, args.c5_length, h0, tol)}
    d.close()
    d = P.make_batch(sizes, 4000, gas=True, use_graph=GRAPH)
    ceil["heldout_batch"] = P.ceilings(d, args.c5_batch_length, dt0, tol)
    d.close()
    f64 = disk3["classical_fp64"]
    ceil["config3_disk_fp64_rejection_free"] = f64["rk_evals"] / (13.0 * f64["passed"])
    achieved = {"config3_disk": disk3["classical_mixed"]["rk_evals"] / disk3["learned_mixed"]["rk_evals"],
                "heldout_batch": batch["classical_mixed"]["rk_evals"] / batch["learned_mixed"]["rk_evals"]}
    for w in ("config3_disk", "heldout_batch"):
        ceil[w]["achieved"] = achieved[w]
        ceil[w]["achieved_over_rejection_free"] = achieved[w] / ceil[w]["rejection_free"]
        ceil[w]["achieved_over_exact_step"] = achieved[w] / ceil[w]["exact_step"]
    value = disk3["classical_fp64"]["rk_evals"] / disk3["learned_mixed"]["rk_evals"]
    out = {"metric": "RK evaluations, plain FP64 RKF78 / learned proposer (config 5)", "value": value,
           "unit": "x fewer RK evaluations", "n_gpus": 1, "steps": disk3["learned_mixed"]["passed"], "warmup": 0,
           "ms_per_step": disk3["learned_mixed"]["seconds"] * 1e3 / max(1, disk3["learned_mixed"]["passed"]),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64 state / f32 pair force / bf16 MLP weights", "data": "synthetic seeded disks",
           "config": {"workload": f"proposer trained on {args.log_steps} logged steps x {len(sizes)} disks of the "
                                  "config-3 physics (seeds 7000+k); ablation to equal simulated time on the config-3 "
                                  f"disk (N=131072, seed 3, {args.c5_length:g} time units) and on the held-out batch "
                                  f"(seeds 4000+k, {args.c5_batch_length:g} time units), RKF7(8) tol {tol:g}",
                      "safety": args.safety, "epochs": args.epochs,
                      "fit": f"quantile {args.c5_quantile:g} of log h*" + (f" - feature {args.c5_base}" if args.c5_base >= 0 else "")},
           "training": {"samples": n_samples, "log_seconds": t_log, "fit_seconds": t_fit, "final_loss": loss,
                        "rms_log_h": rms},
           "config3_disk": disk3, "heldout_batch": batch, "ceilings": ceil,
           "ratios": {"config3_evals_fp64_over_learned": value,
                      "config3_evals_mixed_over_learned": disk3["classical_mixed"]["rk_evals"] / disk3["learned_mixed"]["rk_evals"],
                      "config3_time_mixed_over_learned": disk3["classical_mixed"]["seconds"] / disk3["learned_mixed"]["seconds"],
                      "batch_evals_mixed_over_learned": batch["classical_mixed"]["rk_evals"] / batch["learned_mixed"]["rk_evals"],
                      "batch_time_mixed_over_learned": batch["classical_mixed"]["seconds"] / batch["learned_mixed"]["seconds"]}}
    emit(out)


def energy_drift(args):
    """Long-horizon energy check of the default disk: RKF7(8) at args.tol in
    args.precision until args.energy_orbits orbits (P = --period-days, 365.25
    d: 1 AU), the relative energy error (rd_calc_integral) at every tenth of
    the run.  Not a throughput figure: the run prints its checkpoints and one
    JSON line."""
    r, v, p, b, nbt = R.make_disk(args.n, args.seed)
    a_min = np.min(np.hypot(r["x"][1:], r["y"][1:]))
    h0 = 2 * np.pi * np.sqrt(a_min ** 3 / 1.0) / 1000.0
    # the horizon's orbit: 1 AU (365.25 d) unless --period-days 0 asks for the
    # disk's inner edge
    period = args.period_days * R.K_GAUSS if args.period_days > 0 else 2 * np.pi * np.sqrt(a_min ** 3 / 1.0)
    d = R.Disk(r, v, p, b, nbt)
    d.set_precision(R.RD_PREC_MIXED if args.precision == "mixed" else R.RD_PREC_FP64)
    d.set_tuning(use_graph=GRAPH)
    if args.precision == "mixed" and not args.no_autotune:
        d.autotune()
        d.set_tuning(use_graph=GRAPH)
    it = R.Integrator(d, R.RD_RKF78, h0, adaptive=True, tol=args.tol)
    E0 = d.integral()["E"]
    t0 = time.perf_counter()
    checkpoints, mark = [], 0.1
    while True:
        it.steps(1000)
        t = d.state()[0]
        if t >= mark * args.energy_orbits * period:
            tried, passed, failed = it.counters()
            cp = {"orbits": t / period, "steps": passed, "attempts": tried,
                  "rel_energy_error": abs(d.integral()["E"] / E0 - 1.0), "wall_s": time.perf_counter() - t0}
            checkpoints.append(cp)
            print(f"# {json.dumps(cp)}", file=sys.stderr, flush=True)
            while mark * args.energy_orbits * period <= t:
                mark += 0.1
        if t >= args.energy_orbits * period:
            break
    print(json.dumps({"energy_drift": {"n_bodies": int(len(r)), "seed": args.seed, "tolerance": args.tol,
                                       "precision": args.precision, "period_days": period / R.K_GAUSS,
                                       "checkpoints": checkpoints}}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--precision", choices=["mixed", "fp64"], default="mixed")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-final-state", action="store_true",
                    help="config 2: skip the final-state comparison against FP64 (for profiling runs)")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-graph", action="store_true", help="launch the attempt's kernels directly (no hipGraph)")
    ap.add_argument("--tuning", default="", help="S,CHUNK,BLOCKS_PER_CU[,PAIR_MODE[,SYM_CHUNK_MIN]] for the mixed "
                                                 "kernel (skips autotune; pair mode 1 symmetric, 0 one-sided)")
    ap.add_argument("--config", type=int, default=2, choices=[1, 2, 3, 4, 5],
                    help="1 RK4 N=256, 2 RKF78 N=16384 (default), 3 gas+collisions N=131072, "
                         "4 batch of 512 disks, 5 learned step proposer")
    ap.add_argument("--gas", choices=["analytic", "grid"], default="analytic",
                    help="config 3: the analytic gas disk, or a FARGO frame 0 sampled from it")
    ap.add_argument("--log-steps", type=int, default=2000, help="config 5: logged steps per disk")
    ap.add_argument("--epochs", type=int, default=10, help="config 5: training epochs")
    ap.add_argument("--safety", type=float, default=None, help="config 5: proposer safety factor (default: "
                    "proposer.SAFETY)")
    ap.add_argument("--c5-quantile", type=float, default=None, help="config 5: fitted quantile, 0 = mean "
                    "(default: proposer.QUANTILE)")
    ap.add_argument("--c5-base", type=int, default=None, help="config 5: targets as offsets from this feature, "
                    "-1 = none (default: proposer.BASE)")
    ap.add_argument("--c5-length", type=float, default=1.5, help="config 5: config-3 disk run length (time units)")
    ap.add_argument("--c5-batch-length", type=float, default=2.9, help="config 5: held-out batch run length")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="only run the byte-counter calibration copy of this size and exit")
    ap.add_argument("--energy-orbits", type=float, default=0.0,
                    help="only run the default disk for this many orbits and report the energy drift")
    ap.add_argument("--period-days", type=float, default=YEAR_DAYS,
                    help="orbital period of the energy horizon (365.25: 1 AU; 0: the disk's inner edge)")
    ap.add_argument("--horizon-seconds", type=float, default=20.0,
                    help="config 2: wall seconds of each energy-horizon leg (0: skip)")
    ap.add_argument("--horizon-tol", type=float, default=4e-12,
                    help="config 2: the tighter tolerance of the second energy-horizon leg (4e-12: the loosest "
                         "that keeps the N = 16384 disk within 1e-9 over 10^3 orbits at 1 AU, by a 40 s scan "
                         "over 1e-11, 7e-12, 5e-12 and 4e-12)")
    ap.add_argument("--no-twin", action="store_true", help="config 2: skip the FP64 twin run")
    ap.add_argument("--no-overlap", action="store_true",
                    help="configs 2 and 4: the host-only step sequence (no device-side accept, commit and k0 "
                         "pass while the host computes the step size)")
    ap.add_argument("--profile-run", action="store_true",
                    help="the timed steps only (no CPU leg, FP64 comparisons, twin, energy run or horizon): "
                         "for profilers")
    args = ap.parse_args()
    if args.profile_run:
        args.no_cpu = args.no_final_state = args.no_twin = True
        args.horizon_seconds = 0.0
    global GRAPH
    GRAPH = 0 if args.no_graph else 1
    if args.copy_bytes:
        ms = R.bench_copy(args.copy_bytes, 5)
        print(json.dumps({"copy_bytes": args.copy_bytes, "ms_per_launch": ms,
                          "GBps": 2 * args.copy_bytes / (ms * 1e-3) / 1e9}))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (each GPU box has one GPU)",
              file=sys.stderr)
        sys.exit(2)
    if args.energy_orbits > 0.0:
        return energy_drift(args)
    if args.config == 1:
        return bench_config1(args)
    if args.config == 3:
        return bench_config3(args)
    if args.config == 4:
        return bench_config4(args)
    if args.config == 5:
        import proposer as P
        args.safety = P.SAFETY if args.safety is None else args.safety
        args.c5_quantile = P.QUANTILE if args.c5_quantile is None else args.c5_quantile
        args.c5_base = P.BASE if args.c5_base is None else args.c5_base
        return bench_config5(args)

    r, v, p, b, nbt = R.make_disk(args.n, args.seed)
    # initial step as redinit: min orbital period / 1000 (src/red.cuda.initial/main.cu:317-342)
    a_min = np.min(np.hypot(r["x"][1:], r["y"][1:]))
    h0 = 2 * np.pi * np.sqrt(a_min ** 3 / 1.0) / 1000.0
    d = R.Disk(r, v, p, b, nbt)
    d.set_precision(R.RD_PREC_MIXED if args.precision == "mixed" else R.RD_PREC_FP64)
    d.set_tuning(use_graph=GRAPH)
    if args.tuning:
        tuning = set_tuning_arg(d, args.tuning)
    else:
        tuning = d.autotune() if (args.precision == "mixed" and not args.no_autotune) else d.tuning()
    d.set_tuning(use_graph=GRAPH)
    it = R.Integrator(d, R.RD_RKF78, h0, adaptive=True, tol=args.tol)
    it.set_overlap(not args.no_overlap)
    E_init = d.integral()["E"]
    print(f"# warmup {args.warmup} steps", file=sys.stderr, flush=True)
    it.steps(args.warmup)
    d.sync()
    E0 = d.integral()["E"]
    ev0, pairs0 = it.work()
    tried0, passed0, failed0 = it.counters()
    t_start = d.state()[0]
    print(f"# timed {args.steps} steps", file=sys.stderr, flush=True)
    d.sync()
    t0 = time.perf_counter()
    done = 0
    chunk = 100
    while done < args.steps:
        k = min(chunk, args.steps - done)
        it.steps(k)
        done += k
        print(f"# {done} steps", file=sys.stderr, flush=True)
    d.sync()
    el = time.perf_counter() - t0
    ev1, pairs1 = it.work()
    tried1, passed1, failed1 = it.counters()
    E1 = d.integral()["E"]
    t_end, rr, vv, _, _ = d.state()
    pairs = pairs1 - pairs0
    value = pairs / el
    steps_per_s = (passed1 - passed0) / el
    rel_e = abs(E1 - E0) / abs(E0)

    # pair-kernel launch time (HIP events on the engine stream) for the roofline
    ms, pairs_launch = d.bench_force(20)
    achieved = FLOP_PER_PAIR * pairs_launch / (ms * 1e-3)
    peak = PEAK_FP32 if args.precision == "mixed" else PEAK_FP64
    # the RK passes of one attempt (HIP events on the engine stream)
    stage_rl = roofline_stage(d, args.n + 1, traffic=profiled_stage_traffic(f"k_stage_rk8_disk{args.n}_rkf78.txt")
                              if (args.precision == "mixed" and args.seed == 2) else None)

    # the energy criterion over the workload's run: 1000 accepted steps from t = 0
    run_steps = passed1 if args.profile_run else max(1000, passed1)
    if passed1 < run_steps:
        print(f"# untimed steps {passed1} -> {run_steps} (energy over the run)", file=sys.stderr, flush=True)
        it.steps(run_steps - passed1)
    rel_e_run = abs(d.integral()["E"] / E_init - 1.0)

    # force error of the mixed kernel against the FP64 reference kernel at the final state
    f_err = None
    if args.precision == "mixed" and not args.profile_run:
        am = d.calc_dydx(1, 0, t_end, rr, vv)
        d.set_precision(R.RD_PREC_FP64)
        ae = d.calc_dydx(1, 0, t_end, rr, vv)
        gm = np.stack([am["x"], am["y"], am["z"]], 1)
        ge = np.stack([ae["x"], ae["y"], ae["z"]], 1)
        f_err = float((np.linalg.norm(gm - ge, axis=1) / np.linalg.norm(ge, axis=1))[1:].max())

    # final-state difference against the FP64 kernel: the same fixed-h RKF7(8)
    # steps (h = the timed run's mean accepted step) from the same initial state
    fs_diff = None
    if args.precision == "mixed" and not args.no_final_state:
        h_fix = (t_end - t_start) / args.steps
        ends = {}
        for prec in (R.RD_PREC_MIXED, R.RD_PREC_FP64):
            dd = R.Disk(r, v, p, b, nbt)
            dd.set_precision(prec)
            ii = R.Integrator(dd, R.RD_RKF78, h_fix, adaptive=False, tol=args.tol)
            ii.steps(FS_STEPS)
            dd.sync()
            _, r_end, v_end, _, _ = dd.state()
            ends[prec] = (np.stack([r_end["x"], r_end["y"], r_end["z"]], 1),
                          np.stack([v_end["x"], v_end["y"], v_end["z"]], 1))
            del ii, dd
        (rm, vm), (re, ve) = ends[R.RD_PREC_MIXED], ends[R.RD_PREC_FP64]
        fs_diff = {"steps": FS_STEPS, "h_days": h_fix / R.K_GAUSS,
                   "max_rel_pos": float((np.linalg.norm(rm - re, axis=1) / np.linalg.norm(re, axis=1))[1:].max()),
                   "max_rel_vel": float((np.linalg.norm(vm - ve, axis=1) / np.linalg.norm(ve, axis=1))[1:].max())}

    # the FP64 twin: the same disk and steps with the reference's FP64 pair kernel
    twin = None
    if args.precision == "mixed" and not args.no_twin:
        print("# fp64 twin", file=sys.stderr, flush=True)
        dt_ = R.Disk(r, v, p, b, nbt)
        dt_.set_tuning(use_graph=GRAPH)
        it_ = R.Integrator(dt_, R.RD_RKF78, h0, adaptive=True, tol=args.tol)
        it_.steps(args.warmup)
        _, q0 = it_.work()
        c0 = it_.counters()
        dt_.sync()
        t0 = time.perf_counter()
        it_.steps(args.steps)
        dt_.sync()
        el_ = time.perf_counter() - t0
        _, q1 = it_.work()
        c1 = it_.counters()
        ms64, pl64 = dt_.bench_force(5)
        ach64 = FLOP_PER_PAIR * pl64 / (ms64 * 1e-3)
        twin = {"rkf78_steps_per_s": (c1[1] - c0[1]) / el_, "pair_interactions_per_s": (q1 - q0) / el_,
                "attempts": c1[0] - c0[0], "rejected": c1[2] - c0[2], "ms_per_step": el_ * 1e3 / args.steps,
                "mixed_over_fp64_steps_per_s": steps_per_s / ((c1[1] - c0[1]) / el_),
                "roofline": {"bound": "valu-fp64", "achieved": ach64 / 1e12, "peak": PEAK_FP64 / 1e12,
                             "unit": "TFLOP/s", "frac": ach64 / PEAK_FP64, "kernel": "k_force_exact",
                             "traffic": profiled_traffic("k_force_exact", "k_force_exact_disk16384_rkf78.txt", None)
                             if (args.n == 16384 and args.seed == 2) else None,
                             "ms_per_launch": ms64, "pairs_per_launch": pl64}}
        del it_, dt_

    # energy over 10^3 orbits at 1 AU: bounded runs, drift carried to the horizon
    horizon = None
    if args.precision == "mixed" and args.horizon_seconds > 0:
        horizon = []
        for tol_h in (args.tol, args.horizon_tol):
            print(f"# energy horizon leg, tol {tol_h:g}", file=sys.stderr, flush=True)
            horizon.append(energy_horizon_leg((r, v, p, b, nbt), tol_h, tuning, args.horizon_seconds, h0))

    out = {
        "metric": "pair interactions/sec + RKF78 steps/sec at rel. energy error <=1e-9, 1 GPU",
        "value": value,
        "unit": "pair interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 state / f32 pair force" if args.precision == "mixed" else "f64",
        "data": "synthetic seeded disk (generated in-process)",
        "config": {"workload": f"N={args.n} planetesimals + star, RKF7(8) tol {args.tol:g}, "
                               f"{args.precision} pair kernel, seed {args.seed}",
                   "n_bodies": args.n + 1, "tolerance": args.tol, "precision": args.precision,
                   "tuning": tuning_json(tuning, use_graph=GRAPH)},
        "rkf78_steps_per_s": steps_per_s,
        "attempts": tried1 - tried0,
        "rejected": failed1 - failed0,
        "force_evals": ev1 - ev0,
        "sim_time_days": (t_end - t_start) / R.K_GAUSS,
        "rel_energy_error": rel_e_run,
        "energy_ok": bool(rel_e_run <= 1e-9),
        "energy_basis": f"|E/E0 - 1| over the workload's run: {run_steps} accepted steps from t = 0 "
                        f"(timed steps alone: {rel_e:.2e})",
        "max_rel_force_err_vs_fp64": f_err,
        "final_state_vs_fp64": fs_diff,
        "roofline": {"bound": "valu-fp32" if args.precision == "mixed" else "valu-fp64",
                     "achieved": achieved / 1e12, "peak": peak / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / peak,
                     "traffic": profiled_traffic(pair_kernel(args.precision, tuning),
                                                 f"{pair_kernel(args.precision, tuning)}_disk16384_rkf78.txt", tuning)
                     if (args.n == 16384 and args.seed == 2) else None,
                     "kernel": pair_kernel(args.precision, tuning),
                     "ms_per_launch": ms, "pairs_per_launch": pairs_launch},
        "roofline_stage": stage_rl,
        "fp64_twin": twin,
        "energy_horizon": horizon,
    }
    if not args.no_cpu:
        print("# cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(r, v, p, b, nbt, h0, args.tol)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
