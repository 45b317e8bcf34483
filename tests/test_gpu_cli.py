"""The red-compatible command-line driver (red.cuda_amd/red) on the GPU: a run
from red's own input files must reproduce the oracle's driver loop
(src/red.cuda/main.cu:403-640) bit for bit -- final state and event list --
and write red's output files, in text and in binary, and continue from the
start file it leaves behind."""
import glob
import os
import subprocess

import numpy as np
import pytest

import clifiles as CF
import oracle_lib as O
import redamd as R

pytestmark = pytest.mark.gpu

RED = os.path.join(os.path.dirname(os.path.abspath(R.__file__)), "red")


def _system():
    # a thin ring of inflated bodies: collisions, merges and a rebuild within a few orbits
    return R.make_disk(96, 5, a_min=0.99, a_max=1.01, sig_e=0.02, sig_i=0.001, r_star=1e-5)


PARAMS = {"name": "ring", "integrator": "rkf8", "tolerance": "1.0e-10", "length": "1.0",
          "output interval": "0.05", "hit centrum": "0.2", "ejection": "50.0", "radii enhance factor": "1000",
          "info dt": "1000", "dump dt": "100000"}


def _run(directory, *extra):
    cmd = [RED, "-gpu", "-idir", directory, "-p", "parameters.txt", "-i", "start_files.txt", *extra]
    res = subprocess.run(cmd, capture_output=True, text=True, timeout=240)
    assert res.returncode == 0, res.stderr + res.stdout
    return res


def _oracle(r, v, p, b, nbt, dt, params):
    thr = (float(params["hit centrum"]), float(params["ejection"]), float(params["radii enhance factor"]))
    o = O.ODisk(r, v, p, b, nbt, thr=thr)
    oi = O.OInteg(o, 4, dt, adaptive=True, tol=float(params["tolerance"]))
    rc, steps = oi.run(CF.internal_time(params["length"]), CF.internal_time(params["output interval"]), 2 ** 62)
    assert rc == 0
    return o, steps


@pytest.mark.parametrize("binary,flags", [(False, ()), (True, ("-v", "-pe"))])
def test_cli_run_matches_oracle(tmp_path, binary, flags):
    r, v, p, b, nbt = _system()
    dt = 1.0 * R.K_GAUSS
    params = dict(PARAMS, odr="binary" if binary else "text")
    d = str(tmp_path)
    CF.write_run(d, r, v, p, b, nbt, 0.0, dt, params, binary=binary)
    _run(d, *flags)
    o, steps = _oracle(r, v, p, b, nbt, dt, params)
    ext = "dat" if binary else "txt"
    snaps = sorted(glob.glob(os.path.join(d, f"data_*[0-9].{ext}")))
    assert len(snaps) >= 10
    names, ids, rr, vv, mm = CF.read_data(snaps[-1], R.BODY_MD)
    ot, orr, ovv, opp, obb = o.state()
    act = obb["id"] > 0
    assert np.array_equal(ids, obb["id"][act])
    assert np.array_equal(rr, np.stack([orr["x"], orr["y"], orr["z"]], 1)[act])
    assert np.array_equal(vv, np.stack([ovv["x"], ovv["y"], ovv["z"]], 1)[act])
    assert np.array_equal(mm, opp["mass"][act])
    assert all(n == f"b{i - 1}" for n, i in zip(names, ids))        # names follow their bodies
    # events: one line per handled event, in the oracle's order
    oev = o.event_log()
    assert len(oev) > 16
    lines = open(os.path.join(d, "event.txt")).read().splitlines()
    assert len(lines) == len(oev)
    kinds = {1: "HIT_CENTRUM", 2: "EJECTION", 3: "COLLISION"}
    for line, e in zip(lines, oev):
        f = line.split()
        assert f[0] == kinds[int(e["event_name"])]
        assert (int(f[4]), int(f[5])) == (int(e["id1"]), int(e["id2"]))
        assert float(f[1]) == float(e["t"])
    # the other outputs of main.cu: integrals, info, log, dump, start files
    for fn in ("integral.txt", "integral.event.txt", "info.txt", "log.txt", f"dump.{ext}", f"dump.info.{ext}",
               "start_files.txt"):
        assert os.path.getsize(os.path.join(d, fn)) > 0, fn
    assert len(open(os.path.join(d, "integral.txt")).read().splitlines()) == len(snaps)
    assert len(open(os.path.join(d, "integral.event.txt")).read().splitlines()) % 2 == 0
    if "-pe" in flags:      # the state each colliding step started from
        assert len(glob.glob(os.path.join(d, f"event_*[0-9].{ext}"))) > 0
    if "-v" in flags:
        log = open(os.path.join(d, "log.txt")).read()
        assert "Rebuild the vectors" in log and "COLLISION" in log


@pytest.mark.parametrize("name,method", [("euler", 0), ("rk2", 1), ("rk4", 2)])
def test_cli_fixed_step_integrators(tmp_path, name, method):
    # the ring with collisions, 0.1 yr at h = 0.5 d: the final snapshot equals
    # the oracle's driver loop with the same integrator, bit for bit
    r, v, p, b, nbt = _system()
    dt = 0.5 * R.K_GAUSS
    params = dict(PARAMS, integrator=name, length="0.1")
    del params["tolerance"]
    d = str(tmp_path)
    CF.write_run(d, r, v, p, b, nbt, 0.0, dt, params)
    _run(d)
    thr = (float(params["hit centrum"]), float(params["ejection"]), float(params["radii enhance factor"]))
    o = O.ODisk(r, v, p, b, nbt, thr=thr)
    oi = O.OInteg(o, method, dt, adaptive=True, tol=1e-10)   # the parameter default: adaptive, 1e-10
    rc, _ = oi.run(CF.internal_time(params["length"]), CF.internal_time(params["output interval"]), 2 ** 62)
    assert rc == 0
    last = sorted(glob.glob(os.path.join(d, "data_*[0-9].txt")))[-1]
    _, ids, rr, vv, _ = CF.read_data(last, R.BODY_MD)
    ot, orr, ovv, _, obb = o.state()
    act = obb["id"] > 0
    assert np.array_equal(ids, obb["id"][act])
    assert np.array_equal(rr, np.stack([orr["x"], orr["y"], orr["z"]], 1)[act])
    assert np.array_equal(vv, np.stack([ovv["x"], ovv["y"], ovv["z"]], 1)[act])
    assert float(open(last.replace(".txt", ".info.txt")).read().split()[0]) == ot


def test_cli_backward_run_matches_oracle(tmp_path):
    # a negative length integrates backward: the start step is negated
    # (options.cpp:244-248) and the step is clamped to minus the output
    # interval (main.cu:535-538); collisions and merges still happen
    r, v, p, b, nbt = _system()
    dt = 1.0 * R.K_GAUSS
    params = dict(PARAMS, length="-0.3")
    d = str(tmp_path)
    CF.write_run(d, r, v, p, b, nbt, 0.0, dt, params)
    _run(d)
    o, steps = _oracle(r, v, p, b, nbt, -dt, params)
    last = sorted(glob.glob(os.path.join(d, "data_*[0-9].txt")))[-1]
    _, ids, rr, vv, _ = CF.read_data(last, R.BODY_MD)
    ot, orr, ovv, _, obb = o.state()
    assert ot < 0.0
    act = obb["id"] > 0
    assert np.array_equal(ids, obb["id"][act])
    assert np.array_equal(rr, np.stack([orr["x"], orr["y"], orr["z"]], 1)[act])
    assert np.array_equal(vv, np.stack([ovv["x"], ovv["y"], ovv["z"]], 1)[act])
    assert float(open(last.replace(".txt", ".info.txt")).read().split()[0]) == ot
    assert len(o.event_log()) > 0


@pytest.mark.parametrize("preset,length", [("n_gp", "2.0"), ("solar_system", "1.0"), ("Hansen_2009", "0.5"),
                                           ("coll_stat_run", "0.25"), ("Chambers2001", "0.5"),
                                           ("n_pl_to_test_anal_gd", "1.0"), ("Colliding_bodies", "1.0"),
                                           ("Birgit_scenario", "0.05"), ("GT_scenario_mod", "0.01")])
def test_redinit_then_red(tmp_path, preset, length):
    # red's initial-condition generator writes a run directory that red reads;
    # RKF7(8) over `length` years against the oracle's driver loop, bit for bit
    # (the solar system lists its rocky planets before its giants, as the
    # reference's generator writes it)
    gen = os.path.join(os.path.dirname(RED), "redinit")
    res = subprocess.run([gen, "-preset", preset, "-seed", "12345", "-odir", str(tmp_path)], capture_output=True,
                         text=True, timeout=60)
    assert res.returncode == 0, res.stderr
    r, v, p, b, nbt, t0, dt = CF.read_run(str(tmp_path), R.VAR4, R.PARAM, R.BODY_MD)
    names0 = [line.split()[0] for line in open(os.path.join(str(tmp_path), preset + ".data.txt"))]
    params = {"name": preset, "integrator": "rkf8", "tolerance": "1.0e-10", "length": length,
              "output interval": str(float(length) / 4)}
    with open(os.path.join(str(tmp_path), "parameters.txt"), "w") as f:
        for k, val in params.items():
            f.write(f"{k} = {val}\n")
    _run(str(tmp_path))
    o = O.ODisk(r, v, p, b, nbt, t=t0)
    oi = O.OInteg(o, 4, dt, adaptive=True, tol=1e-10)
    rc, steps = oi.run(CF.internal_time(length), CF.internal_time(str(float(length) / 4)), 2 ** 62)
    assert rc == 0 and steps >= 4     # at least one step per output interval
    last = sorted(glob.glob(os.path.join(str(tmp_path), "data_*[0-9].txt")))[-1]
    names, ids, rr, vv, _ = CF.read_data(last, R.BODY_MD)
    ot, orr, ovv, _, obb = o.state()
    assert names == names0
    assert np.array_equal(ids, obb["id"])
    assert np.array_equal(rr, np.stack([orr["x"], orr["y"], orr["z"]], 1))
    assert np.array_equal(vv, np.stack([ovv["x"], ovv["y"], ovv["z"]], 1))


def test_cli_continues_from_its_start_file(tmp_path):
    r, v, p, b, nbt = _system()
    d = str(tmp_path)
    CF.write_run(d, r, v, p, b, nbt, 0.0, 1.0 * R.K_GAUSS, dict(PARAMS, length="0.1"))
    _run(d)
    first = sorted(glob.glob(os.path.join(d, "data_*[0-9].txt")))
    last_t = float(open(first[-1].replace(".txt", ".info.txt")).read().split()[0])
    start = open(os.path.join(d, "start_files.txt")).read().split()
    assert start[1] == os.path.basename(first[-1])
    # continue for another 0.1 yr from the last snapshot: numbering goes on
    _run(d)
    second = sorted(glob.glob(os.path.join(d, "data_*[0-9].txt")))
    assert len(second) > len(first)
    t_end = float(open(second[-1].replace(".txt", ".info.txt")).read().split()[0])
    assert t_end > last_t


def test_cli_mixed_precision_run(tmp_path):
    # no collisions (radii factor 0): the FP32 pair sum must stay on the FP64 trajectory
    r, v, p, b, nbt = _system()
    d = str(tmp_path)
    params = dict(PARAMS, length="0.1")
    params["radii enhance factor"] = "0"
    CF.write_run(d, r, v, p, b, nbt, 0.0, 1.0 * R.K_GAUSS, params)
    _run(d, "-prec", "mixed")
    o, _ = _oracle(r, v, p, b, nbt, 1.0 * R.K_GAUSS, params)
    last = sorted(glob.glob(os.path.join(d, "data_*[0-9].txt")))[-1]
    _, ids, rr, _, _ = CF.read_data(last, R.BODY_MD)
    t = float(open(last.replace(".txt", ".info.txt")).read().split()[0])
    # the FP32 pair sum moves the step controller's error estimate, so the run
    # ends at another time: carry the oracle there with one short fixed step
    assert abs(t - o.t) < 0.01 * o.t
    O.OInteg(o, 4, t - o.t, adaptive=False).step()
    ot, orr, _, _, obb = o.state()
    assert ot == t
    assert np.array_equal(ids, obb["id"])
    assert np.abs(rr - np.stack([orr["x"], orr["y"], orr["z"]], 1)).max() < 1e-7


def test_cli_fargo_gas_run(tmp_path):
    """red -gf: planetesimals dragged through frame 0 of a FARGO run in -idir
    (written as FARGO writes it, CRLF line ends) follow the oracle's driver loop
    with the same grid -- to 1e-10 AU, the drag taking device atan2, sin and
    cos -- and end far from the same run without gas."""
    from systems import fargo_grid
    r, v, p, b, nbt = R.make_disk(96, 5, body_type=R.BODY_PL, a_min=0.8, a_max=1.2, sig_e=0.02, sig_i=0.001,
                                  size_mode=1, lo=1.0, hi=100.0, cd=1.0)
    d = str(tmp_path)
    params = dict(PARAMS, length="0.2")
    params["radii enhance factor"] = "1"
    dt = 1.0 * R.K_GAUSS
    CF.write_run(d, r, v, p, b, nbt, 0.0, dt, params)
    grid = fargo_grid(48, 128, 12, rho1=2e-3)
    CF.write_fargo(d, *grid, crlf=True)
    _run(d, "-gf", "fargo.par")
    last = sorted(glob.glob(os.path.join(d, "data_*[0-9].txt")))[-1]
    _, ids, rr, vv, _ = CF.read_data(last, R.BODY_MD)
    t = float(open(last.replace(".txt", ".info.txt")).read().split()[0])
    thr = (float(params["hit centrum"]), float(params["ejection"]), float(params["radii enhance factor"]))
    ends = []
    for with_grid in (True, False):
        o = O.ODisk(r, v, p, b, nbt, thr=thr)
        if with_grid:
            o.set_gas_grid(*grid)
        oi = O.OInteg(o, 4, dt, adaptive=True, tol=float(params["tolerance"]))
        rc, _ = oi.run(CF.internal_time(params["length"]), CF.internal_time(params["output interval"]), 2 ** 62)
        assert rc == 0
        # the loop ends with the first step past the length: ulp-level drag
        # differences may move that step, and without gas it differs by steps;
        # carry the oracle to the run's end time with one fixed step
        if with_grid:
            assert abs(t - o.t) < 1e-9 * o.t
        if o.t != t:
            O.OInteg(o, 4, t - o.t, adaptive=False).step()
        ot, orr, ovv, _, obb = o.state()
        assert np.array_equal(ids, obb["id"])
        ends.append((np.stack([orr["x"], orr["y"], orr["z"]], 1), np.stack([ovv["x"], ovv["y"], ovv["z"]], 1)))
    assert np.abs(rr - ends[0][0]).max() < 1e-10 and np.abs(vv - ends[0][1]).max() < 1e-10
    assert np.abs(vv - ends[1][1]).max() > 1e-8


def test_cli_self_test():
    res = subprocess.run([RED, "-t"], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0 and "ok" in res.stdout, res.stdout + res.stderr
