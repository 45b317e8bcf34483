"""red.cuda_amd/redinit (red's initial-condition generator, src/red.cuda.initial)
against the restatement in oracle/redinit_oracle.py, bit for bit, with the
restatement pinned by the C library's rand(), MT19937's published output, the
two-body closed form and the Kepler residual.  CPU only."""
import ctypes
import math
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import redinit_oracle as RO  # noqa: E402

REDINIT = os.path.join(ROOT, "red.cuda_amd", "redinit")
needs_bin = pytest.mark.skipif(not os.path.exists(REDINIT), reason="red.cuda_amd/redinit is not built")


def test_crand_is_the_c_library_rand():
    libc = ctypes.CDLL("libc.so.6")
    for seed in (0, 1, 2, 12345, 1733170000, 2 ** 31 - 1, 2 ** 31, 2 ** 31 + 17, 2 ** 32 - 1):
        libc.srand(ctypes.c_uint(seed))
        want = [libc.rand() for _ in range(400)]
        g = RO.crand(seed)
        assert [next(g) for _ in range(400)] == want, seed


def test_red_random_is_mt19937():
    # the 10000th output of MT19937 with the default seed 5489 (C++ [rand.predef])
    r = RO.RedRandom(5489)
    for _ in range(9999):
        r.rand()
    assert r.rand() == 4123659995


def test_kepler_solver_residual_and_reference_shortcut():
    for e in (0.01, 0.1, 0.3):
        for M in np.linspace(0.1, 3.0, 9):
            E = RO.kepler(e, float(M))
            assert abs(E - e * math.sin(E) - M) < 1e-7
    # tools.cpp:746-749: for M > pi the solution is replaced by (M - pi) + pi
    assert abs(RO.kepler(0.2, 4.0) - 4.0) < 1e-15


def run(tmp_path, preset, seed, stem=None):
    args = [REDINIT, "-preset", preset, "-seed", str(seed), "-odir", str(tmp_path)]
    if stem:
        args += ["-f", stem]
    res = subprocess.run(args, capture_output=True, text=True, timeout=60)
    assert res.returncode == 0, res.stderr
    assert f"The seed number is {seed}" in res.stdout
    return res


def read_outputs(d, stem):
    info = open(os.path.join(d, stem + ".info.txt")).read().split()
    rows = [line.split() for line in open(os.path.join(d, stem + ".data.txt"))]
    oe = open(os.path.join(d, stem + ".oe.txt")).read().splitlines()
    seed = open(os.path.join(d, stem + ".seed.txt")).read()
    start = open(os.path.join(d, "start_files.txt")).read().split()
    return info, rows, oe, seed, start


@needs_bin
@pytest.mark.parametrize("preset,seed", [("two_body", 1), ("n_gp", 12345), ("n_gp", 1733170000), ("n_pp", 7),
                                         ("n_spl", 99), ("n_pl", 2 ** 31 + 5), ("n_tp", 3), ("Dvorak", 7),
                                         ("Dvorak", 1717171717)])
def test_preset_matches_oracle(tmp_path, preset, seed):
    run(tmp_path, preset, seed)
    info, rows, oe, seed_txt, start = read_outputs(str(tmp_path), preset)
    bodies, t0, dt, n_type = RO.generate(preset, seed)
    assert seed_txt == str(seed)
    assert start == [preset + ".info.txt", preset + ".data.txt"]
    assert float(info[0]) == t0 and float(info[1]) == dt and info[2] == "0"
    assert [int(x) for x in info[3:]] == n_type
    assert len(rows) == len(bodies)
    for row, b in zip(rows, bodies):
        assert row[0] == b["name"]
        assert (int(row[1]), int(row[2]), int(row[3])) == (b["id"], b["type"], 0)
        got = [float(x) for x in row[4:]]
        want = [0.0, b["mass"], b["radius"], b["density"], b["cd"], *b["r"], *b["v"]]
        assert got == want, (b["name"], got, want)
    for line, b in zip(oe, bodies):
        want = (f"{t0:15.6e} {b['id']:7d} {b['type']:2d} {b['mass']:15.6e} {b['radius']:15.6e} "
                f"{b['density']:15.6e} {b['cd']:15.6e} " + " ".join(f"{x:15.6e}" for x in b["oe"]))
        assert line == want


@needs_bin
def test_two_body_closed_form(tmp_path):
    # a circular orbit of a = 1 AU at M = 0: relative position (1, 0, 0), relative
    # speed sqrt(mu) AU/day = sqrt(1 + m) in k-day units (K2 = k^2), and
    # dt = 2 pi sqrt(a^3 / mu) / 1000 days
    run(tmp_path, "two_body", 42)
    info, rows, _, _, _ = read_outputs(str(tmp_path), "two_body")
    star, planet = [[float(x) for x in r[9:15]] for r in rows]
    m = float(rows[1][5])
    assert abs(m - 1.0 / 3.3294605e5) < 1e-20
    rel_r = [planet[c] - star[c] for c in range(3)]
    rel_v = [planet[3 + c] - star[3 + c] for c in range(3)]
    assert np.allclose(rel_r, [1.0, 0.0, 0.0], rtol=0, atol=1e-15)
    assert np.allclose(rel_v, [0.0, math.sqrt(1.0 + m), 0.0], rtol=1e-14, atol=1e-16)
    # barycentre at rest at the origin
    assert abs(star[0] * 1.0 + planet[0] * m) < 1e-16 and abs(star[4] * 1.0 + planet[4] * m) < 1e-16
    mu = RO.K2 * (1.0 + m)
    assert abs(float(info[1]) - 2 * math.pi * math.sqrt(1.0 / mu) / 1000.0 * RO.GAUSS) < 1e-15
    assert rows[1][0] == "rocky0"        # Two_body numbers its planet from 0


@needs_bin
def test_dvorak_mass_is_the_mmsn_solid_mass(tmp_path):
    run(tmp_path, "Dvorak", 11)
    _, rows, _, _, _ = read_outputs(str(tmp_path), "Dvorak")
    m = np.array([float(r[5]) for r in rows[1:]])
    assert abs(m.sum() - RO.solid_mass(*RO.DVORAK_NEB)) < 1e-15
    a = np.array([math.hypot(float(r[9]), float(r[10])) for r in rows[1:]])
    assert a.min() > 0.8 and a.max() < 3.0


@needs_bin
def test_same_seed_same_files_and_stem(tmp_path):
    run(tmp_path, "n_pl", 5, stem="input_01")
    a = open(os.path.join(tmp_path, "input_01.data.txt")).read()
    run(tmp_path, "n_pl", 5, stem="input_02")
    assert open(os.path.join(tmp_path, "input_02.data.txt")).read() == a
    run(tmp_path, "n_pl", 6, stem="input_03")
    assert open(os.path.join(tmp_path, "input_03.data.txt")).read() != a
    assert open(os.path.join(tmp_path, "start_files.txt")).read().split() == ["input_03.info.txt",
                                                                              "input_03.data.txt"]


@needs_bin
@pytest.mark.parametrize("args,msg", [(["-preset", "nope"], "Unknown preset"), ([], "No -preset"),
                                      (["-preset", "n_gp", "-seed", "x"], "Invalid number"),
                                      (["-preset", "n_gp", "-seed"], "Missing value"),
                                      (["-bogus"], "Invalid switch")])
def test_redinit_errors(args, msg):
    res = subprocess.run([REDINIT, *args], capture_output=True, text=True, timeout=30)
    assert res.returncode != 0 and msg in res.stderr, res.stderr


@needs_bin
def test_redinit_lists_presets():
    res = subprocess.run([REDINIT, "-l"], capture_output=True, text=True, timeout=30)
    assert res.returncode == 0
    for p in RO.PRESETS:
        assert p in res.stdout
