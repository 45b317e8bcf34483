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
                                         ("Dvorak", 1717171717), ("Chambers2001", 7), ("Chambers2001", 1733170000),
                                         ("coll_stat_run", 12345), ("coll_stat_run", 2 ** 32 - 1),
                                         ("Hansen_2009", 5), ("Hansen_2009", 1733170001),
                                         ("n_pl_to_test_anal_gd", 17), ("solar_system", 1733170002),
                                         ("Colliding_bodies", 3), ("Colliding_bodies", 99),
                                         ("Birgit_2016", 11), ("Birgit_scenario", 12), ("Birgit_scenario", 1733170003),
                                         ("GT_scenario", 13), ("GT_scenario_mod", 14), ("GT_scenario_mod", 2 ** 31 - 1),
                                         ("Elke_binary_test", 15)])
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
    for p in list(RO.PRESETS) + list(RO.EXTRA):
        assert p in res.stdout


def elements(r, v, mu):
    """a, e, i, Omega, omega and the eccentric anomaly of a relative state
    (AU, AU/day), from the angular momentum and eccentricity vectors."""
    r, v = np.asarray(r, float), np.asarray(v, float)
    h = np.cross(r, v)
    rn = np.linalg.norm(r)
    evec = np.cross(v, h) / mu - r / rn
    e = np.linalg.norm(evec)
    a = 1.0 / (2.0 / rn - v.dot(v) / mu)
    inc = math.acos(h[2] / np.linalg.norm(h))
    node = math.atan2(h[0], -h[1]) % (2 * math.pi)
    nvec = np.array([-h[1], h[0], 0.0])
    peri = math.atan2(np.cross(nvec, evec).dot(h) / np.linalg.norm(h), nvec.dot(evec)) % (2 * math.pi)
    E = math.atan2(r.dot(v) / math.sqrt(mu * a), 1.0 - rn / a) % (2 * math.pi)
    return a, e, inc, node, peri, E


@needs_bin
def test_solar_system_recovers_the_horizons_elements(tmp_path):
    """The planets' written states give back the 2015-May-11 elements the
    generator read (a, e, i, Omega, omega; M through Kepler's equation where
    the reference solves it, E = M where it takes its shortcut, M > pi), in
    the reference's order Sun, rocky planets, giants."""
    run(tmp_path, "solar_system", 20150511)
    info, rows, _, _, _ = read_outputs(str(tmp_path), "solar_system")
    assert [int(x) for x in info[3:]] == [1, 4, 4, 0, 0, 0, 0]
    names = [r[0] for r in rows]
    assert names == ["Sun", "Mercury", "Venus", "Earth", "Mars", "Jupiter", "Saturn", "Uranus", "Neptune"]
    assert [int(r[2]) for r in rows] == [0, 2, 2, 2, 2, 1, 1, 1, 1]
    sun = [float(x) for x in rows[0][9:15]]
    for row in rows[1:]:
        st = [float(x) for x in row[9:15]]
        m = float(row[5])
        r = [st[c] - sun[c] for c in range(3)]
        v = [(st[3 + c] - sun[3 + c]) * RO.GAUSS for c in range(3)]      # k-day units back to AU/day
        a, e, inc, node, peri, E = elements(r, v, RO.K2 * (1.0 + m))
        ecc, inc_d, node_d, peri_d, mean_d, sma = RO.HORIZONS[row[0]]
        assert abs(a / sma - 1) < 1e-12 and abs(e - ecc) < 1e-12, row[0]
        assert abs(inc - inc_d * math.pi / 180) < 1e-11, row[0]
        assert abs(node - node_d * math.pi / 180) < 1e-9 and abs(peri - peri_d * math.pi / 180) < 1e-9, row[0]
        M = mean_d * math.pi / 180
        if M < math.pi:
            assert abs((E - e * math.sin(E)) - M) < 1e-8, row[0]
        else:
            assert abs(E - M) < 1e-9, row[0]
        assert abs(float(row[7]) - m / (4.0 / 3.0 * math.pi * float(row[6]) ** 3)) < 1e-12 * float(row[7])


@needs_bin
def test_hansen_jupiter_and_protoplanets(tmp_path):
    run(tmp_path, "Hansen_2009", 9)
    _, rows, _, _, _ = read_outputs(str(tmp_path), "Hansen_2009")
    assert rows[1][0] == "Jupiter" and int(rows[1][2]) == 1
    assert float(rows[1][5]) == 1.0 / 1.0473486e3
    assert abs(float(rows[1][6]) - 71492.0 / 1.495978707e8) < 1e-18
    st = [[float(x) for x in r[9:15]] for r in rows]
    rel = [st[1][c] - st[0][c] for c in range(3)]
    vel = [(st[1][3 + c] - st[0][3 + c]) * RO.GAUSS for c in range(3)]
    a, e, *_ = elements(rel, vel, RO.K2 * (1.0 + float(rows[1][5])))
    assert abs(a - 5.201472759810757) < 1e-11 and abs(e - 4.895926045495331E-02) < 1e-12
    # 400 protoplanets of 0.005 Earth masses on circles of 0.7-1 AU
    m = np.array([float(r[5]) for r in rows[2:]])
    assert len(m) == 400 and np.all(m == 5.0e-3 * (1.0 / 3.3294605e5))
    d = np.array([np.linalg.norm(np.array(s[:3]) - np.array(st[0][:3])) for s in st[2:]])
    assert d.min() > 0.7 - 1e-9 and d.max() < 1.0 + 1e-9


@needs_bin
def test_coll_stat_run_shares_the_integrated_solid_mass(tmp_path):
    """1000 equal protoplanet masses summing to the solid disk's mass, the
    surface density 7 g/cm^2 r^-3/2 (times 4.2 beyond the 2.7 AU snowline)
    integrated over 2.5-3 AU numerically."""
    from scipy.integrate import quad
    run(tmp_path, "coll_stat_run", 4)
    _, rows, _, _, _ = read_outputs(str(tmp_path), "coll_stat_run")
    m = np.array([float(r[5]) for r in rows[1:]])
    assert len(m) == 1000 and np.all(m == m[0])
    sig = lambda r: 7.0 * RO.SIGMA * r ** -1.5 * (4.2 if r > 2.7 else 1.0) * 2 * math.pi * r
    total = quad(sig, 2.5, 2.7, epsrel=1e-13)[0] + quad(sig, 2.7, 3.0, epsrel=1e-13)[0]
    assert abs(m.sum() / total - 1.0) < 1e-12


@needs_bin
def test_chambers2001_names_and_colliding_pair(tmp_path):
    run(tmp_path, "Chambers2001", 8)
    _, rows, _, _, _ = read_outputs(str(tmp_path), "Chambers2001")
    assert len(rows) == 301
    # the reference pushes Chambers2001's 154 names before Dvorak's 301
    assert rows[153][0] == "proto153" and rows[154][0] == "star" and rows[155][0] == "proto1"
    assert rows[300][0] == "proto146"
    run(tmp_path, "Colliding_bodies", 8)
    _, rows, _, _, _ = read_outputs(str(tmp_path), "Colliding_bodies")
    assert [r[0] for r in rows] == ["star", "target", "projectile"]
    t, p = (np.array([float(x) for x in r[9:12]]) for r in rows[1:])
    # 0.01 degree of mean anomaly at periapsis of a = 2.7, e = 0.05: a chord of
    # a (1 + e)^2 / sqrt(1 - e^2)^3 ... to first order r_p * dnu, dnu = dM (1+e)^2/(1-e^2)^1.5
    dM = 0.01 * math.pi / 180
    r_p = 2.7 * (1 - 0.05)
    chord = r_p * dM * (1 + 0.05) ** 2 / (1 - 0.05 ** 2) ** 1.5
    assert abs(np.linalg.norm(p - t) / chord - 1.0) < 1e-4


def test_normal_and_exponential_samplers_have_their_moments():
    """distribution.cpp's normal (mean, 'variance' used as the standard
    deviation, Box-Muller's cosine branch) and exponential (rate lambda)
    samplers, against their closed-form moments over 200000 draws."""
    n = 200000
    d = RO.Dist(4242, "normal", 1.45, 0.35)
    x = np.array([d.next() for _ in range(n)])
    assert abs(x.mean() - 1.45) < 5 * 0.35 / math.sqrt(n)
    assert abs(x.std() / 0.35 - 1.0) < 0.01
    assert abs(np.mean(x < 1.45 - 0.35) - 0.158655) < 0.005        # Phi(-1)
    d = RO.Dist(4243, "exponential", 0.5)
    y = np.array([d.next() for _ in range(n)])
    assert abs(y.mean() - 2.0) < 5 * 2.0 / math.sqrt(n)
    assert abs(np.mean(y > 2.0) - math.exp(-1.0)) < 0.005          # P(Y > 1/lambda) = 1/e


@needs_bin
def test_gt_scenarios_floor_and_giants(tmp_path):
    """Both GT presets: giants replaced by Jupiter, Saturn (and Uranus) in
    order, every other body's a at least 0.4 AU (the driver's floor), and the
    protoplanets' half-Moon masses."""
    for preset, giants in (("GT_scenario", ["Jupiter", "Saturn"]), ("GT_scenario_mod", ["Jupiter", "Saturn", "Uranus"])):
        run(tmp_path, preset, 21)
        _, rows, oe, _, _ = read_outputs(str(tmp_path), preset)
        assert [r[0] for r in rows[1:1 + len(giants)]] == giants
        assert rows[0][0] == ("Sun" if preset == "GT_scenario_mod" else "star")
        sma = np.array([float(line.split()[7]) for line in oe[1:]])
        assert sma.min() >= 0.4
        m = np.array([float(r[5]) for r in rows[1 + len(giants):]])
        assert len(m) == 2000 and np.all(m == 1.0 / 2.0 * (1.0 / 8.130059e1) * (1.0 / 3.3294605e5))
        # GT_scenario's normal(1.45, 0.35) puts ~0.4% below the floor; the exponential's ~18%
        frac = np.mean(sma[len(giants):] == 0.4)
        assert (frac < 0.02) if preset == "GT_scenario" else (0.1 < frac < 0.3)


@needs_bin
def test_elke_binary_and_birgit_comets(tmp_path):
    """Elke_binary_test: the secondary star on its a = 100 AU, e = 0.1 orbit
    about the primary; Birgit_scenario: 100 comets on a = 50-80 AU with
    e = 0.99-0.9999, massless, next to Jupiter and a Saturn of 58232 km."""
    run(tmp_path, "Elke_binary_test", 23)
    _, rows, _, _, _ = read_outputs(str(tmp_path), "Elke_binary_test")
    assert [r[0] for r in rows[:3]] == ["primary_star", "secondary_star", "Jupiter"]
    assert [int(r[2]) for r in rows[:3]] == [0, 0, 1] and len(rows) == 10003
    st = [[float(x) for x in r[9:15]] for r in rows[:2]]
    rel = [st[1][c] - st[0][c] for c in range(3)]
    vel = [(st[1][3 + c] - st[0][3 + c]) * RO.GAUSS for c in range(3)]
    a, e, inc, *_ = elements(rel, vel, RO.K2 * 2.0)
    assert abs(a / 100.0 - 1) < 1e-12 and abs(e - 0.1) < 1e-12 and inc == 0.0
    run(tmp_path, "Birgit_scenario", 23)
    _, rows, oe, _, _ = read_outputs(str(tmp_path), "Birgit_scenario")
    assert [r[0] for r in rows[:3]] == ["star", "Jupiter", "Saturn"]
    assert abs(float(rows[2][6]) - 58232.0 / 1.495978707e8) < 1e-18
    tp = [line.split() for line in oe[3:]]
    assert len(tp) == 100 and all(float(t[3]) == 0.0 for t in tp)
    a = np.array([float(t[7]) for t in tp])
    e = np.array([float(t[8]) for t in tp])
    assert a.min() >= 50.0 and a.max() <= 80.0 and e.min() >= 0.99 and e.max() <= 0.9999
