"""CPU pins of the oracle's FARGO grid drag (calc_linear_index pp_disk.cu:158-193,
the CPU drag loop pp_disk.cu:778-814, rotate_2D_vector util.cu:76-83).

Each check fixes the oracle by something other than its own formula: the
reference's known answers for rotate_2D_vector (red_test.cpp:170-250), cells
built by construction and by numpy's searchsorted, the analytic disk's drag
(a different formula for the gas velocity) on a grid that reproduces it, and
the drag from unit vectors in long double."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_lib as O
from redamd_types import VAR4, GasParams
from systems import fargo_grid, small_system, mixed_types, as3

TWO_PI = 6.2831853071795864769253


# rotate_2D_vector(theta, (0, 1)) known answers, red_test.cpp:170-250
ROT_KAT = [(0.0, (0.0, 1.0)), (math.pi / 2, (-1.0, 0.0)), (math.pi, (0.0, -1.0)), (3 * math.pi / 2, (1.0, 0.0)),
           (math.pi / 4, (-1 / math.sqrt(2.0), 1 / math.sqrt(2.0))),
           (3 * math.pi / 4, (-1 / math.sqrt(2.0), -1 / math.sqrt(2.0)))]


@pytest.mark.parametrize("theta,expected", ROT_KAT)
def test_rotate_2d_known_answers(theta, expected):
    r = O.lib.or_rotate_2d(theta, O.VAR4C(0.0, 1.0, 0.0, 0.0))
    assert abs(r.x - expected[0]) <= 2e-16 and abs(r.y - expected[1]) <= 2e-16
    assert r.z == 0.0 and r.w == 0.0


def test_rotate_2d_is_a_rotation():
    rng = np.random.default_rng(5)
    for _ in range(200):
        th, x, y = rng.uniform(-7, 7), rng.normal(), rng.normal()
        a = O.lib.or_rotate_2d(th, O.VAR4C(x, y, 9.0, 9.0))
        # counter-clockwise by theta, in long double
        c, s = np.cos(np.longdouble(th)), np.sin(np.longdouble(th))
        assert abs(float(c * x - s * y) - a.x) <= 1e-15 * (1 + abs(a.x))
        assert abs(float(s * x + c * y) - a.y) <= 1e-15 * (1 + abs(a.y))
        b = O.lib.or_rotate_2d(-th, a)
        assert abs(b.x - x) <= 1e-14 and abs(b.y - y) <= 1e-14


@pytest.mark.parametrize("n_rad,n_sec", [(1, 1), (3, 8), (17, 64), (64, 257)])
def test_cell_by_construction(n_rad, n_sec):
    ur, de, vr, vt = fargo_grid(n_rad, n_sec, 3)
    g = O.OGrid(ur, de, vr, vt)
    rng = np.random.default_rng(n_rad * 1000 + n_sec)
    dal = TWO_PI / n_sec
    for k in range(n_rad):
        for s in rng.choice(n_sec, size=min(n_sec, 16), replace=False):
            # a point inside ring k and sector s, away from every edge
            r = ur[k] + (ur[k + 1] - ur[k]) * rng.uniform(0.05, 0.95)
            th = (s + rng.uniform(0.05, 0.95)) * dal
            assert g.cell(r * math.cos(th), r * math.sin(th)) == k * n_sec + s


def test_cell_against_searchsorted():
    n_rad, n_sec = 40, 96
    ur, de, vr, vt = fargo_grid(n_rad, n_sec, 4)
    g = O.OGrid(ur, de, vr, vt)
    rng = np.random.default_rng(8)
    x, y = rng.uniform(-2.7, 2.7, 20000), rng.uniform(-2.7, 2.7, 20000)
    n_checked = 0
    for xi, yi in zip(x, y):
        r = math.hypot(xi, yi)
        ang = math.atan2(yi, xi) % TWO_PI
        # stay clear of ring and sector edges, where the two roundings may part
        if np.min(np.abs(ur - r)) < 1e-9 or min(ang % (TWO_PI / n_sec), -ang % (TWO_PI / n_sec)) < 1e-9:
            continue
        if r < ur[0]:
            want = 0
        elif r > ur[-1]:
            want = n_rad * n_sec - 1
        else:
            want = (np.searchsorted(ur, r, side="right") - 1) * n_sec + int(ang // (TWO_PI / n_sec))
        assert g.cell(xi, yi) == want
        n_checked += 1
    assert n_checked > 19000


def test_cell_edges():
    n_rad, n_sec = 5, 12
    ur, de, vr, vt = fargo_grid(n_rad, n_sec, 6)
    g = O.OGrid(ur, de, vr, vt)
    last = n_rad * n_sec - 1
    assert g.cell(0.5 * ur[0], 0.0) == 0                  # inside the inner edge
    assert g.cell(0.0, 0.0) == 0
    assert g.cell(-2.0 * ur[-1], 1.0) == last             # outside the outer edge
    # r on an inner interface belongs to the ring above it
    for k in range(1, n_rad):
        assert g.cell(ur[k], 0.0) == k * n_sec
    # r equal to the last interface fits no half-open ring: the scan leaves ring 0
    assert g.cell(ur[-1], 0.0) == 0
    # y = -0 reads atan2 = -0 (x > 0: sector 0) or -pi (x < 0: held at sector 0)
    rm = 0.5 * (ur[2] + ur[3])
    assert g.cell(rm, -0.0) == 2 * n_sec
    assert g.cell(-rm, -0.0) == 2 * n_sec
    # y just below 0 with x > 0: alpha rounds to 2 pi and the flat index runs
    # into the next ring's sector 0 (held at the last cell in the last ring)
    q = TWO_PI / (TWO_PI / n_sec)
    wrap = n_sec if q >= n_sec else n_sec - 1
    assert g.cell(rm, -1e-300) == 2 * n_sec + wrap
    rl = 0.5 * (ur[-2] + ur[-1])
    assert g.cell(rl, -1e-300) == min((n_rad - 1) * n_sec + wrap, last)


def test_uniform_grid_equals_the_analytic_disk():
    """A grid of constant density rho0 at rest relative to Keplerian rotation
    drags like the analytic disk with rho = (rho0, 0), eta = 0 in the midplane
    (pp_disk.cu:87-155: density rho0 r^0 for r > 0.1, gas at the circular
    velocity) -- a different formula for the gas velocity's direction."""
    n = 500
    r, v, p, b, nbt = small_system(n, 21, mixed_types(n))
    x = as3(r)
    rad = np.hypot(x[:, 0], x[:, 1])
    keep = (rad > 0.45) & (rad < 2.4)
    r["z"] = 0.0
    n_rad, n_sec, rho0 = 30, 48, 3e-9
    ur, _, _, _ = fargo_grid(n_rad, n_sec, 2)
    dens, zero = np.full((n_rad, n_sec), rho0), np.zeros((n_rad, n_sec))
    g = O.OGrid(ur, dens, zero, zero)
    gp = GasParams()
    gp.rho[0], gp.rho[1] = rho0, 0.0
    gp.sch[0], gp.sch[1] = 0.05, 0.0
    gp.eta[0], gp.eta[1] = 0.0, 0.0
    gp.gas_decrease = 0
    for mode in (0, 1):
        a_grid, a_anal = np.zeros(n, VAR4), np.zeros(n, VAR4)
        O.lib.or_grid_drag(C.byref(g.g), mode, O.ptr(nbt), O.ptr(p), O.ptr(r), O.ptr(v), O.ptr(a_grid))
        O.lib.or_drag(0.0, C.byref(gp), mode, n, O.ptr(nbt), O.ptr(p), O.ptr(r), O.ptr(v), O.ptr(a_anal))
        ag, aa = as3(a_grid)[keep], as3(a_anal)[keep]
        assert np.abs(aa).max() > 0
        assert np.all(np.abs(ag - aa) <= 1e-13 * np.abs(aa).max(axis=1, keepdims=True))
        lo = nbt[:4].sum() if mode == 0 else 1
        hi = nbt[:6].sum()
        touched = np.any(as3(a_grid) != 0, axis=1)
        assert not touched[:lo].any() and not touched[hi:].any() and touched[lo:hi][keep[lo:hi]].all()


def grid_drag_ld(ur, dens, vrad, vth, m_star, x, vel, cd, radius, rho_p):
    """Drag a = 3 C_D rho_g |u| u / (8 R rho_p) with the gas velocity
    (sqrt(m/r) + v_theta) e_theta + v_rad e_r from unit vectors, long double."""
    ld = np.longdouble
    n_rad, n_sec = dens.shape
    out = np.zeros((len(x), 3))
    for i in range(len(x)):
        xi, yi = ld(x[i, 0]), ld(x[i, 1])
        rr = np.sqrt(xi * xi + yi * yi)
        k = np.searchsorted(ur, float(rr), side="right") - 1
        s = int((math.atan2(x[i, 1], x[i, 0]) % TWO_PI) // (TWO_PI / n_sec))
        c = k * n_sec + s
        er = np.array([xi / rr, yi / rr, ld(0)])
        et = np.array([-yi / rr, xi / rr, ld(0)])
        vg = (np.sqrt(ld(m_star) / rr) + ld(vth.flat[c])) * et + ld(vrad.flat[c]) * er
        u = vg - np.array([ld(q) for q in vel[i]])
        un = np.sqrt((u * u).sum())
        f = ld(3) * ld(cd[i]) * ld(dens.flat[c]) * un / (ld(8) * ld(radius[i]) * ld(rho_p[i]))
        out[i] = (f * u).astype(np.float64)
    return out


def test_grid_drag_against_unit_vectors():
    n = 400
    r, v, p, b, nbt = small_system(n, 33)
    rng = np.random.default_rng(2)
    rad, ang = rng.uniform(0.45, 2.4, n), rng.uniform(0, TWO_PI, n)
    r["x"], r["y"] = rad * np.cos(ang), rad * np.sin(ang)
    r["z"] = rng.uniform(-0.05, 0.05, n)
    p["radius"] = 10 ** rng.uniform(-8, -6, n)
    p["cd"] = rng.uniform(0.5, 2.0, n)
    ur, de, vr, vt = fargo_grid(24, 40, 9)
    x = as3(r)
    # keep clear of cell edges (there the two roundings may take neighbouring cells)
    rr = np.hypot(x[:, 0], x[:, 1])
    aa = np.arctan2(x[:, 1], x[:, 0]) % TWO_PI
    dal = TWO_PI / 40
    edge = (np.min(np.abs(ur[None, :] - rr[:, None]), axis=1) < 1e-9) | (np.minimum(aa % dal, -aa % dal) < 1e-9)
    g = O.OGrid(ur, de, vr, vt)
    a = np.zeros(n, VAR4)
    O.lib.or_grid_drag(C.byref(g.g), 1, O.ptr(nbt), O.ptr(p), O.ptr(r), O.ptr(v), O.ptr(a))
    want = grid_drag_ld(ur, de, vr, vt, p["mass"][0], x[1:], as3(v)[1:], p["cd"][1:], p["radius"][1:],
                        p["density"][1:])
    got = as3(a)[1:]
    ok = ~edge[1:]
    assert ok.sum() > 390
    assert np.all(np.abs(got[ok] - want[ok]) <= 1e-13 * np.abs(want[ok]).max(axis=1, keepdims=True))
    assert np.all(as3(a)[0] == 0.0)


def test_disk_uses_the_grid_in_place_of_the_analytic_gas():
    """or_calc_dydx with a grid set: gravity + grid drag (the analytic disk it
    was created with is ignored); removing the grid restores the analytic drag."""
    n = 300
    r, v, p, b, nbt = small_system(n, 41, mixed_types(n))
    gp = GasParams()
    gp.rho[0], gp.sch[0], gp.eta[0] = 1e-9, 0.05, 0.002
    ur, de, vr, vt = fargo_grid(12, 20, 1)
    d_none = O.ODisk(r, v, p, b, nbt)
    d = O.ODisk(r, v, p, b, nbt, gas=gp)
    _, base = d_none.calc_dydx(1, 0, 0.0, r, v)
    _, anal = d.calc_dydx(1, 0, 0.0, r, v)
    d.set_gas_grid(ur, de, vr, vt)
    _, grid = d.calc_dydx(1, 0, 0.0, r, v)
    g = O.OGrid(ur, de, vr, vt)
    drag = np.zeros(n, VAR4)
    O.lib.or_grid_drag(C.byref(g.g), 0, O.ptr(nbt), O.ptr(p), O.ptr(r), O.ptr(v), O.ptr(drag))
    assert np.array_equal(as3(grid), as3(base) + as3(drag))
    assert not np.array_equal(as3(grid), as3(anal))
    d.set_gas_grid(None, None, None, None)
    _, back = d.calc_dydx(1, 0, 0.0, r, v)
    assert np.array_equal(as3(back), as3(anal))


def test_bench_grid_reproduces_the_analytic_disk():
    """bench.py --config 3 --gas grid drags through a frame sampled from the
    analytic disk (grid_from_analytic): at the cell centres, in the midplane
    and outside the 0.1 AU taper, its drag equals the analytic disk's."""
    import os
    import sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import bench
    import redamd as R
    g = R.gas_params()
    n_rad, n_sec, r_min, r_max = bench.C3_GRID
    grid = bench.grid_from_analytic(g, *bench.C3_GRID)
    ur = grid[0]
    assert ur.size == n_rad + 1 and grid[1].shape == (n_rad, n_sec)
    n = 300
    rng = np.random.default_rng(4)
    k, s = rng.integers(0, n_rad, n), rng.integers(0, n_sec, n)
    rc, th = np.sqrt(ur[k] * ur[k + 1]), (s + 0.5) * TWO_PI / n_sec
    r, v, p, b, nbt = small_system(n + 1, 5)
    r["x"][1:], r["y"][1:], r["z"] = rc * np.cos(th), rc * np.sin(th), 0.0
    v["x"][1:], v["y"][1:] = -np.sin(th) * np.sqrt(1 / rc) * 1.01, np.cos(th) * np.sqrt(1 / rc) * 1.01
    p["radius"], p["density"] = 1e-8, 3.4e6
    a_grid, a_anal = np.zeros(n + 1, VAR4), np.zeros(n + 1, VAR4)
    og = O.OGrid(*grid)
    O.lib.or_grid_drag(C.byref(og.g), 1, O.ptr(nbt), O.ptr(p), O.ptr(r), O.ptr(v), O.ptr(a_grid))
    O.lib.or_drag(0.0, C.byref(g), 1, n + 1, O.ptr(nbt), O.ptr(p), O.ptr(r), O.ptr(v), O.ptr(a_anal))
    ag, aa = as3(a_grid)[1:], as3(a_anal)[1:]
    assert np.all(np.abs(aa).max(axis=1) > 0)
    assert np.all(np.abs(ag - aa) <= 1e-12 * np.abs(aa).max(axis=1, keepdims=True))
