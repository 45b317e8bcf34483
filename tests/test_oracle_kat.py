"""The oracle against the reference's own known-answer tests and textbook facts.

Reference KATs: src/red.cuda/test.cu:387-448 (interaction bounds),
src/red.cuda.utility/red_test.cpp:22-125 (4-body fixture), 626-950 (tools),
1139-1175 (Kepler).  MT19937 against the C++ standard's published values."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
from redamd_types import VAR4, PARAM, BODY_MD, INTEGRAL

DEG = 1.745329251994329509e-2


def bound(pl, which):
    out = np.zeros(4, np.uint32)
    pl = np.asarray(pl, np.uint32)
    O.lib.or_bound(O.ptr(pl), which, O.ptr(out))
    return tuple(int(x) for x in out)


def test_bounds_kat():
    # test.cu:387-448: n_objects_t(1, 1, 1, 0, 2, 3, 2)
    pl = [1, 1, 1, 0, 2, 3, 2]
    assert bound(pl, 0) == (0, 3, 0, 8)
    assert bound(pl, 1) == (3, 8, 0, 3)
    assert bound(pl, 2) == (8, 10, 0, 8)
    assert bound(pl, 3) == (3, 8, 0, 0)


def fixture4():
    """red_test.cpp:22-125: four unit masses on the unit square."""
    r = np.zeros(4, VAR4)
    v = np.zeros(4, VAR4)
    p = np.zeros(4, PARAM)
    b = np.zeros(4, BODY_MD)
    r["x"] = [0, 1, 1, 0]
    r["y"] = [0, 0, 1, 1]
    v["x"] = [1, 2, 2, 1]
    p["mass"] = 1.0
    p["radius"] = 1.0
    p["density"] = 0.23873241463784300365332564505877
    b["id"] = [1, 1, 2, 3]
    b["body_type"] = [0, 1, 1, 3]
    return r, v, p, b


def test_tools_kat():
    r, v, p, b = fixture4()
    M0 = O.lib.or_get_total_mass(4, O.ptr(b), O.ptr(p))
    assert M0 == 4.0
    R0, V0 = np.zeros(1, VAR4), np.zeros(1, VAR4)
    O.lib.or_calc_bc(4, O.ptr(b), O.ptr(p), O.ptr(r), O.ptr(v), M0, O.ptr(R0), O.ptr(V0))
    assert abs(R0["x"][0] - 0.5) + abs(R0["y"][0] - 0.5) + abs(R0["z"][0]) <= 1e-15
    assert abs(V0["x"][0] - 1.5) + abs(V0["y"][0]) <= 1e-15
    O.lib.or_transform_to_bc(4, O.ptr(b), O.ptr(p), O.ptr(r), O.ptr(v))
    np.testing.assert_allclose(r["x"], [-0.5, 0.5, 0.5, -0.5], atol=1e-15)
    np.testing.assert_allclose(r["y"], [-0.5, -0.5, 0.5, 0.5], atol=1e-15)
    np.testing.assert_allclose(v["x"], [-0.5, 0.5, 0.5, -0.5], atol=1e-15)
    assert abs(O.lib.or_calc_radius(1.0, 1.0) - 0.62035049089940001666800681204778) <= 1e-15
    assert abs(O.lib.or_calc_mass(1.0, 1.0) - 4.1887902047863909846168578443727) <= 1e-15
    assert abs(O.lib.or_calc_density(1.0, 1.0) - 0.23873241463784300365332564505877) <= 1e-15


def test_merge_kat():
    r1, r2, rs = np.zeros(1, VAR4), np.zeros(1, VAR4), np.zeros(1, VAR4)
    r1[0] = (1, 1, 1, 0)
    r2[0] = (4, 3, 1, 0)
    O.lib.or_calc_position_after_collision(1.0, 1.0, O.ptr(r1), O.ptr(r2), O.ptr(rs))
    assert tuple(rs[0])[:3] == (2.5, 2.0, 1.0)
    v1, v2, vs = np.zeros(1, VAR4), np.zeros(1, VAR4), np.zeros(1, VAR4)
    v1[0] = (1, 0, 0, 0)
    v2[0] = (-1, 0, 0, 0)
    O.lib.or_calc_velocity_after_collision(1.0, 1.0, O.ptr(v1), O.ptr(v2), O.ptr(vs))
    assert tuple(vs[0])[:3] == (0.0, 0.0, 0.0)
    v1[0] = (-1, 0, 0, 0)
    O.lib.or_calc_velocity_after_collision(1.0, 1.0, O.ptr(v1), O.ptr(v2), O.ptr(vs))
    assert tuple(vs[0])[:3] == (-1.0, 0.0, 0.0)
    p1, p2, ps = np.zeros(1, PARAM), np.zeros(1, PARAM), np.zeros(1, PARAM)
    p1[0] = (1.0, 1.0, 1.0, 1.0)
    p2[0] = (1.0, 1.0, 1.0, 1.0)
    O.lib.or_calc_physical_properties(O.ptr(p1), O.ptr(p2), O.ptr(ps))
    assert ps["mass"][0] == 2.0 and ps["cd"][0] == 1.0
    assert abs(ps["density"][0] - 0.23873241463784300365332564505877) <= 1e-15
    assert abs(ps["radius"][0] - O.lib.or_calc_radius(2.0, 0.23873241463784300365332564505877)) <= 1e-15


@pytest.mark.parametrize("mean_deg,expected", [(27.0, 48.43417991487915), (225.0, 210.47211284530107)])
def test_kepler_kat(mean_deg, expected):
    E = C.c_double()
    assert O.lib.or_kepler_equation_solver(0.5, mean_deg * DEG, 1.0e-15, C.byref(E)) == 0
    assert abs(E.value / DEG - expected) <= 1e-13


def test_kepler_residual_sweep():
    # E - e sin E = M over the whole circle and a range of eccentricities
    E = C.c_double()
    for e in (0.0, 0.01, 0.3, 0.7, 0.95):
        for M in np.linspace(0.0, 2 * np.pi, 73, endpoint=False):
            assert O.lib.or_kepler_equation_solver(e, M, 1e-14, C.byref(E)) == 0
            assert abs(E.value - e * np.sin(E.value) - M) < 1e-12


def test_mt19937_standard_values():
    g = O.OrRng()
    O.lib.or_rng_seed(C.byref(g), 5489)
    vals = [O.lib.or_rng_u32(C.byref(g)) for _ in range(10000)]
    assert vals[-1] == 4123659995          # C++ [rand.predef]: 10000th output of mt19937
    O.lib.or_rng_seed(C.byref(g), 1)
    assert O.lib.or_rng_u32(C.byref(g)) == 1791095845


def test_samplers_moments():
    g = O.OrRng()
    O.lib.or_rng_seed(C.byref(g), 7)
    n = 40000
    u = np.array([O.lib.or_rng_uniform(C.byref(g), 0.0, 1.0) for _ in range(n)])
    assert u.min() > 0.0 and u.max() <= 1.0 and abs(u.mean() - 0.5) < 0.01
    ray = np.array([O.lib.or_rng_rayleigh(C.byref(g), 0.01) for _ in range(n)])
    assert abs(ray.mean() - 0.01 * np.sqrt(np.pi / 2)) < 2e-4
    a = np.array([O.lib.or_rng_power_law(C.byref(g), 0.5, 5.0, -0.5) for _ in range(n)])
    # pdf ~ a^-0.5 on [0.5, 5]: mean = (5^1.5 - 0.5^1.5) / 1.5 / (2 (5^0.5 - 0.5^0.5))
    mean = (5 ** 1.5 - 0.5 ** 1.5) / 1.5 / (2 * (5 ** 0.5 - 0.5 ** 0.5))
    assert a.min() >= 0.5 and a.max() <= 5.0 and abs(a.mean() - mean) < 0.03


def test_integral_two_body_closed_form():
    # circular binary, G = 1: E = -m1 m2 / (2 a), |C| = mu-reduced form
    m1, m2, a = 1.0, 1e-3, 1.3
    M = m1 + m2
    n = np.sqrt(M / a ** 3)
    r = np.zeros(2, VAR4)
    v = np.zeros(2, VAR4)
    r["x"] = [-m2 / M * a, m1 / M * a]
    v["y"] = [-m2 / M * a * n, m1 / M * a * n]
    p = np.zeros(2, PARAM)
    p["mass"] = [m1, m2]
    b = np.zeros(2, BODY_MD)
    b["id"] = [1, 2]
    out = np.zeros(1, INTEGRAL)
    O.lib.or_calc_integral(2, O.ptr(b), O.ptr(p), O.ptr(r), O.ptr(v), O.ptr(out))
    assert abs(out["E"][0] - (-m1 * m2 / (2 * a))) < 1e-15
    assert abs(out["C"]["z"][0] - m1 * m2 / M * a * a * n) < 1e-15
