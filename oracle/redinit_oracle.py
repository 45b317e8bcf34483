"""Test infrastructure, not product code: a CPU restatement of red's initial-
condition generator (src/red.cuda.initial) used to check red.cuda_amd/redinit.
Only tests import it.

Restated from the reference, step by step:
  rand()                 glibc's random(): TYPE_3 additive feedback, seeded by
                         the Park-Miller LCG (Schrage's method on the signed
                         seed), 310 outputs discarded, each output r >> 1
  red_random             distribution.cpp:11-31 (MT19937, one word per call)
  distributions          distribution.cpp:33-204 (uniform, exponential,
                         Rayleigh, normal, power law, lognormal)
  presets                main.cu:846-1168 (Two_body, n_gp, n_pp, n_spl, n_pl,
                         n_tp), main.cu:571-625 + 1804-1917 (Dvorak), nebula.cpp,
                         main.cu:1467-1688 (Chambers2001), 505-569 + 1690-1802
                         (coll_stat_run), 626-700 + 1917-2003 (Hansen_2009),
                         456-503 + 2863-2911 (n_pl_to_test_anal_gd), 41-249 +
                         438-454 + 2913-2961 (solar_system), 1412-1461 +
                         3055-3153 (Colliding_bodies), 702-844 + 2005-2202
                         (Birgit_2016, Birgit_scenario), 1170-1322 + 2204-2495
                         (GT_scenario, GT_scenario_mod), 1324-1410 + 2963-3053
                         (Elke_binary_test); util_init.cpp:327-366 (the
                         HORIZONS columns)
  populate_disk          main.cu:251-314, util_init.cpp:56-145
  calculate_phase        main.cu:317-342, tools.cpp:715-828, 928-933
  barycentre and units   tools.cpp:351-453

Pins (tests/test_redinit.py): rand() against the C library's own rand() via
ctypes; MT19937 against its published 10000th output; the two-body preset
against the circular-orbit closed form; the Kepler solver against the
residual of E - e sin E = M.
"""
import math

PI = 3.1415926535897932384626
K2 = 0.0002959122082855911025
GAUSS = 1.720209895e-2
DEG = 1.745329251994329509e-2
SOLAR_TO_GRAM = 1.0e3 * 1.98911e30
GRAM_TO_SOLAR = 1.0 / SOLAR_TO_GRAM
METER_TO_AU = 1.0 / 1.495978707e11
KM_TO_AU = 1.0 / (1.495978707e11 / 1.0e3)
R_SUN = 1.0 / 215.094
EARTH = 1.0 / 3.3294605e5
SATURN = 1.0 / 3.497898e3
MERCURY = 1.0 / 6.023600e6
VENUS = 1.0 / 4.0852371e5
MARS = 1.0 / 3.098708e6
JUPITER = 1.0 / 1.0473486e3
URANUS = 1.0 / 2.290298e4
NEPTUNE = 1.0 / 1.941224e4
MOON = 1.0 / 2.7071155e7
CERES = 1.0 / 2.109342e9
MOON_TO_EARTH = 1.0 / 8.130059e1
RHO = GRAM_TO_SOLAR / ((1.0e-2 * METER_TO_AU) * (1.0e-2 * METER_TO_AU) * (1.0e-2 * METER_TO_AU))
SIGMA = GRAM_TO_SOLAR / ((1.0e-2 * METER_TO_AU) * (1.0e-2 * METER_TO_AU))

STAR, GIANT, ROCKY, PROTO, SUPERPL, PL, TESTP = range(7)
TYPE_NAMES = ["star", "giant", "rocky", "proto", "superpl", "pl", "testp"]
M32 = 0xffffffff


def crand(seed):
    """Stream of glibc rand() values after srand(seed)."""
    s = seed & M32
    if s == 0:
        s = 1
    w = s - (1 << 32) if s >= (1 << 31) else s
    r = [w & M32]
    for _ in range(1, 31):
        q = abs(w) // 127773 * (1 if w >= 0 else -1)      # C division truncates
        hi, lo = q, w - q * 127773
        w = 16807 * lo - 2836 * hi
        if w < 0:
            w += 2147483647
        r.append(w & M32)
    r += r[0:3]
    while len(r) < 344:
        r.append((r[-31] + r[-3]) & M32)
    while True:
        x = (r[-31] + r[-3]) & M32
        r.append(x)
        del r[0]
        yield x >> 1


class RedRandom:
    """distribution.cpp:11-31: MT19937 that twists one word per draw."""

    def __init__(self, seed):
        self.I = [seed & M32]
        for i in range(1, 624):
            p = self.I[i - 1]
            self.I.append((1812433253 * (p ^ (p >> 30)) + i) & M32)
        self.idx = 0

    def rand(self):
        I, idx = self.I, self.idx
        j = idx + 1 if idx < 623 else 0
        y = (I[idx] & 0x80000000) | (I[j] & 0x7fffffff)
        y = I[idx + 397 if idx < 227 else idx - 227] ^ (y >> 1) ^ ((y & 1) * 0x9908b0df)
        I[idx] = y
        self.idx = j
        y ^= y >> 11
        y ^= (y << 7) & 0x9d2c5680
        y ^= (y << 15) & 0xefc60000
        y ^= y >> 18
        return y & M32

    def uniform(self, lo=None, hi=None):
        if lo is None:
            return (self.rand() + 1.0) / 4294967296.0
        return lo + (hi - lo) * (self.rand() + 1.0) / 4294967296.0


class Dist:
    def __init__(self, seed, kind, *args):
        self.rr = RedRandom(seed)
        self.kind = kind
        self.args = args

    def _lognormal_pdf(self, x, mu, sigma):
        return (1.0 / (math.sqrt(2.0 * PI) * sigma * x)
                * math.exp(-((math.log(x) - mu) * (math.log(x) - mu)) / (2.0 * (sigma * sigma))))

    def next(self):
        rr, a = self.rr, self.args
        if self.kind == "uniform":
            return rr.uniform(a[0], a[1])
        if self.kind == "rayleigh":
            return a[0] * math.sqrt(-2.0 * math.log(rr.uniform()))
        if self.kind == "exponential":                      # distribution.cpp:90-94
            return -1.0 / a[0] * math.log(rr.uniform())
        if self.kind == "normal":                           # distribution.cpp:120-124: u first, then rand()
            u = rr.uniform()
            return a[0] + a[1] * math.sqrt(-2.0 * math.log(u)) * math.cos(1.4629180792671596E-9 * (rr.rand() + 1.0))
        if self.kind == "power_law":
            x_min, x_max, p = a
            y_min = 0.0 if p == 0.0 else math.pow(x_min, p)
            y_max = 1.0 if p == 0.0 else math.pow(x_max, p)
            if y_min > y_max:
                y_min, y_max = y_max, y_min
            d_y, d_x = y_max - y_min, x_max - x_min
            area = d_x * d_y
            while True:
                x = rr.uniform(0.0, area) / d_y + x_min
                y = rr.uniform(y_min, y_max)
                if not y > math.pow(x, p):
                    return x
        if self.kind == "lognormal":
            x_min, x_max, mu, sigma = a
            y_min = 0.0
            y_max = self._lognormal_pdf(math.exp(mu - sigma * sigma), mu, sigma)
            if y_min > y_max:
                y_min, y_max = y_max, y_min
            d_y, d_x = y_max - y_min, x_max - x_min
            area = d_x * d_y
            while True:
                x = rr.uniform(0.0, area) / d_y + x_min
                y = rr.uniform(y_min, y_max)
                if not y > self._lognormal_pdf(x, mu, sigma):
                    return x
        raise ValueError(self.kind)


SLOTS = ["sma", "ecc", "inc", "peri", "node", "mean", "mass", "radius", "density", "cd"]


def _star():
    return [("mass", "uniform", 1.0, 1.0), ("radius", "uniform", R_SUN, R_SUN), ("cd", "uniform", 0.0, 0.0)]


def _oe10():
    return [("sma", "uniform", 1.0, 10.0), ("ecc", "uniform", 0.0, 0.1), ("inc", "uniform", 0.0, 10.0 * DEG),
            ("peri", "uniform", 0.0, 2.0 * PI), ("node", "uniform", 0.0, 2.0 * PI), ("mean", "uniform", 0.0, 2.0 * PI)]


def solid_mass(r_1, r_2, r_SL, f_neb, Sigma_1, f_ice, p):
    """solid_component::calc_mass (nebula.cpp)."""
    p2 = p + 2.0
    C = 2.0 * PI * f_neb * Sigma_1 / p2 if p2 != 0.0 else 2.0 * PI * f_neb * Sigma_1
    if r_SL <= r_1:
        I = f_ice * (math.pow(r_2, p2) - math.pow(r_1, p2)) if p2 != 0.0 else f_ice * (math.log(r_2) - math.log(r_1))
    elif r_SL <= r_2:
        I = (math.pow(r_SL, p2) - math.pow(r_1, p2) + f_ice * (math.pow(r_2, p2) - math.pow(r_SL, p2)) if p2 != 0.0
             else (math.log(r_SL) - math.log(r_1)) + f_ice * ((math.log(r_2) - math.log(r_SL))))
    else:
        I = math.pow(r_2, p2) - math.pow(r_1, p2) if p2 != 0.0 else (math.log(r_2) - math.log(r_1))
    return C * I


DVORAK_NEB = (0.9, 2.7, 2.7, 3.0, 7.0 * SIGMA, 4.2, -3.0 / 2.0)
CHAMBERS_NEB = (0.9, 2.7, 2.7, 1.0, 8.0 * SIGMA, 4.2, -3.0 / 2.0)
COLL_STAT_NEB = (2.5, 3.0, 2.7, 1.0, 7.0 * SIGMA, 4.2, -3.0 / 2.0)

# Elements of the planets on 2015-May-11 (JPL HORIZONS), in the order of the
# columns util_init.cpp:327-366 reads: e, i, Omega, omega, M (degrees), a (AU)
HORIZONS = {
    "Mercury": (2.056283170656704E-01, 7.004033081870772E+00, 4.831135275416084E+01, 2.917018579009503E+01,
                8.665226794919347E+01, 3.870990540148295E-01),
    "Venus": (6.761247965610893E-03, 3.394467689974794E+00, 7.663950855912078E+01, 5.463888630600258E+01,
              3.625130289912041E+01, 7.233287920706470E-01),
    "Earth": (1.706151645376220E-02, 2.832604859699792E-03, 1.991276585462631E+02, 2.644902808333621E+02,
              1.245850770307868E+02, 9.999423845001243E-01),
    "Mars": (9.345598920376896E-02, 1.848403968432629E+00, 4.951276588949865E+01, 2.865318030175783E+02,
             7.844645806929105E+01, 1.523589291796774E+00),
    "Jupiter": (4.895926045495331E-02, 1.303927595561599E+00, 1.005220570364609E+02, 2.736235124666692E+02,
                1.262463884135522E+02, 5.201472759810757E+00),
    "Saturn": (5.370451604891303E-02, 2.486273622454862E+00, 1.136214918787423E+02, 3.402344971420399E+02,
               1.437153605146956E+02, 9.553843040430950E+00),
    "Uranus": (4.928987553618186E-02, 7.717997843738144E-01, 7.400078543435393E+01, 9.677390970582711E+01,
               2.083879289018873E+02, 1.915532588905842E+01),
    "Neptune": (8.258758489207289E-03, 1.765657591232885E+00, 1.317164628477979E+02, 2.932421741299731E+02,
                2.734992366503029E+02, 2.997852567031729E+01),
}


def horizons_oe(name):
    """[sma, ecc, inc, peri, node, mean] with the degrees in radians."""
    e, i, node, peri, mean, a = HORIZONS[name]
    return [a, e, i * DEG, peri * DEG, node * DEG, mean * DEG]


def _dvorak_types(neb):
    return [(STAR, 1, 0, _star()),
            (PROTO, 300, 1, [("sma", "power_law", neb[0], neb[1], neb[6]), ("ecc", "rayleigh", 0.01),
                             ("inc", "uniform", 0.0, 0.0), ("peri", "uniform", 0.0, 2.0 * PI),
                             ("node", "uniform", 0.0, 0.0), ("mean", "uniform", 0.0, 2.0 * PI),
                             ("mass", "lognormal", 0.1, 3.0, 0.0, 0.25),
                             ("density", "uniform", 2.7 * RHO, 2.7 * RHO), ("cd", "uniform", 0.0, 0.0)])]


# (name, mass, radius in km) of the planets that drivers put in place of their
# first giants (e.g. main.cu:1955-1965, 2265-2292, 2428-2452)
JUPITER_SUB = ("Jupiter", JUPITER, 71492.0)
SATURN_SUB = ("Saturn", SATURN, 60268.0)
URANUS_SUB = ("Uranus", URANUS, 25362.0)


def _giants_and_floor(bodies, giants, sma_floor):
    """The drivers' loop over bodies 1..n-1 before calculate_phase: the k-th
    giant becomes giants[k] (its HORIZONS elements, mass, radius and the
    density from them), then a < sma_floor is raised to sma_floor."""
    g = 0
    for b in bodies[1:]:
        if b["type"] == GIANT and g < len(giants):
            name, m, r_km = giants[g]
            g += 1
            b["name"] = name
            b["mass"] = 1.0 * m
            b["radius"] = r_km * KM_TO_AU
            b["density"] = calc_density(b["mass"], b["radius"])
            b["oe"] = horizons_oe(name)
        if sma_floor > b["oe"][0]:
            b["oe"][0] = sma_floor


def _colliding_post(bodies):
    """main.cu:3090-3115: target and projectile on one orbit."""
    k = 0
    for b in bodies[1:]:
        if b["type"] != PROTO or k >= 2:
            continue
        b["name"] = "target" if k == 0 else "projectile"
        b["oe"] = [2.7, 0.05, 0.0, 0.0, 0.0, 0.0 if k == 0 else 0.01 * DEG]
        k += 1


def _solar_system_bodies():
    """populate_solar_system (main.cu:85-249): Sun, then Mercury..Mars (rocky),
    then Jupiter..Neptune (giant), ids from 1, cd 0."""
    rows = [("Mercury", ROCKY, MERCURY, 2439.7), ("Venus", ROCKY, VENUS, 6051.8), ("Earth", ROCKY, EARTH, 6371.0),
            ("Mars", ROCKY, MARS, 3389.5), ("Jupiter", GIANT, JUPITER, 71492.0), ("Saturn", GIANT, SATURN, 60268.0),
            ("Uranus", GIANT, URANUS, 25559.0), ("Neptune", GIANT, NEPTUNE, 24766.0)]
    sun = {"id": 1, "type": STAR, "name": "Sun", "mass": 1.0, "radius": 1.0 * R_SUN, "cd": 0.0, "oe": [0.0] * 6}
    sun["density"] = calc_density(sun["mass"], sun["radius"])
    bodies = [sun]
    for k, (name, btype, m, r_km) in enumerate(rows):
        b = {"id": k + 2, "type": btype, "name": name, "mass": 1.0 * m, "radius": r_km * KM_TO_AU, "cd": 0.0,
             "oe": horizons_oe(name)}
        b["density"] = calc_density(b["mass"], b["radius"])
        bodies.append(b)
    return bodies


def _chambers_names():
    """set_parameters::Chambers2001 pushes star + proto1..proto153 before
    Dvorak's star + proto1..proto300 (main.cu:402-433, 590-621)."""
    return (["star"] + ["proto" + str(i) for i in range(1, 154)] + ["star"]
            + ["proto" + str(i) for i in range(1, 301)])

# name -> (types [(body type, count, name offset, [(slot, kind, args...)] in construction order)],
#          recompute pass, scale to the MMSN solid mass)
PRESETS = {
    "two_body": ([(STAR, 1, 0, _star()),
                  (ROCKY, 1, 0, [("sma", "uniform", 1.0, 1.0), ("ecc", "uniform", 0.0, 0.0),
                                 ("inc", "uniform", 0.0, 0.0), ("peri", "uniform", 0.0, 0.0),
                                 ("node", "uniform", 0.0, 0.0), ("mean", "uniform", 0.0, 0.0),
                                 ("mass", "uniform", EARTH, EARTH), ("density", "uniform", 2.7 * RHO, 2.7 * RHO),
                                 ("cd", "uniform", 1.0, 1.0)])], False, False),
    "n_gp": ([(STAR, 1, 0, _star()),
              (GIANT, 10, 1, _oe10() + [("mass", "uniform", 0.1 * SATURN, 1.0 * SATURN),
                                        ("density", "uniform", 1.0 * RHO, 2.0 * RHO), ("cd", "uniform", 0.0, 0.0)])],
             True, False),
    "n_pp": ([(STAR, 1, 0, _star()),
              (PROTO, 500, 1, [("sma", "uniform", 0.6, 2.6), ("ecc", "uniform", 0.0, 0.1),
                               ("inc", "uniform", 0.0, 0.0 * DEG), ("peri", "uniform", 0.0, 2.0 * PI),
                               ("node", "uniform", 0.0, 0.0 * PI), ("mean", "uniform", 0.0, 2.0 * PI),
                               ("mass", "uniform", 0.5 * MOON, 2.0 * MOON), ("density", "uniform", 2.7 * RHO, 2.7 * RHO),
                               ("cd", "uniform", 0.0, 0.0)])], True, False),
    "n_spl": ([(STAR, 1, 0, _star()),
               (SUPERPL, 10, 1, _oe10() + [("mass", "uniform", 1.0e-10, 1.0e-9),
                                           ("radius", "uniform", 10.0 * KM_TO_AU, 30.0 * KM_TO_AU),
                                           ("density", "uniform", 1.0 * RHO, 3.0 * RHO), ("cd", "uniform", 1.0, 2.7)])],
              False, False),
    "n_pl": ([(STAR, 1, 0, _star()),
              (PL, 10, 1, _oe10() + [("mass", "uniform", 1.0e-3 * CERES, 1.0e-2 * CERES),
                                     ("density", "uniform", 1.0 * RHO, 3.0 * RHO), ("cd", "uniform", 0.0, 0.0)])],
             True, False),
    "n_tp": ([(STAR, 1, 0, _star()),
              (TESTP, 10, 1, _oe10() + [("mass", "uniform", 0.0, 0.0), ("radius", "uniform", 0.0, 0.0),
                                        ("density", "uniform", 0.0, 0.0), ("cd", "uniform", 0.0, 0.0)])],
             False, False),
    "Dvorak": ([(STAR, 1, 0, _star()),
                (PROTO, 300, 1, [("sma", "power_law", 0.9, 2.7, -3.0 / 2.0), ("ecc", "rayleigh", 0.01),
                                 ("inc", "uniform", 0.0, 0.0), ("peri", "uniform", 0.0, 2.0 * PI),
                                 ("node", "uniform", 0.0, 0.0), ("mean", "uniform", 0.0, 2.0 * PI),
                                 ("mass", "lognormal", 0.1, 3.0, 0.0, 0.25),
                                 ("density", "uniform", 2.7 * RHO, 2.7 * RHO), ("cd", "uniform", 0.0, 0.0)])],
               True, True),
}

# The scenarios whose drivers do more than populate + calculate_phase: MMSN
# mass ("scale" or "share"), the nebula, the reference's name list, the
# driver's changes before calculate_phase, or a table of bodies.
EXTRA = {
    "Chambers2001": dict(types=_dvorak_types(CHAMBERS_NEB), recompute=True, mmsn="scale", neb=CHAMBERS_NEB,
                         names="chambers"),
    "coll_stat_run": dict(types=[(STAR, 1, 0, _star()),
                                 (PROTO, 1000, 1, [("sma", "power_law", 2.5, 3.0, -3.0 / 2.0),
                                                   ("ecc", "rayleigh", 0.02), ("inc", "rayleigh", 0.04),
                                                   ("peri", "uniform", 0.0, 2.0 * PI),
                                                   ("node", "uniform", 0.0, 2.0 * PI),
                                                   ("mean", "uniform", 0.0, 2.0 * PI),
                                                   ("mass", "uniform", CERES, CERES),
                                                   ("density", "uniform", 2.5 * RHO, 2.5 * RHO),
                                                   ("cd", "uniform", 0.0, 0.0)])],
                          recompute=True, mmsn="share", neb=COLL_STAT_NEB),
    "Hansen_2009": dict(types=[(STAR, 1, 0, _star()),
                               (GIANT, 1, 1, [("sma", "uniform", 5.2, 5.2), ("ecc", "uniform", 0.05, 0.05),
                                              ("inc", "uniform", 0.0, 0.0), ("peri", "uniform", 0.0, 2.0 * PI),
                                              ("node", "uniform", 0.0, 0.0), ("mean", "uniform", 0.0, 2.0 * PI),
                                              ("mass", "uniform", JUPITER, JUPITER),
                                              ("density", "uniform", 1.326 * RHO, 1.326 * RHO),
                                              ("cd", "uniform", 0.0, 0.0)]),
                               (PROTO, 400, 1, [("sma", "uniform", 0.7, 1.0), ("ecc", "uniform", 0.0, 0.0),
                                                ("inc", "uniform", 0.0, 0.5 * DEG), ("peri", "uniform", 0.0, 0.0),
                                                ("node", "uniform", 0.0, 2.0 * PI),
                                                ("mean", "uniform", 0.0, 2.0 * PI),
                                                ("mass", "uniform", 5.0e-3 * EARTH, 5.0e-3 * EARTH),
                                                ("density", "uniform", 2.7 * RHO, 2.7 * RHO),
                                                ("cd", "uniform", 0.0, 0.0)])],
                        giants=[JUPITER_SUB]),
    "n_pl_to_test_anal_gd": dict(types=[(STAR, 1, 0, _star()),
                                        (PL, 10, 0, [("sma", "uniform", 1.0, 10.0), ("ecc", "uniform", 0.0, 0.0),
                                                     ("inc", "uniform", 0.0, 0.0), ("peri", "uniform", 0.0, 0.0),
                                                     ("node", "uniform", 0.0, 0.0), ("mean", "uniform", 0.0, 0.0),
                                                     ("radius", "uniform", METER_TO_AU, METER_TO_AU),
                                                     ("density", "uniform", 2.7 * RHO, 2.7 * RHO),
                                                     ("cd", "uniform", 1.0, 1.0)])]),
    "solar_system": dict(table=_solar_system_bodies),
    "Colliding_bodies": dict(types=[(STAR, 1, 0, _star()),
                                    (PROTO, 2, 0, [("sma", "uniform", 1.0, 2.0), ("ecc", "uniform", 0.0, 0.0),
                                                   ("inc", "uniform", 0.0, 0.0), ("peri", "uniform", 0.0, 0.0),
                                                   ("node", "uniform", 0.0, 0.0), ("mean", "uniform", 0.0, 0.0),
                                                   ("mass", "uniform", CERES, CERES),
                                                   ("density", "uniform", 2.7 * RHO, 2.7 * RHO),
                                                   ("cd", "uniform", 0.0, 0.0)])],
                             post=_colliding_post),
    # set_parameters main.cu:702-775, driver 2005-2099
    "Birgit_2016": dict(types=[(STAR, 1, 0, _star()),
                               (GIANT, 2, 1, [("sma", "uniform", 5.2, 5.2), ("ecc", "uniform", 0.05, 0.05),
                                              ("inc", "uniform", 0.0, 0.0), ("peri", "uniform", 0.0, 2.0 * PI),
                                              ("node", "uniform", 0.0, 0.0), ("mean", "uniform", 0.0, 2.0 * PI),
                                              ("mass", "uniform", JUPITER, JUPITER),
                                              ("density", "uniform", 1.326 * RHO, 1.326 * RHO),
                                              ("cd", "uniform", 0.0, 0.0)]),
                               (PROTO, 1000, 1, [("sma", "uniform", 0.3, 2.4), ("ecc", "uniform", 0.02, 0.02),
                                                 ("inc", "uniform", 1.0 * DEG, 1.0 * DEG),
                                                 ("peri", "uniform", 0.0, 2.0 * PI),
                                                 ("node", "uniform", 0.0, 2.0 * PI),
                                                 ("mean", "uniform", 0.0, 2.0 * PI),
                                                 ("mass", "uniform", 0.1 * MOON, 0.1 * MOON),
                                                 ("density", "uniform", 2.0 * RHO, 2.0 * RHO),
                                                 ("cd", "uniform", 0.0, 0.0)])],
                        giants=[JUPITER_SUB, SATURN_SUB]),
    # set_parameters main.cu:777-844, driver 2101-2202 (its Saturn has R = 58232 km)
    "Birgit_scenario": dict(types=[(STAR, 1, 0, _star()),
                                   (GIANT, 2, 1, [("sma", "uniform", 1.0, 2.0), ("ecc", "uniform", 0.0, 0.05),
                                                  ("inc", "uniform", 0.0, 1.0 * DEG),
                                                  ("peri", "uniform", 0.0, 2.0 * PI),
                                                  ("node", "uniform", 0.0, 2.0 * PI),
                                                  ("mean", "uniform", 0.0, 2.0 * PI),
                                                  ("mass", "uniform", JUPITER, JUPITER),
                                                  ("density", "uniform", 2.7 * RHO, 2.7 * RHO),
                                                  ("cd", "uniform", 0.0, 0.0)]),
                                   (TESTP, 100, 1, [("sma", "uniform", 50.0, 80.0), ("ecc", "uniform", 0.99, 0.9999),
                                                    ("inc", "uniform", 0.0, 0.5), ("peri", "uniform", 0.0, 2.0 * PI),
                                                    ("node", "uniform", 0.0, 2.0 * PI),
                                                    ("mean", "uniform", 0.0, 1.0 * PI)])],
                            giants=[JUPITER_SUB, ("Saturn", SATURN, 58232.0)]),
    "GT_scenario": dict(types=None, gt=(2, ("sma", "normal", 1.45, 0.35)), recompute=True,
                        giants=[JUPITER_SUB, SATURN_SUB], sma_floor=0.4),
    "GT_scenario_mod": dict(types=None, gt=(3, ("sma", "exponential", 0.5)), recompute=True,
                            giants=[JUPITER_SUB, SATURN_SUB, URANUS_SUB], sma_floor=0.4, names=["Sun"]),
    # set_parameters main.cu:1324-1410, driver 2963-3053
    "Elke_binary_test": dict(types=[(STAR, 2, 0, [("sma", "uniform", 100.0, 100.0), ("ecc", "uniform", 0.1, 0.1),
                                                  ("inc", "uniform", 0.0, 0.0), ("peri", "uniform", 0.0, 0.0 * PI),
                                                  ("node", "uniform", 0.0, 0.0), ("mean", "uniform", 0.0, 0.0 * PI)]
                                    + _star()),
                                    (GIANT, 1, 1, [("sma", "uniform", 5.2, 5.2), ("ecc", "uniform", 0.0, 0.0),
                                                   ("inc", "uniform", 0.0, 0.0), ("peri", "uniform", 0.0, 2.0 * PI),
                                                   ("node", "uniform", 0.0, 0.0), ("mean", "uniform", 0.0, 2.0 * PI),
                                                   ("mass", "uniform", 1.0 * JUPITER, 1.0 * JUPITER),
                                                   ("density", "uniform", 1.326 * RHO, 1.326 * RHO),
                                                   ("cd", "uniform", 0.0, 0.0)]),
                                    (PROTO, 10000, 1, [("sma", "uniform", 0.5, 4.0), ("ecc", "rayleigh", 0.02),
                                                       ("inc", "uniform", 0.0, 0.0),
                                                       ("peri", "uniform", 0.0, 2.0 * PI),
                                                       ("node", "uniform", 0.0, 0.0),
                                                       ("mean", "uniform", 0.0, 2.0 * PI),
                                                       ("mass", "uniform", 1.0 * CERES, 1.0 * CERES),
                                                       ("density", "uniform", 2.0 * RHO, 2.0 * RHO),
                                                       ("cd", "uniform", 0.0, 0.0)])],
                             names=["primary_star", "secondary_star", "Jupiter"], giants=[JUPITER_SUB], sma_floor=0.4),
}


def _gt_types(n_giant, sma):
    """GT_scenario / GT_scenario_mod (main.cu:1170-1322): the giants' mass and
    every density are uniform(x, 0) -- drawn as x + (0 - x) u, as a build with
    NDEBUG does (distribution.cpp's assert(x_min <= x_max) stops one without)."""
    m_pp = 1.0 / 2.0 * MOON_TO_EARTH * EARTH
    return [(STAR, 1, 0, _star()),
            (GIANT, n_giant, 1, [("sma", "uniform", 1.0, 2.00), ("ecc", "uniform", 0.0, 0.05),
                                 ("inc", "uniform", 0.0, 1.0 * DEG), ("peri", "uniform", 0.0, 2.0 * PI),
                                 ("node", "uniform", 0.0, 2.0 * PI), ("mean", "uniform", 0.0, 2.0 * PI),
                                 ("mass", "uniform", 1.0 * JUPITER, 0.0), ("density", "uniform", 2.7 * RHO, 0.0),
                                 ("cd", "uniform", 0.0, 0.0)]),
            (PROTO, 2000, 1, [sma, ("ecc", "uniform", 0.0, 0.05), ("inc", "uniform", 0.0, 1.0 * DEG),
                              ("peri", "uniform", 0.0, 2.0 * PI), ("node", "uniform", 0.0, 2.0 * PI),
                              ("mean", "uniform", 0.0, 2.0 * PI), ("mass", "uniform", m_pp, m_pp),
                              ("density", "uniform", 2.7 * RHO, 0.0), ("cd", "uniform", 0.0, 0.0)])]


def scenario(name):
    """The generation recipe of a preset, as a dict."""
    if name in EXTRA:
        sc = dict(EXTRA[name])
        if sc.get("gt"):
            sc["types"] = _gt_types(*sc["gt"])
        return sc
    types, recompute, scale = PRESETS[name]
    return dict(types=types, recompute=recompute, mmsn="scale" if scale else None, neb=DVORAK_NEB)


FOUR_PI_3 = 4.188790204786391


def calc_radius(m, rho):
    return math.pow(1.0 / FOUR_PI_3 * m / rho, 1.0 / 3.0)


def calc_density(m, R):
    return 0.0 if R == 0.0 else m / (FOUR_PI_3 * (R * R * R))


def calc_mass(R, rho):
    return FOUR_PI_3 * (R * R * R) * rho


def kepler(ecc, mean, eps=1.0e-8):
    """tools.cpp:715-750 as written (M > pi: E = M - pi solved, then E + ((M' - E) + pi))."""
    if ecc == 0.0 or mean == 0.0 or mean == PI:
        return mean
    s = False
    if PI < mean:
        mean -= PI
        s = True
    E = mean + ecc * math.sin(mean) / (1.0 - math.sin(mean + ecc) + math.sin(mean))
    step = 0
    while True:
        E1 = E - (E - ecc * math.sin(E) - mean) / (1.0 - ecc * math.cos(E))
        err = abs(E1 - E)
        E = E1
        if not err > eps:
            break
        c = step <= 32
        step += 1
        if not c:
            break
    if 32 < step:
        raise RuntimeError("kepler_equation_solver did not converge")
    if s:
        E += (mean - E) + PI
    return E


def calc_phase(mu, oe):
    sma, ecc, inc, peri, node, mean = oe
    E = kepler(ecc, mean)
    v = 2.0 * math.atan(math.sqrt((1.0 + ecc) / (1.0 - ecc)) * math.tan(E / 2.0))
    p = sma * (1.0 - ecc * ecc)
    r = p / (1.0 + ecc * math.cos(v))
    kszi, eta = r * math.cos(v), r * math.sin(v)
    vk, ve = -math.sqrt(mu / p) * math.sin(v), math.sqrt(mu / p) * (ecc + math.cos(v))
    cw, sw, cO, sO, ci, si = (math.cos(peri), math.sin(peri), math.cos(node), math.sin(node), math.cos(inc),
                              math.sin(inc))
    P = (cw * cO - sw * sO * ci, cw * sO + sw * cO * ci, sw * si)
    Q = (-sw * cO - cw * sO * ci, -sw * sO + cw * cO * ci, cw * si)
    rv = [kszi * P[k] + eta * Q[k] for k in range(3)]
    vv = [vk * P[k] + ve * Q[k] for k in range(3)]
    return rv, vv


def generate(name, seed):
    """Bodies (list of dicts), t0, dt and body counts per type, in red's units."""
    sc = scenario(name)
    if "table" in sc:
        bodies = sc["table"]()
        n_type = [0] * 7
        for b in bodies:
            n_type[b["type"]] += 1
        return _phase_bc_units(bodies, n_type)
    types, recompute, mmsn = sc["types"], sc.get("recompute", False), sc.get("mmsn")
    rnd = crand(seed)
    dists = []
    for (_, _, _, entries) in types:
        d = {}
        for e in entries:
            d[e[0]] = Dist(next(rnd), e[1], *e[2:])
        dists.append(d)
    bodies, n_type, bid = [], [0] * 7, 1
    for k, (btype, n, name_from, _) in enumerate(types):
        D = dists[k]
        n_type[btype] += n
        for i in range(n):
            b = {"id": bid, "type": btype,
                 "name": "star" if (btype == STAR and i == 0) else TYPE_NAMES[btype] + str(i + name_from)}
            for s in ("mass", "radius", "density", "cd"):
                b[s] = D[s].next() if s in D else 0.0
            if btype == TESTP:
                b["mass"] = b["radius"] = b["density"] = b["cd"] = 0.0
            elif btype != SUPERPL:
                if "mass" not in D and "radius" in D and "density" in D:
                    b["mass"] = calc_mass(b["radius"], b["density"])
                elif "radius" not in D and "mass" in D and "density" in D:
                    b["radius"] = calc_radius(b["mass"], b["density"])
                elif "density" not in D and "mass" in D and "radius" in D:
                    b["density"] = calc_density(b["mass"], b["radius"])
                else:
                    raise RuntimeError("Missing physical parameters!")
            if btype == STAR and bid == 1:
                b["oe"] = [0.0] * 6
            else:
                b["oe"] = [D[s].next() if s in D else 0.0 for s in SLOTS[:6]]
            b["k"] = k
            bodies.append(b)
            bid += 1
    if mmsn:
        m_solid = solid_mass(*sc["neb"])

        def total():
            m = 0.0
            for b in reversed(bodies):
                if b["type"] == PROTO and b["id"] > 0:
                    m += b["mass"]
            return m
        if mmsn == "scale":
            f = m_solid / total()
            for b in bodies:
                if b["type"] == PROTO:
                    b["mass"] *= f
        else:
            m_pp = m_solid / n_type[PROTO]
            for b in bodies:
                if b["type"] == PROTO:
                    b["mass"] = m_pp
        if abs(total() - m_solid) > 1.0e-15:
            return None
    if recompute:
        for b in bodies:
            D = dists[b["k"]]
            if b["mass"] > 0.0:
                if "density" not in D and "radius" in D:
                    b["density"] = calc_density(b["mass"], b["radius"])
                if "radius" not in D and "density" in D:
                    b["radius"] = calc_radius(b["mass"], b["density"])
    names = sc.get("names")
    for b, nm in zip(bodies, _chambers_names() if names == "chambers" else (names or [])):
        b["name"] = nm
    if sc.get("post"):
        sc["post"](bodies)
    _giants_and_floor(bodies, sc.get("giants", []), sc.get("sma_floor", 0.0))
    return _phase_bc_units(bodies, n_type)


def _phase_bc_units(bodies, n_type):
    """calculate_phase (main.cu:317-342), the barycentre and k-day units."""
    m0 = bodies[0]["mass"]
    bodies[0]["r"], bodies[0]["v"] = [0.0, 0.0, 0.0], [0.0, 0.0, 0.0]
    min_P = 1.7976931348623157e308
    for b in bodies[1:]:
        mu = K2 * (m0 + b["mass"])
        b["r"], b["v"] = calc_phase(mu, b["oe"])
        a = b["oe"][0]
        P = 6.283185307179586 * math.sqrt((a * a * a) / mu)
        if min_P > P:
            min_P = P
    dt = min_P / 1000.0
    M0 = 0.0
    for b in reversed(bodies):
        if b["id"] >= 0:
            M0 += b["mass"]
    R0, V0 = [0.0] * 3, [0.0] * 3
    for b in reversed(bodies):
        for c in range(3):
            R0[c] += b["mass"] * b["r"][c]
            V0[c] += b["mass"] * b["v"][c]
    R0 = [x / M0 for x in R0]
    V0 = [x / M0 for x in V0]
    for b in bodies:
        b["r"] = [b["r"][c] - R0[c] for c in range(3)]
        b["v"] = [(b["v"][c] - V0[c]) / GAUSS for c in range(3)]
    return bodies, 0.0 * GAUSS, dt * GAUSS, n_type
