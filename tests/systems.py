"""Seeded body systems shared by the tests."""
import numpy as np

from redamd_types import VAR4, PARAM, BODY_MD


def small_system(n, seed, types=None, inactive=(), spread=2.0):
    """Star at index 0 plus n-1 bodies with random states, sorted by type."""
    rng = np.random.default_rng(seed)
    r = np.zeros(n, VAR4)
    v = np.zeros(n, VAR4)
    p = np.zeros(n, PARAM)
    b = np.zeros(n, BODY_MD)
    r["x"], r["y"], r["z"] = rng.uniform(-spread, spread, n), rng.uniform(-spread, spread, n), rng.uniform(-0.1, 0.1, n)
    v["x"], v["y"], v["z"] = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n), rng.uniform(-0.1, 0.1, n)
    p["mass"] = 10 ** rng.uniform(-9, -5, n)
    p["mass"][0] = 1.0
    p["radius"] = 1e-5
    p["density"] = 1e6
    p["cd"] = 1.0
    types = list(types) if types is not None else [0] + [3] * (n - 1)
    b["body_type"] = types
    b["id"] = np.arange(1, n + 1)
    for i in inactive:
        b["id"][i] *= -1
    nbt = np.bincount(np.asarray(types), minlength=7).astype(np.uint32)
    return r, v, p, b, nbt


def mixed_types(n):
    """Every body class: star, giants, rocky, protoplanets, super-pl, pl, test particles."""
    k = max(1, n // 10)
    t = [0] + [1] * 2 + [2] * k + [3] * k + [4] * k + [5] * (n - 3 - 4 * k) + [6] * k
    return t[:n]


def bits_equal(a, b):
    """Bit-for-bit equality of every field (padding bytes are not data)."""
    a, b = np.asarray(a), np.asarray(b)
    if a.shape != b.shape:
        return False
    if a.dtype.names:
        if a.dtype.names != b.dtype.names:
            return False
        return all(bits_equal(a[f], b[f]) for f in a.dtype.names)
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    return a.dtype.itemsize == b.dtype.itemsize and np.array_equal(a.view(np.uint8), b.view(np.uint8))


def as3(a):
    return np.stack([a["x"], a["y"], a["z"]], 1)


def fargo_grid(n_rad, n_sec, seed, r_min=0.4, r_max=2.5, rho1=2e-3, vr=1e-3, vt=3e-3):
    """A frame-0 gas grid shaped like a FARGO run: log-spaced ring interfaces
    (n_rad + 1 radii, FARGO's logarithmic RadialSpacing) and (n_rad, n_sec)
    cells of midplane density ~ rho1 r^-2.75 [Msun/AU^3; 2e-3 is a minimum-mass
    nebula's at 1 AU], radial velocity and
    sub-Keplerian azimuthal offset [AU/tau], each with seeded cell noise."""
    rng = np.random.default_rng(seed)
    used_rad = r_min * (r_max / r_min) ** (np.arange(n_rad + 1) / n_rad)
    rc = np.sqrt(used_rad[:-1] * used_rad[1:])[:, None]
    dens = rho1 * rc ** -2.75 * rng.uniform(0.5, 1.5, (n_rad, n_sec))
    vrad = vr * rng.uniform(-1.0, 1.0, (n_rad, n_sec))
    vtheta = -vt * rc ** -0.5 * rng.uniform(0.5, 1.5, (n_rad, n_sec))
    return used_rad, dens, vrad, vtheta
