"""CPU-side checks of the C ABI library: every declared symbol is exported,
input validation errors, the host-only disk generator against the oracle,
and the golden config-1 fixture against the oracle.  No GPU compute."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import oracle_lib as O
import redamd as R
from systems import bits_equal

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "red_abi.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(rd_[a-z0-9_]+)\s*\(", src)))


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) >= 40
    out = subprocess.run(["nm", "-D", "--defined-only", R.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (rd_\w+)", out))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    for n in names:
        getattr(R.lib, n)
    assert R.lib.rd_version() == R.ABI_VERSION == 3


def test_binding_covers_the_header():
    assert set(declared_functions()) == set(R.EXPORTED)


def test_validation_errors():
    r, v, p, b, nbt = R.make_disk(16, 3)
    h = C.c_void_p()
    bad = nbt.copy()
    bad[3] -= 1
    bad[5] += 1          # counts do not match the body types
    assert R.lib.rd_disk_create(17, R._ptr(bad), R._ptr(r), R._ptr(v), R._ptr(p), R._ptr(b), 0.0, None, None, 0,
                                C.byref(h)) == R.RD_EINVAL
    assert "does not match the body types" in R.lib.rd_last_error().decode()
    # a planetesimal (NSI) among the protoplanets (SI): the classes must be contiguous
    b3 = b.copy()
    b3["body_type"][5] = 5
    nbt3 = nbt.copy()
    nbt3[3] -= 1
    nbt3[5] += 1
    assert R.lib.rd_disk_create(17, R._ptr(nbt3), R._ptr(r), R._ptr(v), R._ptr(p), R._ptr(b3), 0.0, None, None, 0,
                                C.byref(h)) == R.RD_EINVAL
    assert "grouped by interaction class" in R.lib.rd_last_error().decode()
    b2 = b.copy()
    b2["id"][4] = 0
    assert R.lib.rd_disk_create(17, R._ptr(nbt), R._ptr(r), R._ptr(v), R._ptr(p), R._ptr(b2), 0.0, None, None, 0,
                                C.byref(h)) == R.RD_EINVAL
    assert R.lib.rd_disk_create(16, R._ptr(nbt), R._ptr(r), R._ptr(v), R._ptr(p), R._ptr(b), 0.0, None, None, 0,
                                C.byref(h)) == R.RD_EINVAL
    with pytest.raises(R.RedError):
        R.check(R.lib.rd_integrator_create(None, 4, 1.0, 1, 1e-10, C.byref(h)))
    # within the SI class any order passes validation (the reference's own
    # solar-system generator writes rocky planets before giants)
    b4 = b.copy()
    b4["body_type"][1:3] = [2, 1]
    nbt4 = nbt.copy()
    nbt4[3] -= 2
    nbt4[1] += 1
    nbt4[2] += 1
    rc = R.lib.rd_disk_create(17, R._ptr(nbt4), R._ptr(r), R._ptr(v), R._ptr(p), R._ptr(b4), 0.0, None, None, 0,
                              C.byref(h))
    assert rc != R.RD_EINVAL, R.lib.rd_last_error().decode()
    if rc == R.RD_OK:
        R.lib.rd_disk_destroy(h)
    # a batch holding an empty disk is refused before any device call
    nbt2 = np.stack([nbt, np.zeros_like(nbt)])
    assert R.lib.rd_batch_create(2, R._ptr(nbt2), R._ptr(r), R._ptr(v), R._ptr(p), R._ptr(b), None, None, None, 0,
                                 C.byref(h)) == R.RD_EINVAL
    assert "at least one body" in R.lib.rd_last_error().decode()


@pytest.mark.parametrize("kw", [dict(n=256, seed=1), dict(n=1000, seed=2),
                                dict(n=300, seed=3, size_mode=1, lo=1.0, hi=100.0, cd=1.0, a_power=-1.5)])
def test_disk_generator_matches_oracle(kw):
    n, seed = kw.pop("n"), kw.pop("seed")
    a = R.make_disk(n, seed, **kw)
    b = O.make_disk(n, seed, **kw)
    for x, y in zip(a, b):
        assert bits_equal(x, y)


def test_disk_generator_distribution():
    r, v, p, b, nbt = R.make_disk(20000, 9)
    rr = np.hypot(r["x"][1:], r["y"][1:])
    assert 0.45 < rr.min() and rr.max() < 5.3
    # dN/da ~ a^-0.5: half the bodies inside a_med = ((sqrt(5)+sqrt(0.5))/2)^2
    a_med = ((np.sqrt(5) + np.sqrt(0.5)) / 2) ** 2
    assert abs(np.mean(rr < a_med) - 0.5) < 0.02
    lm = np.log10(p["mass"][1:])
    assert lm.min() >= -10 and lm.max() <= -8 and abs(lm.mean() + 9) < 0.02
    assert np.all(b["id"] == np.arange(1, 20002)) and b["body_type"][0] == 0


def test_golden_config1_reproduced_by_oracle_prefix():
    """The committed fixture is the oracle's own output: its first 200 steps
    reproduce bit-for-bit from the fixture's initial state (the full 10^4-step
    run is checked on the GPU against the fixture)."""
    g = np.load(os.path.join(ROOT, "tests", "golden", "config1_rk4.npz"))
    r, v, p, b, nbt = O.make_disk(256, 1)
    assert np.array_equal(np.stack([r["x"], r["y"], r["z"]], 1), g["r0"])
    assert np.array_equal(np.stack([v["x"], v["y"], v["z"]], 1), g["v0"])
    assert g["t"] == pytest.approx(10000 * 0.1 * 0.01720209895, rel=1e-12)


def test_binding_refuses_to_load_without_the_native_library(tmp_path):
    """No CPU fallback: the binding, copied next to no librdamd.so, raises at
    import instead of running anything."""
    import shutil
    import sys
    src = os.path.join(ROOT, "red.cuda_amd")
    for f in ("redamd.py", "redamd_types.py"):
        shutil.copy(os.path.join(src, f), tmp_path / f)
    res = subprocess.run([sys.executable, "-c", "import redamd"], cwd=str(tmp_path), capture_output=True, text=True,
                         timeout=120, env={**os.environ, "PYTHONPATH": str(tmp_path)})
    assert res.returncode != 0
    assert "ImportError" in res.stderr and "is not built" in res.stderr
