"""The reference's retry loop and its NaN semantics on the oracle, and the
opt-in guard that bounds the loop (the CPU half of tests/test_gpu_paths.py).

red.cuda retries a step until err <= tol with no bound
(int_rungekutta8.cu:645-709) and takes the error as the maximum of terms
that compare greater than the running maximum (integrator.cu:248-263), so a
NaN term is never the maximum.  A state gone NaN therefore gives err = 0: the
step is accepted and the next step doubled (int_rungekutta8.cu:705-708)."""
import numpy as np
import pytest

import oracle_lib as O
from systems import small_system


def coincident(n=12, seed=3):
    """Bodies 1 and 2 at the same point: their pull is m/0 * 0 = NaN."""
    r, v, p, b, nbt = small_system(n, seed)
    for c in "xyz":
        r[c][2] = r[c][1]
    return r, v, p, b, nbt


@pytest.mark.parametrize("method", [2, 4])
def test_nan_state_is_accepted_and_the_step_doubles(method):
    r, v, p, b, nbt = coincident()
    d = O.ODisk(r, v, p, b, nbt)
    it = O.OInteg(d, method, 0.01, adaptive=True, tol=1e-10)
    rc, dt = it.step_rc()
    assert rc == 0 and dt == 0.01
    assert it.last_error()[0] == 0.0 and it.dt_next() == 0.02
    assert it.counters() == (1, 1, 0)
    _, rr, vv, _, _ = d.state()
    assert np.isnan(vv["x"][1]) and np.isnan(rr["x"][5])


@pytest.mark.parametrize("method", [2, 4])
def test_guard_refuses_a_nonfinite_estimate(method):
    r, v, p, b, nbt = coincident()
    d = O.ODisk(r, v, p, b, nbt)
    it = O.OInteg(d, method, 0.01, adaptive=True, tol=1e-10)
    it.set_guard(nonfinite=True)
    rc, _ = it.step_rc()
    assert rc == 5                                   # OR_ESTATE
    assert it.counters() == (1, 0, 1)
    t, rr, vv, _, _ = d.state()
    assert t == 0.0 and np.array_equal(rr["x"], r["x"]) and np.array_equal(vv["x"], v["x"])


def first_step(method, guard=None, h0=5.0):
    r, v, p, b, nbt = small_system(8, 4)
    d = O.ODisk(r, v, p, b, nbt)
    it = O.OInteg(d, method, h0, adaptive=True, tol=1e-12)
    if guard:
        it.set_guard(**guard)
    rc, dt = it.step_rc()
    return rc, dt, it, d, r


@pytest.mark.parametrize("method", [2, 4])
def test_guard_bounds_the_attempts(method):
    # an oversized first step takes 3 attempts; a bound of 2 stops it
    rc, dt, it, _, _ = first_step(method)
    assert rc == 0 and it.counters() == (3, 1, 2)
    rc, _, it, d, r = first_step(method, dict(max_attempts=2))
    assert rc == 5 and it.counters() == (2, 0, 2)
    t, rr, _, _, _ = d.state()
    assert t == 0.0 and np.array_equal(rr["y"], r["y"])
    rc, dt3, it, _, _ = first_step(method, dict(max_attempts=3))
    assert rc == 0 and dt3 == dt and it.counters() == (3, 1, 2)


def test_guard_minimum_step():
    # the oversized first step tries dt1 > dt2 > dt3 and accepts dt3: a bound
    # at dt2 stops it after the first rejection, one below dt3 lets it pass
    rc, dt3, _, _, _ = first_step(4)
    rc, _, it, _, _ = first_step(4, dict(max_attempts=1))
    dt2 = abs(it.dt_next())
    assert dt2 > dt3
    rc, _, it, d, _ = first_step(4, dict(min_dt=dt2))
    assert rc == 5 and it.counters() == (1, 0, 1) and d.t == 0.0
    rc, dt, it, _, _ = first_step(4, dict(min_dt=0.5 * dt3))
    assert rc == 0 and dt == dt3 and it.counters() == (3, 1, 2)


@pytest.mark.parametrize("method", [2, 4])
def test_zero_step_stalls_without_the_guard(method):
    """A step that has collapsed to 0 (here set so; an infinite estimate
    gives dt *= 0.9 (tol/inf)^(1/q) = 0) has zero error, is accepted and
    doubles to 0 again: t never moves, so a driver loop over simulated time
    would not end.  The minimum-step guard refuses the attempt."""
    r, v, p, b, nbt = small_system(8, 4)
    d = O.ODisk(r, v, p, b, nbt)
    it = O.OInteg(d, method, 0.0, adaptive=True, tol=1e-12)
    for _ in range(3):
        assert it.step_rc() == (0, 0.0)
    assert d.t == 0.0 and it.dt_next() == 0.0 and it.counters() == (3, 3, 0)
    d = O.ODisk(r, v, p, b, nbt)
    it = O.OInteg(d, method, 0.0, adaptive=True, tol=1e-12)
    it.set_guard(min_dt=1e-30)
    assert it.step_rc()[0] == 5 and it.counters() == (0, 0, 0)


def test_guard_leaves_good_steps_alone():
    # a guard that never trips changes nothing, bit for bit
    r, v, p, b, nbt = small_system(16, 5)
    runs = []
    for guard in (False, True):
        d = O.ODisk(r, v, p, b, nbt)
        it = O.OInteg(d, 4, 0.3, adaptive=True, tol=1e-11)
        if guard:
            it.set_guard(max_attempts=50, min_dt=1e-12, nonfinite=True)
        hs = [it.step() for _ in range(8)]
        _, rr, vv, _, _ = d.state()
        runs.append((hs, it.counters(), rr.tobytes(), vv.tobytes()))
    assert runs[0] == runs[1] and runs[0][1][2] > 0
