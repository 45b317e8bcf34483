"""Learned step proposer workflow: log (features -> step) samples from the
engine's own RKF7(8) runs, fit the 12 -> 64 -> 64 -> 1 BF16 MLP on the GPU
with the library's forward/backward/Adam kernels, and run the ablation
(classical controller vs the proposer) on held-out disks.

The reference controller (int_rungekutta8.cu:690-706) still judges every
attempt: the proposer only replaces the first try of the next step, clamped
to [h_c / 8, 8 h_c] around the classical proposal h_c, and a rejected
attempt falls back to the classical shrink.

Features (one row per accepted step, computed from the accepted state):
  0 log h_prev            1 log error ratio of that step   2 0.5 log min r^2/v^2 over pairs
  3 log min periapsis     4 log N                          5 log10 relative Keplerian energy drift
  6 log h two steps back  7 log error ratio two steps back 8 rejections of the last step
  9 log classical h_next 10 log tolerance                 11 change of feature 2 since the last step
Target: log(h * (1 / err_ratio)^(1/8)), the step that would just have met the
tolerance.  By default the network is fitted to a low quantile of that target
taken as an offset from feature 9 (the classical proposal h_c): its output is
then a step that about a fraction q of steps would fail, found from the
features, and it runs with safety 1.  The mean fit (quantile 0) with a safety
factor below 1 is the other mode."""
import time

import numpy as np

import redamd as R


def batch_sizes(n_disk=512, lo=64, hi=4096, seed=4):
    """N_d uniform on [lo, hi] from MT19937(seed) raw outputs, as red_random::uniform(int, int)
    (src/red.cuda.initial/distribution.cpp:43-46)."""
    rs = np.random.RandomState(seed)
    out = []
    for _ in range(n_disk):
        u = int.from_bytes(rs.bytes(4), "little")
        out.append(lo + int(u / 4294967296.0 * (hi - lo + 1)))
    return out


def make_batch(sizes, base_seed, precision=R.RD_PREC_MIXED, gas=False, device=0, use_graph=1):
    """One engine holding len(sizes) disks, disk k seeded base_seed + k. gas:
    the config-3 physics (radii 1-100 km, gas drag on every massive body,
    merging with radius factor 1, hit-centrum 0.1 AU)."""
    if gas:
        parts = [R.make_disk(n, base_seed + k, size_mode=1, lo=1.0, hi=100.0, cd=1.0) for k, n in enumerate(sizes)]
    else:
        parts = [R.make_disk(n, base_seed + k) for k, n in enumerate(sizes)]
    cat = [np.concatenate([q[i] for q in parts]) for i in range(4)]
    nbt = np.stack([q[4] for q in parts])
    if gas:
        d = R.Disk(*cat, nbt, batch=True, threshold=(0.1, 0.0, 1.0), gas=R.gas_params(), device=device)
        d.set_drag_mode(R.RD_DRAG_ALL_MASSIVE)
    else:
        d = R.Disk(*cat, nbt, batch=True, device=device)
    d.set_precision(precision)
    d.set_tuning(use_graph=use_graph)
    return d


def config3_disk(n=131072, seed=3, precision=R.RD_PREC_MIXED, device=0, use_graph=1):
    """The config-3 disk (gas, merging, hit-centrum) and its initial step
    (min orbital period / 1000, src/red.cuda.initial/main.cu:317-342)."""
    r, v, p, b, nbt = R.make_disk(n, seed, size_mode=1, lo=1.0, hi=100.0, cd=1.0)
    d = R.Disk(r, v, p, b, nbt, threshold=(0.1, 0.0, 1.0), gas=R.gas_params(), device=device)
    d.set_drag_mode(R.RD_DRAG_ALL_MASSIVE)
    d.set_precision(precision)
    d.set_tuning(use_graph=use_graph)
    a_min = np.min(np.hypot(r["x"][1:], r["y"][1:]))
    return d, 2 * np.pi * np.sqrt(a_min ** 3) / 1000.0


def collect(disk, steps, dt0=0.1 * R.K_GAUSS, tol=1e-10):
    """Run `steps` accepted RKF7(8) steps on every disk of `disk` with sample
    logging on; returns (X [n, 12] float32, T [n] float32)."""
    it = R.Integrator(disk, R.RD_RKF78, dt0, adaptive=True, tol=tol)
    it.set_logging(True)
    it.steps(steps)
    X, T = it.log()
    it.close()
    return X, T


BASE, QUANTILE, SAFETY = 9, 0.05, 1.0   # the default fit: 5% quantile of log(h*/h_c), used as is


def fit(X, T, seed=5, epochs=15, batch=512, lr=1e-3, device=0, base=BASE, quantile=QUANTILE):
    """A proposer fitted to the samples on the GPU; returns (Mlp, final loss).
    base >= 0: the network fits T - X[:, base] and adds X[:, base] back;
    quantile 0 fits the mean, q in (0, 1) the q-quantile."""
    m = R.Mlp(seed, device=device)
    t = T.astype(np.float64) - (X[:, base] if base >= 0 else 0.0)
    m.set_norm(X.mean(0), X.std(0) + 1e-3, float(t.mean()), float(t.std()) + 1e-6)
    m.set_target(base, quantile)
    loss = m.train(X, T, epochs=epochs, batch=batch, lr=lr)
    return m, loss


def ceilings(disk, length, dt0, tol=1e-10):
    """What a step proposer could at best save on `disk` over `length` of
    simulated time against the classical controller, from one classical run
    with sample logging on (the logging does not change the steps).
    Returns the classical run's counters and two ratios of its RK evaluations
    (12 per attempt + 1 per step, int_rungekutta8.cu:648-656) over:
      rejection_free: the same accepted steps with no rejected attempt,
        (13 A + 12 R) / 13 A;
      exact_step: every step taken at h*, the step its own Fehlberg estimate
        says just meets the tolerance (error ~ h^8), with no rejections --
        13 h_i / h*_i per logged step, where h_i is the first try h_c when it
        passed and the reference retry 0.9 h_c (h*/h_c)^(8/7) when it did not;
        steps without a logged sample (each disk's first) count 13.
    The exact_step figure is an estimate from the h^8 law, not a bound."""
    it = R.Integrator(disk, R.RD_RKF78, dt0, adaptive=True, tol=tol)
    it.set_logging(True)
    it.run(length, 1e30)
    X, T = it.log()
    nd = disk.n_disk()
    c = np.array([it.counters(k) for k in range(nd)], np.int64).sum(0)
    it.close()
    tried, passed = int(c[0]), int(c[1])
    evals = 12 * tried + passed
    r = X[:, 9].astype(np.float64) - T.astype(np.float64)        # log(h_c / h*)
    frac = np.where(r <= 0.0, np.exp(r), 0.9 * np.exp(-r / 7.0))
    exact = 13.0 * float(frac.sum()) + 13.0 * (passed - len(T))
    return {"tried": tried, "passed": passed, "failed": int(c[2]), "rk_evals": evals,
            "rejection_free": evals / (13.0 * passed), "exact_step": evals / exact}


def run(disk, steps=None, length=None, proposer=None, safety=SAFETY, dt0=0.1 * R.K_GAUSS, tol=1e-10,
        energy_disks=None):
    """Integrate `disk` for `steps` accepted steps per disk (or `length` of time),
    with or without the proposer. Returns a dict of work counters, wall time and
    the largest relative energy error over `energy_disks` (default: all)."""
    it = R.Integrator(disk, R.RD_RKF78, dt0, adaptive=True, tol=tol)
    if proposer is not None:
        it.set_proposer(proposer, safety)
    nd = disk.n_disk()
    ks = list(range(nd)) if energy_disks is None else list(energy_disks)
    E0 = [disk.integral(k)["E"] for k in ks]
    disk.sync()
    t0 = time.perf_counter()
    if steps is not None:
        it.steps(steps)
    else:
        it.run(length, 1e30)
    disk.sync()
    el = time.perf_counter() - t0
    c = np.array([it.counters(k) for k in range(nd)], np.int64).sum(0)
    evals, pairs = it.work()
    dE = max(abs(disk.integral(k)["E"] / e - 1.0) for k, e in zip(ks, E0))
    it.close()
    return {"seconds": el, "force_evals": int(evals), "pairs": float(pairs), "tried": int(c[0]),
            "passed": int(c[1]), "failed": int(c[2]), "rel_energy_error": float(dE)}
