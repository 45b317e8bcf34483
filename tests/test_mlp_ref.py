"""The numpy restatement of the proposer MLP, pinned by central finite
differences (a dropped term, wrong sign or transposed weight fails) and by
the BF16 rounding rule."""
import numpy as np

import mlp_ref as M


def test_bf16_round_to_nearest_even():
    vals = np.array([1.0, 1.00390625, 1.0078125, 1.01171875, -3.14159265, 65504.0, 1e-30], np.float32)
    r = M.bf16(vals)
    assert r[0] == 1.0
    assert r[1] == 1.0                   # tie 1 + 2^-8 -> even mantissa (1.0)
    assert r[2] == 1.0078125
    assert r[3] == 1.015625              # tie -> even
    assert abs(r[4] - (-3.140625)) < 1e-7
    assert np.all(M.bf16(r) == r)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    w = rng.normal(0, 0.3, M.NPARAM)
    x = rng.normal(0, 1, (40, M.IN))
    t = rng.normal(0, 1, 40)
    mean, std = np.zeros(M.IN), np.ones(M.IN)
    # gradient at exactly representable BF16 weights, so rounding is the identity
    wb = M.bf16(w).astype(np.float64)
    _, g = M.loss_grad(wb, x, t, mean, std, 0.0, 1.0)
    eps = 1e-6
    for k in list(rng.choice(M.NPARAM, 60, replace=False)) + [M.NPARAM - 1, 0, 768, 832]:
        wp, wm = wb.copy(), wb.copy()
        wp[k] += eps
        wm[k] -= eps
        fd = (M.loss_only(wp, x, t, mean, std, 0.0, 1.0) - M.loss_only(wm, x, t, mean, std, 0.0, 1.0)) / (2 * eps)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])


def test_quantile_offset_gradient_matches_finite_differences():
    """Pinball loss on targets taken as offsets from feature 9: central
    differences again, and the output-bias gradient equals the closed form
    P(y' > t') - q (zero exactly at the q-quantile)."""
    rng = np.random.default_rng(1)
    w = rng.normal(0, 0.3, M.NPARAM)
    x = rng.normal(0, 1, (50, M.IN))
    t = rng.normal(0, 1, 50) + x[:, 9]
    mean, std = rng.normal(0, 0.1, M.IN), 1.0 + rng.random(M.IN)
    q, tm, ts = 0.1, 0.2, 1.5
    wb = M.bf16(w).astype(np.float64)
    loss, g = M.loss_grad(wb, x, t, mean, std, tm, ts, base=9, quantile=q)
    eps = 1e-7
    for k in list(rng.choice(M.NPARAM, 40, replace=False)) + [M.NPARAM - 1, 0, 768, 832]:
        wp, wm = wb.copy(), wb.copy()
        wp[k] += eps
        wm[k] -= eps
        fd = (M.loss_only(wp, x, t, mean, std, tm, ts, 9, q) - M.loss_only(wm, x, t, mean, std, tm, ts, 9, q)) / (2 * eps)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])
    y = (M.forward(wb, x, mean, std, tm, ts, base=9) - x[:, 9] - tm) / ts      # y' in standardised units
    tn = (t - x[:, 9] - tm) / ts
    assert abs(g[M.NPARAM - 1] - (np.mean(y > tn) - q)) < 1e-12
    u = tn - y
    assert abs(loss - np.mean(np.maximum(q * u, (q - 1) * u))) < 1e-12
    # the offset: forward with base 9 is forward without it plus feature 9
    assert np.allclose(M.forward(wb, x, mean, std, tm, ts, base=9) - M.forward(wb, x, mean, std, tm, ts), x[:, 9],
                       rtol=0, atol=1e-12)
