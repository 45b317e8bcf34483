"""float64 numpy restatement of the step-proposer MLP (test infrastructure):
12 -> 64 -> 64 -> 1, tanh hidden layers, weights rounded to BF16 (round to
nearest even) as the GPU uses them, on standardised inputs
x' = (x - mean)/std and targets t' = (t - x[base] - t_mean)/t_std (no x[base]
term for base -1).  Loss: 0.5 mean((y - t')^2) for quantile 0, else the mean
pinball loss max(q u, (q - 1) u), u = t' - y."""
import numpy as np

IN, H = 12, 64
SIZES = [("W1", IN * H), ("b1", H), ("W2", H * H), ("b2", H), ("w3", H), ("b3", 1)]
NPARAM = sum(n for _, n in SIZES)


def bf16(w):
    u = np.asarray(w, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u >> 16) & 1) + 0x7FFF
    return (((u + r) >> 16) << 16).astype(np.uint32).view(np.float32)


def unpack(w):
    out, k = {}, 0
    for name, n in SIZES:
        out[name] = np.asarray(w[k:k + n], np.float64)
        k += n
    out["W1"] = out["W1"].reshape(IN, H)
    out["W2"] = out["W2"].reshape(H, H)
    out["b3"] = out["b3"][0]
    return out


def _base(x, base):
    return np.asarray(x, np.float64)[:, base] if base >= 0 else 0.0


def forward(w, x, mean, std, t_mean, t_std, base=-1):
    p = unpack(bf16(w))
    xn = (np.asarray(x, np.float64) - mean) / std
    h1 = np.tanh(xn @ p["W1"] + p["b1"])
    h2 = np.tanh(h1 @ p["W2"] + p["b2"])
    return (h2 @ p["w3"] + p["b3"]) * t_std + t_mean + _base(x, base)


def _loss_dy(e, quantile):
    """Mean loss and d loss / d y' per sample."""
    n = len(e)
    if quantile > 0:
        w = np.where(e > 0, 1.0 - quantile, -quantile)
        return np.mean(w * e), w / n
    return 0.5 * np.mean(e * e), e / n


def loss_grad(w, x, t, mean, std, t_mean, t_std, base=-1, quantile=0.0):
    p = unpack(bf16(w))
    xn = (np.asarray(x, np.float64) - mean) / std
    h1 = np.tanh(xn @ p["W1"] + p["b1"])
    h2 = np.tanh(h1 @ p["W2"] + p["b2"])
    y = h2 @ p["w3"] + p["b3"]
    e = y - (np.asarray(t, np.float64) - _base(x, base) - t_mean) / t_std
    loss, dy = _loss_dy(e, quantile)
    gw3 = h2.T @ dy
    gb3 = dy.sum()
    d2 = np.outer(dy, p["w3"]) * (1 - h2 * h2)
    gW2 = h1.T @ d2
    gb2 = d2.sum(0)
    d1 = (d2 @ p["W2"].T) * (1 - h1 * h1)
    gW1 = xn.T @ d1
    gb1 = d1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gw3, [gb3]])
    return loss, g


def loss_only(w_unrounded_f64, x, t, mean, std, t_mean, t_std, base=-1, quantile=0.0):
    """Loss with weights used as given (no BF16 rounding), for finite differences."""
    p = unpack(w_unrounded_f64)
    xn = (np.asarray(x, np.float64) - mean) / std
    h1 = np.tanh(xn @ p["W1"] + p["b1"])
    h2 = np.tanh(h1 @ p["W2"] + p["b2"])
    y = h2 @ p["w3"] + p["b3"]
    e = y - (np.asarray(t, np.float64) - _base(x, base) - t_mean) / t_std
    return _loss_dy(e, quantile)[0]
