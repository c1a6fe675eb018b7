"""Host reference of the learned-ordering MLP (test infrastructure: only
tests use it, as the checker of the GPU kernels' gradients).

Same parameter layout as blazingsql_amd/csrc/learned.h: w1 (in x hidden),
b1, w2 (hidden x hidden), b2, w3 (hidden x 2), b3, weights row-major by
input; ReLU after layers 1 and 2; loss = mean over the batch of binary
cross-entropy of output 0 (logit) against target 0 plus cost_weight times
the squared error of output 1 against target 1. Computed in float64.
"""
import numpy as np


def split(params, n_in=24, hidden=64, n_out=2):
    p = np.asarray(params, dtype=np.float64)
    o = 0
    shapes = [(n_in, hidden), (hidden,), (hidden, hidden), (hidden,), (hidden, n_out), (n_out,)]
    out = []
    for s in shapes:
        k = int(np.prod(s))
        out.append(p[o:o + k].reshape(s))
        o += k
    assert o == p.size
    return out


def forward(params, x):
    w1, b1, w2, b2, w3, b3 = split(params)
    x = np.asarray(x, dtype=np.float64)
    h1 = np.maximum(x @ w1 + b1, 0.0)
    h2 = np.maximum(h1 @ w2 + b2, 0.0)
    return h1, h2, h2 @ w3 + b3


def loss(params, x, target, cost_weight=1.0):
    _, _, y = forward(params, x)
    t = np.asarray(target, dtype=np.float64)
    z = y[:, 0]
    bce = np.maximum(z, 0) - z * t[:, 0] + np.log1p(np.exp(-np.abs(z)))
    d = y[:, 1] - t[:, 1]
    return float(np.mean(bce + cost_weight * d * d))


def gradients(params, x, target, cost_weight=1.0):
    """d loss / d params, flattened in the parameter layout."""
    w1, b1, w2, b2, w3, b3 = split(params)
    x = np.asarray(x, dtype=np.float64)
    t = np.asarray(target, dtype=np.float64)
    h1, h2, y = forward(params, x)
    n = x.shape[0]
    dy = np.zeros_like(y)
    dy[:, 0] = (1.0 / (1.0 + np.exp(-y[:, 0])) - t[:, 0]) / n
    dy[:, 1] = cost_weight * 2.0 * (y[:, 1] - t[:, 1]) / n
    dw3, db3 = h2.T @ dy, dy.sum(0)
    dh2 = (dy @ w3.T) * (h2 > 0)
    dw2, db2 = h1.T @ dh2, dh2.sum(0)
    dh1 = (dh2 @ w2.T) * (h1 > 0)
    dw1, db1 = x.T @ dh1, dh1.sum(0)
    return np.concatenate([a.ravel() for a in (dw1, db1, dw2, db2, dw3, db3)])


def sketch(values, valid, rows, sample=1024, bins=16):
    """Column sketch over `sample` evenly spaced rows: null fraction, min,
    max, histogram of the valid sample over [min, max] (fractions)."""
    idx = (np.arange(sample, dtype=np.int64) * rows) // sample if rows else np.zeros(0, dtype=np.int64)
    v = np.asarray(values, dtype=np.float64)[idx] if rows else np.zeros(0)
    ok = np.ones(len(idx), dtype=bool) if valid is None else np.asarray(valid, dtype=bool)[idx]
    out = np.zeros(3 + bins)
    if rows == 0:
        return out
    out[0] = 1.0 - ok.sum() / sample
    vv = v[ok]
    if len(vv) == 0:
        return out
    mn, mx = vv.min(), vv.max()
    out[1], out[2] = mn, mx
    span = mx - mn
    b = np.zeros(len(vv), dtype=np.int64) if span <= 0 else np.clip(((vv - mn) / span * bins).astype(np.int64), 0, bins - 1)
    out[3:] = np.bincount(b, minlength=bins) / len(vv)
    return out
