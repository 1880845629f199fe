"""FP64 NumPy reference of the learned warm-start MLP (24-64-64-64-3, ReLU)
with the same BF16 rounding points as the GPU kernels: inputs, weights,
hidden activations, the output gradient and hidden gradients are rounded to
BF16 (round to nearest even); all sums are FP64."""
import numpy as np

SHAPES = [(24, 64), (64, 64), (64, 64), (64, 3)]


def bf16(a):
    a = np.asarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unflatten(p):
    out, q = [], 0
    for fi, fo in SHAPES:
        W = p[q:q + fi * fo].reshape(fi, fo)
        q += fi * fo
        b = p[q:q + fo]
        q += fo
        out.append((W.astype(np.float64), b.astype(np.float64)))
    return out


def flatten(layers):
    return np.concatenate([np.concatenate([W.reshape(-1), b]) for W, b in layers])


def forward(p, X, round_bf16=True):
    rd = bf16 if round_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    L = unflatten(p)
    h = rd(X)
    acts = [h]
    for i, (W, b) in enumerate(L):
        z = h @ rd(W) + b
        if i < 3:
            h = rd(np.maximum(z, 0.0))
            acts.append(h)
        else:
            h = z
    return h, acts


def gradients(p, X, T, round_bf16=True):
    """d(mean over rows of |y - t|^2)/d(params), flat layout, and the loss."""
    rd = bf16 if round_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    L = unflatten(p)
    n = len(X)
    y, acts = forward(p, X, round_bf16)
    e = y - T
    loss = float(np.sum(e * e) / n)
    dy = rd(2.0 * e / n)
    out, d = [None] * 4, dy
    out[3] = (acts[3].T @ dy, (2.0 * e / n).sum(0))
    for i in range(3, 0, -1):
        W, _ = L[i]
        dh = np.where(acts[i] > 0, d @ rd(W).T, 0.0)
        gb = dh.sum(0)
        d = rd(dh)
        out[i - 1] = (acts[i - 1].T @ d, gb)
    return flatten(out), loss
