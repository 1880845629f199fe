"""Pins the FP64 MLP reference used to check the GPU kernels: with BF16
rounding switched off the loss is smooth, so every layer's bias gradient and
sampled weight gradients are checked against central finite differences; the
forward pass is checked against an explicit per-row evaluation (CPU only)."""
import numpy as np

from tests import mlp_ref


def test_bias_gradients_finite_difference():
    rng = np.random.default_rng(0)
    p = (rng.normal(size=10115) * 0.3).astype(np.float32).astype(np.float64)
    X = rng.normal(size=(40, 24)).astype(np.float32)
    T = rng.normal(size=(40, 3))
    g, _ = mlp_ref.gradients(p, X, T, round_bf16=False)
    offsets, q = [], 0
    for fi, fo in mlp_ref.SHAPES:
        q += fi * fo
        offsets.append((q, fo))
        q += fo
    # sampled weight entries of every layer too
    q = 0
    for fi, fo in mlp_ref.SHAPES:
        for i in rng.choice(fi * fo, size=3, replace=False) + q:
            e = np.zeros_like(p)
            e[i] = 1e-5
            lp = np.sum((mlp_ref.forward(p + e, X, False)[0] - T) ** 2) / len(X)
            lm = np.sum((mlp_ref.forward(p - e, X, False)[0] - T) ** 2) / len(X)
            fd = (lp - lm) / 2e-5
            assert abs(fd - g[i]) <= 1e-4 * max(1.0, abs(fd)), (i, fd, g[i])
        q += fi * fo + fo
    for start, fo in offsets:
        for j in rng.choice(fo, size=min(fo, 3), replace=False):
            i = start + j
            e = np.zeros_like(p)
            e[i] = 1e-5
            lp = np.sum((mlp_ref.forward(p + e, X, False)[0] - T) ** 2) / len(X)
            lm = np.sum((mlp_ref.forward(p - e, X, False)[0] - T) ** 2) / len(X)
            fd = (lp - lm) / 2e-5
            assert abs(fd - g[i]) <= 1e-4 * max(1.0, abs(fd)), (i, fd, g[i])


def test_forward_per_row():
    rng = np.random.default_rng(1)
    p = rng.normal(size=10115) * 0.2
    X = rng.normal(size=(5, 24))
    Y, _ = mlp_ref.forward(p, X)
    L = mlp_ref.unflatten(p)
    for r in range(5):
        h = mlp_ref.bf16(X[r])
        for i, (W, b) in enumerate(L):
            z = np.array([sum(h[k] * mlp_ref.bf16(W[k, j]) for k in range(W.shape[0])) for j in range(W.shape[1])]) + b
            h = mlp_ref.bf16(np.maximum(z, 0)) if i < 3 else z
        np.testing.assert_allclose(Y[r], h, rtol=1e-12, atol=1e-12)
