"""Pins of the FP64 loss / gradient oracle (oracle/loss_oracle.py)."""
import numpy as np

from oracle import loss_oracle as L


def test_single_pair_closed_forms():
    s0 = np.zeros((1, 4))
    s1 = np.zeros((1, 4))
    ctx = np.zeros((1, 1), np.int32)
    out = np.zeros((1, 1), np.int32)
    assert abs(L.loss(s0, s1, ctx, out) - np.log(2.0)) < 1e-15
    s0 = np.array([[1.0, 2.0, 0.0, -1.0]])
    s1 = np.array([[0.5, -0.25, 2.0, 1.0]])
    f = float(s0[0] @ s1[0])          # -1.0
    g0, g1 = L.grad(s0, s1, ctx, out)
    # positive pair: d/df softplus(-f) = sigmoid(f) - 1
    c = 1.0 / (1.0 + np.exp(-f)) - 1.0
    assert np.allclose(g0[0], c * s1[0], rtol=1e-14) and np.allclose(g1[0], c * s0[0], rtol=1e-14)


def test_against_explicit_loops():
    rng = np.random.default_rng(1)
    s0 = rng.normal(0, 0.5, (6, 3))
    s1 = rng.normal(0, 0.5, (5, 3))
    ctx = rng.integers(0, 6, (4, 3)).astype(np.int32)
    out = rng.integers(0, 5, (4, 3)).astype(np.int32)
    tot = 0.0
    g0 = np.zeros_like(s0)
    g1 = np.zeros_like(s1)
    for m in range(4):
        for c in ctx[m]:
            for k, o in enumerate(out[m]):
                f = float(s0[c] @ s1[o])
                y = 1.0 if k == 0 else 0.0
                tot += np.log1p(np.exp(-f)) if k == 0 else np.log1p(np.exp(f))
                sg = 1.0 / (1.0 + np.exp(-f))
                g0[c] += (sg - y) * s1[o]
                g1[o] += (sg - y) * s0[c]
    assert abs(L.loss(s0, s1, ctx, out) - tot) < 1e-12
    a, b = L.grad(s0, s1, ctx, out)
    assert np.allclose(a, g0, rtol=1e-12, atol=1e-14) and np.allclose(b, g1, rtol=1e-12, atol=1e-14)


def test_gradient_matches_central_differences():
    s0, s1, ctx, out = L.gradcheck_workload(rows=40, dim=8, n_mb=16, seed=3)
    g0, g1 = L.grad(s0, s1, ctx, out)
    rng = np.random.default_rng(4)
    h = 1e-3
    for mat, g in ((s0, g0), (s1, g1)):
        for _ in range(25):
            r, d = rng.integers(0, mat.shape[0]), rng.integers(0, mat.shape[1])
            old = mat[r, d]
            mat[r, d] = old + h
            lp = L.loss(s0, s1, ctx, out)
            mat[r, d] = old - h
            lm = L.loss(s0, s1, ctx, out)
            mat[r, d] = old
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[r, d]) <= 1e-6 * max(1.0, abs(g[r, d]))
