"""SGNS logistic-loss oracle (FP64, numpy) — TEST INFRASTRUCTURE ONLY.

The reference has no separate loss or gradient function: its update
(include/dsgl.hpp:1241-1368) is -alpha times the gradient of
    L = sum_minibatch sum_context [ softplus(-s1[o0].s0[c]) + sum_{k>=1} softplus(s1[ok].s0[c]) ]
with a table sigmoid.  These functions state that loss and its exact
gradient in float64 for the minibatch layout of the gradient-check workload
(contexts x (1 target + K shared negatives)).  Only tests/ may import this.

Pins: central finite differences of `loss` itself (h = 1e-3, float64),
closed forms softplus(0) = log 2 and d softplus/dx (0) = 1/2 on a single
pair, and an explicit per-pair Python loop on tiny inputs.
"""
import numpy as np


def softplus(x):
    return np.maximum(x, 0) + np.log1p(np.exp(-np.abs(x)))


def sigmoid(x):
    return 0.5 * (1.0 + np.tanh(0.5 * x))


def loss(syn0, syn1, ctx, out):
    """ctx: [n_mb, n_ctx] rows of syn0; out: [n_mb, n_out] rows of syn1, out[:, 0] positive."""
    x = syn0[ctx]                      # [m, c, d]
    y = syn1[out]                      # [m, o, d]
    f = np.einsum("mcd,mod->mco", x, y)
    sign = np.ones(out.shape[1])
    sign[0] = -1.0
    return float(softplus(f * sign[None, None, :]).sum())


def grad(syn0, syn1, ctx, out):
    x = syn0[ctx]
    y = syn1[out]
    f = np.einsum("mcd,mod->mco", x, y)
    lab = np.zeros(out.shape[1])
    lab[0] = 1.0
    coef = sigmoid(f) - lab[None, None, :]          # dL/df
    g0 = np.zeros_like(syn0)
    g1 = np.zeros_like(syn1)
    np.add.at(g0, ctx, np.einsum("mco,mod->mcd", coef, y))
    np.add.at(g1, out, np.einsum("mco,mcd->mod", coef, x))
    return g0, g1


def gradcheck_workload(rows=1000, dim=128, n_mb=256, n_ctx=21, n_neg=5, seed=0, dtype=np.float64):
    """The gradient-check workload: contexts uniform over rows, negatives
    Zipf(1.0) over rows, syn0 ~ U(-0.5/dim, 0.5/dim), syn1neg ~ N(0, 0.1)."""
    rng = np.random.default_rng(seed)
    syn0 = rng.uniform(-0.5 / dim, 0.5 / dim, (rows, dim)).astype(dtype)
    syn1 = rng.normal(0.0, 0.1, (rows, dim)).astype(dtype)
    ctx = rng.integers(0, rows, (n_mb, n_ctx)).astype(np.int32)
    # Zipf(s=1) over ranks 1..rows: p(k) ~ 1/k
    p = 1.0 / np.arange(1, rows + 1)
    p /= p.sum()
    neg = rng.choice(rows, (n_mb, n_neg), p=p).astype(np.int32)
    tgt = rng.integers(0, rows, (n_mb, 1)).astype(np.int32)
    out = np.concatenate([tgt, neg], axis=1)
    return syn0, syn1, ctx, out
