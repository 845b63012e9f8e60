"""The forward-loss and adjoint kernels (gradient-check workload: 1000 rows x
dim 128, 256 minibatches of 21 contexts x (1 target + 5 Zipf negatives))
against the FP64 oracle and central finite differences, through the C ABI,
plus a test-only torch.autograd.Function wrapping them."""
import numpy as np
import pytest

import dger_sgns as G
from oracle import loss_oracle as L

pytestmark = pytest.mark.gpu


def staged(dtype):
    s0, s1, ctx, out = L.gradcheck_workload(dtype=dtype)
    return (s0, s1, ctx, out, G.DeviceArray(s0), G.DeviceArray(s1), G.DeviceArray(ctx),
            G.DeviceArray(out))


def test_loss_and_gradient_fp64():
    s0, s1, ctx, out, d0, d1, dc, do = staged(np.float64)
    ref = L.loss(s0, s1, ctx, out)
    got = G.loss_fwd(d0, d1, dc, do, 256, 21, 6)
    assert abs(got - ref) <= 1e-12 * abs(ref)
    g0, g1 = G.loss_bwd(d0, d1, dc, do, 256, 21, 6)
    r0, r1 = L.grad(s0, s1, ctx, out)
    assert np.allclose(g0, r0, rtol=1e-10, atol=1e-14)
    assert np.allclose(g1, r1, rtol=1e-10, atol=1e-14)


def test_gradient_central_differences_fp64():
    # central differences (h = 1e-3, FP64) of the GPU forward kernel against
    # the GPU adjoint kernel
    s0, s1, ctx, out, d0, d1, dc, do = staged(np.float64)
    g0, g1 = G.loss_bwd(d0, d1, dc, do, 256, 21, 6)
    rng = np.random.default_rng(7)
    h = 1e-3
    for which, mat, g in ((0, s0, g0), (1, s1, g1)):
        for _ in range(12):
            r = int(ctx[rng.integers(0, 256), 0]) if which == 0 else int(out[rng.integers(0, 256), 0])
            d = int(rng.integers(0, 128))
            vals = []
            for sgn in (1, -1):
                m = mat.copy()
                m[r, d] += sgn * h
                dm = G.DeviceArray(m)
                vals.append(G.loss_fwd(dm if which == 0 else d0, d1 if which == 0 else dm, dc, do,
                                       256, 21, 6))
                dm.free()
            fd = (vals[0] - vals[1]) / (2 * h)
            assert abs(fd - g[r, d]) <= 1e-6 * max(1.0, abs(g[r, d])), (which, r, d, fd, g[r, d])


def test_loss_and_gradient_fp32_vs_fp64():
    s0, s1, ctx, out, d0, d1, dc, do = staged(np.float32)
    ref = L.loss(s0.astype(np.float64), s1.astype(np.float64), ctx, out)
    got = G.loss_fwd(d0, d1, dc, do, 256, 21, 6)
    assert abs(got - ref) <= 1e-5 * abs(ref)
    g0, g1 = G.loss_bwd(d0, d1, dc, do, 256, 21, 6)
    r0, r1 = L.grad(s0.astype(np.float64), s1.astype(np.float64), ctx, out)
    for g, r in ((g0, r0), (g1, r1)):
        assert np.max(np.abs(g - r)) <= 1e-5 * np.max(np.abs(r))
