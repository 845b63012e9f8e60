"""GEMM1 and the sigmoid gradient of the BF16-MFMA general trainer
(mfma_gemm1, the code its minibatches of at most 8 outputs and 64 contexts
run) on staged rows, through dger_sgns_step_mfma_gemm1.

Reference: both operands rounded to bf16 (nearest even), F = I O^T in
float64 (the products of bf16 values are exact in FP32, so FP32 accumulation
differs from it by a few ulps), then the table gradient in float32 as the
kernel computes it: ((label - table[(int)((F + 6) * 83)]) * alpha, or the
saturated (label - 1) * alpha / label * alpha beyond +-6) * meta
(dsgl.hpp:1241-1272).  Entries whose table bin float32 rounding of F could
move are left out; every other coefficient must be equal.  Covers the
output-row counts, context counts around the 32-row tile edges, both label
splits and the row sizes the general MFMA trainer runs."""
import numpy as np
import pytest

import dger_sgns as G
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def bf16(x):
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def reference(Om, Im, fin, meta, alpha, table):
    os_, nin = Om.shape[0], Im.shape[0]
    F = bf16(Im) @ bf16(Om).T                          # contexts x outputs
    lab = np.zeros((nin, os_), np.float32)
    lab[:fin, 0] = 1
    lab[fin:, os_ - 1] = 1
    x = (F + 6.0) * 83.0
    idx = np.clip(np.floor(x), 0, 1000).astype(np.int64)
    a32 = np.float32(alpha)
    mid = (lab - table[idx].astype(np.float32)) * a32
    g = np.where(F > 6, (lab - np.float32(1)) * a32, np.where(F < -6, lab * a32, mid)).astype(np.float32)
    c = (g * meta.astype(np.float32)[None, :]).astype(np.float32)
    # FP32 accumulation of D exact products: within D * 2^-24 of the sum of
    # their magnitudes; a bin is ambiguous if that error could cross its edge
    eps = Im.shape[1] * 2.0 ** -24 * (np.abs(bf16(Im)) @ np.abs(bf16(Om)).T)
    amb = (np.abs(x - np.round(x)) <= 83.0 * eps) | (np.abs(np.abs(F) - 6) <= eps)
    return c.T, amb.T                                  # outputs x contexts


CASES = [(os_, nin, D) for os_ in (2, 5, 8) for nin in (5, 31, 33, 64) for D in (32, 128, 256)]


@pytest.mark.parametrize("os_,nin,D", CASES)
def test_mfma_gemm1_coefficients(os_, nin, D):
    rng = np.random.default_rng(1000 * os_ + 10 * nin + D)
    table = O.exp_table()
    scale = 2.5 / np.sqrt(D)  # dot products spread over the table and past +-6
    Om = rng.normal(0, scale, (os_, D)).astype(np.float32)
    Im = rng.normal(0, scale, (nin, D)).astype(np.float32)
    Im[nin // 3] *= 8.0       # a row whose products saturate
    meta = rng.integers(1, 4, os_).astype(np.int32)
    meta[0] = meta[-1] = 1
    for fin in sorted({0, nin // 2, nin}):
        corr = G.step_mfma_gemm1(Om, Im, fin, meta, 0.0125)
        want, amb = reference(Om, Im, fin, meta, 0.0125, table)
        ok = ~amb
        assert ok.mean() > 0.9, ok.mean()
        assert np.array_equal(corr[ok], want[ok]), (fin, np.argwhere(corr != want)[:5])


def test_mfma_gemm1_rejects_shapes():
    Om = np.zeros((9, 128), np.float32)
    Im = np.zeros((10, 128), np.float32)
    with pytest.raises(G.DgerError) as e:
        G.step_mfma_gemm1(Om, Im, 0, np.ones(9, np.int32), 0.01)   # more than 8 outputs
    assert e.value.code == G.DGER_E_ARG
    with pytest.raises(G.DgerError) as e:
        G.step_mfma_gemm1(Om[:4, :100], Im[:, :100], 0, np.ones(4, np.int32), 0.01)  # size not a multiple of 32
    assert e.value.code == G.DGER_E_ARG
