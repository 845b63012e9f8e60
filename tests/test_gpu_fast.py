"""The BF16-MFMA window-slot trainer (sgns_fast.hip) through the C ABI, at
row lengths 128 (one wavefront per worker) and 256 (two wavefronts, one
half of every row each, GEMM1's half products added through LDS).

* One step (dger_sgns_step_fast_d) against a float64 restatement of the same
  step with the operands rounded to bf16: GEMM1 F = O I^T, the table-sigmoid
  coefficients C = mult * meta * g(F) (dsgl.hpp:1241-1272), md = C^T I
  (dsgl.hpp:1290) and the slot update I += C O (dsgl.hpp:1309), with C
  rounded to bf16 for the last two.  Entries whose sigmoid-table bin is
  ambiguous at float32 precision are excluded from the coefficient check and
  their rows / slots from the update checks.
* Whole runs: the index stream (sentences, windows, negatives, outputs,
  learning rate) is bit-exact with the oracle for one worker, and the model
  stays close to the oracle's FP32 restatement of the window-slot update
  rules (oracle semantics "slots")."""
import numpy as np
import pytest

import dger_sgns as G
from oracle import oracle as O
from corpora import random_corpus

pytestmark = pytest.mark.gpu

TS = 10_000_000


def bf16(x):
    """Round float32 to bf16 (nearest even), returned as float64."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def step_reference(Om, Im, mult, meta, alpha, table):
    os_ = Om.shape[0]
    Ob, Ib = bf16(Om), bf16(Im)
    F = Ib @ Ob.T                                   # slots x outputs
    lab = np.zeros_like(F)
    lab[:24, 0] = 1
    lab[24:, os_ - 1] = 1
    x = (F + 6.0) * 83.0
    idx = np.clip(np.floor(x).astype(np.int64), 0, 1000)
    # float32 arithmetic in the kernel's order: (label - e) * (alpha * count),
    # count = contexts in the slot x output multiplicity; e = 1 above 6, 0
    # below -6 (dsgl.hpp:1249-1270)
    lab32, a32 = lab.astype(np.float32), np.float32(alpha)
    e = np.where(F > 6, np.float32(1), np.where(F < -6, np.float32(0), table[idx])).astype(np.float32)
    cnt = (mult[:, None] * meta[None, :]).astype(np.float32)
    C = ((lab32 - e) * (a32 * cnt)).astype(np.float32).astype(np.float64)
    # bins that float32 rounding of F could change: FP32 accumulation of D
    # exact bf16 products is within D * 2^-24 of the sum of their magnitudes
    eps = Om.shape[1] * 2.0 ** -24 * (np.abs(Ib) @ np.abs(Ob).T)
    amb = (np.abs(x - np.round(x)) <= 83.0 * eps) | (np.abs(np.abs(F) - 6) <= eps)
    amb &= (mult[:, None] > 0)
    Cb = bf16(C.astype(np.float32))
    md = Cb.T @ Ib                                   # outputs x D
    I_new = Im.astype(np.float64) + Cb @ Ob          # slots x D
    return C, md, I_new, amb


@pytest.mark.parametrize("D", [128, 256])
@pytest.mark.parametrize("os_,seed", [(2, 1), (5, 2), (7, 3), (16, 4)])
def test_step_fast_against_float64(os_, seed, D):
    rng = np.random.default_rng(seed)
    t = O.exp_table()
    sd = 0.3 * np.sqrt(128 / D)  # the same spread of dot products at both lengths
    Om = rng.normal(0, sd, (os_, D)).astype(np.float32)
    Im = rng.normal(0, sd, (48, D)).astype(np.float32)
    mult = rng.integers(0, 4, 48).astype(np.int32)
    mult[rng.random(48) < 0.4] = 0
    meta = rng.integers(1, 3, os_).astype(np.int32)
    meta[0] = meta[-1] = 1
    alpha = 0.0125
    corr, md, I2 = G.step_fast(Om, Im, mult, meta, alpha)
    C, md_ref, I_ref, amb = step_reference(Om, Im, mult, meta, alpha, t)
    ok = ~amb
    print(f"D={D} os={os_}: {amb.sum()} ambiguous bins of {amb.size}", flush=True)
    assert amb.sum() <= 0.1 * (mult > 0).sum() * os_
    # coefficients: float32 F against float64 F, same bin
    assert np.array_equal(corr.T[ok], C[ok].astype(np.float32))
    assert np.all(corr.T[mult == 0] == 0)
    # updates: rows / slots touched by an ambiguous bin are left out
    rows_ok = ~amb.any(axis=0)
    slots_ok = ~amb.any(axis=1)
    # float32 accumulation of exact bf16 products: a few ulps of the sum of
    # the terms' magnitudes
    Cb = bf16(C.astype(np.float32))
    mag_md = np.abs(Cb).T @ np.abs(bf16(Im))
    assert np.all(np.abs(md[rows_ok] - md_ref[rows_ok]) <= 1e-6 * mag_md[rows_ok] + 1e-12)
    mag_i = np.abs(Im) + np.abs(Cb) @ np.abs(bf16(Om))
    assert np.all(np.abs(I2[slots_ok] - I_ref[slots_ok]) <= 1e-6 * mag_i[slots_ok] + 1e-12)
    # slots that are no context keep their row exactly
    assert np.array_equal(I2[mult == 0], Im[mult == 0])


def test_step_fast_layout_by_hand():
    # Unit rows make every product exact: F[s][i] = 1 where slot s and output
    # i share their unit vector.  With all F in (-6, 6) and alpha = 1 the
    # coefficients are table values; md and the update follow from them.
    t = O.exp_table()
    os_ = 5
    Om = np.zeros((os_, 128), np.float32)
    Im = np.zeros((48, 128), np.float32)
    for i in range(os_):
        Om[i, 7 * i + 3] = 1.0
    for sl in range(48):
        Im[sl, (7 * (sl % os_) + 3) if sl % 3 else 100 + sl % 20] = 1.0
    mult = np.array([(sl % 4) for sl in range(48)], np.int32)
    meta = np.array([1, 2, 1, 3, 1], np.int32)
    corr, md, I2 = G.step_fast(Om, Im, mult, meta, 1.0)
    F = Im.astype(np.float64) @ Om.T.astype(np.float64)
    for sl in range(48):
        for i in range(os_):
            lab = 1 if (i == 0 and sl < 24) or (i == os_ - 1 and sl >= 24) else 0
            idx = int((F[sl, i] + 6) * 83)
            want = (np.float32(lab) - t[idx]) * np.float32(mult[sl] * meta[i])
            assert corr[i, sl] == pytest.approx(want, rel=1e-6, abs=0), (sl, i)


def test_step_fast_halves_by_hand():
    # Row length 256: every row has one entry in each half, so each F is the
    # sum of one exact product from each wavefront's half (0, 0.5, 1 or 1.5:
    # a half-product lost or counted twice moves it by a whole bin range).
    # The updates then follow exactly: md = C^T I and I += C O touch only the
    # columns of those entries.
    t = O.exp_table()
    os_, D = 6, 256
    Om = np.zeros((os_, D), np.float32)
    Im = np.zeros((48, D), np.float32)
    for i in range(os_):
        Om[i, 5 * i + 1] = 1.0            # first half
        Om[i, 128 + 9 * i + 2] = 0.5      # second half
    for sl in range(48):
        a, b = sl % os_, (sl // 2) % os_
        Im[sl, 5 * a + 1] = 1.0 if sl % 3 else 0.0
        Im[sl, 128 + 9 * b + 2] = 1.0
    mult = np.array([(sl % 4) for sl in range(48)], np.int32)
    meta = np.array([1, 2, 1, 3, 2, 1], np.int32)
    corr, md, I2 = G.step_fast(Om, Im, mult, meta, 1.0)
    F = Im.astype(np.float64) @ Om.T.astype(np.float64)
    assert set(np.unique(F)) <= {0.0, 0.5, 1.0, 1.5}
    C = np.zeros((48, os_))
    for sl in range(48):
        for i in range(os_):
            lab = 1 if (i == 0 and sl < 24) or (i == os_ - 1 and sl >= 24) else 0
            idx = int((F[sl, i] + 6) * 83)
            want = (np.float32(lab) - t[idx]) * np.float32(mult[sl] * meta[i])
            assert corr[i, sl] == pytest.approx(want, rel=1e-6, abs=0), (sl, i)
            C[sl, i] = corr[i, sl]
    # md and the update in float64 from the kernel's own coefficients rounded
    # to bf16 (the contractions' operand type); the rows are exact in bf16
    Cb = bf16(C.astype(np.float32))
    md_ref = Cb.T @ Im.astype(np.float64)
    I_ref = Im.astype(np.float64) + Cb @ Om.astype(np.float64)
    assert np.allclose(md, md_ref, rtol=1e-6, atol=1e-9)
    assert np.allclose(I2, I_ref, rtol=1e-6, atol=1e-9)


def gpu_opts(D, W, K, sample, iter=1, precision="bf16-mfma", alpha=0.025, ts=TS):
    return G.default_opts(size=D, window=W, negative=K, sample=sample, alpha=alpha, iter=iter,
                          threads=2, table_size=ts, precision=precision, debug=0)


FAST_CASES = [
    # W, K, sample, iter, corpus kwargs
    (10, 5, 1e-3, 1, dict(seed=2, lmax=80)),
    (5, 5, 1e-2, 1, dict(seed=4, zipf=True, V=1200)),
    (11, 14, 1e-3, 1, dict(seed=5, n_walks=300)),
    (1, 0, 0.0, 2, dict(seed=3, lmin=1, empty_walks=5)),
    (10, 5, 1e-3, 1, dict(seed=7, n_walks=120, long_walk=2600)),
]


@pytest.mark.parametrize("D", [128, 256])
@pytest.mark.parametrize("W,K,sample,it,ck", FAST_CASES)
def test_fast_index_stream_bit_exact(W, K, sample, it, ck, D):
    corpus, cn = random_corpus(**ck)
    print(f"fast case D={D} W={W} K={K} sample={sample} iter={it} tokens={corpus.size}", flush=True)
    wih, woh, st, rec, ws = O.train(corpus, cn, O.make_opts(dim=D, window=W, negative=K, sample=sample,
                                                             alpha=0.025, iter=it, threads=2, workers=1,
                                                             table_size=TS, semantics="slots"), rec_cap=60000)
    with G.Trainer(gpu_opts(D, W, K, sample, it), cn, corpus) as tr:
        gst, grec, _ = tr.train(workers=1, rec_cap=60000)
        gwih, gwoh = tr.get_model()
    assert gst.steps == st.steps and gst.raw_words == st.raw_words and gst.sentences == st.sentences
    n = int(st.steps)
    assert n > 100
    assert np.array_equal(grec[0, :n], rec[0, :n])
    assert np.isfinite(gwih).all() and np.isfinite(gwoh).all()
    # against the oracle's FP32 restatement of the window-slot update rules:
    # the difference is bf16 storage and bf16 contraction operands
    for g, r, name in ((gwih, wih, "Wih"), (gwoh, woh, "Woh")):
        rel = np.linalg.norm(g - r) / np.linalg.norm(r)
        print(f"  {name}: rel fro vs fp32 window-slot oracle {rel:.4f}, max abs {np.abs(g - r).max():.3e}")
        assert rel < 0.1


@pytest.mark.parametrize("D", [128, 256])
def test_fast_multi_worker_index_streams(D):
    corpus, cn = random_corpus(8, n_walks=1500, zipf=True)
    nw = 7
    _, _, st, rec, ws = O.train(corpus, cn, O.make_opts(dim=D, window=5, negative=5, sample=1e-3, alpha=0.025,
                                                         threads=2, workers=nw, table_size=TS, semantics="slots"),
                                rec_cap=20000)
    with G.Trainer(gpu_opts(D, 5, 5, 1e-3), cn, corpus) as tr:
        gst, grec, gws = tr.train(workers=nw, rec_cap=20000)
    assert np.array_equal(gws, ws)
    idx = np.ones(rec.shape[2], bool)
    idx[6] = False  # the learning rate reads the shared word counter
    for w in range(nw):
        n = int(ws[w])
        assert np.array_equal(grec[w, :n][:, idx], rec[w, :n][:, idx]), w
