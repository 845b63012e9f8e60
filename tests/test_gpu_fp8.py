"""FP8 form (e4m3 rows with per-row FP32 scales, FP32 masters for the top 1 %
ranks, stochastic rounding on every row store).

* storage round trip: master rows exact; coded rows within the e4m3 rounding
  bound (half a step of the 3-bit mantissa, or the subnormal step below
  2^-6 of the row scale);
* one worker: the index stream equals the FP32 form's (indices never depend
  on model values) and the model stays near the FP32 trajectory;
* link prediction on an R-MAT corpus at D = 256 (the precision-stress shape):
  AUC within 0.01 of the FP32 form."""
import numpy as np
import pytest

import dger_sgns as G
from corpora import random_corpus
from dger_sgns import synth

pytestmark = pytest.mark.gpu


def opts(D, W, K, precision, ts=10_000_000, alpha=0.025):
    return G.default_opts(size=D, window=W, negative=K, sample=1e-3, alpha=alpha, iter=1, threads=2,
                          table_size=ts, precision=precision, debug=0)


def test_fp8_storage_round_trip():
    corpus, cn = random_corpus(3, V=1500, n_walks=300)
    rng = np.random.default_rng(4)
    V, D = cn.size, 64
    x = (rng.normal(0, 1, (V, D)) * rng.lognormal(0, 2, (V, 1))).astype(np.float32)
    x[5] = 0.0  # an all-zero row keeps scale 0
    with G.Trainer(opts(D, 5, 5, "fp8"), cn, corpus) as tr:
        tr.set_model(x, x[::-1].copy())
        wih, woh = tr.get_model()
    n_master = -(-V // 100)
    assert np.array_equal(wih[:n_master], x[:n_master])
    rest, got = x[n_master:].astype(np.float64), wih[n_master:].astype(np.float64)
    rowmax = np.abs(rest).max(1, keepdims=True)
    bound = np.maximum(2.0 ** -4 * np.abs(rest), rowmax / 448 * 2.0 ** -10) * (1 + 1e-6) + 1e-30
    assert np.all(np.abs(got - rest) <= bound)
    assert not wih[5].any()
    assert np.array_equal(woh[:n_master], x[::-1][:n_master])


def test_fp8_same_index_stream_close_model():
    corpus, cn = random_corpus(9, n_walks=800)
    with G.Trainer(opts(128, 5, 5, "fp32"), cn, corpus) as tr:
        st, rec, _ = tr.train(workers=1, rec_cap=30000)
        wih, woh = tr.get_model()
    with G.Trainer(opts(128, 5, 5, "fp8"), cn, corpus) as tr:
        gst, grec, _ = tr.train(workers=1, rec_cap=30000)
        gwih, gwoh = tr.get_model()
    n = int(st.steps)
    assert gst.steps == st.steps and np.array_equal(grec[0, :n], rec[0, :n])
    for g, r, bound in ((gwih, wih, 0.25), (gwoh, woh, 0.40)):
        rel = np.linalg.norm(g - r) / np.linalg.norm(r)
        print(f"fp8 vs fp32: rel fro {rel:.4f}", flush=True)
        assert np.isfinite(g).all() and rel <= bound


def test_fp8_link_prediction_close_to_fp32():
    c, pairs, labels = synth.rmat_corpus(scale=15, seed=5, target_tokens=3_000_000, holdout=0.01)
    print(f"fp8 AUC corpus tokens={c.n_tokens} pairs={pairs.shape[0]}", flush=True)
    auc = {}
    for precision in ("fp32", "fp8"):
        with G.Trainer(opts(256, 10, 5, precision, ts=100_000_000, alpha=0.01), c.vertex_cn, c.tokens) as tr:
            tr.train()
            auc[precision] = tr.link_auc(pairs, labels)
    print(f"AUC fp32 {auc['fp32']:.4f} fp8 {auc['fp8']:.4f}", flush=True)
    assert auc["fp32"] > 0.7
    assert abs(auc["fp8"] - auc["fp32"]) <= 0.01
