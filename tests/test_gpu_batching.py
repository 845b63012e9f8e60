"""Batching options of the BF16-MFMA window-slot trainer through the C ABI:
negative sharing (-share 4/8/16: share/2 consecutive steps draw from one
negative group), minibatches smaller than a step (-batch-parity 0
-batch-size m), the level-2 autotuner over them, and a caller-supplied HIP
stream; at row lengths 128 and 256 (two wavefronts per worker), the latter
also with FP32 master rows and stochastic rounding.

The index stream is bit-exact with the oracle's window-slot restatement run
with the same options (one worker), and the model stays as close to the
oracle's FP32 trajectory as the default options do (bf16 storage and
contraction operands are the difference)."""
import ctypes as C

import numpy as np
import pytest

import dger_sgns as G
from corpora import random_corpus
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TS = 10_000_000


def gpu_opts(W, K, D=128, **kw):
    return G.default_opts(size=D, window=W, negative=K, sample=1e-3, alpha=0.025, iter=1, threads=2,
                          table_size=TS, precision="bf16-mfma", debug=0, **kw)


CASES = [
    # W, K, share, minibatch, corpus kwargs, row length
    (10, 5, 4, 0, dict(seed=2, lmax=80), 128),
    (5, 5, 8, 0, dict(seed=4, zipf=True, V=1200), 128),
    (10, 5, 2, 7, dict(seed=2, lmax=80), 128),
    (10, 5, 4, 21, dict(seed=6, lmax=70), 128),
    (3, 1, 16, 1, dict(seed=9, zipf=True, V=900), 128),
    (5, 5, 8, 0, dict(seed=4, zipf=True, V=1200), 256),
    (10, 5, 2, 7, dict(seed=2, lmax=80), 256),
]


@pytest.mark.parametrize("W,K,share,mb,ck,D", CASES)
def test_batching_index_stream_and_model(W, K, share, mb, ck, D):
    corpus, cn = random_corpus(**ck)
    print(f"batching case D={D} W={W} K={K} share={share} minibatch={mb} tokens={corpus.size}", flush=True)
    wih, woh, st, rec, _ = O.train(corpus, cn, O.make_opts(dim=D, window=W, negative=K, sample=1e-3, alpha=0.025,
                                                          iter=1, threads=2, workers=1, table_size=TS,
                                                          semantics="slots", share=share, minibatch=mb),
                                   rec_cap=60000)
    kw = dict(share=share)
    if mb:
        kw.update(batch_parity=0, batch_size=mb)
    with G.Trainer(gpu_opts(W, K, D, **kw), cn, corpus) as tr:
        gst, grec, _ = tr.train(workers=1, rec_cap=60000)
        gwih, gwoh = tr.get_model()
    assert gst.share == share and gst.minibatch == (mb if mb < 4 * W else 0)
    assert gst.steps == st.steps and gst.raw_words == st.raw_words and gst.sentences == st.sentences
    n = int(st.steps)
    assert n > 100
    assert np.array_equal(grec[0, :n], rec[0, :n])
    # every context is trained once, whatever the minibatch size
    assert gst.contexts == int((rec[0, :n, 4]).sum())
    for g, r, name in ((gwih, wih, "Wih"), (gwoh, woh, "Woh")):
        rel = np.linalg.norm(g - r) / np.linalg.norm(r)
        print(f"  {name}: rel fro vs fp32 window-slot oracle {rel:.4f}")
        assert rel < 0.1


def test_masters_and_stochastic_rounding_at_256():
    # the full instance with both row halves: FP32 masters for the top 1 %
    # ranks and stochastic rounding change the arithmetic, not the index
    # stream, and keep the model near the FP32 trajectory
    corpus, cn = random_corpus(seed=4, zipf=True, V=1200)
    print(f"masters + sr at D=256, tokens={corpus.size}", flush=True)
    wih, woh, st, rec, _ = O.train(corpus, cn, O.make_opts(dim=256, window=5, negative=5, sample=1e-3, alpha=0.025,
                                                          iter=1, threads=2, workers=1, table_size=TS,
                                                          semantics="slots"), rec_cap=60000)
    with G.Trainer(gpu_opts(5, 5, 256, masters=1, rounding=0), cn, corpus) as tr:
        gst, grec, _ = tr.train(workers=1, rec_cap=60000)
        gwih, gwoh = tr.get_model()
    n = int(st.steps)
    assert gst.steps == st.steps and n > 100
    assert np.array_equal(grec[0, :n], rec[0, :n])
    for g, r, name in ((gwih, wih, "Wih"), (gwoh, woh, "Woh")):
        rel = np.linalg.norm(g - r) / np.linalg.norm(r)
        print(f"  {name}: rel fro vs fp32 window-slot oracle {rel:.4f}")
        assert rel < 0.1


def test_minibatch_changes_the_update_and_parity_does_not():
    corpus, cn = random_corpus(seed=2, lmax=80)
    out = {}
    for name, kw in (("parity", {}), ("parity-size", dict(batch_size=3)), ("wide", dict(batch_parity=0, batch_size=40)),
                     ("split", dict(batch_parity=0, batch_size=5))):
        with G.Trainer(gpu_opts(10, 5, **kw), cn, corpus) as tr:
            st, _, _ = tr.train(workers=1)
            out[name] = (tr.get_model(), st)
    # the parity switch ignores -batch-size; a minibatch of at least 4*window
    # contexts never splits a step
    for name in ("parity-size", "wide"):
        assert out[name][1].minibatch == 0
        assert all(np.array_equal(a, b) for a, b in zip(out[name][0], out["parity"][0]))
    assert out["split"][1].minibatch == 5
    assert not np.array_equal(out["split"][0][0], out["parity"][0][0])


def test_batching_needs_the_window_slot_trainer():
    cn = np.ones(2000, np.int32)
    for kw in (dict(share=4), dict(batch_parity=0, batch_size=5), dict(autotune=2)):
        for prec, size in (("fp32", 128), ("bf16-mfma", 64)):
            o = G.default_opts(size=size, window=5, negative=5, precision=prec, table_size=1000, **kw)
            with pytest.raises(G.DgerError) as e:
                G.Trainer(o, cn)
            assert e.value.code == G.DGER_E_UNSUPPORTED, (kw, prec)


def test_autotune_level2_picks_a_measured_configuration():
    corpus, cn = random_corpus(seed=11, V=4000, n_walks=6000, lmax=80, zipf=True)
    o = gpu_opts(10, 5, autotune=2)
    with G.Trainer(o, cn, corpus) as tr:
        st, _, _ = tr.train()
        assert st.share in (2, 4, 8) and st.minibatch in (0, 21)
        assert st.steps > 0 and np.isfinite(tr.get_model()[0]).all()
        # the choice sticks for later runs of the context
        st2, _, _ = tr.train()
        assert (st2.share, st2.minibatch, st2.workers) == (st.share, st.minibatch, st.workers)
    print(f"autotune level 2: {st.workers} workers, share {st.share}, minibatch {st.minibatch}")


def _hip():
    h = C.CDLL("libamdhip64.so.7")
    h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    h.hipStreamDestroy.argtypes = [C.c_void_p]
    return h


def test_caller_stream_gives_the_same_run():
    corpus, cn = random_corpus(seed=3, V=1500, n_walks=800, lmax=60)
    hip = _hip()
    s = C.c_void_p()
    assert hip.hipStreamCreate(C.byref(s)) == 0
    try:
        for prec in ("bf16-mfma", "fp32"):
            o = G.default_opts(size=128, window=5, negative=5, sample=1e-3, alpha=0.025, iter=1, threads=2,
                               table_size=TS, precision=prec, debug=0, workers=1)
            e0, r0, st0 = G.train(o, corpus, cn)
            e1, r1, st1 = G.train(o, corpus, cn, stream=s.value)
            assert np.array_equal(e0, e1) and np.array_equal(r0, r1) and st0.steps == st1.steps
            with G.Trainer(o, cn, corpus) as tr:
                tr.set_stream(s.value)
                tr.train(workers=1)
                tr.set_stream(None)
                tr.train(workers=1)   # a second epoch from the first's model, on the own stream
                w2 = tr.get_model()[0]
            with G.Trainer(o, cn, corpus) as tr:
                tr.train(workers=1)
                tr.train(workers=1)
                assert np.array_equal(tr.get_model()[0], w2)
    finally:
        hip.hipStreamDestroy(s)
