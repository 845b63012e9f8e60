"""dger_ctx_train_noalloc, the allocation-free trainer entry point, against
dger_ctx_train on the same contexts and corpora.

* One worker: every run statistic (steps, words, sentences, contexts,
  output-context pairs, final learning rate) is equal, and so is the model
  when dger_ctx_train itself reproduces its model bit for bit on a second
  context (one worker is deterministic unless the device reorders its own
  atomics); otherwise the model is within the run-to-run difference.
* Two calls in a row on one GPU (the static scratch keeps the last call's
  contents), the second on a smaller corpus.
* Several workers: the statistics are equal (each worker's index stream
  depends only on its slice), also at row length 256 (two wavefronts per
  worker, twice the scratch) and in the FP8 form.
* Host-side rejections: a context the window-slot trainer does not run, and a
  worker count whose scratch exceeds the static array."""
import numpy as np
import pytest

import dger_sgns as G
from corpora import random_corpus

pytestmark = pytest.mark.gpu

TS = 10_000_000
STAT_FIELDS = ("steps", "trained_words", "raw_words", "sentences", "contexts", "out_ctx", "workers", "final_alpha")


def opts(precision="bf16-mfma", D=128):
    return G.default_opts(size=D, window=5, negative=5, sample=1e-3, alpha=0.025, iter=1, threads=2,
                          table_size=TS, precision=precision, debug=0)


def trained(corpus, cn, workers, noalloc, **kw):
    with G.Trainer(opts(**kw), cn, corpus) as tr:
        st = tr.train_noalloc(workers=workers) if noalloc else tr.train(workers=workers)[0]
        wih, woh = tr.get_model()
    return st, wih, woh


def check_stats(a, b):
    for f in STAT_FIELDS:
        assert getattr(a, f) == getattr(b, f), f


def test_noalloc_one_worker_two_calls():
    corpus, cn = random_corpus(11, n_walks=900, lmax=80)
    small = corpus[: np.nonzero(corpus == -1)[0][299] + 1]  # the first 300 walks
    for c in (corpus, small):
        print(f"noalloc vs train, one worker, {c.size} tokens", flush=True)
        sa, wa, oa = trained(c, cn, 1, noalloc=False)
        sr, wr, orr = trained(c, cn, 1, noalloc=False)
        sb, wb, ob = trained(c, cn, 1, noalloc=True)
        assert sa.steps > 100
        check_stats(sa, sb)
        check_stats(sa, sr)
        if np.array_equal(wa, wr) and np.array_equal(oa, orr):
            print("  dger_ctx_train repeats its model bit for bit: comparing bit for bit", flush=True)
            assert np.array_equal(wa, wb) and np.array_equal(oa, ob)
        else:
            print("  dger_ctx_train differs run to run: comparing within that difference", flush=True)
            for x, y, z in ((wa, wb, wr), (oa, ob, orr)):
                run_to_run = np.abs(x - z).max()
                assert np.abs(x - y).max() <= 4 * run_to_run + 1e-6


@pytest.mark.parametrize("precision,D", [("bf16-mfma", 128), ("bf16-mfma", 256), ("fp8", 256)])
def test_noalloc_several_workers_stats(precision, D):
    corpus, cn = random_corpus(12, n_walks=2000, zipf=True)
    print(f"noalloc vs train, 16 workers, {precision} D={D}, {corpus.size} tokens", flush=True)
    sa, wa, _ = trained(corpus, cn, 16, noalloc=False, precision=precision, D=D)
    sb, wb, _ = trained(corpus, cn, 16, noalloc=True, precision=precision, D=D)
    check_stats(sa, sb)
    assert np.isfinite(wb).all()
    rel = np.linalg.norm(wa - wb) / np.linalg.norm(wa)
    assert rel < 0.1, rel


def test_noalloc_rejections():
    corpus, cn = random_corpus(13, n_walks=50)
    with G.Trainer(opts("fp32"), cn, corpus) as tr:
        with pytest.raises(G.DgerError) as e:
            tr.train_noalloc(workers=1)
        assert e.value.code == G.DGER_E_UNSUPPORTED
    with G.Trainer(opts(), cn, corpus) as tr:
        with pytest.raises(G.DgerError) as e:
            tr.train_noalloc(workers=1 << 20)  # about 1 KiB of scratch each: past the 16 MiB array
        assert e.value.code == G.DGER_E_ARG
        assert "workers fit" in str(e.value)
