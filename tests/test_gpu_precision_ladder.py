"""BF16 precision ladder of the window-slot trainer (DGER_PRECISION_BF16_MFMA):
stochastic rounding of row stores and Woh updates (-rounding sr) and FP32
master rows for the top 1 % ranks (-masters 1), against the oracle's FP32
restatement of the same update rules (semantics "slots"), one worker.

With a small learning rate over several epochs (alpha0 = 5e-4, 8 epochs,
decaying towards alpha0 * 1e-4) most updates are below half a bf16 step of
the rows they touch.  Nearest rounding then drops them: the model moves
measurably less than the FP32 trajectory ("stalls"); stochastic rounding keeps
them in expectation and tracks the FP32 trajectory; masters remove the
rounding of the hottest rows altogether.  Error is stated as
||W - W_fp32|| / ||W_fp32 - W_init|| (Woh starts at zero), movement as
||W - W_init|| / ||W_fp32 - W_init||."""
import numpy as np
import pytest

import dger_sgns as G
from corpora import random_corpus
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ALPHA, ITER, TS = 5e-4, 8, 1_000_000


def run(corpus, cn, masters, rounding):
    o = G.default_opts(size=128, window=5, negative=5, sample=0, alpha=ALPHA, iter=ITER, threads=2, table_size=TS,
                       precision="bf16-mfma", debug=0, masters=masters, rounding=rounding)
    with G.Trainer(o, cn, corpus) as tr:
        tr.train(workers=1)
        return tr.get_model()


def test_small_alpha_stall_and_compensation():
    corpus, cn = random_corpus(seed=21, V=800, n_walks=600, lmax=60, zipf=True)
    w32, o32, _, _, _ = O.train(corpus, cn, O.make_opts(dim=128, window=5, negative=5, sample=0, alpha=ALPHA,
                                                          iter=ITER, threads=2, workers=1, table_size=TS,
                                                          semantics="slots"))
    wi = O.init_net(cn.size, 128)
    mw, mo = np.linalg.norm(w32 - wi), np.linalg.norm(o32)
    res = {}
    for name, masters, rounding in (("rne", 0, 1), ("sr", 0, 0), ("masters", 1, 1), ("masters+sr", 1, 0)):
        gw, go = run(corpus, cn, masters, rounding)
        res[name] = dict(err_w=np.linalg.norm(gw - w32) / mw, err_o=np.linalg.norm(go - o32) / mo,
                         move_w=np.linalg.norm(gw - wi) / mw, move_o=np.linalg.norm(go) / mo)
        r = res[name]
        print(f"{name:11s}: error Wih {r['err_w']:.3f} Woh {r['err_o']:.3f}; moved Wih {r['move_w']:.3f} "
              f"Woh {r['move_o']:.3f} of the FP32 trajectory", flush=True)
    # nearest rounding stalls
    assert res["rne"]["move_w"] < 0.75 and res["rne"]["err_w"] > 0.3
    # stochastic rounding moves as far as FP32 and stays close to it
    for name in ("sr", "masters+sr"):
        assert 0.95 < res[name]["move_w"] < 1.05
        assert res[name]["err_w"] < 0.12 and res[name]["err_o"] < 0.15
    # masters (exact FP32 rows for the hottest ranks) cut the error further
    assert res["masters+sr"]["err_w"] < res["sr"]["err_w"]
    assert res["masters"]["err_w"] < res["rne"]["err_w"]
