"""Link-prediction evaluation on the device against host float64 scoring.

Scores: dot products of Wih rows, within 1e-5 (relative to the row norms) of
the float64 dot of the same float32 rows.  AUC: the device's Mann-Whitney
count equals the average-rank statistic of the same scores (scipy rankdata),
and the host AUC of float64 scores to 1e-4."""
import numpy as np
import pytest

import dger_sgns as G
from dger_sgns import synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_link_scores_and_auc(precision):
    n, m = 3000, 15000
    u, v = synth.er_graph(n, m, 5)
    tu, tv, pairs, labels = synth.holdout_split(u, v, n, 0.05, 5)
    rowptr, col = synth.csr(n, tu, tv)
    starts = np.repeat(np.arange(n), 4)
    c = synth.walk_corpus(rowptr, col, starts, np.full(starts.size, 40), 5)
    tokens, cn = c.tokens, c.vertex_cn
    print(f"link eval {precision} tokens={tokens.size} pairs={pairs.shape[0]}", flush=True)
    opts = G.default_opts(size=64, window=5, negative=5, sample=1e-3, alpha=0.025, iter=1, threads=2,
                          table_size=1_000_000, precision=precision, debug=0)
    with G.Trainer(opts, cn, tokens) as tr:
        tr.train()
        rank_id, _ = tr.vocab()
        wih, _ = tr.get_model()
        s = tr.link_scores(pairs)
        a = tr.link_auc(pairs, labels)
        # ids outside [0, v_num) are rejected on the host before any upload
        for bad_pairs in ([[0, cn.size]], [[-1, 3]]):
            with pytest.raises(G.DgerError) as e:
                tr.link_scores(np.array(bad_pairs, np.int32))
            assert e.value.code == G.DGER_E_ARG
            with pytest.raises(G.DgerError) as e:
                tr.link_auc(np.array(bad_pairs, np.int32), np.ones(1, np.int8))
            assert e.value.code == G.DGER_E_ARG
    by_id = np.zeros_like(wih)
    by_id[rank_id] = wih
    x, y = by_id[pairs[:, 0]].astype(np.float64), by_id[pairs[:, 1]].astype(np.float64)
    ref = (x * y).sum(1)
    scale = np.linalg.norm(x, axis=1) * np.linalg.norm(y, axis=1) + 1e-30
    err = np.abs(s - ref) / scale
    bad = np.nonzero(err > 1e-5)[0]
    assert bad.size == 0, [(int(i), pairs[i].tolist(), float(s[i]), float(ref[i]), float(scale[i])) for i in bad[:5]]
    assert abs(a - synth.auc(s, labels)) < 1e-12
    assert abs(a - synth.auc(ref, labels)) < 1e-4
    assert 0.0 <= a <= 1.0  # an ER graph has no structure to predict: about 0.5
