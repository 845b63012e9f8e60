"""Pins of the walk oracle (oracle/walk_oracle.c), on the CPU.

* the entropy recursion (include/walk.hpp:300-308) equals the Shannon entropy
  (bits) of the visit counts of every prefix, recomputed from scratch;
* R (walk.hpp:336-344) equals numpy's Pearson correlation of the (H, S)
  samples, the start sample (0, 1) included;
* the trace rate equals 1 - tanh(p_accept) where p_accept is one constant:
  complete graph K_n (deg - common = 1, equal degrees: p_accept = 1) and
  complete bipartite K_{a,a} (no common neighbours: p_accept = 1/a) -- a
  flipped comparison or a dropped factor lands on another constant;
* every walk ends at the first step past min_length where R^2 < threshold or
  R < 0 (numpy R of the emitted path), consecutive vertices are adjacent,
  walk i starts at (first + i) % V and vertex_cn is the corpus histogram."""
import math

import numpy as np
import pytest

from dger_sgns import synth
from oracle import walk_oracle as WO


def entropy_bits(counts):
    c = np.asarray(counts, np.float64)
    p = c / c.sum()
    return float(-(p * np.log2(p)).sum())


def brute_h_r(path):
    """H and R after every visit, from scratch."""
    H, R = [0.0], [1.0]
    for i in range(1, len(path)):
        _, cnt = np.unique(path[: i + 1], return_counts=True)
        H.append(entropy_bits(cnt))
        h = np.array(H)
        s = np.arange(1, i + 2, dtype=np.float64)
        R.append(float(np.corrcoef(h, s)[0, 1]) if h.std() > 0 else float("nan"))
    return np.array(H), np.array(R)


def common_neighbors_np(rowptr, col):
    nb = [set(col[rowptr[u]:rowptr[u + 1]].tolist()) for u in range(rowptr.size - 1)]
    out = np.empty(col.size, np.int32)
    for u in range(rowptr.size - 1):
        for e in range(rowptr[u], rowptr[u + 1]):
            out[e] = len(nb[u] & nb[col[e]])
    return out


def sorted_csr(n, u, v):
    rowptr, col = synth.csr(n, u, v)
    for x in range(n):
        col[rowptr[x]:rowptr[x + 1]].sort()
    return rowptr, col


def complete_graph(n):
    iu = np.triu_indices(n, 1)
    return sorted_csr(n, iu[0].astype(np.int32), iu[1].astype(np.int32))


def complete_bipartite(a):
    u = np.repeat(np.arange(a), a).astype(np.int32)
    v = (a + np.tile(np.arange(a), a)).astype(np.int32)
    return sorted_csr(2 * a, u, v)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_entropy_and_correlation_closed_forms(seed):
    rng = np.random.default_rng(seed)
    path = rng.integers(0, 12, 60).astype(np.int32)   # many repeats
    path[:4] = [20, 21, 22, 23]                       # H rises at first: R defined
    H, R = WO.corr_trace(path)
    bh, br = brute_h_r(path)
    assert np.allclose(H, bh, rtol=1e-10, atol=1e-12)
    ok = ~np.isnan(br)
    assert ok.sum() > 50
    assert np.allclose(R[ok], br[ok], rtol=1e-9, atol=1e-12)


def test_all_new_vertices_entropy_is_log2_count():
    H, R = WO.corr_trace(np.arange(40, dtype=np.int32))
    assert np.allclose(H, np.log2(np.arange(1, 41)), rtol=1e-12, atol=1e-14)
    assert R[1] == 1.0


@pytest.mark.parametrize("graph,p_accept", [("K40", 1.0), ("K6,6", 1.0 / 6)])
def test_trace_rate_is_one_minus_tanh(graph, p_accept):
    rowptr, col = complete_graph(40) if graph == "K40" else complete_bipartite(6)
    cnb = common_neighbors_np(rowptr, col)
    assert np.all(cnb == (38 if graph == "K40" else 0))
    _, _, st = WO.walk(rowptr, col, cnb, n_walkers=4000, seed=7)
    rate = st["traces"] / st["moves"]
    want = 1.0 - math.tanh(p_accept)
    sigma = math.sqrt(want * (1 - want) / st["moves"])
    assert abs(rate - want) < 5 * sigma + 1e-3, (rate, want, st)


def split_walks(corpus):
    ends = np.nonzero(corpus == -1)[0]
    starts = np.concatenate([[0], ends[:-1] + 1])
    return [corpus[s:e] for s, e in zip(starts, ends)]


@pytest.mark.parametrize("first,min_length", [(0, 20), (123, 5)])
def test_walk_rules_on_er_graph(first, min_length):
    n = 400
    u, v = synth.er_graph(n, 2400, seed=3)
    rowptr, col = sorted_csr(n, u, v)
    assert np.diff(rowptr).min() > 0
    cnb = common_neighbors_np(rowptr, col)
    thr = 0.999
    corpus, vcn, st = WO.walk(rowptr, col, cnb, n_walkers=300, first=first, min_length=min_length, seed=11)
    walks = split_walks(corpus)
    assert len(walks) == 300 and st["walks"] == 300 and st["capped"] == 0
    assert np.array_equal(vcn, np.bincount(corpus[corpus >= 0], minlength=n))
    assert st["tokens"] == sum(w.size for w in walks)
    adj = {(int(a), int(b)) for a in range(n) for b in col[rowptr[a]:rowptr[a + 1]]}
    returns = 0
    for i, w in enumerate(walks):
        assert w[0] == (first + i) % n
        assert all((int(a), int(b)) in adj for a, b in zip(w[:-1], w[1:]))
        returns += int(np.sum(w[2:] == w[:-2]))
        _, R = brute_h_r(w)
        last = w.size - 1                      # steps taken
        assert last > min_length
        for s in range(min_length + 1, last + 1):
            r = R[s]
            stop = (r * r < thr) or (r < 0)
            if abs(r * r - thr) < 1e-9:
                continue                       # numpy rounding may differ at a tie
            assert stop == (s == last), (i, s, last, r)
    assert returns > 0 and st["traces"] > 0
