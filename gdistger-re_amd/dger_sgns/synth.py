"""Synthetic graphs and walk corpora for tests and the benchmark (numpy).

The reference has no bundled datasets for its trainer; its trainer input is the
in-memory walk corpus of ``PathSet::dump`` (include/path.hpp:91-96,110): vertex
ids with -1 after every walk, plus per-vertex occurrence counts.  These
generators produce that layout from seeded synthetic graphs:

* ``er_graph``   G(n, m) undirected, no self loops, no duplicates.
* ``rmat_graph`` R-MAT with quadrant probabilities (a, b, c, 1-a-b-c); self
  loops and duplicates removed; vertex ids randomly permuted with the seed so
  id order is unrelated to degree order.
* ``walk_corpus`` uniform random walks on the CSR (the information-oriented
  walker of include/walk.hpp is the corpus *producer* and is not rebuilt
  here; walk lengths are drawn from the requested distribution instead).
* ``holdout_split`` the link-prediction split of src/tools/gtrain.cpp:300-343:
  remove a fraction of edges whose endpoints both still have degree > 1, and
  draw the same number of uniform non-edges.
"""
from __future__ import annotations

import numpy as np

from . import Corpus


def _dedup_undirected(u: np.ndarray, v: np.ndarray):
    keep = u != v
    u, v = u[keep], v[keep]
    lo = np.minimum(u, v).astype(np.int64)
    hi = np.maximum(u, v).astype(np.int64)
    key = np.unique(lo << 32 | hi)
    return (key >> 32).astype(np.int32), (key & 0xFFFFFFFF).astype(np.int32)


def er_graph(n: int, m: int, seed: int):
    """G(n, m): m distinct undirected edges without self loops."""
    rng = np.random.default_rng(seed)
    u = np.empty(0, np.int32)
    v = np.empty(0, np.int32)
    while u.size < m:
        need = int((m - u.size) * 1.2) + 16
        a = rng.integers(0, n, need, dtype=np.int64).astype(np.int32)
        b = rng.integers(0, n, need, dtype=np.int64).astype(np.int32)
        u, v = _dedup_undirected(np.concatenate([u, a]), np.concatenate([v, b]))
    sel = np.sort(rng.choice(u.size, m, replace=False))
    return u[sel], v[sel]


def rmat_graph(scale: int, m: int, a: float, b: float, c: float, seed: int):
    """R-MAT edges (undirected, deduplicated, ids permuted); returns (u, v, n)."""
    rng = np.random.default_rng(seed)
    n = 1 << scale
    target = m
    us, vs = [], []
    have = 0
    gen = int(m * 1.08)
    while True:
        u = np.zeros(gen, np.int64)
        v = np.zeros(gen, np.int64)
        for _ in range(scale):
            r = rng.random(gen, dtype=np.float32)
            bu = (r >= a + b).astype(np.int64)              # quadrants c, d
            bv = ((r >= a) & (r < a + b)) | (r >= a + b + c)  # quadrants b, d
            u = (u << 1) | bu
            v = (v << 1) | bv.astype(np.int64)
        us.append(u.astype(np.int32))
        vs.append(v.astype(np.int32))
        uu, vv = _dedup_undirected(np.concatenate(us), np.concatenate(vs))
        have = uu.size
        if have >= target:
            break
        gen = int((target - have) * 1.3) + 1024
    sel = np.sort(rng.choice(uu.size, target, replace=False))
    uu, vv = uu[sel], vv[sel]
    perm = rng.permutation(n).astype(np.int32)
    return perm[uu], perm[vv], n


def csr(n: int, u: np.ndarray, v: np.ndarray):
    """Symmetric CSR (rowptr int64, col int32)."""
    src = np.concatenate([u, v]).astype(np.int64)
    dst = np.concatenate([v, u]).astype(np.int32)
    order = np.argsort(src, kind="stable")
    col = dst[order]
    deg = np.bincount(src, minlength=n)
    rowptr = np.zeros(n + 1, np.int64)
    np.cumsum(deg, out=rowptr[1:])
    return rowptr, col


def walk_corpus(rowptr: np.ndarray, col: np.ndarray, starts: np.ndarray, lengths: np.ndarray,
                seed: int) -> Corpus:
    """Uniform random walks; walk i starts at starts[i] and has up to lengths[i]
    vertices (it stops early at a vertex without neighbours).  Returns the
    corpus with -1 after every walk and the per-vertex counts."""
    rng = np.random.default_rng(seed)
    n = rowptr.size - 1
    deg = np.diff(rowptr)
    lengths = np.asarray(lengths, np.int64)
    starts = np.asarray(starts, np.int64)
    nw = starts.size
    off = np.zeros(nw + 1, np.int64)
    np.cumsum(lengths + 1, out=off[1:])
    out = np.full(int(off[-1]), -2, np.int32)
    out[off[1:] - 1] = -1
    cur = starts.copy()
    alive = np.ones(nw, bool)
    Lmax = int(lengths.max()) if nw else 0
    for t in range(Lmax):
        idx = np.nonzero(alive & (lengths > t))[0]
        if idx.size == 0:
            break
        c = cur[idx]
        out[off[idx] + t] = c
        d = deg[c]
        ok = d > 0
        alive[idx[~ok]] = False
        idx, c, d = idx[ok], c[ok], d[ok]
        pick = (rng.random(idx.size) * d).astype(np.int64)
        cur[idx] = col[rowptr[c] + pick]
    tokens = out[out != -2]
    cn = np.bincount(tokens[tokens >= 0], minlength=n).astype(np.int32)
    return Corpus(tokens=np.ascontiguousarray(tokens, np.int32), vertex_cn=cn, n_walks=nw)


def holdout_split(u: np.ndarray, v: np.ndarray, n: int, frac: float, seed: int):
    """gtrain-style split (src/tools/gtrain.cpp:300-343).  Returns
    (train_u, train_v, test_pairs [2k, 2], test_labels [2k])."""
    rng = np.random.default_rng(seed)
    m = u.size
    k = int(frac * m)
    deg = np.bincount(np.concatenate([u, v]), minlength=n).astype(np.int64)
    removed = np.zeros(m, bool)
    order = rng.permutation(m)
    pos = []
    for e in order:
        if len(pos) == k:
            break
        a, b = u[e], v[e]
        if deg[a] > 1 and deg[b] > 1:
            deg[a] -= 1
            deg[b] -= 1
            removed[e] = True
            pos.append(e)
    pos = np.asarray(pos, np.int64)
    lo = np.minimum(u, v).astype(np.int64)
    hi = np.maximum(u, v).astype(np.int64)
    edge_keys = np.sort(lo << 32 | hi)
    neg_keys = np.empty(0, np.int64)
    while neg_keys.size < pos.size:
        need = (pos.size - neg_keys.size) * 2 + 64
        a = rng.integers(0, n, need, dtype=np.int64)
        b = rng.integers(0, n, need, dtype=np.int64)
        key = np.minimum(a, b) << 32 | np.maximum(a, b)
        ok = a != b
        i = np.searchsorted(edge_keys, key)
        ok &= ~((i < edge_keys.size) & (edge_keys[np.minimum(i, edge_keys.size - 1)] == key))
        key = key[ok]
        _, first = np.unique(key, return_index=True)
        key = key[np.sort(first)]
        key = key[~np.isin(key, neg_keys)]
        neg_keys = np.concatenate([neg_keys, key])[:pos.size]
    neg = np.stack([neg_keys >> 32, neg_keys & 0xFFFFFFFF], 1)
    test = np.concatenate([np.stack([u[pos], v[pos]], 1).astype(np.int64), neg.astype(np.int64)])
    labels = np.concatenate([np.ones(pos.size, np.int8), np.zeros(neg.shape[0], np.int8)])
    keep = ~removed
    return u[keep], v[keep], test, labels


def corpus_er(n=10_000, m=50_000, walks_per_vertex=10, length=80, seed=1) -> Corpus:
    """Erdos-Renyi n=10k, avg degree 10, 10 uniform walks of length 80 per vertex."""
    u, v = er_graph(n, m, seed)
    rowptr, col = csr(n, u, v)
    starts = np.repeat(np.arange(n, dtype=np.int64), walks_per_vertex)
    lengths = np.full(starts.size, length, np.int64)
    return walk_corpus(rowptr, col, starts, lengths, seed + 1000)


def lognormal_lengths(rng, count: int, mu: float, sigma: float, lo: int, hi: int) -> np.ndarray:
    return np.clip(np.rint(rng.lognormal(mu, sigma, count)), lo, hi).astype(np.int64)


def rmat_corpus(scale=21, avg_degree=16, abc=(0.57, 0.19, 0.19), seed=3, target_tokens=100_000_000,
                mu=3.7, sigma=0.5, lmin=5, lmax=120, holdout=0.0):
    """R-MAT graph and a corpus of ~target_tokens walk tokens with log-normal
    lengths clipped to [lmin, lmax], walks started at random non-isolated
    vertices.  With holdout > 0 the walks run on the training graph and the
    held-out test pairs are returned too.  Returns (corpus, test_pairs, labels)."""
    n = 1 << scale
    m = avg_degree * n // 2
    u, v, n = rmat_graph(scale, m, *abc, seed=seed)
    test = labels = None
    if holdout > 0:
        u, v, test, labels = holdout_split(u, v, n, holdout, seed + 7)
    rowptr, col = csr(n, u, v)
    rng = np.random.default_rng(seed + 11)
    mean = float(np.exp(mu + sigma * sigma / 2))
    count = int(target_tokens / min(mean, lmax) * 1.05) + 16
    lengths = lognormal_lengths(rng, count, mu, sigma, lmin, lmax)
    keep = np.searchsorted(np.cumsum(lengths), target_tokens, side="right")
    lengths = lengths[:max(1, keep)]
    live = np.nonzero(np.diff(rowptr) > 0)[0]
    starts = live[rng.integers(0, live.size, lengths.size)]
    corpus = walk_corpus(rowptr, col, starts, lengths, seed + 13)
    return corpus, test, labels


def auc(scores: np.ndarray, labels: np.ndarray) -> float:
    """ROC AUC by the Mann-Whitney U statistic with average ranks for ties."""
    from scipy.stats import rankdata
    scores = np.asarray(scores, np.float64)
    labels = np.asarray(labels).astype(bool)
    npos = int(labels.sum())
    nneg = labels.size - npos
    if npos == 0 or nneg == 0:
        return float("nan")
    r = rankdata(scores, method="average")
    return float((r[labels].sum() - npos * (npos + 1) / 2.0) / (npos * nneg))


def link_auc(emb_rank: np.ndarray, rank_id: np.ndarray, pairs: np.ndarray, labels: np.ndarray) -> float:
    """Link-prediction AUC of dot-product scores of rank-ordered embeddings."""
    n = int(max(rank_id.max() + 1, pairs.max() + 1))
    by_id = np.zeros((n, emb_rank.shape[1]), np.float32)
    by_id[rank_id] = emb_rank
    a = by_id[pairs[:, 0]]
    b = by_id[pairs[:, 1]]
    return auc(np.einsum("ij,ij->i", a, b, dtype=np.float64), labels)
