"""Hand-worked traces of the oracle trainer's index stream (dsgl.hpp:786-1408).

Random numbers come from the LCG in closed form (Python integers, pinned in
test_oracle_lcg.py); the structure of each expected step (which sentence is
loaded when, window contexts, negative groups, write-back timing, epoch end)
is worked out by hand from the reference loop and written as literals."""
import numpy as np

from oracle import oracle as O

A = 25214903917
M = 1 << 64


def lcg_seq(seed, n):
    out, s = [], seed
    for _ in range(n):
        s = (s * A + 11) % M
        out.append(s)
    return out


def run(corpus, cn, **kw):
    opts = O.make_opts(**kw)
    wih, woh, st, rec, ws = O.train(np.array(corpus, np.int32), np.array(cn, np.int32), opts,
                                    rec_cap=1000)
    recs = [O.decode_records(rec[w], int(ws[w]), opts.window, opts.negative)
            for w in range(opts.workers)]
    return st, recs


# Vertex 0 is heavy (p > 0.9 on a 10-slot table), so every table draw is rank
# 0 and each negative is s % (V-1) + 1 (dsgl.hpp:1064-1067).  Ties among ids
# 1..7 keep id order (insertion sort), so rank == id.
CN8 = [10000, 1, 1, 1, 1, 1, 1, 1]


def test_trace_two_sentence_schedule():
    corpus = [1, 2, 3, -1, 4, 5, -1, 6, 7, -1, 1, 3, -1]
    st, (recs,) = run(corpus, CN8, dim=8, window=1, negative=1, sample=0.0, table_size=10)
    s = lcg_seq(0, 6)
    neg = [x % 7 + 1 for x in s[:3]]  # refresh: max(3, 2) groups of K=1 draws
    # window draws s4, s5, s6 (b = s % 1 = 0); negative group = step % 3
    expect = [
        dict(t1=1, t2=4, inputs=[2, 5], fin=1, grp=1),
        dict(t1=2, t2=5, inputs=[1, 3, 4], fin=2, grp=2),   # sen2 [4,5] ends here
        dict(t1=3, t2=6, inputs=[2, 7], fin=1, grp=0),      # sen2 reloaded [6,7]; sen1 ends
    ]
    # then sen1 reloads [1,3] and consumes the last token: the slice is done and
    # both in-flight sentences are dropped (dsgl.hpp:1031-1050)
    assert len(recs) == 3
    assert st.steps == 3 and st.raw_words == 13 and st.sentences == 2
    for r, e in zip(recs, expect):
        assert (r["t1"], r["t2"], r["inputs"], r["fin"], r["grp"], r["b"]) == \
            (e["t1"], e["t2"], e["inputs"], e["fin"], e["grp"], 0)
        n = neg[e["grp"]]
        outs = [e["t1"]] + ([n] if n != e["t1"] else []) + [e["t2"]]
        assert r["outs"] == outs and r["meta"] == [1] * len(outs)


def test_trace_window_draws():
    # window 3: b = s % 3, half-width 3 - b (dsgl.hpp:1100-1132)
    walk1 = [1, 2, 3, 4, 5, 6, 7]
    walk2 = [7, 6, 5, 4, 3, 2, 1]
    corpus = walk1 + [-1] + walk2 + [-1] + [1, 2, -1]
    st, (recs,) = run(corpus, CN8, dim=8, window=3, negative=1, sample=0.0, table_size=10)
    s = lcg_seq(0, 7 + 7)
    wdraws = s[7:]  # 7 negative draws (max(7,7) x K=1), then one window draw per step
    assert len(recs) == 7
    for p, r in enumerate(recs):
        b = wdraws[p] % 3
        hw = 3 - b
        ctx1 = [walk1[c] for c in range(max(0, p - hw), min(6, p + hw) + 1) if c != p]
        ctx2 = [walk2[c] for c in range(max(0, p - hw), min(6, p + hw) + 1) if c != p]
        assert r["b"] == b
        assert r["inputs"] == ctx1 + ctx2 and r["fin"] == len(ctx1)


def test_trace_subsampling_draws():
    # Rank w >= 1 has ratio 1e-5 * 10007 / 1 and keep threshold sqrt(r)+r =
    # 0.4164; a token is kept iff threshold >= (s & 0xFFFF)/65536 with one
    # draw per non-separator token (dsgl.hpp:921-926).
    rng = np.random.default_rng(0)
    wa = rng.integers(1, 8, 40).tolist()
    wb = rng.integers(1, 8, 40).tolist()
    corpus = wa + [-1] + wb + [-1] + [1, 2, 3, -1]
    thr = O.keep_thresholds(np.array(CN8, np.int32), sum(CN8), 1e-5)
    s = lcg_seq(0, 80)
    kept_a = [w for w, x in zip(wa, s[:40]) if not thr[w] < np.float32(x & 0xFFFF) / np.float32(65536)]
    kept_b = [w for w, x in zip(wb, s[40:]) if not thr[w] < np.float32(x & 0xFFFF) / np.float32(65536)]
    assert 5 < len(kept_a) < 35 and 5 < len(kept_b) < 35
    st, (recs,) = run(corpus, CN8, dim=8, window=2, negative=2, sample=1e-5, table_size=10)
    n = min(len(kept_a), len(kept_b))
    assert [r["t1"] for r in recs[:n]] == kept_a[:n]
    assert [r["t2"] for r in recs[:n]] == kept_b[:n]


def test_trace_duplicate_negatives_and_target_drop():
    # V = 3: every negative is s % 2 + 1, so K = 5 draws repeat; repeats raise
    # the multiplicity, a negative equal to t1 is dropped, t2 is appended
    # without a duplicate check (dsgl.hpp:1136-1158).
    cn = [10000, 1, 1]
    corpus = [1, 2, 1, 2, 1, -1, 2, 1, 2, 1, 2, -1, 1, 1, -1]
    st, (recs,) = run(corpus, cn, dim=8, window=2, negative=5, sample=0.0, table_size=10)
    s = lcg_seq(0, 25)
    negs = [x % 2 + 1 for x in s[:25]]  # max(5,5) groups of 5
    assert len(recs) == 5
    for i, r in enumerate(recs):
        grp = (i + 1) % 5
        assert r["grp"] == grp
        g = negs[grp * 5:(grp + 1) * 5]
        uniq = []
        for v in g:
            if v not in uniq and v != r["t1"]:
                uniq.append(v)
        assert r["outs"] == [r["t1"]] + uniq + [r["t2"]]
        assert r["meta"] == [1] + [g.count(v) for v in uniq] + [1]


def test_trace_worker_slices():
    # Three workers: slices of N//3 tokens, the last takes the remainder; seed
    # = worker index; a slice that starts mid-walk starts with that partial
    # walk (dsgl.hpp:791, :811-821).
    walks = [[1, 2, 3, 4], [5, 6, 7], [2, 4, 6, 1], [3, 5, 7, 2], [4, 6, 1, 3], [7, 5, 3, 1]]
    corpus = []
    for w in walks:
        corpus += w + [-1]
    N = len(corpus)  # 29
    st, recs = run(corpus, CN8, dim=8, window=1, negative=1, sample=0.0, table_size=10, workers=3,
                   threads=4)
    per = N // 3
    starts = [0, per, 2 * per]
    for w in range(3):
        seg = corpus[starts[w]:(starts[w] + per if w < 2 else N)]
        first = seg[:seg.index(-1)] if -1 in seg else seg
        if len(first) >= 2 and recs[w]:
            assert recs[w][0]["t1"] == first[0]


def test_alpha_decay_chunks():
    rng = np.random.default_rng(9)
    V = 1200
    walks = [rng.integers(0, V, 30) for _ in range(1500)]
    corpus = np.concatenate([np.append(w, -1) for w in walks]).astype(np.int32)
    cn = np.bincount(corpus[corpus >= 0], minlength=V).astype(np.int32)
    opts = O.make_opts(dim=8, window=3, negative=2, sample=0.0, table_size=100000, alpha=0.025)
    _, _, st, rec, ws = O.train(corpus, cn, opts, rec_cap=40000)
    recs = O.decode_records(rec[0], int(ws[0]), 3, 2)
    alphas = np.array([r["alpha"] for r in recs], np.float32)
    assert alphas[0] == np.float32(0.025)
    assert np.all(np.diff(alphas) <= 0)
    changes = np.nonzero(np.diff(alphas))[0]
    # one update per >10000 consumed tokens (dsgl.hpp:886-897)
    assert abs(len(changes) - corpus.size // 10031) <= 2
    tw = int(cn.sum())
    a1 = alphas[changes[0] + 1]
    cands = [np.float32(0.025) * (np.float32(1) - np.float32(x) / np.float32(tw + 1))
             for x in range(10001, 10001 + 2 * 31 + 2)]  # up to two loads of 31 tokens
    assert any(a1 == c for c in cands)


def test_step_math_against_float64_products():
    rng = np.random.default_rng(11)
    t = O.exp_table()
    for os_, nin, fin, D in [(7, 40, 20, 128), (3, 5, 5, 64), (6, 13, 0, 100)]:
        Om = rng.normal(0, 0.3, (os_, D)).astype(np.float32)
        Im = rng.normal(0, 0.3, (nin, D)).astype(np.float32)
        meta = rng.integers(1, 3, os_).astype(np.int32)
        corr64, md64, gin64, sig64 = O.step_f64(Om, Im, fin, meta, 0.01, t)
        # float64 matrix products of the same coefficients
        assert np.allclose(md64, corr64 @ Im.astype(np.float64), rtol=1e-12, atol=1e-15)
        assert np.allclose(gin64, corr64.T @ Om.astype(np.float64), rtol=1e-12, atol=1e-15)
        f = Om.astype(np.float64) @ Im.astype(np.float64).T
        lab = np.zeros((os_, nin))
        lab[0, :fin] = 1
        lab[os_ - 1, fin:] = 1
        idx = np.clip(((f + 6) * 83).astype(np.int64), 0, 1000)
        g = np.where(f > 6, lab - 1, np.where(f < -6, lab, lab - t[idx].astype(np.float64))) * 0.01
        assert np.allclose(corr64, g * meta[:, None], rtol=1e-12, atol=1e-15)
        corr32, md32, gin32, sig32 = O.step_f32(Om, Im, fin, meta, 0.01, t)
        ok = sig32 == sig64
        assert ok.mean() > 0.99
        scale = np.abs(md64).max()
        assert np.max(np.abs(md32 - md64)) <= 1e-5 * scale + 1e-6 * np.abs(corr64).max()


def test_step_labels_by_hand():
    # Zero output rows give f = 0 -> table index 498; contexts before fin have
    # row 0 positive, the rest have the last row positive (dsgl.hpp:1249-1270).
    t = O.exp_table()
    Om = np.zeros((4, 8), np.float32)
    Im = np.ones((5, 8), np.float32)
    corr, _, _, sig = O.step_f32(Om, Im, 2, np.array([1, 2, 1, 1], np.int32), 0.5, t)
    assert np.all(sig == 498)
    s = t[498]
    expect = np.array([
        [(1 - s) * 0.5] * 2 + [(0 - s) * 0.5] * 3,
        [(0 - s) * 0.5 * 2] * 5,
        [(0 - s) * 0.5] * 5,
        [(0 - s) * 0.5] * 2 + [(1 - s) * 0.5] * 3,
    ], np.float32)
    assert np.array_equal(corr, expect)


def test_trace_negative_sharing():
    # share 4 (two steps per negative group): the refresh draws
    # ceil(max(3, 2) / 2) = 2 groups of K=1, then one window draw per step;
    # step n (neg_count n) uses group (n >> 1) mod 2 (window-slot rules, whose
    # index stream is the reference's at share 2)
    corpus = [1, 2, 3, -1, 4, 5, -1, 6, 7, -1, 1, 3, -1]
    for share, ngrp in ((2, 3), (4, 2), (8, 1)):
        st, (recs,) = run(corpus, CN8, dim=8, window=1, negative=1, sample=0.0, table_size=10,
                          semantics="slots", share=share)
        s = lcg_seq(0, ngrp + 3)
        neg = [x % 7 + 1 for x in s[:ngrp]]
        shift = {2: 0, 4: 1, 8: 2}[share]
        assert len(recs) == 3 and st.steps == 3
        for n, (r, t1, t2) in enumerate(zip(recs, (1, 2, 3), (4, 5, 6)), start=1):
            g = (n >> shift) % ngrp
            assert r["grp"] == g and r["b"] == 0 and (r["t1"], r["t2"]) == (t1, t2)
            v = neg[g]
            assert r["outs"] == [t1] + ([v] if v != t1 else []) + [t2]
