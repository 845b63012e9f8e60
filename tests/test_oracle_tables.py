"""Model init (dsgl.hpp:433-457), sigmoid table (dsgl.hpp:1557-1562),
subsampling thresholds (dsgl.hpp:921-926) and the unigram^0.75 noise table
(dsgl.hpp:121-147) of the oracle."""
import numpy as np

from oracle import oracle as O

A = 25214903917
M = 1 << 64


def test_init_net_first_values_by_hand():
    D = 128
    w = O.init_net(3, D)
    s = 1
    expect = []
    for _ in range(3 * D):
        s = (s * A + 11) % M
        x = np.float32(s & 0xFFFF) / np.float32(65536.0)
        expect.append((x - np.float32(0.5)) / np.float32(D))
    assert np.array_equal(w.reshape(-1), np.array(expect, np.float32))
    # s1 = A + 11 = 25214903928, whose low 16 bits are 59000
    assert (25214903928 & 0xFFFF) == 59000
    assert w[0, 0] == np.float32((59000 / 65536 - 0.5) / 128)


def test_init_net_range_and_mean():
    D = 64
    w = O.init_net(2000, D)
    assert w.min() >= -0.5 / D and w.max() < 0.5 / D
    assert abs(float(w.mean())) < 1e-4


def test_exp_table_closed_form():
    t = O.exp_table()
    x = (np.arange(1001) / 1000.0 * 2 - 1) * 6
    ref = 1.0 / (1.0 + np.exp(-x))
    assert np.max(np.abs(t.astype(np.float64) - ref)) < 2e-7
    assert t[500] == np.float32(0.5)
    assert np.all(np.diff(t) >= 0)
    assert np.max(np.abs(t[:501].astype(np.float64) + t[::-1][:501] - 1.0)) < 2e-7


def test_keep_thresholds_closed_form_and_rank_indexing():
    rng = np.random.default_rng(3)
    cn = rng.integers(1, 5000, 300).astype(np.int32)
    tw = int(cn.sum())
    thr = O.keep_thresholds(cn, tw, 1e-3)
    r = 1e-3 * tw / cn.astype(np.float64)
    assert np.allclose(thr, np.sqrt(r) + r, rtol=3e-7)
    # The threshold of rank w reads cn_vocab[w] (a per-id array), not the
    # count of the vertex holding rank w.
    cn2 = cn.copy()
    cn2[7] = 1
    thr2 = O.keep_thresholds(cn2, tw, 1e-3)
    assert thr2[7] != thr[7] and np.array_equal(np.delete(thr2, 7), np.delete(thr, 7))


def test_unigram_table_hand_examples():
    # V=2, cn=[3,1], T=10: p0 = 3^.75/(3^.75+1) = 0.6951; slots 0..7 hold
    # rank 0 (the advance happens at slot 7 where 0.7 > 0.6951).
    assert O.unigram_table(np.array([3, 1], np.int32), 10).tolist() == [0] * 8 + [1] * 2
    # p0 = 31.62/33.62 = 0.9405 > 0.9: rank 0 fills the whole table.
    assert O.unigram_table(np.array([100, 1, 1], np.int32), 10).tolist() == [0] * 10
    # Five equal counts (p = 0.2): the float comparison a/T > d1 is strict and
    # the table advances one rank per slot at most, so rank 4 never appears.
    assert O.unigram_table(np.array([1] * 5, np.int32), 10).tolist() == [0, 0, 0, 0, 1, 1, 2, 2, 3, 3]


def test_unigram_table_invariants_and_shares():
    rng = np.random.default_rng(4)
    cn = np.sort(rng.integers(50, 5000, 200))[::-1].astype(np.int32)
    T = 1_000_000
    t = O.unigram_table(cn, T)
    assert t[0] == 0 and t[-1] == cn.size - 1
    d = np.diff(t)
    assert np.all((d == 0) | (d == 1))
    p = cn.astype(np.float64) ** 0.75
    p /= p.sum()
    cum = np.cumsum(p)
    # first slot holding rank k+1 equals floor(T*cum_k)+2 up to float rounding
    starts = np.searchsorted(t, np.arange(1, cn.size), side="left")
    assert np.max(np.abs(starts - (np.floor(T * cum[:-1]) + 2))) <= 2


def test_unigram_pow_sum_split():
    cn = np.arange(1, 1001, dtype=np.int32)
    exact = float(np.sum(cn.astype(np.float64) ** 0.75))
    for th in (1, 2, 3, 12):
        assert abs(O.unigram_pow_sum(cn, th) - exact) / exact < 1e-14
