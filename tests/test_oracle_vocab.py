"""Vocabulary ranking (dsgl.hpp:314-382, :246-252): std::sort by count,
descending, over (cn, id) records in id order."""
import numpy as np

from oracle import oracle as O


def test_permutation_and_order():
    rng = np.random.default_rng(0)
    cn = rng.integers(0, 50, 5000).astype(np.int32)
    rank_id, id_rank, tw = O.build_vocab(cn)
    assert sorted(rank_id.tolist()) == list(range(cn.size))
    assert np.all(np.diff(cn[rank_id]) <= 0)
    assert np.array_equal(id_rank[rank_id], np.arange(cn.size))
    assert tw == int(cn.astype(np.int64).sum())


def test_distinct_counts_match_python_sort():
    rng = np.random.default_rng(1)
    cn = rng.permutation(3000).astype(np.int32)
    rank_id, _, _ = O.build_vocab(cn)
    assert rank_id.tolist() == sorted(range(cn.size), key=lambda i: -cn[i])


def test_small_inputs_keep_id_order_on_ties():
    # libstdc++ sorts 16 or fewer elements by insertion sort, which is stable:
    # equal counts keep increasing vertex id.
    rng = np.random.default_rng(2)
    for n in (2, 5, 9, 16):
        for _ in range(20):
            cn = rng.integers(0, 3, n).astype(np.int32)
            rank_id, _, _ = O.build_vocab(cn)
            expect = sorted(range(n), key=lambda i: (-cn[i], i))
            assert rank_id.tolist() == expect
