"""The oracle's model-update semantics against hand-worked traces.

The cases (tests/update_cases.py) pin what the index-stream traces cannot:
which writes reach Wih / Woh and in which order (dsgl.hpp:1314-1407), and
what the end of a corpus slice throws away (dsgl.hpp:1031-1050).  All
arithmetic in them is exact, so the comparison is exact."""
import numpy as np
import pytest

from oracle import oracle as O
import update_cases as U


@pytest.mark.parametrize("name", sorted(U.CASES))
def test_oracle_update_semantics(name):
    case = U.CASES[name]()
    wih0, woh0 = U.initial_model()
    opts = O.make_opts(dim=U.D, window=1, negative=case["negative"], sample=0.0,
                       alpha=float(U.ALPHA), table_size=10, init_model=0)
    wih, woh, st, _, _ = O.train(np.array(case["corpus"], np.int32), U.CN, opts, wih=wih0, woh=woh0)
    assert st.steps == case["steps"] and st.sentences == case["sentences"]
    assert np.array_equal(wih, case["wih"]), (wih, case["wih"])
    assert np.array_equal(woh, case["woh"]), (woh, case["woh"])


def test_update_cases_stay_saturated():
    # Every row keeps a first component in [3, 4] and a second below 1 in
    # magnitude, so every dot product stays above 9 > 6 (the saturated
    # branch the traces assume).
    for name, make in U.CASES.items():
        case = make()
        for m in (case["wih"], case["woh"]):
            rows = m[1:]
            assert np.all((rows[:, 0] >= 3) & (rows[:, 0] <= 4)), name
            assert np.all(np.abs(rows[:, 1]) < 1), name


def test_mutated_oracle_semantics_fail_the_traces():
    # A restatement that wrote target updates after the negative write-back
    # (no clobber) or kept sen1's copy of a shared word would disagree with
    # the hand traces; check that the traces tell those apart.
    W, d = U.initial_model()
    a = U.ALPHA
    case = U.case_target_is_pending_negative()
    assert not np.array_equal(case["woh"][5], d[5] - a * W[6] - a * (W[5] + W[3]))
    case = U.case_shared_word_second_sentence_wins()
    assert not np.array_equal(case["wih"][2], W[2] - a * d[3])
    case = U.case_slice_end_drops_buffers()
    assert not np.array_equal(case["woh"][5], d[5] - a * (W[2] + W[4]))
