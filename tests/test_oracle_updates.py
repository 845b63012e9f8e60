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


def test_slots_minibatch_trace():
    """Window-slot update rules (the BF16-MFMA trainer's), K = 1, window 1,
    sen1 = [1, 2], sen2 = [3, 4]; negatives 5 and 5 (groups 1 then 0).  Every
    coefficient is 0 (label 1, saturated) or -a.

    One minibatch per step: step 1 (t1 = 1, negative 5, t2 = 3; contexts 2
    and 4) reads every output row before any update; step 2 (t1 = 2, 5,
    t2 = 4; contexts 1 and 3) reads Woh after step 1.  Minibatches of one
    context: the second context of a step reads the output rows the first one
    updated, so context 4 sees Woh[5] = d5 - a W2 and context 3 sees
    Woh[5] and Woh[2], Woh[4] after context 1's update.  Woh ends the same
    (saturated coefficients do not depend on the rows)."""
    W, d = U.initial_model()
    a = U.ALPHA
    corpus = np.array([1, 2, -1, 3, 4, -1] + U.PAD, np.int32)
    woh = d.copy()
    woh[1] = d[1] - a * W[4]
    woh[3] = d[3] - a * W[2]
    woh[5] = d[5] - a * (W[2] + W[4]) - a * (W[1] + W[3])
    woh[2] = d[2] - a * W[3]
    woh[4] = d[4] - a * W[1]
    whole = W.copy()
    whole[2] = W[2] - a * (d[5] + d[3])
    whole[4] = W[4] - a * (d[1] + d[5])
    whole[1] = W[1] - a * ((d[5] - a * (W[2] + W[4])) + d[4])
    whole[3] = W[3] - a * (d[2] + (d[5] - a * (W[2] + W[4])))
    split = whole.copy()
    split[4] = W[4] - a * (d[1] + (d[5] - a * W[2]))
    split[3] = W[3] - a * (d[2] + (d[5] - a * (W[2] + W[4]) - a * W[1]))
    for mb, want in ((0, whole), (4, whole), (1, split)):
        opts = O.make_opts(dim=U.D, window=1, negative=1, sample=0.0, alpha=float(a), table_size=10,
                           init_model=0, semantics="slots", minibatch=mb)
        wih, woh_got, st, _, _ = O.train(corpus, U.CN, opts, wih=W, woh=d)
        assert st.steps == 2 and st.sentences == 2
        assert np.array_equal(wih, want), (mb, wih, want)
        assert np.array_equal(woh_got, woh), (mb, woh_got, woh)
    assert not np.array_equal(split, whole)


@pytest.mark.parametrize("name", sorted(U.SLOT_CASES))
def test_oracle_slot_semantics(name):
    """The oracle's window-slot rules (semantics "slots", the BF16-MFMA
    window-slot trainer's) against the hand-worked traces of
    tests/update_cases.py."""
    case = U.SLOT_CASES[name]()
    wih0, woh0 = case["model"]()
    opts = O.make_opts(dim=U.D, window=1, negative=case["negative"], sample=0.0,
                       alpha=float(U.ALPHA), table_size=10, init_model=0, semantics="slots")
    wih, woh, st, _, _ = O.train(np.array(case["corpus"], np.int32), case["cn"], opts, wih=wih0, woh=woh0)
    assert st.steps == case["steps"] and st.sentences == case["sentences"]
    assert np.array_equal(wih, case["wih"]), (wih, case["wih"])
    assert np.array_equal(woh, case["woh"]), (woh, case["woh"])


def test_slot_traces_stay_saturated_and_tell_rules_apart():
    for name, make in U.SLOT_CASES.items():
        case = make()
        for m in (case["wih"], case["woh"]):
            rows = m[1:]
            assert np.all((rows[:, 0] >= 3) & (rows[:, 0] <= 4)), name
            assert np.all(np.abs(rows[:, 1]) < 2), name
    # keeping sen1's own copy of rank 1 across its absence from the window
    # (instead of reading Wih again) would end elsewhere
    case = U.slots_case_reentry_reads_wih()
    assert not np.array_equal(case["wih"][1], case["wrong_wih1"])
    # the reference's buffers on the slice-end corpus drop the negative
    # updates and rank 3's row; the window-slot rules keep them
    ref, slots = U.case_slice_end_drops_buffers(), U.slots_case_slice_end()
    assert not np.array_equal(ref["woh"][5], slots["woh"][5])
    assert not np.array_equal(ref["wih"][3], slots["wih"][3])
