"""Hand-worked model-update traces of the DSGL trainer (include/dsgl.hpp).

Each case is a tiny corpus trained by one worker from a chosen model, with
every dot product far above 6, so that the table sigmoid saturates
(dsgl.hpp:1249-1270: g = (label - 1) * alpha) and every coefficient is 0 or
-alpha.  Rows are (4, c) with small dyadic c and alpha = 1/16, so every sum in
the step is exact in float32 and the expected Wih / Woh below can be written
down step by step from the reference loop, independently of any summation
order.  The expectations were worked out by hand from:

  * step assembly and labels         dsgl.hpp:1089-1160, :1241-1272
  * GEMM2 / GEMM3 on staged rows     dsgl.hpp:1290, :1309
  * scatter: contexts into the sentence buffers, targets straight into Woh,
    negatives into the negative buffer                  dsgl.hpp:1314-1368
  * write-backs: a finished sentence's buffer overwrites its Wih rows, the
    negative buffer overwrites its Woh rows after two finished sentences
                                                        dsgl.hpp:1377-1407
  * slice end: in-flight sentences and the negative buffer are dropped
                                                        dsgl.hpp:1031-1050

Vertex 0 is heavy, so on a 10-slot noise table every draw is rank 0 and each
negative is s % 7 + 1 (dsgl.hpp:1064-1067); from seed 0 the draws give
negatives 5, 5, 3, 4, 1, 7, ... .  Ties among ids 1..7 keep id order, so
rank == vertex id.
"""
import numpy as np

V = 8
D = 2
CN = np.array([10000, 1, 1, 1, 1, 1, 1, 1], np.int32)
ALPHA = np.float32(1.0 / 16.0)
# A one-token walk at the end of every corpus: loading the second sentence
# must not reach the slice end (that would drop both sentences before the
# first step, dsgl.hpp:1031); the walk itself is too short to train
# (dsgl.hpp:935-939).
PAD = [7, -1]


def initial_model():
    """Wih row r = (4, r/8), Woh row r = (4, -r/16); row 0 is never touched."""
    wih = np.zeros((V, D), np.float32)
    woh = np.zeros((V, D), np.float32)
    for r in range(1, V):
        wih[r] = (4.0, r / 8.0)
        woh[r] = (4.0, -r / 16.0)
    return wih, woh


def case_target_is_pending_negative():
    """K = 1.  sen1 = [5, 2], sen2 = [3, 6]; negative groups [5], [5].

    Step 1 (group 1 = 5): t1 = 5, so the negative equal to t1 is dropped;
    Woh[5] += -a W6 directly.  Step 2 (group 0 = 5): outputs [2, 5, 6]; the
    staged row of negative 5 is the refresh snapshot d5 (step 1's direct
    update is not in it).  After step 2 both sentences end and the negative
    buffer overwrites Woh[5]: step 1's direct update is lost."""
    W, d = initial_model()
    a = ALPHA
    corpus = [5, 2, -1, 3, 6, -1] + PAD
    wih, woh = W.copy(), d.copy()
    woh[3] = d[3] - a * W[2]
    woh[2] = d[2] - a * W[3]
    woh[6] = d[6] - a * W[5]
    woh[5] = d[5] - a * (W[5] + W[3])          # not d5 - a W6 - a (W5 + W3)
    wih[2] = W[2] - a * d[3]
    wih[5] = W[5] - a * (d[5] + d[6])
    wih[3] = W[3] - a * (d[2] + d[5])
    wih[6] = W[6] - a * d[5]
    return dict(corpus=corpus, negative=1, wih=wih, woh=woh, steps=2, sentences=2)


def case_repeated_context():
    """K = 0.  sen1 = [1, 2, 1], sen2 = [3, 4, 7].  At step 2 the window of
    t1 = 2 holds rank 1 twice: its buffer row takes both updates in context
    order (W1 - a d4 - a d4).  Step 3 reads the rows step 1 updated."""
    W, d = initial_model()
    a = ALPHA
    corpus = [1, 2, 1, -1, 3, 4, 7, -1] + PAD
    wih, woh = W.copy(), d.copy()
    # step 1: t1 = 1 (ctx 2), t2 = 3 (ctx 4)
    b2 = W[2] - a * d[3]
    b4 = W[4] - a * d[1]
    o1 = d[1] - a * W[4]
    woh[3] = d[3] - a * W[2]
    # step 2: t1 = 2 (ctx 1, 1), t2 = 4 (ctx 3, 7)
    wih[1] = W[1] - a * d[4] - a * d[4]
    wih[3] = W[3] - a * d[2]
    wih[7] = W[7] - a * d[2]
    woh[2] = d[2] - a * (W[3] + W[7])
    woh[4] = d[4] - a * (W[1] + W[1])
    # step 3: t1 = 1 (ctx 2 from the buffer), t2 = 7 (ctx 4 from the buffer)
    woh[1] = o1 - a * b4
    woh[7] = d[7] - a * b2
    wih[2] = b2 - a * d[7]
    wih[4] = b4 - a * o1
    return dict(corpus=corpus, negative=0, wih=wih, woh=woh, steps=3, sentences=2)


def case_shared_word_second_sentence_wins():
    """K = 0.  sen1 = [1, 2], sen2 = [3, 2]: rank 2 sits in both sentence
    buffers.  Step 2 has t1 = t2 = 2 (Woh[2] takes both updates), both
    sentences end on it, and sen2's write-back of rank 2 comes last."""
    W, d = initial_model()
    a = ALPHA
    corpus = [1, 2, -1, 3, 2, -1] + PAD
    wih, woh = W.copy(), d.copy()
    woh[1] = d[1] - a * W[2]
    woh[3] = d[3] - a * W[2]
    woh[2] = d[2] - a * W[3] - a * W[1]
    wih[1] = W[1] - a * d[2]
    wih[3] = W[3] - a * d[2]
    wih[2] = W[2] - a * d[1]                   # sen2's buffer; sen1's was W2 - a d3
    return dict(corpus=corpus, negative=0, wih=wih, woh=woh, steps=2, sentences=2)


def case_slice_end_drops_buffers():
    """K = 1.  sen1 = [1, 2], sen2 = [3, 4, 5, 6]; 4 negative groups
    [5], [5], [3], [4] from draws 1-4.  Step 1 (group 1): outputs [1, 5, 3];
    step 2 (group 2): outputs [2, 3, 4] with negative 3 staged from the
    refresh snapshot d3.  sen1 ends after step 2 and is written back; then
    the corpus slice ends: sen2's buffer (ranks 3-6) and the negative buffer
    (ranks 5, 3) are dropped, while the direct target updates stay."""
    W, d = initial_model()
    a = ALPHA
    corpus = [1, 2, -1, 3, 4, 5, 6, -1] + PAD
    wih, woh = W.copy(), d.copy()
    woh[1] = d[1] - a * W[4]
    woh[3] = d[3] - a * W[2]                   # target update kept, negative update dropped
    woh[2] = d[2] - a * (W[3] + W[5])
    woh[4] = d[4] - a * W[1]
    # woh[5] stays d5: its negative-buffer update is dropped
    wih[1] = W[1] - a * (d[3] + d[4])
    wih[2] = W[2] - a * (d[5] + d[3])
    # wih[3..6] stay: sen2 never finished
    return dict(corpus=corpus, negative=1, wih=wih, woh=woh, steps=2, sentences=1)


CASES = {
    "target_is_pending_negative": case_target_is_pending_negative,
    "repeated_context": case_repeated_context,
    "shared_word_second_sentence_wins": case_shared_word_second_sentence_wins,
    "slice_end_drops_buffers": case_slice_end_drops_buffers,
}


# ---- window-slot semantics (the BF16-MFMA window-slot trainer's rules) -------
# Rules that differ from the reference's buffers (the oracle's "slots"
# semantics): a window row is written back to Wih when the last occurrence of
# its rank leaves the window (or at the sentence end), and a rank entering the
# window reads its row from Wih at the step before it enters; negative rows
# take their updates in Woh directly, like targets; a rank twice in one window
# has one row.  The expectations below are worked by hand from these rules and
# the step of dsgl.hpp:1241-1333 (same saturated coefficients as above).

V16 = 16
# Sixteen ids, id 0 heavy and ids 1..15 tied: libstdc++ sorts at most 16
# elements by insertion sort, which keeps the tie in id order, so rank == id.
CN16 = np.array([100000] + [1] * 15, np.int32)


def initial_model16():
    """As initial_model() over 16 rows: Wih row r = (4, r/8), Woh row r = (4, -r/16)."""
    wih = np.zeros((V16, D), np.float32)
    woh = np.zeros((V16, D), np.float32)
    for r in range(1, V16):
        wih[r] = (4.0, r / 8.0)
        woh[r] = (4.0, -r / 16.0)
    return wih, woh


def slots_case_reentry_reads_wih():
    """Window 1, K = 0.  sen1 = [1, 2, 3, 4, 5, 1], sen2 = [6, 1, 7, 8, 9, 10].

    Rank 1 leaves sen1's window at the transition after step 2 and is
    written back; sen2 holds its own row of rank 1 and writes it back one
    transition later (after step 3).  At step 4 rank 1 enters sen1's window
    again and its row is read from Wih: sen2's version.  Step p of the
    sentence pair trains t1 = sen1[p-1], t2 = sen2[p-1]; with K = 0 the
    outputs are [t1, t2], sen1's contexts take -a O[t2] and sen2's -a O[t1],
    and md[t1] = -a (sen2's contexts), md[t2] = -a (sen1's contexts)."""
    W, d = initial_model16()
    a = ALPHA
    corpus = [1, 2, 3, 4, 5, 1, -1, 6, 1, 7, 8, 9, 10, -1] + PAD
    wih, woh = W.copy(), d.copy()
    o1 = d[1] - a * W[1]                             # Woh[1] after step 1
    o1 = o1 - a * (W[1] + W[3])                      # after step 2 (t2 = 1)
    b1 = W[1] - a * d[1] - a * d[3]                  # sen2's rank-1 row, written back after step 3
    reread = b1                                      # what sen1 reads for rank 1 at step 4
    wih[1] = reread - a * d[9]                       # step 5 (context of t1 = 5, t2 = 9), sentence end
    wih[2] = W[2] - a * d[6] - a * d[7]              # steps 1, 3; left after step 3
    wih[3] = W[3] - a * (d[1] - a * W[1]) - a * d[8]  # steps 2, 4; left after step 4
    wih[4] = W[4] - a * d[7] - a * d[9]              # steps 3, 5; left after step 5
    wih[5] = W[5] - a * d[8] - a * d[10]             # steps 4, 6; sentence end
    wih[6] = W[6] - a * d[2]                         # step 2; left after step 2
    wih[7] = W[7] - a * d[2] - a * d[4]              # steps 2, 4; left after step 4
    wih[8] = W[8] - a * d[3] - a * d[5]              # steps 3, 5; left after step 5
    wih[9] = W[9] - a * d[4] - a * o1                # steps 4, 6 (t1 = 1 at step 6 reads o1)
    wih[10] = W[10] - a * d[5]                       # step 5; sentence end
    woh[1] = o1 - a * (W[9] - a * d[4])              # step 6: t1 = 1, sen2's context is rank 9
    woh[2] = d[2] - a * (W[6] + W[7])                # step 2: t1 = 2
    woh[3] = d[3] - a * (W[1] - a * d[1] + W[8])     # step 3: t1 = 3, sen2's contexts 1 and 8
    woh[4] = d[4] - a * (W[7] - a * d[2] + W[9])     # step 4: t1 = 4
    woh[5] = d[5] - a * (W[8] - a * d[3] + W[10])    # step 5: t1 = 5
    woh[6] = d[6] - a * W[2]                         # step 1: t2 = 6
    woh[7] = d[7] - a * (W[2] - a * d[6] + W[4])     # step 3: t2 = 7
    woh[8] = d[8] - a * (W[3] - a * (d[1] - a * W[1]) + W[5])  # step 4: t2 = 8
    woh[9] = d[9] - a * (W[4] - a * d[7] + reread)   # step 5: t2 = 9, sen1's contexts 4 and 1
    woh[10] = d[10] - a * (W[5] - a * d[8])          # step 6: t2 = 10
    kept = W[1] - a * (d[1] - a * W[1]) - a * d[9]   # what keeping sen1's own row would give
    return dict(corpus=corpus, negative=0, cn=CN16, model=initial_model16, wih=wih, woh=woh,
                steps=6, sentences=2, wrong_wih1=kept)


def slots_case_slice_end():
    """Window 1, K = 1: the corpus of case_slice_end_drops_buffers under the
    window-slot rules.  sen1 = [1, 2], sen2 = [3, 4, 5, 6]; groups [5], [5],
    [3], [4].  Step 1 (group 1): outputs [1, 5, 3]; step 2 (group 2):
    outputs [2, 3, 4].  The negative updates of Woh[5] and Woh[3] land in
    Woh directly and stay; rank 3 leaves sen2's window after step 2 and is
    written back; sen1 ends and is written back; then the slice ends and
    sen2's rows still in its window (ranks 4, 5, 6) are dropped."""
    W, d = initial_model()
    a = ALPHA
    corpus = [1, 2, -1, 3, 4, 5, 6, -1] + PAD
    wih, woh = W.copy(), d.copy()
    o3 = d[3] - a * W[2]                              # Woh[3] after step 1 (t2 = 3)
    woh[1] = d[1] - a * W[4]
    woh[5] = d[5] - a * (W[2] + W[4])                 # negative of step 1: both sentences' contexts
    woh[3] = o3 - a * (W[1] + W[3] + W[5])            # negative of step 2
    woh[2] = d[2] - a * (W[3] + W[5])
    woh[4] = d[4] - a * W[1]
    wih[1] = W[1] - a * (o3 + d[4])                   # step 2: context of t1 = 2, outputs 3 and 4
    wih[2] = W[2] - a * (d[5] + d[3])                 # step 1: context of t1 = 1, outputs 5 and 3
    wih[3] = W[3] - a * (d[2] + o3)                   # step 2: context of t2 = 4; left after step 2
    # wih[4..6] stay: sen2 never finished and its window rows are dropped
    return dict(corpus=corpus, negative=1, cn=CN, model=initial_model, wih=wih, woh=woh, steps=2, sentences=1)


# Cases whose rules coincide under both semantics: the same expectations.
SLOT_CASES = {
    "reentry_reads_wih": slots_case_reentry_reads_wih,
    "slice_end": slots_case_slice_end,
    "repeated_context": lambda: dict(case_repeated_context(), cn=CN, model=initial_model),
    "shared_word_second_sentence_wins": lambda: dict(case_shared_word_second_sentence_wins(), cn=CN,
                                                     model=initial_model),
}
