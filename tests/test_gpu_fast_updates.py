"""The BF16-MFMA window-slot trainer (sgns_fast.hip) against the hand-worked
window-slot traces of tests/update_cases.py, through the C ABI, one worker.

The traces' two-element rows sit in the first two elements of 128-element
rows (the trainer's row length), the rest zero.  Every dot product stays far
above 6, so each coefficient is exactly 0 or -alpha (or -2 alpha for a rank
twice in a window); what remains is the rounding of bf16 row storage and bf16
contraction operands, at most a few bf16 steps of rows near 4 (0.016 each).
The rules the traces tell apart move a row by a quarter or more (a row read
again from Wih on re-entry instead of kept, negative updates kept instead of
dropped at the slice end, a row written back when it leaves the window), so a
0.02 tolerance checks the rules, not the rounding."""
import numpy as np
import pytest

import dger_sgns as G
import update_cases as U

pytestmark = pytest.mark.gpu

TOL = 0.02


def widen(m):
    out = np.zeros((m.shape[0], 128), np.float32)
    out[:, :m.shape[1]] = m
    return out


@pytest.mark.parametrize("name", sorted(U.SLOT_CASES))
def test_window_slot_trainer_follows_slot_traces(name):
    case = U.SLOT_CASES[name]()
    print(f"slot trace {name}", flush=True)
    wih0, woh0 = case["model"]()
    opts = G.default_opts(size=128, window=1, negative=case["negative"], sample=0.0, alpha=float(U.ALPHA),
                          iter=1, threads=2, table_size=10, precision="bf16-mfma", debug=0)
    with G.Trainer(opts, case["cn"], np.array(case["corpus"], np.int32)) as tr:
        tr.set_model(widen(wih0), widen(woh0))
        st, _, _ = tr.train(workers=1)
        wih, woh = tr.get_model()
    assert st.steps == case["steps"] and st.sentences == case["sentences"]
    assert not wih[:, 2:].any() and not woh[:, 2:].any()
    err_i = np.abs(wih[:, :2] - case["wih"]).max()
    err_o = np.abs(woh[:, :2] - case["woh"]).max()
    print(f"  max |GPU - trace|: Wih {err_i:.2e}, Woh {err_o:.2e}", flush=True)
    assert err_i <= TOL and err_o <= TOL
    if "wrong_wih1" in case:
        # the row of rank 1 is the one read again from Wih, not sen1's own copy
        assert np.abs(wih[1, :2] - case["wrong_wih1"]).max() > 10 * TOL
