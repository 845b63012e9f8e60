"""The gfx950 trainer's model-update semantics against the hand-worked traces
of tests/update_cases.py (one worker, FP32 rows), through the C ABI: which
writes reach Wih / Woh and in which order (dsgl.hpp:1314-1407) and what the
end of a corpus slice drops (dsgl.hpp:1031-1050).  Exact arithmetic, exact
comparison."""
import numpy as np
import pytest

import dger_sgns as G
import update_cases as U

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", sorted(U.CASES))
def test_gpu_update_semantics(name):
    case = U.CASES[name]()
    print(f"case {name}", flush=True)
    wih0, woh0 = U.initial_model()
    opts = G.default_opts(size=U.D, window=1, negative=case["negative"], sample=0.0,
                          alpha=float(U.ALPHA), iter=1, threads=2, table_size=10, debug=0)
    with G.Trainer(opts, U.CN, np.array(case["corpus"], np.int32)) as tr:
        tr.set_model(wih0, woh0)
        st, _, _ = tr.train(workers=1)
        wih, woh = tr.get_model()
    assert st.steps == case["steps"] and st.sentences == case["sentences"]
    assert np.array_equal(wih, case["wih"]), (wih, case["wih"])
    assert np.array_equal(woh, case["woh"]), (woh, case["woh"])
