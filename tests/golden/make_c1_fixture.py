#!/usr/bin/env python3
"""Writes c1_er10k_fp32.json: the oracle's single-worker FP32 run on the
Erdos-Renyi 10k-vertex corpus (10 uniform walks of length 80 per vertex, seed
1; dim 128, window 10, negative 5, sample 1e-3, alpha 0.01, iter 1) reduced to
its counts and SHA-256 digests of the final Wih / Woh and of the first
REC_STEPS index-stream records.  The full run takes the oracle several
minutes on one core, so the GPU test checks the digests instead of rerunning
it.  Usage: python3 tests/golden/make_c1_fixture.py"""
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "gdistger-re_amd"))
from dger_sgns import synth  # noqa: E402
from oracle import oracle as O  # noqa: E402

REC_STEPS = 200_000
OPTS = dict(dim=128, window=10, negative=5, sample=1e-3, alpha=0.01, iter=1, threads=2, workers=1,
            table_size=100_000_000)


def digest(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    c = synth.corpus_er()
    t0 = time.time()
    wih, woh, st, rec, _ = O.train(c.tokens, c.vertex_cn, O.make_opts(**OPTS), rec_cap=REC_STEPS)
    out = {
        "corpus": "synth.corpus_er() (n=10000, m=50000, 10 walks of length 80 per vertex, seed 1)",
        "tokens": int(c.tokens.size),
        "opts": OPTS,
        "steps": int(st.steps), "raw_words": int(st.raw_words), "sentences": int(st.sentences),
        "final_alpha_bits": int(np.float32(st.final_alpha).view(np.uint32)),
        "rec_steps": REC_STEPS,
        "rec_sha256": digest(rec[0, :REC_STEPS]),
        "wih_sha256": digest(wih), "woh_sha256": digest(woh),
        "oracle_seconds": round(time.time() - t0, 1),
    }
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "c1_er10k_fp32.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
