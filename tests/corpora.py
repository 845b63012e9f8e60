"""Small seeded corpora in the reference's layout (ids, -1 after each walk)
with the edge cases the trainer has to handle."""
import numpy as np


def random_corpus(seed, V=1500, n_walks=600, lmin=1, lmax=60, long_walk=0, empty_walks=0, zipf=False):
    rng = np.random.default_rng(seed)
    parts = []
    for _ in range(n_walks):
        L = int(rng.integers(lmin, lmax + 1))
        if zipf:
            w = (rng.zipf(1.3, L) - 1) % V
        else:
            w = rng.integers(0, V, L)
        parts.append(np.append(w, -1))
    if long_walk:
        # longer than MAX_SENTENCE_LENGTH (dsgl.hpp:23): split at 1000 kept tokens
        parts.insert(len(parts) // 2, np.append(rng.integers(0, V, long_walk), -1))
    for _ in range(empty_walks):
        parts.insert(int(rng.integers(0, len(parts))), np.array([-1]))
    corpus = np.concatenate(parts).astype(np.int32)
    cn = np.bincount(corpus[corpus >= 0], minlength=V).astype(np.int32)
    return corpus, cn
