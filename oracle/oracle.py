"""SGNS oracle — TEST INFRASTRUCTURE ONLY.

ctypes front end of ``oracle/sgns_oracle.cpp``, a CPU restatement of DistGER's
DSGL trainer (``include/dsgl.hpp``).  Only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` may import this module, and only as
the checker (or the timed CPU baseline).  The product library never uses it.

Pins (what checks each part, see tests/test_oracle_*.py):
  lcg_steps / lcg_jump    closed form of the affine map; big-integer Python LCG;
                          jump-ahead vs sequential.
  build_vocab             permutation + non-increasing counts + tie order of a
                          small input sorted by libstdc++'s insertion sort
                          (stable for n <= 16) + Python sorted() on distinct counts.
  init_net                first values worked by hand from the LCG (seed 1).
  exp_table               closed form 1/(1+exp(-x)) in float64, symmetry, t[500]=1/2.
  keep_thresholds         closed form sqrt(r)+r in float64 (within float rounding)
                          and the rank-indexing quirk.
  unigram_table           an independent Python re-derivation of the lagged-CDF
                          boundaries on small tables (exact), monotonicity,
                          coverage, per-rank share vs cn^0.75 probability.
  train (index stream)    hand-worked traces on tiny corpora (window/negative
                          draws, subsampling, separator handling, slice split,
                          duplicate negatives, t1-drop), conservation of
                          tokens, and per-step invariants over random corpora.
  step_f32 / step_f64     float64 numpy matrix products of the same staged rows
                          (relative 1e-5), label layout checked on a hand case.
  train (model updates)   hand-worked traces with exact dyadic arithmetic
                          (tests/update_cases.py): reference semantics -- target
                          updates in Woh, negative-buffer clobber, second
                          sentence's write-back winning, slice-end drops;
                          window-slot semantics ("slots") -- write-back when a
                          rank leaves the window and re-read from Wih when it
                          re-enters, negative updates in Woh, slice-end drop of
                          the rows still in a window, minibatch order, negative
                          sharing; traces that tell the two semantics apart.
Parity unpinned: the std::sort tie order for more than 16 equal counts (it is
libstdc++ introsort's, which the reference shares but which cannot be checked
without a reference build), and the reference's MKL summation order (dsgl.hpp
calls cblas_sgemm; MKL's order is unspecified, so only the 1e-5 tolerance is
checked, not bit equality with MKL).
"""
from __future__ import annotations

import ctypes as C
import os
import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))

i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")

TABLE_SIZE = 100_000_000  # dsgl.hpp:82


class OrcOpts(C.Structure):
    _fields_ = [
        ("dim", C.c_int32), ("window", C.c_int32), ("negative", C.c_int32),
        ("iter", C.c_int32), ("threads", C.c_int32), ("workers", C.c_int32),
        ("sample", C.c_float), ("alpha", C.c_float), ("table_size", C.c_int64),
        ("init_model", C.c_int32), ("parallel", C.c_int32), ("semantics", C.c_int32),
        ("share_shift", C.c_int32), ("minibatch", C.c_int32),
    ]


class OrcStats(C.Structure):
    _fields_ = [
        ("steps", C.c_uint64), ("trained_words", C.c_uint64), ("raw_words", C.c_uint64),
        ("sentences", C.c_uint64), ("final_alpha", C.c_float), ("vocab_size", C.c_int32),
        ("train_seconds", C.c_double),
    ]


def build(fast: bool = True) -> None:
    """Compile the oracle libraries with oracle/Makefile."""
    import subprocess
    targets = ["liborc_sgns.so"] + (["liborc_sgns_fast.so"] if fast else [])
    subprocess.run(["make", "-C", _HERE, "-s"] + targets, check=True)


def _load(name: str):
    path = os.path.join(_HERE, name)
    if not os.path.exists(path):
        build()
    lib = C.CDLL(path)
    lib.orc_lcg_steps.restype = C.c_uint64
    lib.orc_lcg_steps.argtypes = [C.c_uint64, C.c_uint64]
    lib.orc_lcg_jump.restype = C.c_uint64
    lib.orc_lcg_jump.argtypes = [C.c_uint64, C.c_uint64]
    lib.orc_build_vocab.restype = C.c_int
    lib.orc_build_vocab.argtypes = [i32p, C.c_int32, i32p, i32p, C.POINTER(C.c_uint64)]
    lib.orc_init_net.restype = None
    lib.orc_init_net.argtypes = [C.c_int32, C.c_int32, f32p]
    lib.orc_exp_table.restype = None
    lib.orc_exp_table.argtypes = [f32p]
    lib.orc_keep_thresholds.restype = None
    lib.orc_keep_thresholds.argtypes = [i32p, C.c_int32, C.c_uint64, C.c_float, f32p]
    lib.orc_unigram_pow_sum.restype = C.c_double
    lib.orc_unigram_pow_sum.argtypes = [i32p, C.c_int32, C.c_int32]
    lib.orc_unigram_table.restype = C.c_int
    lib.orc_unigram_table.argtypes = [i32p, C.c_int32, C.c_int64, C.c_int32, i32p]
    lib.orc_record_stride.restype = C.c_int32
    lib.orc_record_stride.argtypes = [C.c_int32, C.c_int32]
    lib.orc_train.restype = C.c_int
    lib.orc_train.argtypes = [C.POINTER(OrcOpts), i32p, C.c_uint64, i32p, C.c_int32, f32p, f32p,
                              C.c_void_p, C.c_int64, C.c_void_p, C.POINTER(OrcStats)]
    lib.orc_step_f32.restype = None
    lib.orc_step_f32.argtypes = [C.c_int32] * 4 + [i32p, C.c_float, f32p, f32p, f32p, f32p, f32p,
                                                    f32p, C.c_void_p]
    lib.orc_step_f64.restype = None
    lib.orc_step_f64.argtypes = [C.c_int32] * 4 + [i32p, C.c_double, f32p, f64p, f64p, f64p, f64p,
                                                    f64p, C.c_void_p]
    return lib


_LIBS: dict = {}


def lib(fast: bool = False):
    name = "liborc_sgns_fast.so" if fast else "liborc_sgns.so"
    if name not in _LIBS:
        _LIBS[name] = _load(name)
    return _LIBS[name]


# ------------------------------------------------------------------ wrappers
def lcg_steps(s: int, n: int) -> int:
    return int(lib().orc_lcg_steps(s, n))


def lcg_jump(s: int, n: int) -> int:
    return int(lib().orc_lcg_jump(s, n))


def build_vocab(cn: np.ndarray):
    cn = np.ascontiguousarray(cn, dtype=np.int32)
    V = cn.size
    rank_id = np.empty(V, np.int32)
    id_rank = np.empty(V, np.int32)
    tw = C.c_uint64(0)
    rc = lib().orc_build_vocab(cn, V, rank_id, id_rank, C.byref(tw))
    assert rc == 0
    return rank_id, id_rank, int(tw.value)


def init_net(V: int, D: int) -> np.ndarray:
    out = np.empty(V * D, np.float32)
    lib().orc_init_net(V, D, out)
    return out.reshape(V, D)


def exp_table() -> np.ndarray:
    t = np.empty(1001, np.float32)
    lib().orc_exp_table(t)
    return t


def keep_thresholds(cn_by_id: np.ndarray, train_words: int, sample: float) -> np.ndarray:
    cn_by_id = np.ascontiguousarray(cn_by_id, dtype=np.int32)
    thr = np.empty(cn_by_id.size, np.float32)
    lib().orc_keep_thresholds(cn_by_id, cn_by_id.size, train_words, sample, thr)
    return thr


def unigram_pow_sum(cn_rank: np.ndarray, threads: int) -> float:
    cn_rank = np.ascontiguousarray(cn_rank, dtype=np.int32)
    return float(lib().orc_unigram_pow_sum(cn_rank, cn_rank.size, threads))


def unigram_table(cn_rank: np.ndarray, table_size: int = TABLE_SIZE, threads: int = 2) -> np.ndarray:
    cn_rank = np.ascontiguousarray(cn_rank, dtype=np.int32)
    t = np.empty(table_size, np.int32)
    rc = lib().orc_unigram_table(cn_rank, cn_rank.size, table_size, threads, t)
    assert rc == 0
    return t


def record_stride(window: int, negative: int) -> int:
    return int(lib().orc_record_stride(window, negative))


SEMANTICS = {"reference": 0, "slots": 1}


def make_opts(dim=128, window=10, negative=5, iter=1, threads=2, workers=1, sample=1e-3,
              alpha=0.01, table_size=TABLE_SIZE, init_model=1, parallel=0, semantics="reference",
              share=2, minibatch=0) -> OrcOpts:
    """semantics: "reference" (dsgl.hpp's buffers) or "slots" (the update
    rules of the library's BF16-MFMA window-slot trainer).  share (2, 4, 8,
    16: target positions per negative group) and minibatch (contexts per
    minibatch, 0 = a step's at once) apply to "slots" only."""
    shift = {2: 0, 4: 1, 8: 2, 16: 3}[share]
    return OrcOpts(dim, window, negative, iter, threads, workers, sample, alpha, table_size,
                   init_model, parallel, SEMANTICS[semantics], shift, minibatch)


def train(corpus: np.ndarray, cn_by_id: np.ndarray, opts: OrcOpts, wih=None, woh=None,
          rec_cap: int = 0, fast: bool = False):
    """Run the oracle trainer.  Returns (wih, woh, stats, records, worker_steps).

    records: int32 array [workers, rec_cap, stride] (or None); see
    sgns_oracle.cpp orc_record_stride for the layout."""
    corpus = np.ascontiguousarray(corpus, dtype=np.int32)
    cn_by_id = np.ascontiguousarray(cn_by_id, dtype=np.int32)
    V, D = cn_by_id.size, opts.dim
    if wih is None:
        wih = np.zeros((V, D), np.float32)
    if woh is None:
        woh = np.zeros((V, D), np.float32)
    wih = np.ascontiguousarray(wih, dtype=np.float32).copy()
    woh = np.ascontiguousarray(woh, dtype=np.float32).copy()
    stride = record_stride(opts.window, opts.negative)
    rec = None
    if rec_cap > 0:
        rec = np.full((opts.workers, rec_cap, stride), -1, np.int32)
    wsteps = np.zeros(opts.workers, np.uint64)
    st = OrcStats()
    rc = lib(fast).orc_train(C.byref(opts), corpus, corpus.size, cn_by_id, V, wih.reshape(-1),
                             woh.reshape(-1),
                             rec.ctypes.data if rec is not None else None, rec_cap,
                             wsteps.ctypes.data, C.byref(st))
    if rc != 0:
        raise ValueError(f"orc_train failed: {rc}")
    return wih, woh, st, rec, wsteps


def step_f32(O, I, fin, meta, alpha, table):
    O = np.ascontiguousarray(O, np.float32)
    I = np.ascontiguousarray(I, np.float32)
    os_, D = O.shape
    nin = I.shape[0]
    corr = np.empty((os_, nin), np.float32)
    md = np.empty((os_, D), np.float32)
    gin = np.empty((nin, D), np.float32)
    sig = np.empty((os_, nin), np.int32)
    lib().orc_step_f32(D, os_, nin, fin, np.ascontiguousarray(meta, np.int32), alpha,
                       np.ascontiguousarray(table, np.float32), O, I, corr, md, gin,
                       sig.ctypes.data)
    return corr, md, gin, sig


def step_f64(O, I, fin, meta, alpha, table):
    O = np.ascontiguousarray(O, np.float64)
    I = np.ascontiguousarray(I, np.float64)
    os_, D = O.shape
    nin = I.shape[0]
    corr = np.empty((os_, nin), np.float64)
    md = np.empty((os_, D), np.float64)
    gin = np.empty((nin, D), np.float64)
    sig = np.empty((os_, nin), np.int32)
    lib().orc_step_f64(D, os_, nin, fin, np.ascontiguousarray(meta, np.int32), alpha,
                       np.ascontiguousarray(table, np.float32), O, I, corr, md, gin,
                       sig.ctypes.data)
    return corr, md, gin, sig


def decode_records(rec: np.ndarray, n: int, window: int, negative: int):
    """Split raw records into a list of dicts (first n rows)."""
    out = []
    W, K = window, negative
    for r in rec[:n]:
        nin, os_ = int(r[4]), int(r[5])
        out.append(dict(
            t1=int(r[0]), t2=int(r[1]), b=int(r[2]), fin=int(r[3]), nin=nin, os=os_,
            alpha=float(np.array([r[6]], np.int32).view(np.float32)[0]), grp=int(r[7]),
            inputs=[int(x) for x in r[8:8 + nin]],
            outs=[int(x) for x in r[8 + 4 * W:8 + 4 * W + os_]],
            meta=[int(x) for x in r[8 + 4 * W + K + 2:8 + 4 * W + K + 2 + os_]],
        ))
    return out
