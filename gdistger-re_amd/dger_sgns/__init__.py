"""Python front end of libdger_sgns.so (include/dger_sgns.h) over ctypes.

The library is the product: a gfx950 HIP implementation of DistGER's DSGL
skip-gram negative-sampling trainer (reference: include/dsgl.hpp).  This module
only marshals arguments; there is no Python or CPU fallback.  If the shared
library is missing the import of any entry point raises ``LibraryMissing``.

The function ``dsgl`` mirrors the reference entry point
``int dsgl(int argc, char **argv, vector<int> *vertex_cn, ...)``
(include/dsgl.hpp:1464): argv-style options with the same flag names, the
corpus as vertex ids with -1 after every walk, and per-vertex counts.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("DGER_SGNS_LIB") or os.path.join(_HERE, "libdger_sgns.so")

DGER_OK = 0
DGER_E_ARG = -1
DGER_E_CORPUS = -2
DGER_E_VOCAB = -3
DGER_E_HIP = -4
DGER_E_ALLOC = -5
DGER_E_IO = -6
DGER_E_STATE = -7
DGER_E_UNSUPPORTED = -8
PRECISION = {"fp32": 0, "bf16": 1, "bf16-mfma": 2, "fp8": 3}


class LibraryMissing(RuntimeError):
    pass


class DgerError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"dger error {code}: {msg}")
        self.code = code


class Opts(C.Structure):
    _fields_ = [
        ("size", C.c_int32), ("window", C.c_int32), ("negative", C.c_int32),
        ("batch_size", C.c_int32), ("sample", C.c_float), ("alpha", C.c_float),
        ("min_count", C.c_int32), ("iter", C.c_int32), ("threads", C.c_int32),
        ("debug", C.c_int32), ("binary", C.c_int32), ("disk", C.c_int32),
        ("sync_period", C.c_float), ("min_sync_words", C.c_int32),
        ("full_sync_times", C.c_int32), ("message_size", C.c_int32),
        ("eoutput", C.c_char * 512), ("save_vocab", C.c_char * 512),
        ("read_vocab", C.c_char * 512), ("save_emb", C.c_char * 512),
        ("workers", C.c_int32), ("precision", C.c_int32), ("table_size", C.c_int64),
        ("device", C.c_int32), ("graph", C.c_int32), ("autotune", C.c_int32),
        ("rounding", C.c_int32), ("masters", C.c_int32), ("batch_parity", C.c_int32),
        ("share", C.c_int32),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("steps", C.c_uint64), ("trained_words", C.c_uint64), ("raw_words", C.c_uint64),
        ("sentences", C.c_uint64), ("final_alpha", C.c_float), ("workers", C.c_int32),
        ("vocab_size", C.c_int32), ("max_sentence", C.c_int32), ("kernel_ms", C.c_double),
        ("words_per_sec", C.c_double), ("contexts", C.c_uint64), ("out_ctx", C.c_uint64),
        ("share", C.c_int32), ("minibatch", C.c_int32),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class WalkOpts(C.Structure):
    _fields_ = [
        ("n_walkers", C.c_uint64), ("first_walker", C.c_uint64), ("min_length", C.c_uint32),
        ("max_length", C.c_uint32), ("r_threshold", C.c_double), ("seed", C.c_uint64),
        ("device", C.c_int32), ("chunk_walkers", C.c_uint32),
    ]


class WalkStats(C.Structure):
    _fields_ = [
        ("walks", C.c_uint64), ("tokens", C.c_uint64), ("moves", C.c_uint64), ("traces", C.c_uint64),
        ("capped", C.c_uint64), ("walk_ms", C.c_double),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


_vp = C.c_void_p
_i32 = C.c_int32
_SIGS = {
    "dger_last_error": (C.c_char_p, []),
    "dger_record_stride": (_i32, [_i32, _i32]),
    "dger_sgns_default_opts": (C.c_int, [C.POINTER(Opts)]),
    "dger_sgns_parse_args": (C.c_int, [C.c_int, C.POINTER(C.c_char_p), C.POINTER(Opts)]),
    "dger_sgns_train": (C.c_int, [C.POINTER(Opts), _vp, C.c_uint64, _vp, C.c_uint32, _vp, _vp,
                                  C.POINTER(Stats)]),
    "dger_sgns_train_stream": (C.c_int, [C.POINTER(Opts), _vp, C.c_uint64, _vp, C.c_uint32, _vp, _vp,
                                         C.POINTER(Stats), _vp]),
    "dger_ctx_set_stream": (C.c_int, [_vp, _vp]),
    "dger_sgns_main": (C.c_int, [C.c_int, C.POINTER(C.c_char_p), _vp, C.c_uint64, _vp, C.c_uint32]),
    "dger_ctx_create": (C.c_int, [C.POINTER(Opts), _vp, C.c_uint32, C.POINTER(_vp)]),
    "dger_ctx_set_corpus": (C.c_int, [_vp, _vp, C.c_uint64]),
    "dger_ctx_reset_model": (C.c_int, [_vp]),
    "dger_ctx_train": (C.c_int, [_vp, _i32, _vp, C.c_int64, _vp, C.POINTER(Stats)]),
    "dger_ctx_train_noalloc": (C.c_int, [_vp, _i32, C.POINTER(Stats)]),
    "dger_sgns_step_mfma_gemm1": (C.c_int, [_i32, _i32, _i32, _i32, _vp, C.c_float, _vp, _vp, _vp]),
    "dger_ctx_get_model": (C.c_int, [_vp, _vp, _vp]),
    "dger_ctx_set_model": (C.c_int, [_vp, _vp, _vp]),
    "dger_ctx_get_vocab": (C.c_int, [_vp, _vp, C.POINTER(C.c_uint64)]),
    "dger_ctx_get_table": (C.c_int, [_vp, _vp]),
    "dger_ctx_save": (C.c_int, [_vp, C.c_char_p, _i32]),
    "dger_emb_write": (C.c_int, [C.c_char_p, C.c_uint32, _i32, _vp, _vp, _i32]),
    "dger_ctx_destroy": (None, [_vp]),
    "dger_sgns_step": (C.c_int, [_i32, _i32, _i32, _i32, _vp, C.c_float, _vp, _vp, _vp, _vp, _vp]),
    "dger_sgns_step_fast": (C.c_int, [_i32, _vp, _vp, C.c_float, _vp, _vp, _vp, _vp, _vp]),
    "dger_sgns_step_fast_d": (C.c_int, [_i32, _i32, _vp, _vp, C.c_float, _vp, _vp, _vp, _vp, _vp]),
    "dger_loss_fwd_f32": (C.c_int, [_vp, _vp, _i32, _vp, _vp, _i32, _i32, _i32, _vp]),
    "dger_loss_bwd_f32": (C.c_int, [_vp, _vp, _i32, _vp, _vp, _i32, _i32, _i32, C.c_float, _vp, _vp]),
    "dger_loss_fwd_f64": (C.c_int, [_vp, _vp, _i32, _vp, _vp, _i32, _i32, _i32, _vp]),
    "dger_loss_bwd_f64": (C.c_int, [_vp, _vp, _i32, _vp, _vp, _i32, _i32, _i32, C.c_double, _vp, _vp]),
    "dger_dev_alloc": (C.c_int, [C.POINTER(_vp), C.c_uint64]),
    "dger_dev_free": (C.c_int, [_vp]),
    "dger_dev_upload": (C.c_int, [_vp, _vp, C.c_uint64]),
    "dger_dev_download": (C.c_int, [_vp, _vp, C.c_uint64]),
    "dger_dev_zero": (C.c_int, [_vp, C.c_uint64]),
    "dger_dev_copy": (C.c_int, [_vp, _vp, C.c_uint64]),
    "dger_ctx_link_scores": (C.c_int, [_vp, _vp, C.c_uint64, _vp]),
    "dger_ctx_link_auc": (C.c_int, [_vp, _vp, _vp, C.c_uint64, C.POINTER(C.c_double)]),
    "dger_device_count": (C.c_int, [C.POINTER(_i32)]),
    "dger_walk_default_opts": (None, [C.POINTER(WalkOpts)]),
    "dger_walk_common_neighbors": (C.c_int, [_vp, _vp, C.c_uint32, _vp, _i32]),
    "dger_walk_run": (C.c_int, [_vp, _vp, _vp, C.c_uint32, C.POINTER(WalkOpts), _vp, C.c_uint64,
                                C.POINTER(C.c_uint64), _vp, C.POINTER(WalkStats)]),
    "dger_walk_run_device": (C.c_int, [_vp, _vp, _vp, C.c_uint32, C.POINTER(WalkOpts), C.POINTER(_vp), _vp,
                                       C.POINTER(WalkStats)]),
    "dger_walk_corpus_info": (C.c_int, [_vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint32)]),
    "dger_walk_corpus_download": (C.c_int, [_vp, _vp, C.c_uint64]),
    "dger_walk_corpus_destroy": (None, [_vp]),
    "dger_ctx_set_corpus_walks": (C.c_int, [_vp, _vp]),
}
EXPORTS = tuple(_SIGS)

_lib = None


def lib():
    """Load libdger_sgns.so (raises LibraryMissing if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LibraryMissing(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc: int) -> None:
    if rc != DGER_OK:
        raise DgerError(rc, lib().dger_last_error().decode(errors="replace"))


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data


def device_count() -> int:
    n = _i32(0)
    _check(lib().dger_device_count(C.byref(n)))
    return int(n.value)


def record_stride(window: int, negative: int) -> int:
    return int(lib().dger_record_stride(window, negative))


def _argv(args):
    arr = (C.c_char_p * (len(args) + 1))()
    for i, a in enumerate(args):
        arr[i] = a.encode() if isinstance(a, str) else a
    arr[len(args)] = None
    return arr


def parse_args(argv) -> Opts:
    """dger_sgns_parse_args over a list of strings (argv[0] is the program)."""
    o = Opts()
    _check(lib().dger_sgns_parse_args(len(argv), _argv(argv), C.byref(o)))
    return o


def default_opts(**kw) -> Opts:
    o = Opts()
    _check(lib().dger_sgns_default_opts(C.byref(o)))
    for k, v in kw.items():
        if k == "precision" and isinstance(v, str):
            v = PRECISION[v]
        if k == "eoutput":
            v = v.encode()
        setattr(o, k, v)
    return o


def _i32arr(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


def walk_default_opts(**kw) -> WalkOpts:
    o = WalkOpts()
    lib().dger_walk_default_opts(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def common_neighbors(rowptr, col, device: int = 0) -> np.ndarray:
    """|N(u) ∩ N(col[e])| per edge (dger_walk_common_neighbors); rows sorted."""
    rowptr = np.ascontiguousarray(rowptr, np.int64)
    col = _i32arr(col)
    out = np.zeros(col.size, np.int32)
    _check(lib().dger_walk_common_neighbors(_ptr(rowptr), _ptr(col), rowptr.size - 1, _ptr(out), device))
    return out


class WalkCorpus:
    """Walks left in HBM by dger_walk_run_device."""

    def __init__(self, h):
        self._h = h

    def info(self):
        n, m = C.c_uint64(0), C.c_uint32(0)
        _check(lib().dger_walk_corpus_info(self._h, C.byref(n), C.byref(m)))
        return int(n.value), int(m.value)

    def download(self) -> np.ndarray:
        n, _ = self.info()
        out = np.empty(max(n, 1), np.int32)
        _check(lib().dger_walk_corpus_download(self._h, _ptr(out), n))
        return out[:n]

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().dger_walk_corpus_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def walk_device(rowptr, col, cnb, opts: WalkOpts):
    """Like walk(), but the corpus stays in HBM: returns (WalkCorpus, vertex_cn, stats)."""
    rowptr = np.ascontiguousarray(rowptr, np.int64)
    col = _i32arr(col)
    cnb = _i32arr(cnb)
    vcn = np.zeros(rowptr.size - 1, np.int32)
    h = _vp()
    st = WalkStats()
    _check(lib().dger_walk_run_device(_ptr(rowptr), _ptr(col), _ptr(cnb), rowptr.size - 1, C.byref(opts),
                                      C.byref(h), _ptr(vcn), C.byref(st)))
    return WalkCorpus(h), vcn, st


def walk(rowptr, col, cnb, opts: WalkOpts):
    """Information-oriented walks on the GPU.  Returns the corpus (vertex
    ids, -1 after every walk), vertex_cn and the stats.  The walks are made
    in device memory (dger_walk_run_device) and copied out at their real
    length, so the host never holds a buffer sized for the longest possible
    corpus (n_walkers * (max_length + 1) ids)."""
    wc, vcn, st = walk_device(rowptr, col, cnb, opts)
    with wc:
        corpus = wc.download()
    return corpus, vcn, st


class Trainer:
    """Device-resident trainer (the dger_ctx_* API)."""

    def __init__(self, opts: Opts, vertex_cn, corpus=None):
        self.opts = opts
        self.cn = _i32arr(vertex_cn)
        self.V = int(self.cn.size)
        self.D = int(opts.size)
        h = _vp()
        _check(lib().dger_ctx_create(C.byref(opts), _ptr(self.cn), self.V, C.byref(h)))
        self._h = h
        if corpus is not None:
            self.set_corpus(corpus)

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().dger_ctx_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_corpus(self, corpus) -> None:
        c = _i32arr(corpus)
        _check(lib().dger_ctx_set_corpus(self._h, _ptr(c), c.size))

    def set_corpus_walks(self, walks: "WalkCorpus") -> None:
        """Corpus from walks left in HBM (dger_ctx_set_corpus_walks)."""
        _check(lib().dger_ctx_set_corpus_walks(self._h, walks._h))

    def set_stream(self, stream=None) -> None:
        """Queue the context's device work on a caller's HIP stream (an
        integer handle such as torch.cuda.Stream().cuda_stream; None = the
        context's own stream)."""
        _check(lib().dger_ctx_set_stream(self._h, stream))

    def reset_model(self) -> None:
        _check(lib().dger_ctx_reset_model(self._h))

    def train(self, workers: int = 0, rec_cap: int = 0):
        """Returns (stats, records [workers, rec_cap, stride] or None, worker_steps)."""
        st = Stats()
        nw = workers if workers > 0 else (self.opts.workers or 0)
        rec = None
        ws = None
        if nw > 0:
            ws = np.zeros(nw, np.uint64)
            if rec_cap > 0:
                rec = np.full((nw, rec_cap, record_stride(self.opts.window, self.opts.negative)), -1,
                              np.int32)
        elif rec_cap > 0:
            raise ValueError("records need an explicit worker count")
        _check(lib().dger_ctx_train(self._h, workers, _ptr(rec), rec_cap, _ptr(ws), C.byref(st)))
        return st, rec, ws

    def train_noalloc(self, workers: int = 0) -> "Stats":
        """dger_ctx_train_noalloc: one run with no device allocation (the
        window-slot trainer, scratch in the library's static device array)."""
        st = Stats()
        _check(lib().dger_ctx_train_noalloc(self._h, workers, C.byref(st)))
        return st

    def link_scores(self, pairs) -> np.ndarray:
        """Dot-product scores of vertex-id pairs on the device-resident model."""
        p = np.ascontiguousarray(pairs, np.int32).reshape(-1, 2)
        out = np.empty(p.shape[0], np.float32)
        _check(lib().dger_ctx_link_scores(self._h, _ptr(p), p.shape[0], _ptr(out)))
        return out

    def link_auc(self, pairs, labels) -> float:
        """Link-prediction ROC AUC computed on the device."""
        p = np.ascontiguousarray(pairs, np.int32).reshape(-1, 2)
        lab = np.ascontiguousarray(labels, np.int8)
        if lab.size != p.shape[0]:
            raise ValueError("one label per pair")
        r = C.c_double()
        _check(lib().dger_ctx_link_auc(self._h, _ptr(p), _ptr(lab), p.shape[0], C.byref(r)))
        return float(r.value)

    def get_model(self):
        wih = np.empty((self.V, self.D), np.float32)
        woh = np.empty((self.V, self.D), np.float32)
        _check(lib().dger_ctx_get_model(self._h, _ptr(wih), _ptr(woh)))
        return wih, woh

    def set_model(self, wih=None, woh=None) -> None:
        a = None if wih is None else np.ascontiguousarray(wih, np.float32)
        b = None if woh is None else np.ascontiguousarray(woh, np.float32)
        _check(lib().dger_ctx_set_model(self._h, _ptr(a), _ptr(b)))

    def vocab(self):
        rank_id = np.empty(self.V, np.int32)
        tw = C.c_uint64(0)
        _check(lib().dger_ctx_get_vocab(self._h, _ptr(rank_id), C.byref(tw)))
        return rank_id, int(tw.value)

    def table(self) -> np.ndarray:
        t = np.empty(int(self.opts.table_size), np.int32)
        _check(lib().dger_ctx_get_table(self._h, _ptr(t)))
        return t

    def save(self, path: str, binary: int = 0) -> None:
        _check(lib().dger_ctx_save(self._h, path.encode(), binary))


def emb_write(path: str, ids, rows, binary: int = 0) -> None:
    """dger_emb_write: rows (V x D float32) with their vertex ids in the
    reference's embedding file format (text or binary)."""
    r = np.ascontiguousarray(rows, np.float32)
    i = _i32arr(ids)
    if r.ndim != 2 or i.size != r.shape[0]:
        raise ValueError("one id per row")
    _check(lib().dger_emb_write(path.encode(), r.shape[0], r.shape[1], _ptr(i), _ptr(r), binary))


def emb_read(path: str, binary: int = 0):
    """Reads an embedding file written with the given -binary: returns
    (ids int64 [V], rows float32 [V, D])."""
    with open(path, "rb") as f:
        data = f.read()
    nl = data.index(b"\n")
    V, D = (int(x) for x in data[:nl].split())
    ids = np.empty(V, np.int64)
    rows = np.empty((V, D), np.float32)
    p = nl + 1
    for r in range(V):
        sp = data.index(b" ", p)
        ids[r] = int(data[p:sp])
        p = sp + 1
        if binary:
            rows[r] = np.frombuffer(data, np.float32, D, p)
            p += 4 * D
        else:
            nxt = data.index(b"\n", p)
            vals = data[p:nxt].split()
            if len(vals) != D:
                raise ValueError(f"row {r}: {len(vals)} values")
            rows[r] = np.array(vals).astype(np.float32)
            p = nxt
        if data[p:p + 1] != b"\n":
            raise ValueError(f"row {r}: no newline")
        p += 1
    if p != len(data):
        raise ValueError("trailing bytes")
    return ids, rows


def train(opts: Opts, corpus, vertex_cn, stream=None):
    """dger_sgns_train (dger_sgns_train_stream when a HIP stream handle is
    given): returns (embeddings V x D in rank order, rank_id, stats)."""
    c = _i32arr(corpus)
    cn = _i32arr(vertex_cn)
    emb = np.empty((cn.size, opts.size), np.float32)
    rank_id = np.empty(cn.size, np.int32)
    st = Stats()
    if stream is None:
        _check(lib().dger_sgns_train(C.byref(opts), _ptr(c), c.size, _ptr(cn), cn.size, _ptr(emb),
                                     _ptr(rank_id), C.byref(st)))
    else:
        _check(lib().dger_sgns_train_stream(C.byref(opts), _ptr(c), c.size, _ptr(cn), cn.size, _ptr(emb),
                                            _ptr(rank_id), C.byref(st), stream))
    return emb, rank_id, st


def dsgl(argv, vertex_cn, local_corpus) -> int:
    """Mirror of dsgl(argc, argv, &vertex_cn, &new_sort, &graph) (include/dsgl.hpp:1464).

    Returns 0 on success; raises DgerError with the library's code otherwise
    (the reference exits or asserts instead)."""
    c = _i32arr(local_corpus)
    cn = _i32arr(vertex_cn)
    _check(lib().dger_sgns_main(len(argv), _argv(argv), _ptr(c), c.size, _ptr(cn), cn.size))
    return 0


def step(O, I, fin: int, meta, alpha: float):
    """One minibatch step on the device (dger_sgns_step): returns corr, md, gin."""
    O = np.ascontiguousarray(O, np.float32)
    I = np.ascontiguousarray(I, np.float32)
    meta = _i32arr(meta)
    os_, D = O.shape
    nin = I.shape[0]
    corr = np.empty((os_, nin), np.float32)
    md = np.empty((os_, D), np.float32)
    gin = np.empty((nin, D), np.float32)
    _check(lib().dger_sgns_step(D, os_, nin, fin, _ptr(meta), alpha, _ptr(O), _ptr(I), _ptr(corr),
                                _ptr(md), _ptr(gin)))
    return corr, md, gin


def step_mfma_gemm1(O, I, fin: int, meta, alpha: float) -> np.ndarray:
    """GEMM1 + sigmoid gradient of the BF16-MFMA general trainer on the device
    (dger_sgns_step_mfma_gemm1): returns corr (n_out x n_ctx)."""
    O = np.ascontiguousarray(O, np.float32)
    I = np.ascontiguousarray(I, np.float32)
    meta = _i32arr(meta)
    os_, D = O.shape
    nin = I.shape[0]
    corr = np.empty((os_, nin), np.float32)
    _check(lib().dger_sgns_step_mfma_gemm1(D, os_, nin, fin, _ptr(meta), alpha, _ptr(O), _ptr(I), _ptr(corr)))
    return corr


FAST_DIM = 128
FAST_DIMS = (128, 256)  # row lengths of the window-slot trainer
FAST_SLOTS = 48


def step_fast(O, I, mult, meta, alpha: float):
    """One step of the window-slot BF16-MFMA trainer on the device
    (dger_sgns_step_fast_d): O (n_out x D), I (48 x D slot rows), mult (48
    contexts per slot), meta (n_out), D = 128 or 256.  Returns corr
    (n_out x 48), md (n_out x D) and the updated slot rows (48 x D)."""
    O = np.ascontiguousarray(O, np.float32)
    I = np.ascontiguousarray(I, np.float32)
    mult = _i32arr(mult)
    meta = _i32arr(meta)
    os_, D = O.shape
    if D not in FAST_DIMS or I.shape != (FAST_SLOTS, D) or mult.size != FAST_SLOTS or meta.size != os_:
        raise ValueError("shapes: O (n_out, D), I (48, D), mult (48,), meta (n_out,), D = 128 or 256")
    corr = np.empty((os_, FAST_SLOTS), np.float32)
    md = np.empty((os_, D), np.float32)
    I2 = np.empty((FAST_SLOTS, D), np.float32)
    _check(lib().dger_sgns_step_fast_d(D, os_, _ptr(meta), _ptr(mult), alpha, _ptr(O), _ptr(I), _ptr(corr),
                                       _ptr(md), _ptr(I2)))
    return corr, md, I2


class DeviceArray:
    """A device allocation made through the library (for the loss kernels)."""

    def __init__(self, host: np.ndarray):
        self.dtype = host.dtype
        self.shape = host.shape
        self.nbytes = int(host.nbytes)
        p = _vp()
        _check(lib().dger_dev_alloc(C.byref(p), self.nbytes))
        self.ptr = p
        host = np.ascontiguousarray(host)
        _check(lib().dger_dev_upload(self.ptr, _ptr(host), self.nbytes))

    @classmethod
    def zeros(cls, shape, dtype):
        return cls(np.zeros(shape, dtype))

    def zero(self):
        _check(lib().dger_dev_zero(self.ptr, self.nbytes))

    def copy_from(self, other: "DeviceArray"):
        """Device-to-device copy of other's bytes into this allocation."""
        if other.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        _check(lib().dger_dev_copy(self.ptr, other.ptr, self.nbytes))

    def get(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        _check(lib().dger_dev_download(_ptr(out), self.ptr, self.nbytes))
        return out

    def free(self):
        if self.ptr is not None and self.ptr.value:
            lib().dger_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def loss_fwd(syn0: DeviceArray, syn1: DeviceArray, ctx: DeviceArray, out: DeviceArray,
             n_mb: int, n_ctx: int, n_out: int) -> float:
    f64 = syn0.dtype == np.float64
    acc = DeviceArray.zeros((1,), np.float64 if f64 else np.float32)
    fn = lib().dger_loss_fwd_f64 if f64 else lib().dger_loss_fwd_f32
    _check(fn(syn0.ptr, syn1.ptr, syn0.shape[1], ctx.ptr, out.ptr, n_mb, n_ctx, n_out, acc.ptr))
    v = float(acc.get()[0])
    acc.free()
    return v


def loss_bwd(syn0: DeviceArray, syn1: DeviceArray, ctx: DeviceArray, out: DeviceArray,
             n_mb: int, n_ctx: int, n_out: int, scale: float = 1.0):
    f64 = syn0.dtype == np.float64
    g0 = DeviceArray.zeros(syn0.shape, syn0.dtype)
    g1 = DeviceArray.zeros(syn1.shape, syn1.dtype)
    fn = lib().dger_loss_bwd_f64 if f64 else lib().dger_loss_bwd_f32
    _check(fn(syn0.ptr, syn1.ptr, syn0.shape[1], ctx.ptr, out.ptr, n_mb, n_ctx, n_out, scale,
              g0.ptr, g1.ptr))
    r = (g0.get(), g1.get())
    g0.free()
    g1.free()
    return r


@dataclass
class Corpus:
    """A walk corpus in the reference's in-memory layout (include/path.hpp:91-96,110)."""
    tokens: np.ndarray      # int32 vertex ids, -1 after every walk
    vertex_cn: np.ndarray   # int32 occurrences per vertex id
    n_walks: int

    @property
    def n_tokens(self) -> int:
        return int(self.tokens.size)
