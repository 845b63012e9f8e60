"""Walk oracle -- TEST INFRASTRUCTURE ONLY.

ctypes front end of ``oracle/walk_oracle.c``, a CPU restatement of DistGER's
information-oriented random walk (``include/walk.hpp:260-387``, ``:789-1043``)
for one partition.  Only ``tests/`` may import it, as the checker of the GPU
corpus producer.  What pins each function is listed in walk_oracle.c.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


def _load():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liborc_walk.so")
        if not os.path.exists(path):
            subprocess.run(["make", "-C", _HERE, "-s", "liborc_walk.so"], check=True)
        L = C.CDLL(path)
        vp = C.c_void_p
        L.orc_corr_trace.restype = None
        L.orc_corr_trace.argtypes = [vp, C.c_int64, vp, vp]
        L.orc_walk_run.restype = C.c_int64
        L.orc_walk_run.argtypes = [vp, vp, vp, C.c_uint32, C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32,
                                   C.c_double, C.c_uint64, vp, C.c_int64, vp, vp]
        _lib = L
    return _lib


def corr_trace(path) -> tuple[np.ndarray, np.ndarray]:
    """(H, R) after every visit of the given vertex sequence."""
    p = np.ascontiguousarray(path, np.int32)
    H = np.zeros(p.size, np.float64)
    R = np.zeros(p.size, np.float64)
    _load().orc_corr_trace(p.ctypes.data, p.size, H.ctypes.data, R.ctypes.data)
    return H, R


def walk(rowptr, col, cnb, n_walkers, first=0, min_length=20, max_length=1024, r_threshold=0.999, seed=1,
         chunk=100_000):
    """Returns (corpus int32 with -1 after each walk, vertex_cn, stats dict);
    walkers run in chunks so the host buffer stays chunk * (max_length + 1)."""
    rowptr = np.ascontiguousarray(rowptr, np.int64)
    col = np.ascontiguousarray(col, np.int32)
    cnb = np.ascontiguousarray(cnb, np.int32)
    V = rowptr.size - 1
    keys = ("walks", "tokens", "moves", "traces", "capped")
    total = {k: 0 for k in keys}
    vcn = np.zeros(V, np.int64)
    parts = []
    part_vcn = np.zeros(V, np.int32)
    st = np.zeros(5, np.uint64)
    for lo in range(0, int(n_walkers), chunk):
        n = min(chunk, int(n_walkers) - lo)
        cap = n * (max_length + 1)
        corpus = np.empty(max(cap, 1), np.int32)
        got = _load().orc_walk_run(rowptr.ctypes.data, col.ctypes.data, cnb.ctypes.data, V, first + lo, n,
                                   min_length, max_length, r_threshold, seed, corpus.ctypes.data, cap,
                                   part_vcn.ctypes.data, st.ctypes.data)
        assert got >= 0
        parts.append(corpus[:got].copy())
        vcn += part_vcn
        for k, x in zip(keys, st):
            total[k] += int(x)
    corpus = np.concatenate(parts) if parts else np.empty(0, np.int32)
    return corpus, vcn.astype(np.int32), total
