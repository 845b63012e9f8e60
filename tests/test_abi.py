"""C ABI checks that need no GPU: the library loads, exports every function
include/dger_sgns.h declares, parses options like dsgl() (dsgl.hpp:1496-1535),
and rejects bad inputs with error codes before touching the device."""
import os
import re
import subprocess

import numpy as np
import pytest

import dger_sgns as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "dger_sgns.h")
HEADERS = [os.path.join(ROOT, "include", h) for h in ("dger_sgns.h", "dger_walk.h")]


def header_functions():
    names = set()
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names |= set(re.findall(r"\b(dger_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_function():
    lib = G.lib()
    declared = header_functions()
    assert len(declared) >= 25
    for name in declared:
        assert hasattr(lib, name), name
    assert declared == set(G.EXPORTS)


def test_default_options_are_the_reference_defaults():
    o = G.default_opts()
    # dsgl.hpp:67-76
    assert (o.size, o.window, o.negative, o.batch_size, o.min_count, o.iter, o.threads) == \
        (100, 5, 5, 11, 5, 5, 12)
    assert np.float32(o.alpha) == np.float32(0.1) and np.float32(o.sample) == np.float32(1e-3)
    assert o.table_size == 100_000_000 and o.precision == 0 and o.workers == 0


def test_parse_args_flags():
    o = G.parse_args(["prog", "-size", "128", "-window", "10", "-negative", "5", "-batch-size", "21",
                      "-sample", "1e-3", "-alpha", "0.01", "-min-count", "0", "-iter", "1",
                      "-eoutput", "emb.txt", "-threads", "9", "-unknown-flag", "3", "-disk",
                      "-binary", "1", "-precision", "bf16", "-workers", "64"])
    assert (o.size, o.window, o.negative, o.batch_size, o.min_count, o.iter, o.threads) == \
        (128, 10, 5, 21, 0, 1, 9)
    assert o.eoutput == b"emb.txt" and o.disk == 1 and o.binary == 1
    assert o.precision == 1 and o.workers == 64
    assert np.float32(o.alpha) == np.float32(0.01)


def test_parse_args_gpu_options():
    o = G.parse_args(["prog", "-precision", "fp8", "-graph", "0", "-autotune", "1", "-table-size", "1000"])
    assert o.precision == G.PRECISION["fp8"] == 3 and o.graph == 0 and o.autotune == 1 and o.table_size == 1000
    d = G.default_opts()
    assert d.graph == 1 and d.autotune == 0 and d.precision == 0
    for name, code in G.PRECISION.items():
        assert G.parse_args(["prog", "-precision", name]).precision == code
        assert G.parse_args(["prog", "-precision", str(code)]).precision == code


def test_parse_args_batching_options():
    d = G.default_opts()
    # parity defaults: -batch-size replaced by 2*window+1 (dsgl.hpp:832), two
    # target positions per negative group (dsgl.hpp:1136-1158)
    assert d.batch_parity == 1 and d.share == 2 and d.rounding == 1 and d.masters == 0
    o = G.parse_args(["prog", "-batch-parity", "0", "-batch-size", "7", "-share", "4", "-autotune", "2",
                      "-rounding", "sr", "-masters", "1"])
    assert (o.batch_parity, o.batch_size, o.share, o.autotune, o.rounding, o.masters) == (0, 7, 4, 2, 0, 1)


def test_parse_args_missing_value():
    with pytest.raises(G.DgerError) as e:
        G.parse_args(["prog", "-window"])
    assert e.value.code == G.DGER_E_ARG
    with pytest.raises(G.DgerError) as e:
        G.parse_args(["prog", "-precision", "fp16"])
    assert e.value.code == G.DGER_E_ARG


def test_bad_options_rejected_before_device():
    cn = np.ones(10, np.int32)
    for kw, code in [(dict(size=7), G.DGER_E_UNSUPPORTED), (dict(size=2048), G.DGER_E_UNSUPPORTED),
                     (dict(window=0), G.DGER_E_UNSUPPORTED), (dict(negative=-1), G.DGER_E_UNSUPPORTED),
                     (dict(iter=0), G.DGER_E_ARG), (dict(table_size=0), G.DGER_E_ARG),
                     # FP8 rows need 4-byte aligned scales and the whole-row step paths
                     (dict(size=66, precision="fp8"), G.DGER_E_UNSUPPORTED),
                     (dict(size=64, window=17, precision="fp8"), G.DGER_E_UNSUPPORTED),
                     (dict(size=64, negative=7, precision="fp8"), G.DGER_E_UNSUPPORTED),
                     (dict(precision=9), G.DGER_E_ARG),
                     (dict(share=3), G.DGER_E_ARG), (dict(share=32), G.DGER_E_ARG),
                     (dict(batch_parity=2), G.DGER_E_ARG), (dict(batch_parity=0, batch_size=0), G.DGER_E_ARG),
                     (dict(autotune=3), G.DGER_E_ARG)]:
        with pytest.raises(G.DgerError) as e:
            G.Trainer(G.default_opts(**kw), cn)
        assert e.value.code == code, kw
    with pytest.raises(G.DgerError) as e:
        G.Trainer(G.default_opts(size=8), np.ones(1, np.int32))
    assert e.value.code == G.DGER_E_VOCAB
    with pytest.raises(G.DgerError) as e:
        G.Trainer(G.default_opts(size=8), np.array([1, -1, 2], np.int32))
    assert e.value.code == G.DGER_E_ARG


def test_dsgl_small_vocabulary_rejected():
    # dsgl.hpp:370 asserts a vocabulary of more than 1000 vertices.
    cn = np.ones(1000, np.int32)
    corpus = np.array([1, 2, 3, -1], np.int32)
    with pytest.raises(G.DgerError) as e:
        G.dsgl(["prog", "-size", "8"], cn, corpus)
    assert e.value.code == G.DGER_E_VOCAB


def test_no_device_reports_hip_error():
    if G.device_count() > 0:
        pytest.skip("a HIP device is present; this checks the no-device error path")
    with pytest.raises(G.DgerError) as e:
        G.Trainer(G.default_opts(size=8), np.ones(10, np.int32))
    assert e.value.code in (G.DGER_E_HIP, G.DGER_E_ARG)


def test_walk_bad_graphs_rejected_before_device():
    rowptr = np.array([0, 1, 2], np.int64)
    with pytest.raises(G.DgerError) as e:      # neighbour id out of range
        G.walk(rowptr, np.array([1, 5], np.int32), np.zeros(2, np.int32), G.walk_default_opts(n_walkers=2))
    assert e.value.code == G.DGER_E_ARG
    with pytest.raises(G.DgerError) as e:      # decreasing rowptr
        G.walk(np.array([0, 2, 1], np.int64), np.array([1, 0], np.int32), np.zeros(2, np.int32),
               G.walk_default_opts(n_walkers=2))
    assert e.value.code == G.DGER_E_ARG
    with pytest.raises(G.DgerError) as e:      # max_length outside [2, 2^20]
        G.walk(rowptr, np.array([1, 0], np.int32), np.zeros(2, np.int32),
               G.walk_default_opts(n_walkers=2, max_length=1))
    assert e.value.code == G.DGER_E_UNSUPPORTED
    with pytest.raises(G.DgerError) as e:      # unsorted row for common neighbours
        G.common_neighbors(np.array([0, 2, 3, 4], np.int64), np.array([2, 1, 0, 0], np.int32))
    assert e.value.code == G.DGER_E_ARG


def test_record_stride():
    assert G.record_stride(10, 5) == 8 + 40 + 14


def test_cxx_shim_links(tmp_path):
    """A C++ caller of the header-only dsgl() shim compiles, links and gets the
    vocabulary error code back without a device."""
    src = tmp_path / "caller.cpp"
    src.write_text(r'''
#include "dger_sgns.h"
#include <cstdio>
int main() {
  std::vector<int> cn(10, 1), corpus = {1, 2, 3, -1};
  std::vector<uint32_t> new_sort;
  const char* argv[] = {"prog", "-size", "8", nullptr};
  int rc = dsgl(3, (char**)argv, &cn, &new_sort, &corpus);
  std::printf("%d %s\n", rc, dger_last_error());
  return rc == DGER_E_VOCAB ? 0 : 1;
}
''')
    libdir = os.path.dirname(G.LIB_PATH)
    exe = tmp_path / "caller"
    subprocess.run(["g++", "-std=c++17", str(src), "-I", os.path.join(ROOT, "include"), "-L", libdir,
                    "-ldger_sgns", f"-Wl,-rpath,{libdir}", "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr


def test_walk_header_is_plain_c(tmp_path):
    """A C99 caller of dger_walk.h and dger_sgns.h compiles, links, reads the
    walk defaults (minLength 20, R 0.999: include/walk.hpp:404, :24) and gets
    DGER_E_ARG for a graph with an out-of-range neighbour, without a device."""
    src = tmp_path / "walk_caller.c"
    src.write_text(r'''
#include "dger_sgns.h"
#include "dger_walk.h"
#include <stdio.h>
int main(void) {
  dger_walk_opts o;
  dger_walk_default_opts(&o);
  if (o.min_length != 20 || o.r_threshold != 0.999 || o.max_length != 1024) return 2;
  int64_t rowptr[3] = {0, 1, 2};
  int32_t col[2] = {1, 7}, cnb[2] = {0, 0}, vcn[2], corpus[8];
  uint64_t n = 0;
  o.n_walkers = 2;
  int rc = dger_walk_run(rowptr, col, cnb, 2, &o, corpus, 8, &n, vcn, NULL);
  printf("%d %s\n", rc, dger_last_error());
  return rc == DGER_E_ARG ? 0 : 1;
}
''')
    libdir = os.path.dirname(G.LIB_PATH)
    exe = tmp_path / "walk_caller"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", str(src), "-I", os.path.join(ROOT, "include"),
                    "-L", libdir, "-ldger_sgns", f"-Wl,-rpath,{libdir}", "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
