"""Embedding files in my_saveModel's format (include/dsgl.hpp:1443-1461),
written by dger_emb_write (host only).

The text path must give the bytes printf("%f ") gives for every value: it is
checked against Python's "%f" formatting, which rounds the exact binary value
to six decimals, half to even, like glibc's printf; the values include exact
ties (odd multiples of 1/128 and of 2^-k), denormals, signed zeros, values
rounding to zero from below, large magnitudes, infinities and NaN.  The binary
path must round-trip every bit."""
import numpy as np
import pytest

import dger_sgns as G


def expected_text(ids, rows):
    out = [f"{rows.shape[0]} {rows.shape[1]}\n"]
    for i, r in zip(ids, rows):
        out.append(f"{int(i) & 0xFFFFFFFF} " + "".join("%f " % float(x) for x in r) + "\n")
    return "".join(out).encode()


def edge_values():
    v = [0.0, -0.0, 1.0, -1.0, 0.5, 1 / 128, 3 / 128, -5 / 128, 2.0 ** -21, 2.0 ** -20, 5e-7, -5e-7,
         4.9999997e-7, 1e-7, -1e-9, 0.0000015, 0.9999995, 9.9999995, 123456.789, -98765.4321,
         2.0 ** 23 + 1, 2.0 ** 43 - 2 ** 20, 2.0 ** 43, 2.0 ** 44, 3.4e38, -3.4e38, 1e-45, -1e-45,
         1.17549435e-38, np.inf, -np.inf]
    # ties at the sixth decimal: x * 1e6 = n + 1/2 exactly
    v += [k / 128 for k in range(1, 400, 2)] + [k * 2.0 ** -13 for k in range(1, 200, 3)]
    return np.array(v, np.float32)


def test_text_matches_printf_on_edge_values(tmp_path):
    vals = edge_values()
    D = 8
    vals = np.concatenate([vals, np.zeros((-vals.size) % D, np.float32)])
    rows = vals.reshape(-1, D)
    ids = np.arange(rows.shape[0], dtype=np.int32) * 7919
    p = str(tmp_path / "e.txt")
    G.emb_write(p, ids, rows, 0)
    assert open(p, "rb").read() == expected_text(ids, rows)


def test_text_matches_printf_on_random_values(tmp_path):
    rng = np.random.default_rng(5)
    V, D = 3000, 37          # several formatting blocks and a ragged last one
    bits = rng.integers(0, 2 ** 32, (V, D), dtype=np.uint64).astype(np.uint32)
    rows = bits.view(np.float32)
    rows = np.where(np.isfinite(rows) & (np.abs(rows) < 1e30), rows, np.float32(0.25))
    rows[: V // 2] = rng.normal(0, 0.05, (V // 2, D)).astype(np.float32)   # embedding-like values
    ids = rng.permutation(V).astype(np.int32)
    p = str(tmp_path / "r.txt")
    G.emb_write(p, ids, rows, 0)
    assert open(p, "rb").read() == expected_text(ids, rows)
    rid, back = G.emb_read(p, 0)
    assert np.array_equal(rid, ids)
    # six decimals (5e-7), then the float32 rounding of the parsed value
    assert np.all(np.abs(back.astype(np.float64) - rows) <= 6e-7 * np.maximum(1, np.abs(rows)))


def test_nan_is_written_like_printf(tmp_path):
    # glibc prints the sign of a NaN (Python's "%f" does not)
    rows = np.array([[np.nan, -np.nan, 1.5]], np.float32)
    assert rows.view(np.uint32)[0, 1] >> 31 == 1
    p = str(tmp_path / "n.txt")
    G.emb_write(p, np.array([3], np.int32), rows, 0)
    assert open(p, "rb").read() == b"1 3\n3 nan -nan 1.500000 \n"


def test_binary_round_trip(tmp_path):
    rng = np.random.default_rng(6)
    V, D = 5000, 128
    rows = rng.normal(0, 0.1, (V, D)).astype(np.float32)
    rows[7, 3] = -0.0
    rows[9, 0] = 1e-40
    ids = rng.permutation(2 * V)[:V].astype(np.int32)
    p = str(tmp_path / "b.bin")
    G.emb_write(p, ids, rows, 1)
    data = open(p, "rb").read()
    # header, then "id " + raw float32 + newline per row (dsgl.hpp:1448-1453)
    assert data.startswith(b"5000 128\n%d " % ids[0])
    assert len(data) == len(b"5000 128\n") + sum(len(b"%d " % i) + 4 * D + 1 for i in ids)
    rid, back = G.emb_read(p, 1)
    assert np.array_equal(rid, ids)
    assert np.array_equal(back.view(np.uint32), rows.view(np.uint32))


def test_empty_and_bad_arguments(tmp_path):
    p = str(tmp_path / "z.txt")
    G.emb_write(p, np.zeros(0, np.int32), np.zeros((0, 4), np.float32), 0)
    assert open(p, "rb").read() == b"0 4\n"
    with pytest.raises(G.DgerError) as e:
        G.emb_write(str(tmp_path / "no" / "dir.txt"), np.zeros(1, np.int32), np.zeros((1, 4), np.float32), 0)
    assert e.value.code == G.DGER_E_IO
