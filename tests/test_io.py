"""Binary system files (include/xamg_io.h) in the layout of the reference's reader
(src/io/detail/reader.inl:47-177), checked against files written byte by byte with numpy
and against the reference's block split and fallbacks."""
import os

import numpy as np
import pytest


def numpy_file(path, row, col, val, b=None, x=None):
    """the layout written independently: u32 n, u32 nnz, u32 row, u32 col, f64 val, f64 b, f64 x"""
    parts = [np.array([len(row) - 1, len(col)], np.uint32).tobytes(),
             np.asarray(row, np.uint32).tobytes(), np.asarray(col, np.uint32).tobytes(),
             np.asarray(val, np.float64).tobytes()]
    for v in (b, x):
        if v is not None:
            parts.append(np.asarray(v, np.float64).tobytes())
    with open(path, "wb") as f:
        f.write(b"".join(parts))


@pytest.fixture
def system(X):
    row, col, val = X.cube(6, 5, 4)
    n, nv = len(row) - 1, 3
    rng = np.random.default_rng(0)
    return row, col, val, rng.standard_normal(n * nv), rng.standard_normal(n * nv), nv


def test_reader_matches_numpy_layout(X, tmp_path, system):
    row, col, val, b, x, nv = system
    p = str(tmp_path / "sys.bin")
    numpy_file(p, row, col, val, b, x)
    r = X.read_system(p, nv)
    assert r["nrows"] == len(row) - 1 and r["found"] == 3 and r["row_begin"] == 0
    for k, v in (("row", row), ("col", col), ("val", val), ("b", b), ("x", x)):
        np.testing.assert_array_equal(r[k], v)


def test_writer_bytes_equal_numpy_layout(X, tmp_path, system):
    row, col, val, b, x, nv = system
    p1, p2 = str(tmp_path / "a.bin"), str(tmp_path / "b.bin")
    X.write_system(p1, row, col, val, nv, b, x)
    numpy_file(p2, row, col, val, b, x)
    assert open(p1, "rb").read() == open(p2, "rb").read()
    X.write_system(p1, row, col, val)
    numpy_file(p2, row, col, val)
    assert open(p1, "rb").read() == open(p2, "rb").read()


@pytest.mark.parametrize("nprocs", [1, 2, 3, 7])
def test_rank_blocks_reassemble_the_system(X, tmp_path, system, nprocs):
    # reader.inl:57-66: nrows/nprocs rows per rank, the last takes the remainder; each
    # block's row pointers are rebased to its first non-zero, columns stay global
    row, col, val, b, x, nv = system
    n = len(row) - 1
    p = str(tmp_path / "sys.bin")
    numpy_file(p, row, col, val, b, x)
    rows, cols, vals, bs, xs = [0], [], [], [], []
    for q in range(nprocs):
        rb, nr = X.io_block_range(n, nprocs, q)
        assert rb == (n // nprocs) * q
        assert nr == (n // nprocs if q < nprocs - 1 else n - rb)
        r = X.read_system(p, nv, nprocs, q)
        assert r["row"][0] == 0 and r["row_begin"] == rb
        rows += (r["row"][1:] + rows[-1]).tolist()
        cols.append(r["col"])
        vals.append(r["val"])
        bs.append(r["b"])
        xs.append(r["x"])
    np.testing.assert_array_equal(rows, row)
    np.testing.assert_array_equal(np.concatenate(cols), col)
    np.testing.assert_array_equal(np.concatenate(vals), val)
    np.testing.assert_array_equal(np.concatenate(bs), b)
    np.testing.assert_array_equal(np.concatenate(xs), x)


def test_missing_vectors_fall_back_like_the_reference(X, tmp_path, system):
    # reader.inl:148-168: no b -> b = 1 (and x is not read); b but no x -> x = 0
    row, col, val, b, x, nv = system
    p = str(tmp_path / "sys.bin")
    numpy_file(p, row, col, val)
    r = X.read_system(p, nv)
    assert r["found"] == 0 and (r["b"] == 1.0).all() and (r["x"] == 0.0).all()
    numpy_file(p, row, col, val, b)
    r = X.read_system(p, nv)
    assert r["found"] == 1 and np.array_equal(r["b"], b) and (r["x"] == 0.0).all()
    # a file holding b for fewer RHS than asked has no complete b section
    r = X.read_system(p, nv + 1)
    assert r["found"] == 0 and (r["b"] == 1.0).all()


def test_corrupt_files_are_rejected(X, tmp_path, system):
    row, col, val, b, x, nv = system
    p = str(tmp_path / "sys.bin")
    numpy_file(p, row, col, val, b, x)
    data = open(p, "rb").read()
    with open(p, "wb") as f:  # cut inside the value array
        f.write(data[: 8 + 4 * len(row) + 4 * len(col) + 8])
    with pytest.raises(X.XamgError):
        X.read_system(p, nv)
    bad = col.copy()
    bad[5] = len(row) + 10
    numpy_file(p, row, bad, val)
    with pytest.raises(X.XamgError):
        X.read_system(p, nv)
    badrow = row.copy()
    badrow[3] = badrow[4] + 1
    numpy_file(p, badrow, col, val)
    with pytest.raises(X.XamgError):
        X.read_system(p, nv)
    with pytest.raises(X.XamgError):
        X.read_system(str(tmp_path / "missing.bin"), nv)


def test_empty_matrix_round_trip(X, tmp_path):
    p = str(tmp_path / "e.bin")
    X.write_system(p, np.zeros(1, np.uint32), np.zeros(0, np.uint32), np.zeros(0))
    r = X.read_system(p, 1)
    assert r["nrows"] == 0 and len(r["col"]) == 0 and len(r["row"]) == 1
