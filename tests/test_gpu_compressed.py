"""The reference's compact-index block (csr_matrix::compress, csr_matrix.inl:86-170) on the
GPU: empty rows and columns dropped, compact column ids of 1, 2 or 4 bytes (the width
define_int_type picks for the compact column count) decoded in the SpMM kernel, output
rows through row_ind.  Products and their order are those of the uncompressed block, so
y is bit-identical to the oracle's SpMV of the uncompressed block on the full operand."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

WIDTH = {0: np.uint8, 1: np.uint16, 2: np.uint32}


def block(rng, nrows, ncols_full, used, per_row):
    # rows with per_row nonzeros over `used` distinct columns; every 3rd row empty
    cols_pool = np.sort(rng.choice(ncols_full, size=used, replace=False))
    row = [0]
    col = []
    for i in range(nrows):
        k = 0 if i % 3 == 0 else int(rng.integers(1, per_row + 1))
        col.extend(rng.choice(cols_pool, size=k, replace=False).tolist())
        row.append(len(col))
    col = np.array(col, np.uint32)
    # make sure every pooled column is used once, so the compact count is `used`
    col[:used] = cols_pool[rng.permutation(used)] if len(col) >= used else col[:used]
    return np.array(row, np.uint32), col, rng.standard_normal(len(col))


@pytest.mark.parametrize("ncols_full,used,per_row,width", [
    (400, 200, 12, 0), (9000, 3000, 40, 1), (150000, 70000, 400, 2)])
@pytest.mark.parametrize("nv", [1, 4, 32])
@pytest.mark.parametrize("prec", [0, 1])
def test_compressed_block_bit_exact(gpu, O, ncols_full, used, per_row, width, nv, prec):
    X = gpu
    rng = np.random.default_rng(ncols_full + nv + prec)
    nrows, off = 900, 37
    row, col, val = block(rng, nrows, ncols_full, used, per_row)
    cmp = O.compress(row, col, val, ncols_full)
    assert cmp["width"] == width and cmp["ncols"] == used
    assert cmp["nrows"] == sum(1 for i in range(nrows) if row[i + 1] > row[i])
    ccol = cmp["col"].astype(WIDTH[width])
    M = X.Matrix.compressed(nrows, cmp["row"], ccol, val, row_ind=cmp["row_ind"],
                            col_ind=cmp["col_ind"], col_offset=off,
                            x_rows=ncols_full + off, prec=prec)
    assert M.info()["index_bits"] == 8 * np.dtype(WIDTH[width]).itemsize
    dt = np.float64 if prec == 0 else np.float32
    xf = rng.standard_normal((ncols_full + off) * nv).astype(dt)
    dxf = X.DeviceArray.from_host(xf)
    dxc = X.DeviceArray(used * nv, dt)
    M.gather(nv, dxf, dxc)
    xc = dxc.to_host()
    assert np.array_equal(xc.reshape(-1, nv), xf.reshape(-1, nv)[cmp["col_ind"] + off])
    xblk = xf.reshape(-1, nv)[off:off + ncols_full].reshape(-1)
    y0 = rng.standard_normal(nrows * nv).astype(dt)
    dy = X.DeviceArray.from_host(y0)
    M.spmv(nv, dxc, dy, accumulate=True)
    ref = O.spmv(row, col, val, xblk, nv, y=y0, ncols=ncols_full, f32=prec == 1)
    assert np.array_equal(dy.to_host(), ref)
    # y = A x writes the block's rows only (csr_Ax_y on an indexed block)
    dy.upload(y0)
    M.spmv(nv, dxc, dy)
    full = O.spmv(row, col, val, xblk, nv, ncols=ncols_full, f32=prec == 1).reshape(-1, nv)
    expect = y0.reshape(-1, nv).copy()
    expect[cmp["row_ind"]] = full[cmp["row_ind"]]
    assert np.array_equal(dy.to_host().reshape(-1, nv), expect)


def test_compressed_block_rejects_bad_input(gpu):
    X = gpu
    row = np.array([0, 2, 3], np.uint32)
    col = np.array([0, 5, 1], np.uint8)
    with pytest.raises(X.XamgError):  # compact column 5 >= ncols 2
        X.Matrix.compressed(4, row, col, np.ones(3), row_ind=[0, 2], col_ind=[3, 9],
                            x_rows=10)
    with pytest.raises(X.XamgError):  # row_ind not ascending
        X.Matrix.compressed(4, row, np.array([0, 1, 1], np.uint8), np.ones(3),
                            row_ind=[2, 1], col_ind=[3, 9], x_rows=10)
    with pytest.raises(X.XamgError):  # col_ind + offset beyond the full operand
        X.Matrix.compressed(4, row, np.array([0, 1, 1], np.uint8), np.ones(3),
                            row_ind=[0, 2], col_ind=[3, 9], col_offset=1, x_rows=10)
