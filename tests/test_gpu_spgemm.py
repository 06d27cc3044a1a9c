"""General bit-BSR SpGEMM C = A * B (gbo_bsrbmm, gbo_bsrbmm_symbolic /
_numeric / _values) against the oracle's element-level product.

The oracle's C = A * B (or_spgemm: rows, columns ascending, integer counts;
pinned against scipy in test_oracle.py) is packed by the oracle's packer:
the GPU's C must equal it byte for byte (rowptr, colind, tiles -- so C has
exactly the tiles the product's pattern has), and its values at C's set
bits, read in tile bit order, must equal the oracle's counts.  Cases: R-MAT
A*A, L*L^T (the bsrbmm command line's product), the king grid, rectangular
operands with partial blocks, empty operands, every block size.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
BS = (8, 16, 32, 64)
WORD = {8: np.uint8, 16: np.uint16, 32: np.uint32, 64: np.uint64}


def _rect(gbo, m, n, p, seed):
    """A random m x n pattern (rows sorted, no duplicates)."""
    rng = np.random.default_rng(seed)
    d = rng.random((m, n)) < p
    rp = np.concatenate([[0], np.cumsum(d.sum(1))]).astype(np.int32)
    ci = np.nonzero(d)[1].astype(np.int32)
    return gbo.Csr(m, n, rp, ci)


def _tile_triples(bs, nbr, rp, ci, tw, counts=None, off=None):
    """(row, col[, count]) of every set bit of a row-packed bit-BSR, in CSR
    order (the counts taken in tile bit order)."""
    T = len(ci)
    if T == 0:
        return np.zeros(0, np.int64), np.zeros(0, np.int64), np.zeros(0, np.int64)
    words = tw.reshape(T, bs).astype(np.uint64)
    shifts = (bs - 1 - np.arange(bs)).astype(np.uint64)
    bits = ((words[:, :, None] >> shifts[None, None, :]) & np.uint64(1)).astype(bool)  # [T, r, c]
    t_idx, r_idx, c_idx = np.nonzero(bits)  # tile bit order (C order)
    trow = np.repeat(np.arange(nbr), np.diff(rp))
    rows = trow[t_idx].astype(np.int64) * bs + r_idx
    cols = ci[t_idx].astype(np.int64) * bs + c_idx
    if counts is not None:
        assert off[-1] == len(counts) == len(rows)
        assert np.array_equal(off[:-1], np.concatenate([[0], np.cumsum(bits.reshape(T, -1).sum(1))[:-1]]))
    o = np.lexsort((cols, rows))
    return rows[o], cols[o], (counts[o] if counts is not None else None)


def _check_product(gbo, po, Acsr, Bcsr, bs, counts=True):
    A = gbo.BsrMatrix.from_csr(Acsr, bs, 0)
    B = gbo.BsrMatrix.from_csr(Bcsr, bs, 0)
    rp, ci, val = po.spgemm(Acsr.nrows, Acsr.rowptr, Acsr.colind, Bcsr.rowptr, Bcsr.colind, Bcsr.ncols)
    wrp, wci, wtw = po.pack(bs, Acsr.nrows, rp, ci)
    if counts:
        C, off, cnt = gbo.bsrbmm(A, B, counts=True)
    else:
        C = gbo.bsrbmm(A, B)
    assert (C.nrows, C.ncols, C.bs) == (Acsr.nrows, Bcsr.ncols, bs)
    grp, gci, gtw = C.host_arrays()
    assert np.array_equal(grp, wrp), "C block rowptr"
    assert np.array_equal(gci, wci), "C block columns"
    assert np.array_equal(gtw, wtw), "C tiles"
    if counts:
        rows, cols, cv = _tile_triples(bs, C.nblockrows, grp, gci, gtw, cnt, off)
        want_rows = np.repeat(np.arange(Acsr.nrows), np.diff(rp))
        assert np.array_equal(rows, want_rows) and np.array_equal(cols, ci)
        assert np.array_equal(cv, val), "C values"
    # the two-phase caller-allocates form gives the same C
    prp, pci, ptw, poff, pnnz = gbo.bsrbmm_phases(A, B)
    assert np.array_equal(prp, wrp) and np.array_equal(pci, wci) and np.array_equal(ptw, wtw)
    assert pnnz == len(ci) and poff[-1] == pnnz
    for M in (A, B, C):
        M.free()


@pytest.mark.parametrize("bs", BS)
def test_rmat_square(gbo, po, bs):
    g = gbo.gen_rmat(10, 8, 0.57, 0.19, 0.19, seed=11)
    _check_product(gbo, po, g, g, bs)


@pytest.mark.parametrize("bs", BS)
def test_lower_times_transpose(gbo, po, bs):
    """L * L^T, the product the bsrbmm<N> command line computes."""
    g = gbo.gen_rmat(12, 8, 0.57, 0.19, 0.19, seed=2)
    L = g.tril(strict=False)
    _check_product(gbo, po, L, L.transpose(), bs)


@pytest.mark.parametrize("bs", BS)
def test_king_grid(gbo, po, bs):
    g = gbo.gen_king(200)
    _check_product(gbo, po, g, g, bs)


@pytest.mark.parametrize("bs", BS)
def test_rectangular_partial_blocks(gbo, po, bs):
    A = _rect(gbo, 301, 517, 0.02, 1)
    B = _rect(gbo, 517, 203, 0.03, 2)
    _check_product(gbo, po, A, B, bs)
    # a directed pair: A * A^T != A^T * A
    _check_product(gbo, po, A.transpose(), A, bs)


@pytest.mark.parametrize("bs", (8, 64))
def test_empty_operands(gbo, po, bs):
    A = _rect(gbo, 100, 130, 0.0, 3)
    B = _rect(gbo, 130, 90, 0.05, 4)
    _check_product(gbo, po, A, B, bs)
    _check_product(gbo, po, B.transpose(), A.transpose(), bs)


@pytest.mark.parametrize("bs", (8, 16))
def test_rmat13_square(gbo, po, bs):
    """R-MAT scale 13, edge factor 16: skewed rows (hub block rows pair with
    thousands of tiles; ~10^7 candidate tile pairs at bs 8)."""
    g = gbo.gen_rmat(13, 16, 0.57, 0.19, 0.19, seed=1)
    _check_product(gbo, po, g, g, bs)


def test_operand_checks(gbo):
    A = gbo.BsrMatrix.from_csr(_rect(gbo, 64, 80, 0.1, 5), 16, 0)
    B = gbo.BsrMatrix.from_csr(_rect(gbo, 64, 80, 0.1, 6), 16, 0)
    with pytest.raises(gbo.GboError):  # inner dimensions differ
        gbo.bsrbmm(A, B)
    B2 = gbo.BsrMatrix.from_csr(_rect(gbo, 80, 64, 0.1, 6), 32, 0)
    with pytest.raises(gbo.GboError):  # block sizes differ
        gbo.bsrbmm(A, B2)
    B3 = gbo.BsrMatrix.from_csr(_rect(gbo, 80, 64, 0.1, 6), 16, 1)
    with pytest.raises(gbo.GboError):  # column-packed right operand
        gbo.bsrbmm(A, B3)
