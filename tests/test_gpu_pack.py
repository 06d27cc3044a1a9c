"""GPU packer parity: gbo_csr2bsr_* / gbo_pack_vec / gbo_csr2bsr_values_u32
against the oracle's restatement of csr2bsr_batch_{N} (+ _row/_col) and
ToBit{N}Row, byte for byte."""
import numpy as np
import pytest

from graphs import cases

pytestmark = pytest.mark.gpu
BS = (8, 16, 32, 64)


@pytest.fixture(scope="module")
def graph_cases(gbo):
    return cases(gbo)


@pytest.mark.parametrize("bs", BS)
@pytest.mark.parametrize("direction", (0, 1))
def test_pack_matches_oracle(gbo, po, graph_cases, bs, direction):
    for name, g in graph_cases:
        A = gbo.BsrMatrix.from_csr(g, bs, direction)
        rp, ci, tw = A.host_arrays()
        orp, oci, otw = po.pack(bs, g.nrows, g.rowptr, g.colind, direction)
        assert np.array_equal(rp, orp), name
        assert np.array_equal(ci, oci), name
        assert np.array_equal(tw, otw), name
        # every nonzero is one bit (the cases have no duplicates)
        assert sum(bin(int(w)).count("1") for w in tw) == g.nnz, name


@pytest.mark.parametrize("bs", BS)
def test_reference_shaped_batch_call(gbo, po, bs):
    g = gbo.gen_er(1024, 0.02, 0)
    d_rp, d_ci, d_t, nbr, nb = gbo.csr2bsr_batch(g, bs, 0)
    orp, oci, otw = po.pack(bs, g.nrows, g.rowptr, g.colind, 0)
    assert nb == len(oci)
    assert np.array_equal(d_rp.to_numpy(np.int32, nbr + 1), orp)
    assert np.array_equal(d_ci.to_numpy(np.int32, nb), oci)
    assert np.array_equal(d_t.to_numpy(gbo.WORD_DTYPE[bs], nb * bs), otw)


def test_unsorted_rows_and_duplicates(gbo, po):
    # mmio_highlevel semantics keep duplicates and file order inside a row;
    # the packer ORs duplicates and needs no sorted input.
    rng = np.random.default_rng(5)
    n = 300
    rows = rng.integers(0, n, 4000)
    cols = rng.integers(0, n, 4000)
    order = np.argsort(rows, kind="stable")
    rp = np.zeros(n + 1, np.int32)
    np.add.at(rp, rows + 1, 1)
    rp = np.cumsum(rp).astype(np.int32)
    ci = cols[order].astype(np.int32)  # unsorted inside rows, with repeats
    g = gbo.Csr(n, n, rp, ci)
    for bs in BS:
        A = gbo.BsrMatrix.from_csr(g, bs)
        got = A.host_arrays()
        want = po.pack(bs, n, rp, ci)
        for a, b in zip(got, want):
            assert np.array_equal(a, b)


def test_bad_column_rejected(gbo):
    g = gbo.Csr(4, 4, np.array([0, 1, 2, 2, 2], np.int32), np.array([1, 9], np.int32))
    with pytest.raises(gbo.GboError) as e:
        gbo.BsrMatrix.from_csr(g, 8)
    assert e.value.code == -1


def test_nblocks_mismatch_rejected(gbo):
    g = gbo.gen_er(200, 0.05, 1)
    L = gbo.lib()
    nbr = (200 + 31) // 32
    d_rp = gbo.DeviceBuffer(4 * (nbr + 1))
    d_ci = gbo.DeviceBuffer(4 * 1000)
    d_t = gbo.DeviceBuffer(4 * 32 * 1000)
    rc = L.gbo_csr2bsr_batch(32, 0, g.rowptr.ctypes.data, g.colind.ctypes.data, 200, 200, g.nnz,
                             d_rp.ptr, d_ci.ptr, d_t.ptr, nbr, 3, None)
    assert rc == -1 and b"nblocks" in L.gbo_last_error()


@pytest.mark.parametrize("bs", BS)
def test_pack_vec(gbo, po, bs):
    rng = np.random.default_rng(bs)
    for n in (1, bs - 1, bs, 1000, 4099):
        x = (rng.random(n) < 0.3).astype(np.float32) * rng.random(n).astype(np.float32)
        x[rng.random(n) < 0.05] = -1.0  # only > 0 sets a bit
        assert np.array_equal(gbo.pack_vec(bs, x), po.pack_vec(bs, x))


@pytest.mark.parametrize("bs", BS)
def test_values_in_tile_bit_order(gbo, po, graph_cases, bs):
    for name, g in graph_cases[:4]:
        w = gbo.edge_weights(g, 1, 64, seed=11)
        A = gbo.BsrMatrix.from_csr(g, bs)
        tv = gbo.TileValues(A, g, w)
        rp, ci, tw = A.host_arrays()
        otv, owoff = po.pack_values(bs, g.nrows, g.rowptr, g.colind, w, rp, ci, tw)
        assert np.array_equal(tv.woff.to_numpy(np.int64, A.nblocks + 1), owoff), name
        assert np.array_equal(tv.tval.to_numpy(np.uint32, tv.count), otv), name
