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


def _hub_graph(gbo, n=40000, hubs=(0, 5, 17, 39990), seed=3):
    """Random sparse graph plus hub vertices joined to most vertices and a
    100-clique: block rows far above 2048 nonzeros, packed in pieces."""
    from graphs import sym_csr
    rng = np.random.default_rng(seed)
    u = [rng.integers(0, n, 4 * n)]
    v = [rng.integers(0, n, 4 * n)]
    for h in hubs:
        nb = rng.choice(n, int(0.6 * n), replace=False)
        u.append(np.full(len(nb), h))
        v.append(nb)
    cu, cv = np.triu_indices(100, 1)
    u.append(cu + 20000)
    v.append(cv + 20000)
    rp, ci = sym_csr(n, np.concatenate(u), np.concatenate(v))
    return gbo.Csr(n, n, rp, ci)


@pytest.mark.parametrize("bs", BS)
@pytest.mark.parametrize("direction", (0, 1))
def test_pack_hub_rows_in_pieces(gbo, po, bs, direction):
    """Block rows of more than 2048 nonzeros are packed by several groups over
    block-column ranges; the result is the oracle's byte for byte."""
    g = _hub_graph(gbo)
    nnz_per_block_row = np.add.reduceat(np.diff(g.rowptr), np.arange(0, g.nrows, bs))
    assert nnz_per_block_row.max() > 10 * 2048
    A = gbo.BsrMatrix.from_csr(g, bs, direction)
    rp, ci, tw = A.host_arrays()
    orp, oci, otw = po.pack(bs, g.nrows, g.rowptr, g.colind, direction)
    assert np.array_equal(rp, orp) and np.array_equal(ci, oci) and np.array_equal(tw, otw)
    if direction == 0:
        w = gbo.edge_weights(g, 1, 64, seed=2)
        tv = gbo.TileValues(A, g, w)
        otv, owoff = po.pack_values(bs, g.nrows, g.rowptr, g.colind, w, rp, ci, tw)
        assert np.array_equal(tv.woff.to_numpy(np.int64, A.nblocks + 1), owoff)
        assert np.array_equal(tv.tval.to_numpy(np.uint32, tv.count), otv)


@pytest.mark.parametrize("bs", (8, 64))
def test_pack_device_unsorted_heavy_row_rejected(gbo, bs):
    """On the device path (no host sort) a row whose block columns go down is
    refused, also when the row is long enough to be packed in pieces and the
    out-of-order pair sits anywhere in it."""
    n = 20000
    cols = np.arange(0, n, 2, dtype=np.int32)  # 10000 sorted entries in row 0
    rng = np.random.default_rng(bs)
    for a, b in [(0, 9999), (100, 4000), (2047, 2048)] + [tuple(sorted(rng.choice(10000, 2, replace=False))) for _ in range(4)]:
        c = cols.copy()
        c[a], c[b] = c[b], c[a]
        if (c[a] // bs) == (c[b] // bs):
            continue  # same block column: no order is broken at this block size
        rp = np.zeros(n + 1, np.int32)
        rp[1:] = len(c)
        d_rp, d_ci = gbo.DeviceBuffer.from_numpy(rp), gbo.DeviceBuffer.from_numpy(c)
        with pytest.raises(gbo.GboError) as e:
            gbo.BsrMatrix.from_device_csr(d_rp, d_ci, n, n, bs)
        assert "sorted" in str(e.value), (a, b)
        d_rp.free()
        d_ci.free()
    rp = np.zeros(n + 1, np.int32)
    rp[1:] = len(cols)
    A = gbo.BsrMatrix.from_device_csr(gbo.DeviceBuffer.from_numpy(rp), gbo.DeviceBuffer.from_numpy(cols), n, n, bs)
    assert A.nblocks == (n + bs - 1) // bs


def test_fill_refuses_rowptr_of_another_matrix(gbo):
    """gbo_csr2bsr_fill_dev writes as many tiles as the CSR has; a rowptr whose
    total disagrees (buffers sized for another matrix) is refused."""
    g = gbo.gen_er(500, 0.05, 2)
    L = gbo.lib()
    d_rp, d_ci = gbo.DeviceBuffer.from_numpy(g.rowptr), gbo.DeviceBuffer.from_numpy(g.colind)
    nbr = (500 + 31) // 32
    brp = np.zeros(nbr + 1, np.int32)
    brp[-1] = 3
    d_brp = gbo.DeviceBuffer.from_numpy(brp)
    d_bci, d_t = gbo.DeviceBuffer(4 * 3), gbo.DeviceBuffer(4 * 32 * 3)
    rc = L.gbo_csr2bsr_fill_dev(32, 0, d_rp.ptr, d_ci.ptr, 500, 500, d_brp.ptr, d_bci.ptr, d_t.ptr, None)
    assert rc == -1 and b"rowptr" in L.gbo_last_error()
