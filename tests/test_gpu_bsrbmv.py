"""GPU bsrbmv parity: count / BFS step / min-plus / sum / min through the C ABI
against the oracle (bit-exact for integer ops, FP32 sum within a stated
tolerance of the FP64 oracle and bitwise reproducible).  Each op runs with
the default plan and with tiny chunks so that most block rows are split
across wavefronts (atomic and ordered-fix-up combine paths)."""
import numpy as np
import pytest

from conftest import bernoulli_x
from graphs import cases

pytestmark = pytest.mark.gpu
BS = (8, 16, 32, 64)
CHUNKS = (0, 3)  # 0 = default chunk; 3 tiles = split rows everywhere
# count also runs with a chunk that is not a multiple of the streamed
# kernel's tiles per lane and spans several of its wave steps
COUNT_CHUNKS = (0, 3, 1029)


@pytest.fixture(scope="module")
def graph_cases(gbo):
    return cases(gbo)


def packed(gbo, g, bs, chunk, variant=0):
    A = gbo.BsrMatrix.from_csr(g, bs, 0, plan=False)
    A.plan(chunk, variant)
    return A, A.host_arrays()


@pytest.mark.parametrize("bs", BS)
@pytest.mark.parametrize("chunk", COUNT_CHUNKS)
@pytest.mark.parametrize("variant", (0, 1, 2, 3))
def test_count(gbo, po, graph_cases, bs, chunk, variant):
    if variant and bs > 16:
        pytest.skip("kernel variants exist for bs 8/16 only")
    for name, g in graph_cases:
        A, (rp, ci, tw) = packed(gbo, g, bs, chunk, variant)
        x = bernoulli_x(g.ncols, 0.1, 3)
        xb = gbo.pack_vec(bs, x)
        y = gbo.bsrbmv_count(A, xb)
        assert np.array_equal(y, po.count(bs, g.nrows, rp, ci, tw, xb)), name
        # accumulate (+=) semantics of the reference kernels
        y0 = np.arange(g.nrows, dtype=np.int32)
        assert np.array_equal(gbo.bsrbmv_count(A, xb, y_init=y0), y + y0), name


@pytest.mark.parametrize("bs", (8, 16, 32))
@pytest.mark.parametrize("variant", (0, 1, 2, 3))
def test_count_long_rows(gbo, po, bs, variant):
    """Block rows of >300k tiles, each row count > 2^16, in one chunk: the
    per-lane 16-bit field accumulators of the bs 8/16 variants must be
    flushed before they wrap."""
    if variant and bs > 16:
        pytest.skip("kernel variants exist for bs 8/16 only")
    ncols, nrows = 3_000_000, 16
    cols = [np.arange(r % 8, ncols, 7, dtype=np.int32) for r in range(nrows)]
    rowptr = np.zeros(nrows + 1, np.int32)
    rowptr[1:] = np.cumsum([len(c) for c in cols])
    g = gbo.Csr(nrows, ncols, rowptr, np.concatenate(cols))
    x = bernoulli_x(ncols, 0.5, 11)
    xb = gbo.pack_vec(bs, x)
    want = np.array([int(x[c].sum()) for c in cols])
    assert want.min() > 1 << 16
    for chunk in (0, 1 << 30):
        A = gbo.BsrMatrix.from_csr(g, bs, 0, plan=False)
        A.plan(chunk, variant)
        y = gbo.bsrbmv_count(A, xb)
        assert np.array_equal(y, want), chunk
        A.free()


def test_count_dense_bruteforce(gbo):
    g = gbo.gen_er(1024, 0.02, 0)
    D = np.zeros((1024, 1024), np.int64)
    for r in range(1024):
        D[r, g.colind[g.rowptr[r]:g.rowptr[r + 1]]] = 1
    x = bernoulli_x(1024, 0.1, 0)
    want = D @ x.astype(np.int64)
    for bs in BS:
        A = gbo.BsrMatrix.from_csr(g, bs)
        assert np.array_equal(gbo.bsrbmv_count(A, gbo.pack_vec(bs, x)), want)


@pytest.mark.parametrize("bs", BS)
@pytest.mark.parametrize("chunk", COUNT_CHUNKS)
@pytest.mark.parametrize("variant", (0, 1, 3))
def test_bfs_step(gbo, po, graph_cases, bs, chunk, variant):
    """Variant 3 at bs 8 is the dense step (count kernel, then the rows with
    a nonzero count packed and the visited ones removed)."""
    if variant and bs > 16:
        pytest.skip("kernel variants exist for bs 8/16 only")
    rng = np.random.default_rng(bs + chunk)
    for name, g in graph_cases:
        A, (rp, ci, tw) = packed(gbo, g, bs, chunk, variant)
        for pf, pv in ((0.01, 0.0), (0.2, 0.5), (0.5, 0.95)):
            f = gbo.pack_vec(bs, (rng.random(g.ncols) < pf).astype(np.float32))
            v = gbo.pack_vec(bs, (rng.random(g.nrows) < pv).astype(np.float32))
            nx, c = gbo.bsrbmv_bfs_step(A, f, v)
            onx, oc = po.bfs_step(bs, g.nrows, rp, ci, tw, f, v)
            assert np.array_equal(nx, onx), name
            assert c == oc, name


@pytest.mark.parametrize("bs", BS)
@pytest.mark.parametrize("chunk", COUNT_CHUNKS)
@pytest.mark.parametrize("variant", (0, 3))
def test_minplus(gbo, po, graph_cases, bs, chunk, variant):
    """Variant 3 at bs 8/16 is the set-bit list kernel (k_value_lpt)."""
    if variant and bs > 16:
        pytest.skip("variant 3 changes the min-plus kernel at bs 8/16 only")
    rng = np.random.default_rng(100 + bs)
    for name, g in graph_cases:
        A, (rp, ci, tw) = packed(gbo, g, bs, chunk, variant)
        dist = rng.integers(0, 1000, g.nrows).astype(np.uint32)
        dist[rng.random(g.nrows) < 0.3] = gbo.INF_U32
        dist[rng.random(g.nrows) < 0.01] = gbo.INF_U32 - 3  # saturation
        assert np.array_equal(gbo.bsrbmv_minplus(A, dist), po.minplus(bs, g.nrows, rp, ci, tw, dist)), name
        w = gbo.edge_weights(g, 1, 64, seed=9)
        tv = gbo.TileValues(A, g, w)
        otv, owoff = po.pack_values(bs, g.nrows, g.rowptr, g.colind, w, rp, ci, tw)
        got = gbo.bsrbmv_minplus(A, dist, tv)
        assert np.array_equal(got, po.minplus(bs, g.nrows, rp, ci, tw, dist, otv, owoff)), name


@pytest.mark.parametrize("bs", BS)
@pytest.mark.parametrize("chunk", COUNT_CHUNKS)
@pytest.mark.parametrize("variant", (0, 3))
def test_sum_f32(gbo, po, graph_cases, bs, chunk, variant):
    if variant and bs > 16:
        pytest.skip("variant 3 changes the sum kernel at bs 8/16 only")
    rng = np.random.default_rng(200 + bs)
    for name, g in graph_cases:
        A, (rp, ci, tw) = packed(gbo, g, bs, chunk, variant)
        x = rng.random(g.ncols).astype(np.float32)
        y = gbo.bsrbmv_sum_f32(A, x)
        want = po.sum_f64(bs, g.nrows, rp, ci, tw, x.astype(np.float64))
        # FP32 accumulation of <= max-degree terms in (0,1): relative 1e-5
        np.testing.assert_allclose(y, want, rtol=1e-5, atol=1e-6, err_msg=name)
        # reproducible: the same bits on every call
        for _ in range(3):
            assert np.array_equal(y.view(np.uint32), gbo.bsrbmv_sum_f32(A, x).view(np.uint32)), name


@pytest.mark.parametrize("bs", BS)
@pytest.mark.parametrize("chunk", COUNT_CHUNKS)
@pytest.mark.parametrize("variant", (0, 3))
def test_min_i32(gbo, po, graph_cases, bs, chunk, variant):
    if variant and bs > 16:
        pytest.skip("variant 3 changes the min kernel at bs 8/16 only")
    rng = np.random.default_rng(300 + bs)
    for name, g in graph_cases:
        A, (rp, ci, tw) = packed(gbo, g, bs, chunk, variant)
        x = rng.integers(-2**31, 2**31 - 1, g.ncols).astype(np.int32)
        assert np.array_equal(gbo.bsrbmv_min_i32(A, x), po.min_i32(bs, g.nrows, rp, ci, tw, x)), name


def test_autotune(gbo, po):
    """The autotuner returns a valid configuration whose results are the
    same as the default's (variants and chunk sizes never change results)."""
    g = gbo.gen_rmat(12, 16, seed=1)
    d_rp = gbo.DeviceBuffer.from_numpy(g.rowptr)
    d_ci = gbo.DeviceBuffer.from_numpy(g.colind)
    t = gbo.autotune(d_rp, d_ci, g.nrows, g.ncols, 0)
    assert t.bs in (8, 16, 32, 64) and t.chunk > 0 and t.ms > 0
    t2 = gbo.autotune(d_rp, d_ci, g.nrows, g.ncols, 0)  # cached
    assert (t2.bs, t2.chunk, t2.variant) == (t.bs, t.chunk, t.variant)
    A = gbo.BsrMatrix.from_csr(g, t.bs, plan=False)
    A.plan(t.chunk, t.variant)
    x = (np.arange(g.ncols) % 3 == 0).astype(np.float32)
    xb = gbo.pack_vec(t.bs, x)
    rp, ci, tw = A.host_arrays()
    assert np.array_equal(gbo.bsrbmv_count(A, xb), po.count(t.bs, g.nrows, rp, ci, tw, xb))


def test_autotune_bfs_step(gbo, po):
    """BFS-step tuning (bs 8 and 16 include the dense variant 3): the picked
    plan's step equals the oracle's."""
    g = gbo.gen_rmat(12, 16, seed=2)
    d_rp = gbo.DeviceBuffer.from_numpy(g.rowptr)
    d_ci = gbo.DeviceBuffer.from_numpy(g.colind)
    t = gbo.autotune(d_rp, d_ci, g.nrows, g.ncols, 1)
    assert t.bs in (8, 16, 32, 64) and t.chunk > 0 and t.ms > 0
    assert t.variant in ((0, 1, 3) if t.bs <= 16 else (0,))
    A = gbo.BsrMatrix.from_csr(g, t.bs, plan=False)
    A.plan(t.chunk, t.variant)
    rng = np.random.default_rng(5)
    f = gbo.pack_vec(t.bs, (rng.random(g.ncols) < 0.3).astype(np.float32))
    v = gbo.pack_vec(t.bs, (rng.random(g.nrows) < 0.3).astype(np.float32))
    rp, ci, tw = A.host_arrays()
    nx, c = gbo.bsrbmv_bfs_step(A, f, v)
    onx, oc = po.bfs_step(t.bs, g.nrows, rp, ci, tw, f, v)
    assert np.array_equal(nx, onx) and c == oc


class _TileView:
    """A gbo_bsr that shares a matrix's arrays and plan but reads its tiles
    from another address; never freed (it owns nothing)."""

    def __init__(self, gbo, A, tiles_ptr):
        import ctypes
        self.s = gbo.BsrStruct()
        ctypes.memmove(ctypes.addressof(self.s), ctypes.addressof(A.s), ctypes.sizeof(self.s))
        self.s.tiles = tiles_ptr
        self.ref = ctypes.byref(self.s)
        self.bs, self.nrows, self.nblockrows, self.nblocks = A.bs, A.nrows, A.nblockrows, A.nblocks


@pytest.mark.parametrize("bs", BS)
def test_unaligned_tiles_give_the_same_results(gbo, bs):
    """The kernels that read tiles as 8- or 16-byte vectors (bs 8/16 variant 3,
    the bs 32 count / min / min-plus, the dense BFS step) need 16-byte-aligned
    tiles; tiles 8 bytes into a buffer take the row-group kernels and give the
    same results (the FP32 sum within rounding: its order of additions
    differs)."""
    g = gbo.gen_rmat(11, 16, seed=4)
    A = gbo.BsrMatrix.from_csr(g, bs)
    nb = A.tile_bytes()
    buf = gbo.DeviceBuffer(nb + 16)
    gbo._check(gbo.lib().gbo_memcpy_d2d(buf.ptr + 8, A.s.tiles, nb))
    V = _TileView(gbo, A, buf.ptr + 8)
    rng = np.random.default_rng(11)
    xb = gbo.pack_vec(bs, (rng.random(g.ncols) < 0.3).astype(np.float32))
    assert np.array_equal(gbo.bsrbmv_count(V, xb), gbo.bsrbmv_count(A, xb))
    f = gbo.pack_vec(bs, (rng.random(g.ncols) < 0.5).astype(np.float32))
    v = gbo.pack_vec(bs, (rng.random(g.nrows) < 0.1).astype(np.float32))
    nx, c = gbo.bsrbmv_bfs_step(V, f, v)
    nx0, c0 = gbo.bsrbmv_bfs_step(A, f, v)
    assert np.array_equal(nx, nx0) and c == c0
    dist = rng.integers(0, 1000, g.nrows).astype(np.uint32)
    dist[rng.random(g.nrows) < 0.3] = gbo.INF_U32
    assert np.array_equal(gbo.bsrbmv_minplus(V, dist), gbo.bsrbmv_minplus(A, dist))
    tv = gbo.TileValues(A, g, gbo.edge_weights(g, 1, 64, seed=9))
    assert np.array_equal(gbo.bsrbmv_minplus(V, dist, tv), gbo.bsrbmv_minplus(A, dist, tv))
    xi = rng.integers(-1000, 1000, g.ncols).astype(np.int32)
    assert np.array_equal(gbo.bsrbmv_min_i32(V, xi), gbo.bsrbmv_min_i32(A, xi))
    xf = rng.random(g.ncols).astype(np.float32)
    np.testing.assert_allclose(gbo.bsrbmv_sum_f32(V, xf), gbo.bsrbmv_sum_f32(A, xf), rtol=1e-6)
    tv.free()
    buf.free()
    A.free()
