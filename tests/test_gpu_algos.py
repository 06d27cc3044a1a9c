"""GPU driver parity: BFS levels, SSSP distances, PageRank, CC labels and
triangle counts against textbook CPU algorithms and closed forms."""
import math

import numpy as np
import pytest

from graphs import cases, complete_graph

pytestmark = pytest.mark.gpu
BS = (8, 16, 32, 64)


@pytest.fixture(scope="module")
def graph_cases(gbo):
    return cases(gbo)


def sources(g):
    deg = np.diff(g.rowptr)
    s = [0, g.nrows - 1]
    nz = np.nonzero(deg)[0]
    if len(nz):
        s.append(int(nz[len(nz) // 2]))
    return sorted(set(s))


@pytest.mark.parametrize("bs", BS)
def test_bfs_levels(gbo, po, graph_cases, bs):
    for name, g in graph_cases:
        A = gbo.BsrMatrix.from_csr(g, bs)
        rp, ci, tw = A.host_arrays()
        for s in sources(g):
            lv, sizes = gbo.bfs(A, s)
            want = po.bfs(g.nrows, g.rowptr, g.colind, s)
            assert np.array_equal(lv, want), (name, s)
            # per-level discovered counts == histogram of levels (the
            # reference's per-level "succ" line)
            hist = np.bincount(want[want >= 0])
            assert np.array_equal(sizes, hist), (name, s)
            olv, osz = po.bfs_bsr(bs, g.nrows, rp, ci, tw, s)
            assert np.array_equal(olv, want) and np.array_equal(osz, hist), (name, s)


@pytest.mark.parametrize("bs", (8, 16))
@pytest.mark.parametrize("variant", (0, 1, 3))
def test_bfs_levels_variants(gbo, po, graph_cases, bs, variant):
    """Every bs 8/16 BFS step kernel (row-group, lane-per-tile, and at bs 8
    the dense step) gives the oracle's levels; at bs 16 the levels with a
    sparse frontier take the row-group kernel's path that reads only the
    tiles under nonzero frontier words.  Chunk 3 splits most rows."""
    for name, g in graph_cases:
        for chunk in (0, 3):
            A = gbo.BsrMatrix.from_csr(g, bs, 0, plan=False)
            A.plan(chunk, variant)
            for s in sources(g):
                lv, _ = gbo.bfs(A, s)
                assert np.array_equal(lv, po.bfs(g.nrows, g.rowptr, g.colind, s)), (name, chunk, s)
            A.free()


@pytest.mark.parametrize("bs", BS)
def test_bfs_direction_optimizing(gbo, po, graph_cases, bs):
    """gbo_bfs_do (push for frontiers of at most n/64 vertices, pull over AT
    otherwise) == the oracle's queue BFS: undirected cases with A = AT, and
    a directed graph with A and AT packed separately."""
    for name, g in graph_cases:
        A = gbo.BsrMatrix.from_csr(g, bs)
        for s in sources(g):
            lv, sizes = gbo.bfs(A, s, A=A)
            want = po.bfs(g.nrows, g.rowptr, g.colind, s)
            assert np.array_equal(lv, want), (name, s)
            assert np.array_equal(sizes, np.bincount(want[want >= 0])), (name, s)
    rng = np.random.default_rng(bs)
    n = 3000
    src = rng.integers(0, n, 12000)
    dst = rng.integers(0, n, 12000)
    keep = src != dst
    order = np.lexsort((dst[keep], src[keep]))
    e = np.unique(np.stack([src[keep][order], dst[keep][order]], 1), axis=0)
    rowptr = np.zeros(n + 1, np.int32)
    rowptr[1:] = np.cumsum(np.bincount(e[:, 0], minlength=n))
    g = gbo.Csr(n, n, rowptr, e[:, 1].astype(np.int32))
    Ad = gbo.BsrMatrix.from_csr(g, bs)
    ATd = gbo.BsrMatrix.from_csr(g.transpose(), bs)
    for s in (0, 17, 2999):
        want = po.bfs(n, g.rowptr, g.colind, s)
        lv, _ = gbo.bfs(ATd, s, A=Ad)
        assert np.array_equal(lv, want), ("directed", s)
        lv2, _ = gbo.bfs(ATd, s)
        assert np.array_equal(lv2, want), ("directed pull", s)


def test_bfs_closed_forms(gbo):
    from graphs import path_graph, star_graph
    rp, ci = path_graph(1000)
    lv, _ = gbo.bfs(gbo.BsrMatrix.from_csr(gbo.Csr(1000, 1000, rp, ci), 64), 0)
    assert np.array_equal(lv, np.arange(1000))
    rp, ci = star_graph(513)
    lv, sizes = gbo.bfs(gbo.BsrMatrix.from_csr(gbo.Csr(513, 513, rp, ci), 8), 0)
    assert lv[0] == 0 and np.all(lv[1:] == 1) and list(sizes) == [1, 512]


@pytest.mark.parametrize("bs", BS)
def test_sssp(gbo, po, graph_cases, bs):
    for name, g in graph_cases:
        A = gbo.BsrMatrix.from_csr(g, bs)
        s = sources(g)[-1]
        d, _ = gbo.sssp(A, s)
        lv = po.bfs(g.nrows, g.rowptr, g.colind, s)
        want_unit = np.where(lv < 0, gbo.INF_U32, lv).astype(np.uint32)
        assert np.array_equal(d, want_unit), name
        w = gbo.edge_weights(g, 1, 64, seed=5)
        tv = gbo.TileValues(A, g, w)
        dw, _ = gbo.sssp(A, s, tv)
        assert np.array_equal(dw, po.dijkstra(g.nrows, g.rowptr, g.colind, w, s)), name


@pytest.mark.parametrize("bs", BS)
def test_pagerank(gbo, po, graph_cases, bs):
    for name, g in graph_cases:
        if g.nnz == 0:
            continue
        A = gbo.BsrMatrix.from_csr(g, bs)
        p, it, delta = gbo.pagerank(A, g.outdeg(), 0.85, 1e-8, 300)
        want, _ = po.pagerank(g.nrows, g.rowptr, g.colind, g.outdeg(), 0.85, 1e-14, 100000)
        # FP32 iterate vs FP64 fixpoint of the same formula: L1 <= 1e-6
        assert np.abs(p.astype(np.float64) - want).sum() <= 1e-6, (name, it, delta)


def test_pagerank_closed_forms(gbo):
    from graphs import sym_csr
    n = 640
    u = np.arange(n)
    rp, ci = sym_csr(n, u, (u + 1) % n)  # cycle: uniform
    p, _, _ = gbo.pagerank(gbo.BsrMatrix.from_csr(gbo.Csr(n, n, rp, ci), 32), np.diff(rp), 0.85, 1e-10, 1000)
    np.testing.assert_allclose(p, 1.0 / n, rtol=1e-5)
    rp, ci = complete_graph(40)
    p, _, _ = gbo.pagerank(gbo.BsrMatrix.from_csr(gbo.Csr(40, 40, rp, ci), 16), np.diff(rp), 0.85, 1e-10, 1000)
    np.testing.assert_allclose(p, 1.0 / 40, rtol=1e-5)


@pytest.mark.parametrize("bs", (8, 32))
def test_driver_results_in_device_memory(gbo, po, bs):
    """out= leaves a driver's result in the caller's device buffer (the
    oracle's answer, and the same bits as the host-array return); PageRank
    takes its out-degrees from device memory too.  A buffer too small for
    n results is refused before any launch."""
    g = gbo.gen_rmat(12, 16, seed=3)
    n = g.nrows
    A = gbo.BsrMatrix.from_csr(g, bs)
    d = gbo.DeviceBuffer(4 * n)
    lv = po.bfs(n, g.rowptr, g.colind, 0)
    r, _ = gbo.bfs(A, 0, out=d)
    assert r is d
    assert np.array_equal(d.to_numpy(np.int32, n), lv)
    gbo.bfs(A, 0, A=A, out=d)
    assert np.array_equal(d.to_numpy(np.int32, n), lv)
    gbo.sssp(A, 0, out=d)
    assert np.array_equal(d.to_numpy(np.uint32, n), np.where(lv < 0, gbo.INF_U32, lv).astype(np.uint32))
    gbo.cc(A, out=d)
    assert np.array_equal(d.to_numpy(np.int32, n), po.cc(n, g.rowptr, g.colind))
    d_od = gbo.DeviceBuffer.from_numpy(g.outdeg().astype(np.int32))
    _, it, _ = gbo.pagerank(A, d_od, 0.85, 1e-8, 300, out=d)
    p_host, it_host, _ = gbo.pagerank(A, g.outdeg(), 0.85, 1e-8, 300)
    assert it == it_host
    assert np.array_equal(d.to_numpy(np.float32, n).view(np.uint32), p_host.view(np.uint32))
    with pytest.raises(ValueError):
        gbo.cc(A, out=gbo.DeviceBuffer(4 * n - 4))
    # multi-source levels are [nsrc][n]: a buffer for one source fewer is refused
    src = np.array([0, 5, 9], np.int32)
    with pytest.raises(ValueError):
        gbo.bfs_multisource(A, src, out=gbo.DeviceBuffer(4 * n * 2))
    dm = gbo.DeviceBuffer(4 * n * 3)
    assert gbo.bfs_multisource(A, src, out=dm) >= 1
    lvm = dm.to_numpy(np.int32, 3 * n).reshape(3, n)
    for k, s in enumerate(src):
        assert np.array_equal(lvm[k], po.bfs(n, g.rowptr, g.colind, int(s)))


@pytest.mark.parametrize("bs", BS)
def test_cc(gbo, po, graph_cases, bs):
    for name, g in graph_cases:
        A = gbo.BsrMatrix.from_csr(g, bs)
        lab, _ = gbo.cc(A)
        want = po.cc(g.nrows, g.rowptr, g.colind)
        assert np.array_equal(lab, want), name


@pytest.mark.parametrize("bs", BS)
def test_tc_closed_forms(gbo, bs):
    for m in (2, 3, 17, 70):
        assert gbo.tc(gbo.gen_king(m), bs) == 4 * (m - 1) ** 2
    for n in (3, 10, 65):
        rp, ci = complete_graph(n)
        assert gbo.tc(gbo.Csr(n, n, rp, ci), bs) == math.comb(n, 3)


@pytest.mark.parametrize("bs", BS)
def test_tc_vs_oracle(gbo, po, graph_cases, bs):
    for name, g in graph_cases:
        want = po.tc(g.nrows, g.rowptr, g.colind)
        assert gbo.tc(g, bs) == want, name


def _clique_on_rmat(gbo, scale, ef, k, seed):
    """R-MAT (sparse tiles) with a k-clique on vertices 0..k-1 (dense tiles)."""
    from graphs import sym_csr
    g = gbo.gen_rmat(scale, ef, seed=seed)
    u = np.repeat(np.arange(g.nrows), np.diff(g.rowptr))
    cu, cv = np.triu_indices(k, 1)
    rp, ci = sym_csr(g.nrows, np.concatenate([u, cu]), np.concatenate([g.colind, cv]))
    return gbo.Csr(g.nrows, g.nrows, rp, ci)


@pytest.mark.parametrize("bs", (32, 64))
def test_tc_lane_and_group_matches(gbo, po, bs):
    """bs 32/64 count a match whose A tile has at most 16 set bits in one lane
    (from the tile's bit list, built when the tiles average at most 8 bits)
    and a denser one with a group of bs lanes.  An R-MAT graph with a 100-clique
    has both kinds in one call; the king grid's tiles are dense (no lists)."""
    g = _clique_on_rmat(gbo, 15, 4, 100, 3)
    L = g.tril(strict=True)
    Lr = gbo.BsrMatrix.from_csr(L, bs, 0)
    _, _, tw = Lr.host_arrays()
    per_tile = np.unpackbits(tw.view(np.uint8)).reshape(Lr.nblocks, -1).sum(axis=1)
    assert per_tile.mean() <= 8 and (per_tile > 16).any() and (per_tile <= 16).any()
    want = po.tc(g.nrows, g.rowptr, g.colind)
    assert gbo.tc(g, bs) == want
    king = gbo.gen_king(300)
    assert gbo.tc(king, bs) == 4 * 299 ** 2


@pytest.mark.parametrize("bs", (8, 16, 32))
def test_tc_bitmap_and(gbo, po, bs):
    """Mask tiles whose row and column both have block-column bitmaps (rows
    of more than 64 tiles) intersect them by an AND of the two maps.  A dense
    random graph makes every row heavy; R-MAT makes only the hubs heavy."""
    dense = gbo.gen_er(2400 if bs < 32 else 4800, 0.12, 9)
    L = dense.tril(strict=True)
    r = np.repeat(np.arange(L.nrows), np.diff(L.rowptr))
    nb = (L.nrows + bs - 1) // bs
    tiles_per_row = np.bincount(np.unique((r // bs) * nb + L.colind // bs) // nb, minlength=nb)
    assert (tiles_per_row > 64).sum() >= nb // 3  # heavy rows: bitmaps on both sides
    for g in (dense, gbo.gen_rmat(14, 16, seed=4)):
        assert gbo.tc(g, bs) == po.tc(g.nrows, g.rowptr, g.colind)


def test_tc_bs16_unaligned_tiles(gbo, po):
    """bs 16 reads A's tiles as 16-byte vectors when they are 16-byte aligned;
    tiles at an address that is not (a view 2 bytes into a buffer) take the
    row-by-row form and count the same."""
    import ctypes
    g = gbo.gen_rmat(11, 16, seed=5)
    L = g.tril(strict=True)
    Lr = gbo.BsrMatrix.from_csr(L, 16, 0)
    Lc = gbo.BsrMatrix.from_csr(L.transpose(), 16, 1)
    want = po.tc(g.nrows, g.rowptr, g.colind)
    assert gbo.bsrbmm_tc(Lr, Lc) == want
    nb = Lr.tile_bytes()
    buf = gbo.DeviceBuffer(nb + 16)
    gbo._check(gbo.lib().gbo_memcpy_d2d(buf.ptr + 2, Lr.s.tiles, nb))
    s2 = gbo.BsrStruct()
    ctypes.memmove(ctypes.addressof(s2), ctypes.addressof(Lr.s), ctypes.sizeof(s2))
    s2.tiles = buf.ptr + 2
    s2.plan = None
    d = gbo.DeviceBuffer(8)
    gbo._check(gbo.lib().gbo_bsrbmm_tc(ctypes.byref(s2), Lc.ref, d.ptr, None))
    assert int(d.to_numpy(np.uint64, 1)[0]) == want
    d.free()
    buf.free()


def test_tc_full_size_workloads(gbo, po):
    """The full-size triangle-counting workloads: the 2048x2048 king grid
    (closed form 4*2047^2 = 16,760,836) and R-MAT s20 (seed 2) against the
    oracle's merge count, at the smallest and largest block sizes."""
    king = gbo.gen_king(2048)
    for bs in (8, 64):
        assert gbo.tc(king, bs) == 4 * 2047 ** 2
    g = gbo.gen_rmat(20, 16, seed=2)
    want = po.tc(g.nrows, g.rowptr, g.colind)
    for bs in (8, 64):
        assert gbo.tc(g, bs) == want


@pytest.mark.parametrize("bs", BS)
def test_bsrbmm_operands(gbo, po, bs):
    """gbo_bsrbmm_tc / gbo_bsrbmm_sum on explicit (L row, L^T col) operands,
    as the reference's tc/bsrbmm drivers build them, against the oracle's
    restatement of the v5 traversal and the closed form sum_c colcount_c(L)^2."""
    for g in (gbo.gen_rmat(10, 16, seed=2), gbo.gen_er(700, 0.03, 4)):
        L = g.tril(strict=True)
        LT = L.transpose()
        Lr = gbo.BsrMatrix.from_csr(L, bs, 0)
        Lc = gbo.BsrMatrix.from_csr(LT, bs, 1)
        arp, aci, atw = Lr.host_arrays()
        brp, bci, btw = Lc.host_arrays()
        want = po.bsrbmm_masked(bs, arp, aci, atw, brp, bci, btw)
        assert want == po.tc(g.nrows, g.rowptr, g.colind)
        assert gbo.bsrbmm_tc(Lr, Lc) == want
        colcount = np.bincount(L.colind, minlength=g.nrows).astype(np.int64)
        s = int((colcount ** 2).sum())
        assert po.bsrbmm_sum(bs, arp, aci, atw, brp, bci, btw) == s
        assert gbo.bsrbmm_sum(Lr, Lc) == s


@pytest.mark.parametrize("bs", BS)
def test_bfs_multisource(gbo, po, graph_cases, bs):
    """Batched BFS == per-source BFS (oracle queue BFS), bit-exact, for 1, 64,
    100, 150 and 256 sources (one to four 64-bit frontier words per vertex)
    including duplicates and isolated vertices."""
    rng = np.random.default_rng(bs)
    for name, g in graph_cases:
        A = gbo.BsrMatrix.from_csr(g, bs)
        for k in (1, 64, 100, 150, 256):
            src = rng.integers(0, g.nrows, k).astype(np.int32)
            src[0] = src[-1]  # a duplicate
            lv, nl = gbo.bfs_multisource(A, src)
            for s in range(0, k, max(1, k // 9)):
                want = po.bfs(g.nrows, g.rowptr, g.colind, int(src[s]))
                assert np.array_equal(lv[s], want), (name, k, s)
            assert nl == int(lv.max()) + 1, (name, k)


def _fan_graph(gbo, m=300):
    """Vertex 0 joined to vertices 1..m, vertex m+1 joined to the same m:
    pulling m+1 reads m set bits of one tile row inside one 64-tile block
    (more than the 128 a per-row list holds in one pass), plus a tail."""
    from graphs import sym_csr
    a = np.arange(1, m + 1)
    u = np.concatenate([np.zeros(m, np.int64), np.full(m, m + 1), [m + 1, m + 2]])
    v = np.concatenate([a, a, [m + 2, m + 3]])
    rp, ci = sym_csr(m + 4, u, v)
    return gbo.Csr(m + 4, m + 4, rp, ci)


def test_bfs_multisource_long_rows(gbo, po):
    """bs 8 batched BFS with a tile row of 300 frontier bits in one block
    (three passes over its list) and 1, 150 and 256 sources."""
    g = _fan_graph(gbo)
    A = gbo.BsrMatrix.from_csr(g, 8)
    rng = np.random.default_rng(4)
    for k in (1, 150, 256):
        src = np.concatenate([[0], rng.integers(0, g.nrows, k - 1)]).astype(np.int32)
        lv, _ = gbo.bfs_multisource(A, src)
        for s in range(k):
            assert np.array_equal(lv[s], po.bfs(g.nrows, g.rowptr, g.colind, int(src[s]))), (k, s)


def _ppr_case(gbo, seed, n_seeds=20, K=8):
    g = gbo.gen_rmat(9, 8, seed=seed)
    rng = np.random.default_rng(seed)
    S = np.zeros((g.nrows, K), np.float64)
    for k in range(K):
        idx = rng.choice(g.nrows, n_seeds, replace=False)
        S[idx, k] = rng.dirichlet(np.full(n_seeds, 0.1))
    W = rng.standard_normal((g.nrows, K))
    return g, S, W


def test_ppr_bs8_kernels_give_the_same_bits(gbo):
    """At bs 8 with K = 64 and 16-byte aligned rows the SpMM lists the set bits
    per tile row (k_ppr_spmm8r); otherwise it uses one flat list
    (k_ppr_spmm8).  Both add each row's terms in the same order, so a
    fixed number of forward and adjoint iterations gives the same bits when
    S (forward) or W (adjoint) is moved 4 bytes off alignment."""
    g, S, W = _ppr_case(gbo, 7, n_seeds=30, K=64)
    _ppr_kernels_agree(gbo, g, S, W)
    # a tile row with 300 set bits in one 64-tile block (several list passes)
    f = _fan_graph(gbo)
    rng = np.random.default_rng(8)
    Sf = rng.dirichlet(np.full(f.nrows, 0.5), 64).T.copy()
    Wf = rng.standard_normal((f.nrows, 64))
    _ppr_kernels_agree(gbo, f, Sf, Wf)


def _ppr_kernels_agree(gbo, g, S, W):
    n, K = g.nrows, 64
    A = gbo.BsrMatrix.from_csr(g, 8)
    d_od = gbo.DeviceBuffer.from_numpy(g.outdeg().astype(np.int32))
    S32, W32 = S.astype(np.float32), W.astype(np.float32)
    nb = S32.nbytes

    def placed(arr, off):
        buf = gbo.DeviceBuffer(nb + 16)
        gbo._check(gbo.lib().gbo_memcpy_h2d(buf.ptr + off, arr.ctypes.data, nb))
        return buf, buf.ptr + off

    out = {}
    for off in (0, 4):
        bs_, ps = placed(S32, off)
        bw_, pw = placed(W32, off)
        d_p, d_g = gbo.DeviceBuffer(nb), gbo.DeviceBuffer(nb)
        it, _ = gbo.ppr_forward_ptr(A, d_od.ptr, K, 0.85, ps, d_p.ptr, 0.0, 9, strict=False)
        ga, ita = gbo.ppr_adjoint_ptr(A, A, d_od.ptr, K, 0.85, ps, d_p.ptr, pw, d_g.ptr, 0.0, 9, strict=False)
        assert it == 9 and ita == 9
        out[off] = (d_p.to_numpy(np.float32, n * K), d_g.to_numpy(np.float32, n * K), ga)
        for b in (bs_, bw_, d_p, d_g):
            b.free()
    assert np.array_equal(out[0][0].view(np.uint32), out[4][0].view(np.uint32))
    assert np.array_equal(out[0][1].view(np.uint32), out[4][1].view(np.uint32))
    assert out[0][2] == out[4][2]
    assert np.abs(out[0][0]).sum() > 0 and np.abs(out[0][1]).sum() > 0


@pytest.mark.parametrize("bs", (8, 16, 32))
def test_ppr_forward_adjoint(gbo, bs):
    """GPU PPR forward (FP32) within 1e-6 L1 per column of the FP64 fixed
    point; adjoint gradients within 1e-5 relative of the FP64 gradient."""
    import ppr_oracle as pp
    g, S, W = _ppr_case(gbo, 3)
    A = gbo.BsrMatrix.from_csr(g, bs)
    deg = g.outdeg()
    P, it, delta = gbo.ppr_forward(A, deg, S, 0.85, 1e-8, 1000)
    T = pp.transition(g.nrows, g.rowptr, g.colind)
    want = pp.forward(T, S, 0.85)
    assert np.abs(P - want).sum(0).max() <= 1e-6, (it, delta)
    gS, ga, _ = gbo.ppr_adjoint(A, A, deg, S, P, W, 0.85, 1e-9, 1000)
    wS, wa = pp.adjoint(T, S, W, 0.85)
    assert np.abs(gS - wS).max() <= 1e-5 * np.abs(wS).max()
    assert abs(ga - wa) <= 1e-5 * abs(wa)
    # bitwise reproducible
    P2, _, _ = gbo.ppr_forward(A, deg, S, 0.85, 1e-8, 1000)
    assert np.array_equal(P.view(np.uint32), P2.view(np.uint32))


def test_ppr_autograd(gbo):
    """torch.autograd.Function: backward equals the FP64 adjoint and central
    finite differences (h = 1e-3, FP64 forward) of the loss."""
    torch = pytest.importorskip("torch")
    import gbo_torch
    import ppr_oracle as pp
    g, S, W = _ppr_case(gbo, 5, K=4)
    A = gbo.BsrMatrix.from_csr(g, 16)
    St = torch.tensor(S, dtype=torch.float32, requires_grad=True)
    at = torch.tensor(0.85, dtype=torch.float32, requires_grad=True)
    # the wrapper solves strictly: the tolerance must be reachable in FP32
    # (this case's iterates stop changing near 4e-9 relative L1)
    P = gbo_torch.personalized_pagerank(St, at, A, A, g.outdeg(), 2e-8, 2000)
    loss = (P * torch.tensor(W, dtype=torch.float32)).sum()
    loss.backward()
    T = pp.transition(g.nrows, g.rowptr, g.colind)
    wS, wa = pp.adjoint(T, S, W, 0.85)
    h = 1e-3
    L = lambda a_: float((W * pp.forward(T, S, a_)).sum())
    fd_a = (L(0.85 + h) - L(0.85 - h)) / (2 * h)
    assert abs(at.grad.item() - wa) <= 1e-5 * abs(wa)
    assert abs(at.grad.item() - fd_a) <= 1e-4 * abs(fd_a)
    assert np.abs(St.grad.numpy() - wS).max() <= 1e-5 * np.abs(wS).max()


@pytest.mark.parametrize("bs", BS)
def test_graph_batch(gbo, po, bs):
    """One persistent launch (direct and hipGraph-captured, replayed) gives
    the same BFS levels and CC labels as per-graph plain runs and the oracle."""
    graphs = gbo.gen_er_batch(count=24, n_lo=50, n_hi=3000, d_lo=0.5, d_hi=6.0, seed=4)
    graphs.append(gbo.gen_rmat(10, 4, seed=6))
    # ragged members: one isolated vertex, one edge, a 700-vertex path (many
    # more BFS levels than the rest of the batch), n = 65 (a partial block row)
    path = np.arange(699)
    pr = np.concatenate([path, path + 1])
    pc = np.concatenate([path + 1, path])
    o = np.lexsort((pc, pr))
    pr, pc = pr[o], pc[o]
    graphs.insert(0, gbo.Csr(1, 1, np.array([0, 0]), np.zeros(0, np.int32)))
    graphs.insert(3, gbo.Csr(2, 2, np.array([0, 1, 2]), np.array([1, 0])))
    graphs.insert(7, gbo.Csr(700, 700, np.searchsorted(pr, np.arange(701)), pc))
    graphs.append(gbo.gen_er_m(65, 100, seed=9))
    rng = np.random.default_rng(bs)
    src = np.array([rng.integers(0, g.nrows) for g in graphs], np.int32)
    B = gbo.GraphBatch(graphs, bs)
    for use_graph in (False, True, True):
        B.run(src, use_graph)
        levels, labels = B.results()
        for i, g in enumerate(graphs):
            assert np.array_equal(levels[i], po.bfs(g.nrows, g.rowptr, g.colind, int(src[i]))), (i, use_graph)
            assert np.array_equal(labels[i], po.cc(g.nrows, g.rowptr, g.colind)), (i, use_graph)
    # plain per-graph runs agree too
    for i in (0, len(graphs) - 1):
        A = gbo.BsrMatrix.from_csr(graphs[i], bs)
        lv, _ = gbo.bfs(A, int(src[i]))
        lb, _ = gbo.cc(A)
        assert np.array_equal(lv, levels[i]) and np.array_equal(lb, labels[i])


def test_graph_batch_barrier_timeout(gbo, po, monkeypatch):
    """A zero barrier wait limit (set from the host) makes every grid barrier
    wait expire at once: all workgroups leave, the call reports GBO_EDEVICE,
    and the next call with the normal limit is exact again."""
    graphs = gbo.gen_er_batch(count=6, n_lo=50, n_hi=800, d_lo=1.0, d_hi=4.0, seed=11)
    src = np.zeros(len(graphs), np.int32)
    B = gbo.GraphBatch(graphs, 8)
    monkeypatch.setenv("GBO_BATCH_SPIN_MS", "0")
    with pytest.raises(gbo.GboError, match="GBO_EDEVICE"):
        B.run(src, use_graph=False)
    monkeypatch.delenv("GBO_BATCH_SPIN_MS")
    B.run(src, use_graph=False)
    levels, labels = B.results()
    for i, g in enumerate(graphs):
        assert np.array_equal(levels[i], po.bfs(g.nrows, g.rowptr, g.colind, 0)), i
        assert np.array_equal(labels[i], po.cc(g.nrows, g.rowptr, g.colind)), i
    # the captured launch: a graph captured with the normal limit is captured
    # again when the limit changes (the error names the limit that applied),
    # and once more when it changes back
    B.run(src, use_graph=True)
    monkeypatch.setenv("GBO_BATCH_SPIN_MS", "0")
    with pytest.raises(gbo.GboError, match="exceeded 0 ms"):
        B.run(src, use_graph=True)
    monkeypatch.delenv("GBO_BATCH_SPIN_MS")
    B.run(src, use_graph=True)
    levels, labels = B.results()
    for i, g in enumerate(graphs):
        assert np.array_equal(levels[i], po.bfs(g.nrows, g.rowptr, g.colind, 0)), i
        assert np.array_equal(labels[i], po.cc(g.nrows, g.rowptr, g.colind)), i
    B.free()
