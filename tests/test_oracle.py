"""Pins the oracle (oracle/oracle.c) against closed forms, brute force,
independent libraries (scipy, networkx counts in tests/golden/counts.json)
and hand-computed tiles -- so that a dropped term, wrong sign or transposed
operand in the restatement fails here, on the CPU."""
import json
import math
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.csgraph as cg

from conftest import GOLDEN
from graphs import complete_graph, path_graph, random_sym, star_graph, sym_csr

BS = (8, 16, 32, 64)


@pytest.fixture(scope="module")
def counts():
    with open(os.path.join(GOLDEN, "counts.json")) as f:
        return json.load(f)


def dense(n, rp, ci):
    D = np.zeros((n, n), np.int64)
    for r in range(n):
        D[r, ci[rp[r]:rp[r + 1]]] = 1
    return D


def test_hand_computed_tiles(po):
    """64x64 symmetric pattern: diagonals 0,1,2,32,33,34 and edge (0,44).
    bs=32: tiles (0,0) rows 0-2 diagonal, (0,1) row 0 column 12, (1,0) row 12
    column 0, (1,1) rows 0-2 diagonal -> T = 4; bs=64: T = 1."""
    rp, ci = sym_csr(64, [0, 1, 2, 32, 33, 34, 0], [0, 1, 2, 32, 33, 34, 44])
    # sym_csr drops loops; put the diagonal back
    rows = [set(ci[rp[r]:rp[r + 1]]) for r in range(64)]
    for d in (0, 1, 2, 32, 33, 34):
        rows[d].add(d)
    rp = np.cumsum([0] + [len(s) for s in rows]).astype(np.int32)
    ci = np.array([c for s in rows for c in sorted(s)], np.int32)
    brp, bci, tw = po.pack(32, 64, rp, ci)
    assert list(brp) == [0, 2, 4] and list(bci) == [0, 1, 0, 1]
    assert [int(w) for w in tw[0:3]] == [1 << 31, 1 << 30, 1 << 29]
    assert int(tw[32 + 0]) == 1 << (31 - 12)           # tile (0,1) word 0
    assert int(tw[64 + 12]) == 1 << 31                 # tile (1,0) word 12
    assert [int(w) for w in tw[96:99]] == [1 << 31, 1 << 30, 1 << 29]
    brp, bci, tw = po.pack(64, 64, rp, ci)
    assert len(bci) == 1 and int(tw[0]) == (1 << 63) | (1 << (63 - 44))


@pytest.mark.parametrize("bs", BS)
@pytest.mark.parametrize("direction", (0, 1))
def test_pack_roundtrip(po, bs, direction):
    for n, m, seed in ((100, 400, 1), (257, 2000, 2), (64, 0, 3)):
        rp, ci = random_sym(n, m, seed) if m else (np.zeros(n + 1, np.int32), np.zeros(0, np.int32))
        brp, bci, tw = po.pack(bs, n, rp, ci, direction)
        assert np.array_equal(po.unpack(bs, n, n, brp, bci, tw, direction), dense(n, rp, ci))
        assert int(np.unpackbits(tw.view(np.uint8)).sum()) == len(ci)
        for I in range(len(brp) - 1):
            assert np.all(np.diff(bci[brp[I]:brp[I + 1]]) > 0)


@pytest.mark.parametrize("bs", BS)
def test_pack_vec_closed_form(po, bs):
    x = np.zeros(3 * bs + 5, np.float32)
    x[[0, bs - 1, bs, 3 * bs + 4]] = 1.0
    w = po.pack_vec(bs, x)
    assert int(w[0]) == (1 << (bs - 1)) | 1
    assert int(w[1]) == 1 << (bs - 1)
    assert int(w[3]) == 1 << (bs - 1 - 4)


def test_c1_golden(po):
    """Config ER n=1024 p=0.02 seed 0: oracle count == dense matvec and BSR
    BFS == scipy BFS (tests/golden/c1_expected.npz)."""
    z = np.load(os.path.join(GOLDEN, "c1_expected.npz"))
    rp, ci = z["rowptr"], z["colind"]
    rp2, ci2 = po.gen_er(1024, 0.02, 0)
    assert np.array_equal(rp, rp2) and np.array_equal(ci, ci2)
    for bs in BS:
        brp, bci, tw = po.pack(bs, 1024, rp, ci)
        xb = po.pack_vec(bs, z["x"])
        assert np.array_equal(po.count(bs, 1024, brp, bci, tw, xb), z["y"])
        lv, sizes = po.bfs_bsr(bs, 1024, brp, bci, tw, 0)
        assert np.array_equal(lv, z["level0"])
        assert np.array_equal(sizes, np.bincount(z["level0"][z["level0"] >= 0]))
    assert np.array_equal(po.bfs(1024, rp, ci, 0), z["level0"])


def test_generators_match_engine(gbo, po, counts):
    g = gbo.gen_er(1024, 0.02, 0)
    rp, ci = po.gen_er(1024, 0.02, 0)
    assert np.array_equal(g.rowptr, rp) and np.array_equal(g.colind, ci)
    import hashlib
    assert hashlib.sha1(g.colind.tobytes()).hexdigest() == counts["er1024_p0.02_seed0_colind_sha1"]
    for s in (10, 14):
        r = gbo.gen_rmat(s, 16, seed=2)
        rp, ci = po.gen_rmat(s, 16, seed=2)
        assert np.array_equal(r.rowptr, rp) and np.array_equal(r.colind, ci)
        assert r.nnz == counts[f"rmat{s}_seed2_nnz"]
        assert hashlib.sha1(r.colind.tobytes()).hexdigest() == counts[f"rmat{s}_seed2_colind_sha1"]


def test_king_structure(gbo):
    for m in (1, 2, 5, 33):
        k = gbo.gen_king(m)
        assert k.nnz == 2 * (2 * m * (m - 1) + 2 * (m - 1) ** 2)
        D = dense(m * m, k.rowptr, k.colind)
        assert np.array_equal(D, D.T)


def test_tc_oracle(po, gbo, counts):
    for m in (2, 3, 17, 64):
        k = gbo.gen_king(m)
        assert po.tc(k.nrows, k.rowptr, k.colind) == counts[f"king{m}_triangles"] == 4 * (m - 1) ** 2
    for n in (3, 10, 65):
        rp, ci = complete_graph(n)
        assert po.tc(n, rp, ci) == math.comb(n, 3)
    for s in (10, 14):
        r = gbo.gen_rmat(s, 16, seed=2)
        assert po.tc(r.nrows, r.rowptr, r.colind) == counts[f"rmat{s}_seed2_triangles_networkx"]
    rp, ci = random_sym(150, 1500, 9)
    D = dense(150, rp, ci)
    assert po.tc(150, rp, ci) == int(np.trace(D @ D @ D)) // 6


@pytest.mark.parametrize("bs", BS)
def test_bsrbmm_oracle(po, gbo, bs):
    r = gbo.gen_rmat(10, 16, seed=2)
    L = r.tril(True)
    LT = L.transpose()
    a = po.pack(bs, L.nrows, L.rowptr, L.colind, 0)
    b = po.pack(bs, LT.nrows, LT.rowptr, LT.colind, 1)
    assert po.bsrbmm_masked(bs, *a, *b) == po.tc(r.nrows, r.rowptr, r.colind)
    D = dense(r.nrows, L.rowptr, L.colind)
    assert po.bsrbmm_sum(bs, *a, *b) == int((D @ D.T).sum())


@pytest.mark.parametrize("bs", (8, 32))
def test_bfs_oracles(po, bs):
    for rp, ci, n in ((*path_graph(300), 300), (*star_graph(100), 100), (*random_sym(500, 800, 4), 500)):
        A = sp.csr_matrix((np.ones(len(ci)), ci, rp), shape=(n, n))
        d = cg.shortest_path(A, unweighted=True, indices=0)
        want = np.where(np.isinf(d), -1, d).astype(np.int32)
        assert np.array_equal(po.bfs(n, rp, ci, 0), want)
        lv, _ = po.bfs_bsr(bs, n, *po.pack(bs, n, rp, ci), 0)
        assert np.array_equal(lv, want)


@pytest.mark.parametrize("bs", (16, 64))
def test_minplus_and_dijkstra(po, gbo, bs):
    rp, ci = random_sym(400, 1600, 5)
    g = gbo.Csr(400, 400, rp, ci)
    w = gbo.edge_weights(g, 1, 64, 7)
    A = sp.csr_matrix((w.astype(np.float64), ci, rp), shape=(400, 400))
    d = cg.dijkstra(A, indices=3)
    want = np.where(np.isinf(d), 0xFFFFFFFF, d).astype(np.uint32)
    assert np.array_equal(po.dijkstra(400, rp, ci, w, 3), want)
    # Bellman-Ford fixpoint of the oracle's tile min-plus product
    pk = po.pack(bs, 400, rp, ci)
    tv, woff = po.pack_values(bs, 400, rp, ci, w, *pk)
    dist = np.full(400, 0xFFFFFFFF, np.uint32)
    dist[3] = 0
    for _ in range(400):
        nd = po.minplus(bs, 400, *pk, dist, tv, woff)
        if np.array_equal(nd, dist):
            break
        dist = nd
    assert np.array_equal(dist, want)
    # weights symmetric per undirected edge
    Wd = sp.csr_matrix((w.astype(np.int64), ci, rp), shape=(400, 400)).toarray()
    assert np.array_equal(Wd, Wd.T) and w.min() >= 1 and w.max() <= 64


def test_pagerank_oracle():
    import pyoracle as po
    for n, (rp, ci) in ((50, complete_graph(50)), (200, random_sym(200, 600, 8))):
        deg = np.diff(rp).astype(np.int32)
        p, _ = po.pagerank(n, rp, ci, deg, 0.85, 1e-15, 100000)
        # direct solve of (I - alpha A D^-1) p = (1-alpha)/n
        A = dense(n, rp, ci).astype(np.float64)
        Dinv = np.where(deg > 0, 1.0 / np.maximum(deg, 1), 0.0)
        M = np.eye(n) - 0.85 * A * Dinv[None, :]
        want = np.linalg.solve(M, np.full(n, 0.15 / n))
        assert np.abs(p - want).sum() < 1e-12
        if deg.min() > 0:
            assert abs(p.sum() - 1.0) < 1e-12


def test_cc_oracle(po):
    rp, ci = random_sym(3000, 1500, 6)
    A = sp.csr_matrix((np.ones(len(ci)), ci, rp), shape=(3000, 3000))
    nc, lab = cg.connected_components(A)
    first = {}
    for v in range(3000):
        first.setdefault(lab[v], v)
    assert np.array_equal(po.cc(3000, rp, ci), np.array([first[lab[v]] for v in range(3000)], np.int32))


def test_count_sample_rows(po):
    """The cpu_baseline sample path (block-row range) equals the full sweep."""
    rp, ci = random_sym(1000, 5000, 11)
    brp, bci, tw = po.pack(16, 1000, rp, ci)
    xb = po.pack_vec(16, (np.arange(1000) % 3 == 0).astype(np.float32))
    full = po.count(16, 1000, brp, bci, tw, xb)
    part = po.count(16, 1000, brp, bci, tw, xb, 10, 40)
    assert np.array_equal(part[160:640], full[160:640]) and not part[:160].any()


def test_ppr_oracle_finite_differences():
    """FP64 adjoint formulas vs central differences (h = 1e-3) of the forward."""
    import ppr_oracle as pp
    rp, ci = random_sym(60, 150, 12)
    rp = rp.copy()
    P = pp.transition(60, rp, ci)
    rng = np.random.default_rng(1)
    S = rng.dirichlet(np.full(60, 0.3), size=3).T  # [n][K]
    W = rng.standard_normal((60, 3))
    alpha, h = 0.85, 1e-3
    gS, ga = pp.adjoint(P, S, W, alpha)
    L = lambda S_, a_: float((W * pp.forward(P, S_, a_)).sum())
    fd_a = (L(S, alpha + h) - L(S, alpha - h)) / (2 * h)
    assert abs(fd_a - ga) <= 1e-5 * max(1.0, abs(ga))
    # second-order convergence of the central difference towards the formula
    h2 = h / 10
    fd_a2 = (L(S, alpha + h2) - L(S, alpha - h2)) / (2 * h2)
    assert abs(fd_a2 - ga) <= abs(fd_a - ga) / 50 + 1e-10
    for (i, k) in ((0, 0), (7, 1), (33, 2)):
        Sp, Sm = S.copy(), S.copy()
        Sp[i, k] += h
        Sm[i, k] -= h
        fd = (L(Sp, alpha) - L(Sm, alpha)) / (2 * h)
        assert abs(fd - gS[i, k]) <= 1e-8 + 1e-6 * abs(fd)
    # p is a fixed point of the PageRank update
    p = pp.forward(P, S, alpha)
    assert np.abs(p - (alpha * P.T @ p + (1 - alpha) * S)).max() < 1e-12


# ------------------------------------------------------------------------
# Value products (sum, min, min-plus) and the BFS step, each against a
# dense brute force built from the CSR directly (not from the tiles).  One
# case is directed, so a transposed operand (a row/column mix-up in the
# tile loops) fails; two sizes are not multiples of 64, so the padded last
# block row and column are covered; one matrix is empty.
def _value_cases(po):
    rp, ci = po.gen_er(1024, 0.02, 0)
    yield "er1024", 1024, rp, ci
    rp, ci = random_sym(1000, 3000, 7)
    yield "rand1000", 1000, rp, ci
    rng = np.random.default_rng(21)
    u, v = rng.integers(0, 517, 4000), rng.integers(0, 517, 4000)
    rp, ci = po.csr_from_edges(517, u, v, symmetric=False, drop_loops=False)
    yield "directed517", 517, rp, ci
    yield "empty130", 130, np.zeros(131, np.int32), np.zeros(0, np.int32)


@pytest.mark.parametrize("bs", BS)
def test_sum_f64_dense(po, bs):
    """or_bsrbmv_sum_f64: y = A x in FP64 (signed x, so a dropped or doubled
    term shows)."""
    rng = np.random.default_rng(bs)
    for name, n, rp, ci in _value_cases(po):
        D = dense(n, rp, ci).astype(np.float64)
        x = rng.standard_normal(n)
        pk = po.pack(bs, n, rp, ci)
        y = po.sum_f64(bs, n, *pk, x)
        want = D @ x
        bound = 1e-12 * (np.abs(D) @ np.abs(x)) + 1e-300
        assert np.all(np.abs(y - want) <= bound), name
        if name == "directed517":  # the transpose gives a different answer
            assert not np.allclose(y, D.T @ x)


@pytest.mark.parametrize("bs", BS)
def test_min_i32_dense(po, bs):
    """or_bsrbmv_min_i32: y[i] = min over the row's neighbours j of x[j],
    INT32_MAX for a row without neighbours; x spans negative values."""
    rng = np.random.default_rng(bs + 100)
    imax = np.iinfo(np.int32).max
    for name, n, rp, ci in _value_cases(po):
        D = dense(n, rp, ci).astype(bool)
        x = rng.integers(-2**31, 2**31 - 1, n, dtype=np.int64).astype(np.int32)
        x[: n // 3] = rng.integers(-50, 50, n // 3)
        pk = po.pack(bs, n, rp, ci)
        y = po.min_i32(bs, n, *pk, x)
        want = np.where(D, x[None, :].astype(np.int64), imax).min(axis=1) if n else np.zeros(0)
        assert np.array_equal(y, want.astype(np.int32)), name
        deg = np.diff(rp)
        assert np.all(y[deg == 0] == imax), name


def _sat_add(a, b):
    return np.minimum(a.astype(np.int64) + b.astype(np.int64), 0xFFFFFFFF)


@pytest.mark.parametrize("bs", BS)
def test_minplus_unit_dense(po, bs):
    """Unit-weight branch of or_bsrbmv_minplus (w = NULL): out[i] =
    min(dist[i], min_j dist[j] + 1) saturating at 0xFFFFFFFF; its fixpoint
    from a single source equals the BFS levels (scipy)."""
    rng = np.random.default_rng(bs + 200)
    for name, n, rp, ci in _value_cases(po):
        D = dense(n, rp, ci).astype(bool)
        dist = rng.integers(0, 5000, n).astype(np.uint32)
        dist[rng.random(n) < 0.3] = 0xFFFFFFFF
        dist[rng.random(n) < 0.05] = 0xFFFFFFFE  # +1 lands exactly on INF
        pk = po.pack(bs, n, rp, ci)
        out = po.minplus(bs, n, *pk, dist)
        cand = np.where(D, _sat_add(dist, np.ones(n, np.uint32))[None, :], 0xFFFFFFFF).min(axis=1) if n else dist
        want = np.minimum(dist.astype(np.int64), cand).astype(np.uint32)
        assert np.array_equal(out, want), name
        if n == 0 or name == "directed517":
            continue
        d = np.full(n, 0xFFFFFFFF, np.uint32)
        d[0] = 0
        for _ in range(n + 1):
            nd = po.minplus(bs, n, *pk, d)
            if np.array_equal(nd, d):
                break
            d = nd
        A = sp.csr_matrix((np.ones(len(ci)), ci, rp), shape=(n, n))
        lv = cg.shortest_path(A, unweighted=True, indices=0)
        assert np.array_equal(d, np.where(np.isinf(lv), 0xFFFFFFFF, lv).astype(np.uint32)), name


@pytest.mark.parametrize("bs", (8, 64))
def test_minplus_weighted_dense(po, gbo, bs):
    """Weighted branch, one product: out[i] = min(dist[i], min_j dist[j] +
    W[i,j]) with the weights addressed through the tile bit order (woff +
    rank); directed, so a weight read from the mirrored edge fails."""
    rng = np.random.default_rng(bs + 300)
    for name, n, rp, ci in _value_cases(po):
        if len(ci) == 0:
            continue
        w = rng.integers(1, 65, len(ci)).astype(np.uint32)
        W = np.zeros((n, n), np.int64)
        for r in range(n):
            W[r, ci[rp[r]:rp[r + 1]]] = w[rp[r]:rp[r + 1]]
        dist = rng.integers(0, 100, n).astype(np.uint32)
        dist[rng.random(n) < 0.3] = 0xFFFFFFFF
        pk = po.pack(bs, n, rp, ci)
        tv, woff = po.pack_values(bs, n, rp, ci, w, *pk)
        out = po.minplus(bs, n, *pk, dist, tv, woff)
        cand = np.where(W > 0, np.minimum(dist.astype(np.int64)[None, :] + W, 0xFFFFFFFF), 0xFFFFFFFF).min(axis=1)
        want = np.minimum(dist.astype(np.int64), cand).astype(np.uint32)
        assert np.array_equal(out, want), name


@pytest.mark.parametrize("bs", BS)
def test_bfs_step_dense(po, bs):
    """or_bsrbmv_bfs_step with arbitrary frontier and visited sets (not only
    those a BFS reaches): next = (A f > 0) & ~visited, count = |next|."""
    rng = np.random.default_rng(bs + 400)
    for name, n, rp, ci in _value_cases(po):
        D = dense(n, rp, ci)
        f = (rng.random(n) < 0.1).astype(np.float32)
        vis = (rng.random(n) < 0.4).astype(np.float32)
        pk = po.pack(bs, n, rp, ci)
        nx, cnt = po.bfs_step(bs, n, *pk, po.pack_vec(bs, f), po.pack_vec(bs, vis))
        want = ((D @ f.astype(np.int64)) > 0) & (vis == 0)
        assert np.array_equal(nx, po.pack_vec(bs, want.astype(np.float32))), name
        assert cnt == int(want.sum()), name


# ------------------------------------------------------------------------
# Generators: the engine's generators are checked against the oracle's, so
# the oracle's must be pinned by what the distributions fix.
def _rmat_raw(po, scale, ef, a, b, c, seed):
    m = ef << scale
    u = np.empty(m, np.int32)
    v = np.empty(m, np.int32)
    po.lib().or_gen_rmat_edges(scale, ef, a, b, c, seed, u.ctypes.data, v.ctypes.data)
    return u.astype(np.int64), v.astype(np.int64)


@pytest.mark.parametrize("abc", ((0.57, 0.19, 0.19), (0.45, 0.25, 0.20)))
def test_rmat_quadrant_frequencies(po, abc):
    """Every recursion level of the raw R-MAT edge list (before
    symmetrisation) picks quadrant (row half, column half) = (0,0), (0,1),
    (1,0), (1,1) with probability a, b, c, d: each frequency within 5 sigma.
    The second parameter set has b != c, so swapping the row and column bit
    or two quadrants fails."""
    a, b, c = abc
    probs = {(0, 0): a, (0, 1): b, (1, 0): c, (1, 1): 1 - a - b - c}
    scale, ef = 14, 16
    u, v = _rmat_raw(po, scale, ef, a, b, c, 7)
    m = len(u)
    assert u.min() >= 0 and u.max() < 1 << scale and v.max() < 1 << scale
    for level in range(scale):
        sh = scale - 1 - level  # level 0 decides the top bit
        bu, bv = (u >> sh) & 1, (v >> sh) & 1
        for (qu, qv), p in probs.items():
            freq = np.count_nonzero((bu == qu) & (bv == qv)) / m
            assert abs(freq - p) <= 5 * math.sqrt(p * (1 - p) / m), (level, qu, qv, freq, p)
    # vertex 0 as a source: every level picks the top row half, (a + b)^scale
    p0 = (a + b) ** scale
    k0 = np.count_nonzero(u == 0)
    assert abs(k0 - m * p0) <= 5 * math.sqrt(m * p0 * (1 - p0))


def test_er_edge_count_and_uniformity(po):
    """G(n, p): the number of undirected edges is Binomial(n(n-1)/2, p)
    (within 5 sigma), every pair u < v, and the degrees have the binomial
    mean and variance within sampling error."""
    n, p = 2000, 0.01
    m = po.lib().or_gen_er_edges(n, p, 3, None, None, 0)
    pairs = n * (n - 1) // 2
    assert abs(m - pairs * p) <= 5 * math.sqrt(pairs * p * (1 - p))
    u = np.empty(m, np.int32)
    v = np.empty(m, np.int32)
    po.lib().or_gen_er_edges(n, p, 3, u.ctypes.data, v.ctypes.data, m)
    assert np.all(u < v) and len(np.unique(u.astype(np.int64) * n + v)) == m
    deg = np.bincount(np.concatenate([u, v]), minlength=n)
    mu = (n - 1) * p
    assert abs(deg.mean() - mu) <= 5 * math.sqrt(mu * (1 - p) / n) * 2
    assert 0.8 * mu * (1 - p) <= deg.var() <= 1.2 * mu * (1 - p)


# ------------------------------------------------------------------------
# FP64 personalized PageRank (or_ppr_solve / or_ppr_check): the fixed points
# equal the dense solves of oracle/ppr_oracle.py (themselves pinned by
# finite differences above), on a directed graph with vertices without
# out-edges, so a swapped pull/push CSR or a dropped teleport term fails.
def _directed_case():
    import pyoracle as po
    rng = np.random.default_rng(33)
    n = 90
    u, v = rng.integers(0, n, 400), rng.integers(0, n, 400)
    u[u < 6] = 6  # vertices 0..5 have no out-edges
    rp, ci = po.csr_from_edges(n, u, v, symmetric=False)
    return n, rp, ci


def test_ppr_solve_dense():
    import ppr_oracle as pp
    import pyoracle as po
    n, rp, ci = _directed_case()
    deg = np.diff(rp).astype(np.int32)
    T = pp.transition(n, rp, ci)
    rng = np.random.default_rng(4)
    S = rng.dirichlet(np.full(n, 0.3), size=3).T
    W = rng.standard_normal((n, 3))
    # forward pulls over A^T: row i of the transpose lists i's in-neighbours
    tr = sp.csr_matrix((np.ones(len(ci)), ci, rp), shape=(n, n)).T.tocsr()
    trp, tci = tr.indptr.astype(np.int32), tr.indices.astype(np.int32)
    P, it = po.ppr_solve(n, trp, tci, deg, 0.85, 0, S)
    assert it > 0 and np.abs(P - pp.forward(T, S, 0.85)).max() < 1e-12
    lam, it = po.ppr_solve(n, rp, ci, deg, 0.85, 1, W)
    wl = np.linalg.solve(np.eye(n) - 0.85 * T, W)
    assert it > 0 and np.abs(lam - wl).max() < 1e-11 * np.abs(wl).max()
    # the wrong CSR for either map gives a different answer
    P2, _ = po.ppr_solve(n, rp, ci, deg, 0.85, 0, S)
    assert np.abs(P2 - P).max() > 1e-3


def test_ppr_check_bounds():
    """or_ppr_check: the residual of the exact fixed point is at rounding
    level; for a perturbed X the residual bound |X - T(X)|/(1-alpha) covers
    the true distance (1-norm for the forward, max-norm for the adjoint);
    dot and y_l1 equal lambda^T (P^T p - s) and |P^T p - s|_1 computed
    densely."""
    import ppr_oracle as pp
    import pyoracle as po
    n, rp, ci = _directed_case()
    deg = np.diff(rp).astype(np.int32)
    T = pp.transition(n, rp, ci)
    tr = sp.csr_matrix((np.ones(len(ci)), ci, rp), shape=(n, n)).T.tocsr()
    trp, tci = tr.indptr.astype(np.int32), tr.indices.astype(np.int32)
    rng = np.random.default_rng(5)
    S = rng.dirichlet(np.full(n, 0.3), size=2).T
    W = rng.standard_normal((n, 2))
    a = 0.85
    P = pp.forward(T, S, a)
    lam = np.linalg.solve(np.eye(n) - a * T, W)
    l1, linf, dot, yl1 = po.ppr_check(n, trp, tci, deg, a, 0, P, S, lam)
    assert l1.max() < 1e-6  # FP32 rounding of P only
    y = T.T @ P.astype(np.float32).astype(np.float64)
    lam32 = lam.astype(np.float32).astype(np.float64)
    np.testing.assert_allclose(dot, (lam32 * (y - S.astype(np.float32))).sum(0), rtol=1e-9)
    np.testing.assert_allclose(yl1, np.abs(y - S.astype(np.float32)).sum(0), rtol=1e-9)
    Xp = P + 1e-3 * rng.standard_normal(P.shape)
    l1, _, _, _ = po.ppr_check(n, trp, tci, deg, a, 0, Xp, S)
    true = np.abs(Xp.astype(np.float32) - P).sum(0)
    assert np.all(true <= l1 / (1 - a) + 1e-6) and np.all(l1 > 0.1 * true)
    Lp = lam + 1e-3 * rng.standard_normal(lam.shape)
    _, linf, _, _ = po.ppr_check(n, rp, ci, deg, a, 1, Lp, W)
    true = np.abs(Lp.astype(np.float32) - lam).max(0)
    assert np.all(true <= linf / (1 - a) + 1e-5)


@pytest.mark.parametrize("case", ["rect", "rmat", "empty", "directed"])
def test_spgemm_against_scipy(po, case):
    """or_spgemm (the general product's checker): pattern and integer counts
    equal scipy's A @ B on rectangular, skewed, empty and directed operands
    (a transposed operand or a dropped term changes the counts)."""
    import scipy.sparse as sp
    if case == "rect":
        A = sp.random(70, 90, 0.05, format="csr", random_state=1)
        B = sp.random(90, 40, 0.08, format="csr", random_state=2)
    elif case == "rmat":
        import gbo
        g = gbo.gen_rmat(9, 8, 0.57, 0.19, 0.19, seed=4)
        A = B = sp.csr_matrix((np.ones(g.nnz), g.colind, g.rowptr), shape=(g.nrows, g.nrows))
    elif case == "empty":
        A = sp.csr_matrix((30, 50))
        B = sp.random(50, 20, 0.2, format="csr", random_state=3)
    else:
        A = sp.random(60, 60, 0.06, format="csr", random_state=5)
        B = A.T.tocsr()
    A, B = A.tocsr(), B.tocsr()
    A.data[:] = 1
    B.data[:] = 1
    A.sort_indices()
    B.sort_indices()
    rp, ci, val = po.spgemm(A.shape[0], A.indptr, A.indices, B.indptr, B.indices, B.shape[1])
    C = (A @ B).tocsr()
    C.eliminate_zeros()
    C.sort_indices()
    assert np.array_equal(rp, C.indptr) and np.array_equal(ci, C.indices)
    assert np.array_equal(val, C.data.astype(np.int32))
    if case == "directed":  # A A^T differs from A^T A: the operands are not swapped
        rp2, ci2, _ = po.spgemm(60, B.indptr, B.indices, A.indptr, A.indices, 60)
        assert not (np.array_equal(rp, rp2) and np.array_equal(ci, ci2))
