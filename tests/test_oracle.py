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
