"""Parity at the full workload sizes: every graph configuration of the
benchmark, at its real shape, against the oracle.

- R-MAT scale 22 (4,194,304 vertices, edge factor 16, a/b/c = .57/.19/.19,
  symmetrized, deduplicated, seed 1) at block sizes 8/16/32/64: BFS levels
  (pull and direction-optimizing) against the oracle's queue BFS, SSSP with
  uint32 weights uniform in [1, 64] against Dijkstra, CC labels against
  union-find, PageRank (damping 0.85) within 1e-6 L1 of the FP64 fixed
  point -- both against the oracle's FP64 power iteration and by the FP64
  residual bound |p - T(p)|_1 / (1 - alpha), which holds for any p because
  the PageRank map T is an alpha-contraction in the 1-norm.
- 256-source batched BFS on the same graph (sources uniform, seed 3): every
  source's levels against the oracle's queue BFS.
- 512 Erdos-Renyi graphs (n log-uniform in [1k, 64k], mean degree uniform
  in [8, 32], seed 4), BFS + CC of all of them in one hipGraph-replayed
  launch: every graph against the oracle.
- Personalized PageRank on the same R-MAT graph with 64 Dirichlet(0.1)
  vectors over 1000 seed vertices each (seed 5), w ~ N(0, 1): every
  column's forward error bounded by its FP64 residual (<= 1e-6 in L1), the
  adjoint's by its max-norm residual, grad_alpha by the FP64 product plus
  those bounds (<= 1e-5 relative); two columns against an FP64 CPU adjoint
  and central differences (h = 1e-3, FP64 forward solves).

The oracle runs on the host cores (threads for the many independent BFS
runs).  Progress lines go to the real stderr so a long case is visibly
alive.
"""
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

# minutes of host-side oracle work at these sizes: a limit of their own
pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]
BS = (8, 16, 32, 64)
ALPHA = 0.85
WORKERS = min(16, os.cpu_count() or 4)


def note(msg):
    sys.__stderr__.write(f"[fullsize {time.strftime('%H:%M:%S')}] {msg}\n")
    sys.__stderr__.flush()


class S22:
    """The R-MAT s22 graph and the oracle's answers (computed once, lazily)."""

    def __init__(self, gbo, po):
        self.gbo, self.po = gbo, po
        t = time.time()
        self.g = gbo.gen_rmat(22, 16, 0.57, 0.19, 0.19, seed=1)
        self.n = self.g.nrows
        self.deg = self.g.outdeg()
        self.src = 0 if self.deg[0] > 0 else int(np.nonzero(self.deg)[0][0])
        self.d_rp = gbo.DeviceBuffer.from_numpy(self.g.rowptr)
        self.d_ci = gbo.DeviceBuffer.from_numpy(self.g.colind)
        self._cache = {}
        note(f"R-MAT s22 generated: n={self.n} nnz={self.g.nnz} ({time.time() - t:.1f}s)")

    def oracle(self, key):
        if key not in self._cache:
            g, po, t = self.g, self.po, time.time()
            if key == "bfs":
                v = po.bfs(self.n, g.rowptr, g.colind, self.src)
            elif key == "weights":
                v = self.gbo.edge_weights(g, 1, 64, 1)
            elif key == "dijkstra":
                v = po.dijkstra(self.n, g.rowptr, g.colind, self.oracle("weights"), self.src)
            elif key == "cc":
                v = po.cc(self.n, g.rowptr, g.colind)
            elif key == "pagerank":
                v, it = po.pagerank(self.n, g.rowptr, g.colind, self.deg, ALPHA, 1e-14, 100000)
            self._cache[key] = v
            note(f"oracle {key} ({time.time() - t:.1f}s)")
        return self._cache[key]

    def pack(self, bs):
        return self.gbo.BsrMatrix.from_device_csr(self.d_rp, self.d_ci, self.n, self.n, bs, 0)


@pytest.fixture(scope="module")
def s22(gbo, po):
    return S22(gbo, po)


@pytest.mark.parametrize("bs", BS)
def test_rmat22_drivers(gbo, po, s22, bs):
    """BFS / weighted SSSP / CC / PageRank on R-MAT s22 at each block size."""
    n, g = s22.n, s22.g
    A = s22.pack(bs)
    note(f"bs {bs}: packed {A.nblocks} tiles (plan variant {A.s.plan_variant})")
    want = s22.oracle("bfs")
    hist = np.bincount(want[want >= 0])
    lv, sizes = gbo.bfs(A, s22.src)
    assert np.array_equal(lv, want) and np.array_equal(sizes, hist), "pull BFS"
    lv, sizes = gbo.bfs(A, s22.src, A=A)
    assert np.array_equal(lv, want) and np.array_equal(sizes, hist), "direction-optimizing BFS"
    note(f"bs {bs}: BFS ok ({len(hist)} levels)")
    tv = gbo.TileValues(A, g, s22.oracle("weights"))
    d, it = gbo.sssp(A, s22.src, tv)
    tv.free()
    assert np.array_equal(d, s22.oracle("dijkstra")), "weighted SSSP"
    note(f"bs {bs}: SSSP ok ({it} iterations)")
    lab, it = gbo.cc(A)
    assert np.array_equal(lab, s22.oracle("cc")), "CC"
    note(f"bs {bs}: CC ok ({it} iterations)")
    # FP32 iterates stop changing somewhere near 1e-8 L1 at this size: run to
    # 1e-9 or 300 iterations, then bound the error directly
    p, it, delta = gbo.pagerank(A, s22.deg, ALPHA, 1e-9, 300)
    p64 = s22.oracle("pagerank")
    l1 = float(np.abs(p.astype(np.float64) - p64).sum())
    # residual bound: needs no reference solution
    res, _, _, _ = po.ppr_check(n, g.rowptr, g.colind, s22.deg, ALPHA, 0, p.reshape(n, 1),
                                np.full((n, 1), 1.0 / n, np.float32))
    bound = float(res[0]) / (1 - ALPHA)
    note(f"bs {bs}: PageRank {it} iterations, L1 vs FP64 {l1:.2e}, residual bound {bound:.2e}")
    assert l1 <= 1e-6 and bound <= 1e-6
    A.free()


def _compare_levels(po, g, sources, levels_list):
    """Every source's levels (for each GPU result) equal the queue BFS."""
    def check(i):
        want = po.bfs(g.nrows, g.rowptr, g.colind, int(sources[i]))
        return [i for lv in levels_list if not np.array_equal(lv[i], want)]
    with ThreadPoolExecutor(WORKERS) as ex:
        bad = [b for r in ex.map(check, range(len(sources))) for b in r]
    return bad


def test_rmat22_multisource_bfs(gbo, po, s22):
    """256 sources (uniform, seed 3) in one batched BFS, bs 8 and bs 64."""
    src = np.random.default_rng(3).integers(0, s22.n, 256).astype(np.int32)
    results = []
    for bs in (8, 64):
        A = s22.pack(bs)
        t = time.time()
        lv, nl = gbo.bfs_multisource(A, src)
        note(f"multi-source bs {bs}: {nl} levels ({time.time() - t:.2f}s incl. readback)")
        assert nl == int(lv.max()) + 1
        results.append(lv)
        A.free()
    t = time.time()
    bad = _compare_levels(po, s22.g, src, results)
    note(f"multi-source: 256 sources checked ({time.time() - t:.1f}s)")
    assert not bad, f"sources differing: {sorted(set(bad))[:10]}"


def test_er512_graph_batch(gbo, po):
    """512 ER graphs, BFS + CC of all of them in one replayed hipGraph launch
    (bs 8, the autotuner's pick for this batch, and bs 32)."""
    t = time.time()
    graphs = gbo.gen_er_batch(count=512, seed=4)
    rng = np.random.default_rng(4)
    src = np.array([rng.integers(0, g.nrows) for g in graphs], np.int32)
    note(f"512 ER graphs generated: {sum(g.nrows for g in graphs)} vertices, "
         f"{sum(g.nnz for g in graphs)} nnz ({time.time() - t:.1f}s)")
    res = []
    for bs in (8, 32):
        B = gbo.GraphBatch(graphs, bs)
        B.run(src, use_graph=True)
        B.run(src, use_graph=True)  # a replay of the captured launch
        res.append(B.results())
        B.free()
        note(f"graph batch bs {bs}: done")

    def check(i):
        g = graphs[i]
        lv = po.bfs(g.nrows, g.rowptr, g.colind, int(src[i]))
        lb = po.cc(g.nrows, g.rowptr, g.colind)
        return [(i, k) for k, (levels, labels) in enumerate(res)
                if not (np.array_equal(levels[i], lv) and np.array_equal(labels[i], lb))]
    with ThreadPoolExecutor(WORKERS) as ex:
        bad = [b for r in ex.map(check, range(len(graphs))) for b in r]
    assert not bad, f"graphs differing (graph, bs index): {bad[:10]}"


def _ppr_inputs(n, K=64, nseeds=1000, seed=5):
    rng = np.random.default_rng(seed)
    S = np.zeros((n, K), np.float32)
    for k in range(K):
        idx = rng.choice(n, nseeds, replace=False)
        S[idx, k] = rng.dirichlet(np.full(nseeds, 0.1)).astype(np.float32)
    W = rng.standard_normal((n, K)).astype(np.float32)
    return S, W


def test_rmat22_ppr_forward_adjoint(gbo, po, s22):
    """Personalized PageRank, K = 64, bs 8 (the benchmark's block size)."""
    n, g, deg, a = s22.n, s22.g, s22.deg, ALPHA
    rp, ci = g.rowptr, g.colind  # symmetric: the pull CSR is the out-edge CSR
    S, W = _ppr_inputs(n)
    A = s22.pack(8)
    t = time.time()
    # solves run to 1e-8 (relative L1 change) or 400 iterations, whichever
    # comes first: the FP32 floor of the change is not known in advance, and
    # the FP64 residual bounds below measure the error itself
    P, itf, df = gbo.ppr_forward(A, deg, S, a, 1e-8, 400)
    gS, ga, ita = gbo.ppr_adjoint(A, A, deg, S, P, W, a, 1e-8, 400)
    note(f"PPR K=64: forward {itf} iterations, adjoint {ita} ({time.time() - t:.1f}s incl. copies)")
    lam = (gS / np.float32(1 - a)).astype(np.float32)
    # forward: per-column 1-norm error bound
    t = time.time()
    rf, _, dot, yl1 = po.ppr_check(n, rp, ci, deg, a, 0, P, S, lam)
    ef = rf / (1 - a)
    # adjoint: per-column max-norm error bound
    _, ra, _, _ = po.ppr_check(n, rp, ci, deg, a, 1, lam, W)
    ea = ra / (1 - a)
    note(f"PPR FP64 residuals ({time.time() - t:.1f}s): forward L1 bound max {ef.max():.2e}, "
         f"adjoint max-norm bound max {ea.max():.2e}")
    assert ef.max() <= 1e-6, "forward: a column is not within 1e-6 L1 of the FP64 fixed point"
    lam_inf = np.abs(lam).max(0)
    # grad_S = (1 - alpha) lambda: max-norm error relative to the largest
    # exact gradient entry (|lambda*| >= |lambda| - bound)
    assert np.all(ea <= 1e-5 * (lam_inf - ea)), "adjoint: grad_S not within 1e-5"
    # grad_alpha = sum_k lambda_k . (P^T p_k - s_k): the GPU's product and
    # reduction against the same product in FP64 on the GPU's lambda and p.
    # (The worst-case effect of the lambda and p errors bounded above,
    # sum_k |dl_k|_inf |P^T p_k - s_k|_1 + |l*_k|_inf |dp_k|_1, is printed
    # for information: it adds the rounding of all 64 x 4M entries with one
    # sign, so it is far above the error; the two-column FP64 solves below
    # check grad_alpha against an independent FP64 gradient.)
    g64 = float(dot.sum())
    err = float(np.sum(ea * yl1 + (lam_inf + ea) * ef))
    note(f"PPR grad_alpha {ga:.6e} vs FP64 product {g64:.6e} (worst-case solve error {err:.1e})")
    assert abs(ga - g64) <= 1e-5 * abs(g64), (ga, g64)

    # Two columns against FP64 CPU solves warm-started from the GPU result.
    cols = [3, 41]
    S2, W2 = S[:, cols].astype(np.float64), W[:, cols].astype(np.float64)
    t = time.time()
    lam64, it = po.ppr_solve(n, rp, ci, deg, a, 1, W2, lam[:, cols], tol=1e-13, max_iter=1000)
    assert it > 0
    p64, it = po.ppr_solve(n, rp, ci, deg, a, 0, S2, P[:, cols], tol=1e-13, max_iter=1000)
    assert it > 0
    note(f"PPR FP64 two-column adjoint and forward ({time.time() - t:.1f}s)")
    assert np.abs(P[:, cols] - p64).sum(0).max() <= 1e-6
    gs64 = (1 - a) * lam64
    scale = np.abs(gs64).max()
    assert np.abs(gS[:, cols] - gs64).max() <= 1e-5 * scale
    # entries: the largest gradient, a seed vertex of a column, vertex 0
    i_max = np.unravel_index(np.argmax(np.abs(gs64)), gs64.shape)
    seed_v = int(np.nonzero(S[:, cols[1]])[0][0])
    entries = [(int(i_max[0]), int(i_max[1])), (seed_v, 1), (0, 0)]
    for i, k in entries:
        assert abs(gS[i, cols[k]] - gs64[i, k]) <= 1e-5 * abs(gs64[i, k]) + 1e-7 * scale, (i, k)
    # central differences of the FP64 forward, h = 1e-3: in S for the three
    # entries (columns s_k +- h e_i solved together) ...
    h = 1e-3
    R = np.zeros((n, 2 * len(entries)))
    for m, (i, k) in enumerate(entries):
        R[:, 2 * m] = S2[:, k]
        R[:, 2 * m + 1] = S2[:, k]
        R[i, 2 * m] += h
        R[i, 2 * m + 1] -= h
    X0 = np.repeat(p64[:, [k for _, k in entries]], 2, axis=1)
    t = time.time()
    Pfd, it = po.ppr_solve(n, rp, ci, deg, a, 0, R, X0, tol=1e-13, max_iter=1000)
    assert it > 0
    for m, (i, k) in enumerate(entries):
        fd = float(W2[:, k] @ (Pfd[:, 2 * m] - Pfd[:, 2 * m + 1])) / (2 * h)
        assert abs(gS[i, cols[k]] - fd) <= 1e-5 * abs(fd) + 1e-7 * scale, (i, k, fd)
    # ... and in alpha, for the loss restricted to the two columns (the GPU
    # adjoint of the two-column problem)
    _, ga2, _ = gbo.ppr_adjoint(A, A, deg, S[:, cols], P[:, cols], W[:, cols], a, 1e-8, 400)
    Lp = po.ppr_solve(n, rp, ci, deg, a + h, 0, S2, p64, tol=1e-13, max_iter=1000)[0]
    Lm = po.ppr_solve(n, rp, ci, deg, a - h, 0, S2, p64, tol=1e-13, max_iter=1000)[0]
    fd_a = float((W2 * (Lp - Lm)).sum()) / (2 * h)
    # the FP64 adjoint formula on the same two columns: lambda . (P^T p - s)
    import scipy.sparse as sp
    M = sp.csr_matrix((np.ones(len(ci)), ci, rp), shape=(n, n))
    inv = np.divide(1.0, deg, out=np.zeros(n), where=deg > 0)
    ga64 = float((lam64 * (M @ (p64 * inv[:, None]) - S2)).sum())
    note(f"PPR two-column grad_alpha: GPU {ga2:.6e}, FP64 adjoint {ga64:.6e}, central difference {fd_a:.6e} "
         f"({time.time() - t:.1f}s)")
    assert abs(ga2 - ga64) <= 1e-5 * abs(ga64)
    assert abs(ga2 - fd_a) <= 1e-4 * abs(fd_a)
    A.free()
