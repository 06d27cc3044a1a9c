#!/usr/bin/env python3
"""bsrbmv GTEPS + HBM-roofline fraction on R-MAT scale 22, one MI355X.

Workload: R-MAT scale 22 (n = 4,194,304, edge factor 16, a/b/c/d =
.57/.19/.19/.05, symmetrized, self-loops dropped, deduplicated, seed 1),
packed on the GPU into bit-BSR at block sizes 8/16/32/64.  One "step" is
one bsrbmv count sweep y = A x over all tiles (x = seeded Bernoulli(0.5)
bits, as in the reference's bsrbmv benchmark).  Inputs are resident in HBM
before timing starts; steps are timed with HIP events on the stream the
kernels run on.  The headline `value` is GTEPS = nnz / sweep time at the
fastest block size; per-block-size sweeps, BFS and PageRank on the same
graph are reported alongside.  `cpu_baseline` times the oracle's OpenMP
restatement of the same sweep on a bounded sample of block rows.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--scale 22]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "graph-bit-ops_amd"))

HBM_PEAK_GBS = 8000.0
METRIC = "bsrbmv GTEPS + HBM roofline frac, R-MAT s22; batched/adjoint vs plain on 1 GPU"


def gbo_default_chunk(bs):
    return {8: 256, 16: 256, 32: 256, 64: 256}[bs]


PROFILE_MD = os.path.join(ROOT, "profiles", "bsrbmv_count_rmat_s22.md")


def measured_traffic(bs, variant, scale):
    """HBM bytes per count-sweep launch (calibrated FETCH_SIZE + WRITE_SIZE)
    for this block size and kernel variant, from the committed rocprofv3
    summary of the same workload (R-MAT s22 only), or None."""
    if scale != 22 or not os.path.exists(PROFILE_MD):
        return None
    for line in open(PROFILE_MD):
        if line.startswith(f"traffic: bs={bs} variant={variant} "):
            f = dict(kv.split("=") for kv in line.split()[1:])
            return int(f["fetch_bytes"]) + int(f["write_bytes"])
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def sweep_bytes(A, n):
    """Algorithmic bytes of one count sweep: tiles + colind + rowptr + x bits + y."""
    T, bs, nbr = A.nblocks, A.bs, A.nblockrows
    return T * bs * bs // 8 + 4 * T + 4 * (nbr + 1) + (n + 7) // 8 + 4 * n


class Timer:
    """HIP events on the null stream (the stream every call below uses)."""

    def __init__(self, gbo):
        import ctypes
        self.gbo = gbo
        self.L = gbo.lib()
        self.a, self.b = ctypes.c_void_p(), ctypes.c_void_p()
        gbo._check(self.L.gbo_event_create(ctypes.byref(self.a)))
        gbo._check(self.L.gbo_event_create(ctypes.byref(self.b)))

    def start(self):
        self.gbo._check(self.L.gbo_event_record(self.a, None))

    def stop_ms(self):
        import ctypes
        self.gbo._check(self.L.gbo_event_record(self.b, None))
        ms = ctypes.c_float()
        self.gbo._check(self.L.gbo_event_elapsed(ctypes.byref(ms), self.a, self.b))
        return float(ms.value)


def time_count(gbo, A, dx, dy, steps, warmup, timer):
    L = gbo.lib()
    for _ in range(warmup):
        gbo._check(L.gbo_bsrbmv_count(A.ref, dx.ptr, dy.ptr, 0, None))
    gbo.synchronize()
    timer.start()
    for _ in range(steps):
        gbo._check(L.gbo_bsrbmv_count(A.ref, dx.ptr, dy.ptr, 0, None))
    return timer.stop_ms() / steps


def time_bfs_step(gbo, A, dx, steps, warmup, timer):
    """One pull-BFS step over every tile: frontier = the count sweep's x bits
    (half the vertices), nothing visited."""
    L = gbo.lib()
    nwb = (A.nblockcols * A.bs) // 8
    vis = gbo.DeviceBuffer(nwb)
    gbo._check(L.gbo_memset(vis.ptr, 0, nwb))
    nxt = gbo.DeviceBuffer(max(A.nblockrows, 1) * A.bs // 8)
    cnt = gbo.DeviceBuffer(8)
    for _ in range(warmup):
        gbo._check(L.gbo_bsrbmv_bfs_step(A.ref, dx.ptr, vis.ptr, nxt.ptr, cnt.ptr, None))
    gbo.synchronize()
    timer.start()
    for _ in range(steps):
        gbo._check(L.gbo_bsrbmv_bfs_step(A.ref, dx.ptr, vis.ptr, nxt.ptr, cnt.ptr, None))
    ms = timer.stop_ms() / steps
    vis.free()
    nxt.free()
    cnt.free()
    return ms


def pick_plan(gbo, A, count_plan, run, steps, warmup, timer):
    """(chunk, variant) for one set-bit driver: the count-tuned plan or a
    default chunk (and at bs 8/16 variant 3, the set-bit list kernel),
    whichever gives the fastest sweep of that driver's product `run`.  Plans
    never change integer results; the sum's order of additions follows the
    chunk."""
    bc, bv = count_plan
    dc = gbo_default_chunk(A.bs)
    cands = {(c, v) for c in (bc, dc, 2 * dc) for v in ({bv, 3} if A.bs <= 16 else {bv})}
    times = {}
    for c, v in sorted(cands):
        A.plan(c, v)
        for _ in range(warmup):
            gbo._check(run())
        gbo.synchronize()
        timer.start()
        for _ in range(steps):
            gbo._check(run())
        times[(A.s.plan_chunk, v)] = timer.stop_ms() / steps
    return min(times, key=times.get), times


def time_value_sweeps(gbo, A, g, w_sssp, steps, warmup, timer, plans=None):
    """One full sweep of each value op over every tile: sum (PageRank's
    product, x = 1/n), min (CC's product, x = vertex ids) and min-plus
    (SSSP's product with the uint32 [1,64] weights, half the distances
    finite).  Algorithmic bytes: tiles + colind + rowptr + 4n in + 4n out,
    and for min-plus also 4 bytes per nonzero of weights + 8 per tile of
    weight offsets."""
    L = gbo.lib()
    n = A.nrows
    base = sweep_bytes(A, n) - (n + 7) // 8 - 4 * n  # tiles, colind, rowptr
    out = {}
    dy = gbo.DeviceBuffer(4 * max(n, 1))
    rng = np.random.default_rng(7)
    xs = gbo.DeviceBuffer.from_numpy(np.full(n, 1.0 / n, np.float32))
    xi = gbo.DeviceBuffer.from_numpy(np.arange(n, dtype=np.int32))
    dist = rng.integers(0, 1000, n).astype(np.uint32)
    dist[rng.random(n) < 0.5] = 0xFFFFFFFF
    xd = gbo.DeviceBuffer.from_numpy(dist)
    tv = gbo.TileValues(A, g, w_sssp)
    calls = {
        "sum_f32": (lambda: L.gbo_bsrbmv_sum_f32(A.ref, xs.ptr, dy.ptr, None), base + 8 * n),
        "min_i32": (lambda: L.gbo_bsrbmv_min_i32(A.ref, xi.ptr, dy.ptr, None), base + 8 * n),
        "minplus_weighted": (lambda: L.gbo_bsrbmv_minplus(A.ref, tv.tval.ptr, tv.woff.ptr, xd.ptr, dy.ptr, None),
                             base + 8 * n + 4 * g.nnz + 8 * A.nblocks),
    }
    for name, (fn, byt) in calls.items():
        if plans:  # each op on its own plan
            A.plan(*plans[name])
        for _ in range(warmup):
            gbo._check(fn())
        gbo.synchronize()
        timer.start()
        for _ in range(steps):
            gbo._check(fn())
        ms = timer.stop_ms() / steps
        out[name] = {"ms": round(ms, 4), "bytes": int(byt),
                     "frac_hbm": round(byt / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    tv.free()
    for b in (dy, xs, xi, xd):
        b.free()
    return out


def hbm_ceilings(gbo, timer, nbytes=1 << 30, reps=10):
    """Measured HBM ceilings on this GPU: a plain copy kernel ((read + write)
    bytes / time) and a read-only stream (read bytes / time), 1 GiB each."""
    L = gbo.lib()
    a = gbo.DeviceBuffer(nbytes)
    b = gbo.DeviceBuffer(nbytes)
    sink = gbo.DeviceBuffer(16)
    gbo._check(L.gbo_memset(a.ptr, 1, nbytes))  # equal bytes: the XOR fold is 0, never 0xFFFFFFFF
    out = {}
    for name, fn, moved in (("copy", lambda: L.gbo_stream_copy(b.ptr, a.ptr, nbytes, None), 2 * nbytes),
                            ("read", lambda: L.gbo_stream_read(a.ptr, nbytes, 0xFFFFFFFF, sink.ptr, None), nbytes)):
        gbo._check(fn())
        gbo.synchronize()
        timer.start()
        for _ in range(reps):
            gbo._check(fn())
        ms = timer.stop_ms() / reps
        out[name] = round(moved / (ms * 1e-3) / 1e9, 1)
    for buf in (a, b, sink):
        buf.free()
    return out


def gather_probes(gbo, A, dx, d_ci, nnz, n, timer, reps=10):
    """Random-gather rate of this GPU for the two gather patterns of the
    sweeps, each fed with the sweep's own indices and nothing else:
    - count: one x byte per tile, indexed by the tiles' block columns (the
      count sweep's x gathers plus its colind stream, without the tiles);
    - value: one 4-byte value of an n-long vector per nonzero, indexed by
      the CSR column of every nonzero (the value sweeps' gathers)."""
    L = gbo.lib()
    sink = gbo.DeviceBuffer(16)
    vec = gbo.DeviceBuffer.from_numpy(np.full(n, 1.0 / n, np.float32))
    out = {}
    for name, table, elem, idx, cnt in (("count_x_gathers", dx.ptr, 1, A.s.colind, A.nblocks),
                                        ("value_gathers", vec.ptr, 4, d_ci.ptr, nnz)):
        fn = lambda: L.gbo_gather_probe(table, elem, idx, cnt, 0xFFFFFFFF, sink.ptr, None)
        gbo._check(fn())
        gbo.synchronize()
        timer.start()
        for _ in range(reps):
            gbo._check(fn())
        ms = timer.stop_ms() / reps
        out[name] = {"gathers": int(cnt), "ms": round(ms, 4), "G_per_s": round(cnt / (ms * 1e-3) / 1e9, 1)}
    sink.free()
    vec.free()
    return out


def bench_multisource(gbo, A, g, timer, nsrc=256, seed=3):
    """C4a: 256-source batched BFS vs 256 plain BFS runs on the same GPU, all
    results in device memory (the plain runs write their levels into their
    slice of a second [nsrc][n] buffer); the comparison runs after both
    timings.  Plain runs: pull BFS (gbo_bfs) and direction-optimizing BFS
    (gbo_bfs_do), each over all 256 sources.  Throughput = sources * nnz /
    time.  Roofline of the batched pass: per level the tiles, colind and
    rowptr once plus the frontier and visited words (2 * K * n / 8 bytes)."""
    rng = np.random.default_rng(seed)
    src = rng.integers(0, g.nrows, nsrc).astype(np.int32)
    n = g.nrows
    isolated = int((g.outdeg()[src] == 0).sum())
    out = gbo.DeviceBuffer(4 * nsrc * n)
    plain = gbo.DeviceBuffer(4 * nsrc * n)
    gbo.bfs_multisource(A, src[:64], out)  # warm-up
    gbo.synchronize()
    t = time.perf_counter()
    nl = gbo.bfs_multisource(A, src, out)
    gbo.synchronize()
    tb = time.perf_counter() - t
    res = {"sources": nsrc, "isolated_sources": isolated, "levels": nl, "batched_ms": round(tb * 1e3, 2), "bs": A.bs}
    for name, kw in (("pull", {}), ("direction_optimizing", {"A": A})):
        gbo.bfs(A, int(src[0]), out=plain.view(0, 4 * n), **kw)  # warm-up
        gbo.synchronize()
        t = time.perf_counter()
        for i in range(nsrc):
            gbo.bfs(A, int(src[i]), out=plain.view(4 * i * n, 4 * n), **kw)
        gbo.synchronize()
        tp = time.perf_counter() - t
        exact = bool(np.array_equal(out.to_numpy(np.int32, nsrc * n), plain.to_numpy(np.int32, nsrc * n)))
        res[f"plain_{name}_ms_per_source"] = round(tp * 1e3 / nsrc, 3)
        res[f"plain_{name}_GTEPS"] = round(nsrc * g.nnz / tp / 1e9, 3)
        res[f"speedup_vs_{name}"] = round(tp / tb, 2)
        res[f"bit_exact_vs_{name}"] = exact
    out.free()
    plain.free()
    res["batched_GTEPS"] = round(nsrc * g.nnz / tb / 1e9, 2)
    level_bytes = A.tile_bytes() + 4 * A.nblocks + 4 * (A.nblockrows + 1) + 2 * nsrc * n // 8
    res["bytes_per_level"] = int(level_bytes)
    res["frac_hbm"] = round(nl * level_bytes / tb / 1e9 / HBM_PEAK_GBS, 4)
    return res


def bench_graph_batch(gbo, count=512, bs=None, seed=4):
    """C4b: `count` ER graphs (n log-uniform in [1k, 64k], mean degree in
    [8, 32]) -- BFS + CC of all of them in one persistent hipGraph launch vs
    per-graph plain gbo_bfs + gbo_cc.  The block size is the autotuner's pick
    for a BFS-step sweep of the block-diagonal union unless `bs` is given."""
    t = time.perf_counter()
    graphs = gbo.gen_er_batch(count=count, seed=seed)
    gen_s = time.perf_counter() - t
    nnz = sum(g.nnz for g in graphs)
    rng = np.random.default_rng(seed)
    src = np.array([rng.integers(0, g.nrows) for g in graphs], np.int32)
    tuned = None
    if bs is None:
        nu = sum(g.nrows for g in graphs)
        rp, ci, off, base = [np.zeros(1, np.int64)], [], 0, 0
        for g in graphs:
            rp.append(g.rowptr[1:].astype(np.int64) + off)
            ci.append(g.colind + base)
            off += g.nnz
            base += g.nrows
        d_rp = gbo.DeviceBuffer.from_numpy(np.concatenate(rp).astype(np.int32))
        d_ci = gbo.DeviceBuffer.from_numpy(np.concatenate(ci).astype(np.int32))
        tu = gbo.autotune(d_rp, d_ci, nu, nu, op=1)
        d_rp.free()
        d_ci.free()
        bs = tu.bs
        tuned = {"op": "bfs step over the union", "bs": tu.bs, "chunk": tu.chunk, "variant": tu.variant,
                 "ms": round(tu.ms, 4)}
    t = time.perf_counter()
    B = gbo.GraphBatch(graphs, bs)
    pack_s = time.perf_counter() - t
    B.run(src, use_graph=True)  # capture + first replay
    t = time.perf_counter()
    reps = 3
    for _ in range(reps):
        B.run(src, use_graph=True)
    tb = (time.perf_counter() - t) / reps
    levels, labels = B.results()
    # per-graph plain runs: results stay in device memory (each graph's slice
    # of two concatenated buffers), compared after the timing
    off = np.concatenate([[0], np.cumsum([gg.nrows for gg in graphs])])
    d_lv = gbo.DeviceBuffer(4 * int(off[-1]))
    d_lb = gbo.DeviceBuffer(4 * int(off[-1]))
    mats = [gbo.BsrMatrix.from_csr(gg, bs) for gg in graphs]
    gbo.synchronize()
    t = time.perf_counter()
    for i, (gg, A) in enumerate(zip(graphs, mats)):
        gbo.bfs(A, int(src[i]), out=d_lv.view(4 * int(off[i]), 4 * gg.nrows))
        gbo.cc(A, out=d_lb.view(4 * int(off[i]), 4 * gg.nrows))
    gbo.synchronize()
    tp = time.perf_counter() - t
    plv, plb = d_lv.to_numpy(np.int32), d_lb.to_numpy(np.int32)
    exact = all(np.array_equal(plv[off[i]:off[i + 1]], levels[i]) and np.array_equal(plb[off[i]:off[i + 1]], labels[i])
                for i in range(len(graphs)))
    for A in mats:
        A.free()
    d_lv.free()
    d_lb.free()
    B.free()
    return {"graphs": count, "vertices": int(sum(g.nrows for g in graphs)), "nnz": int(nnz), "bs": bs,
            "autotuned": tuned,
            "gen_s": round(gen_s, 1), "pack_s": round(pack_s, 2), "batched_ms": round(tb * 1e3, 2),
            "plain_ms": round(tp * 1e3, 1), "speedup": round(tp / tb, 2), "bit_exact_vs_plain": exact}


def bench_ppr(gbo, A, g, timer, d_ci=None, K=64, nseeds=1000, seed=5, tol=1e-6, max_iter=200):
    """C5: PPR forward and adjoint, 64 Dirichlet(0.1) personalization
    vectors over 1000 seed vertices each; w ~ N(0, 1).  Inputs and outputs
    are device-resident ([n][K] float32); each solve is timed with HIP
    events around the C call on the null stream.  Roofline of one SpMM
    iteration (compulsory bytes): tiles + colind + rowptr + X read once +
    Y written once + S read once (3 * n * K * 4); the row gathers actually
    read nnz * K * 4 bytes, reported as gathered_GBs."""
    rng = np.random.default_rng(seed)
    n = g.nrows
    S = np.zeros((n, K), np.float32)
    for k in range(K):
        idx = rng.choice(n, nseeds, replace=False)
        S[idx, k] = rng.dirichlet(np.full(nseeds, 0.1)).astype(np.float32)
    W = rng.standard_normal((n, K)).astype(np.float32)
    d_s, d_w = gbo.DeviceBuffer.from_numpy(S), gbo.DeviceBuffer.from_numpy(W)
    d_p, d_g = gbo.DeviceBuffer(S.nbytes), gbo.DeviceBuffer(S.nbytes)
    d_od = gbo.DeviceBuffer.from_numpy(g.outdeg().astype(np.int32))
    gbo.ppr_forward_ptr(A, d_od.ptr, K, 0.85, d_s.ptr, d_p.ptr, 1e-2, 5, strict=False)  # warm-up
    gbo.synchronize()
    timer.start()
    itf, df = gbo.ppr_forward_ptr(A, d_od.ptr, K, 0.85, d_s.ptr, d_p.ptr, tol, max_iter, strict=False)
    tf = timer.stop_ms()
    timer.start()
    ga, ita = gbo.ppr_adjoint_ptr(A, A, d_od.ptr, K, 0.85, d_s.ptr, d_p.ptr, d_w.ptr, d_g.ptr, tol, max_iter,
                                  strict=False)
    ta = timer.stop_ms()
    probe = None
    if d_ci is not None and K == 64:
        # the SpMM's row reads alone: one 256-byte row of P per nonzero
        sink = gbo.DeviceBuffer(16)
        fn = lambda: gbo.lib().gbo_gather_probe(d_p.ptr, 256, d_ci.ptr, g.nnz, 0xFFFFFFFF, sink.ptr, None)
        gbo._check(fn())
        gbo.synchronize()
        timer.start()
        for _ in range(3):
            gbo._check(fn())
        probe = round(timer.stop_ms() / 3, 3)
        sink.free()
    for b in (d_s, d_w, d_p, d_g, d_od):
        b.free()
    it_bytes = A.tile_bytes() + 4 * A.nblocks + 4 * (A.nblockrows + 1) + 3 * n * K * 4
    ms_it = tf / max(itf, 1)
    return {"K": K, "bs": A.bs, "tol": tol, "forward_ms": round(tf, 1), "forward_iters": itf, "forward_delta": df,
            "adjoint_ms": round(ta, 1), "adjoint_iters": ita, "adjoint_over_forward": round(ta / tf, 3),
            "ms_per_spmm_iter": round(ms_it, 3), "bytes_per_iter": int(it_bytes),
            "frac_hbm_per_iter": round(it_bytes / (ms_it * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "gathered_GBs": round(g.nnz * K * 4 / (ms_it * 1e-3) / 1e9, 1), "row_gather_probe_ms": probe,
            "spmm_iter_over_probe": round(ms_it / probe, 3) if probe else None, "grad_alpha": ga}


def cpu_baseline(gbo, A, xbits, n, target_s):
    """Oracle OpenMP sweep over a contiguous sample of block rows."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po

    bs = A.bs
    rp = np.empty(A.nblockrows + 1, np.int32)
    gbo._check(gbo.lib().gbo_memcpy_d2h(rp.ctypes.data, A.s.rowptr, rp.nbytes))
    # Sample: block rows from the middle of the matrix holding ~1/16 of the
    # tiles (skips the hub rows at the top, which would flatter nobody).
    T = A.nblocks
    want = max(T // 16, 1)
    r0 = int(np.searchsorted(rp, T // 2) - 1)
    r0 = max(0, min(r0, A.nblockrows - 1))
    r1 = int(np.searchsorted(rp, rp[r0] + want))
    r1 = max(r0 + 1, min(r1, A.nblockrows))
    t0, t1 = int(rp[r0]), int(rp[r1])
    ci = np.empty(t1 - t0, np.int32)
    tw = np.empty((t1 - t0) * bs, gbo.WORD_DTYPE[bs])
    wb = bs // 8
    gbo._check(gbo.lib().gbo_memcpy_d2h(ci.ctypes.data, A.s.colind + 4 * t0, ci.nbytes))
    gbo._check(gbo.lib().gbo_memcpy_d2h(tw.ctypes.data, A.s.tiles + t0 * bs * wb, tw.nbytes))
    lrp = (rp[r0:r1 + 1] - t0).astype(np.int32)
    nrows = (r1 - r0) * bs
    bits = int(np.unpackbits(tw.view(np.uint8)).sum())
    po.count(bs, nrows, lrp, ci, tw, xbits)  # warm-up
    reps, el = 0, 0.0
    while el < target_s or reps < 3:
        t = time.perf_counter()
        po.count(bs, nrows, lrp, ci, tw, xbits)
        el += time.perf_counter() - t
        reps += 1
        if reps >= 50:
            break
    sec = el / reps
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {
        "value": round(bits / sec / 1e9, 4), "unit": "GTEPS", "cores": threads, "kind": "port",
        "sample": (f"oracle/oracle.c or_bsrbmv_count (OpenMP, {threads} threads) over block rows "
                   f"[{r0},{r1}) = {t1 - t0} of {T} tiles ({bits} nonzeros) at bs={bs}, "
                   f"{reps} reps, {sec * 1e3:.1f} ms/rep"),
    }


def time_spgemm(gbo, Lr, LTr, timer):
    """The general SpGEMM C = L * L^T (symbolic + numeric, bit offsets,
    integer values at C's set bits), timed with HIP events on the null
    stream around the C calls; returns (ms, C tiles, C nnz, sum of values)."""
    import ctypes
    L = gbo.lib()
    C = gbo.BsrStruct()
    timer.start()
    gbo._check(L.gbo_bsrbmm(Lr.ref, LTr.ref, ctypes.byref(C), None))
    off = gbo.DeviceBuffer(8 * (C.nblocks + 1))
    nnz = ctypes.c_int64()
    gbo._check(L.gbo_bsr_bit_offsets(ctypes.byref(C), off.ptr, ctypes.byref(nnz), None))
    cnt = gbo.DeviceBuffer(4 * max(nnz.value, 1))
    gbo._check(L.gbo_bsrbmm_values(Lr.ref, LTr.ref, ctypes.byref(C), off.ptr, cnt.ptr, None))
    ms = timer.stop_ms()
    total = int(cnt.to_numpy(np.int32, nnz.value).astype(np.int64).sum())
    nb = C.nblocks
    L.gbo_bsr_free(ctypes.byref(C))
    off.free()
    cnt.free()
    return ms, nb, nnz.value, total


def bench_tc(gbo, timer, reps=3):
    """C3: triangle counting by bsrbmm (L row tiles x L^T column tiles,
    masked by L, popcount, uint64 total) on the 2048x2048 king grid (closed
    form 4(m-1)^2) and R-MAT s20 (seed 2).  The whole gbo_bsrbmm_tc call
    (block transpose of L^T, row form of its tiles, the masked product) is
    timed with HIP events on prepacked operands; strict tril/transposition
    run on the host.  bytes = both operands' tiles + colind + rowptr read
    once.  On the king grid the general SpGEMM C = L * L^T (sum of C's values
    = sum_c colcount_c(L)^2) is timed as well."""
    out = {}
    for name, g in (("king2048", gbo.gen_king(2048)), ("rmat20", gbo.gen_rmat(20, 16, seed=2))):
        L = g.tril(strict=True)
        LT = L.transpose()
        rec = {"n": g.nrows, "nnz": g.nnz, "L_nnz": L.nnz}
        if name == "king2048":
            rec["closed_form"] = 4 * 2047 ** 2
        per = {}
        for bs in (8, 16, 32, 64):
            Lr = gbo.BsrMatrix.from_csr(L, bs, 0)
            Lc = gbo.BsrMatrix.from_csr(LT, bs, 1)
            cnt = gbo.bsrbmm_tc(Lr, Lc)  # warm-up
            timer.start()
            for _ in range(reps):
                cnt = gbo.bsrbmm_tc(Lr, Lc)
            ms = timer.stop_ms() / reps
            byt = 2 * (Lr.tile_bytes() + 4 * Lr.nblocks + 4 * (Lr.nblockrows + 1))
            per[bs] = {"triangles": cnt, "ms": round(ms, 3), "L_tiles": Lr.nblocks,
                       "GTEPS_L": round(L.nnz / (ms * 1e-3) / 1e9, 3), "bytes": int(byt),
                       "frac_hbm": round(byt / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
            Lc.free()
            if name == "king2048":
                LTr = gbo.BsrMatrix.from_csr(LT, bs, 0)
                time_spgemm(gbo, Lr, LTr, timer)  # warm-up
                sms, cnb, cnnz, csum = time_spgemm(gbo, Lr, LTr, timer)
                want = int((np.bincount(L.colind, minlength=g.ncols).astype(np.int64) ** 2).sum())
                per[bs]["spgemm_LLT"] = {"ms": round(sms, 3), "C_tiles": cnb, "C_nnz": cnnz, "values_sum": csum,
                                         "equals_colcount_squares": csum == want}
                LTr.free()
            Lr.free()
        rec["per_block_size"] = per
        counts = {v["triangles"] for v in per.values()}
        rec["consistent_across_bs"] = len(counts) == 1
        if "closed_form" in rec:
            rec["matches_closed_form"] = counts == {rec["closed_form"]}
        out[name] = rec
        log(f"[bench] tc {name}: {rec}")
    return out


def bench_ingest(gbo, g, workdir):
    """.mtx ingest at scale: the graph written as a pattern .mtx file, then
    read with both reader semantics, then through the binary CSR cache
    (first call parses and writes it, second loads it).  Host work only."""
    import tempfile
    rec = {"threads": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))}
    with tempfile.TemporaryDirectory(dir=workdir) as d:
        p = os.path.join(d, "graph.mtx")
        t = time.perf_counter()
        gbo.mtx_write(p, g)
        rec["write_s"] = round(time.perf_counter() - t, 2)
        rec["mtx_GB"] = round(os.path.getsize(p) / 1e9, 3)
        for sem in (0, 1):
            t = time.perf_counter()
            c = gbo.mtx_read(p, sem, 0)
            rec[f"read_semantics{sem}_s"] = round(time.perf_counter() - t, 2)
            rec[f"semantics{sem}_equal"] = bool(np.array_equal(c.rowptr, g.rowptr) and
                                                np.array_equal(c.colind, g.colind))
            del c
        for key in ("cache_write_s", "cache_load_s"):
            t = time.perf_counter()
            c = gbo.mtx_read(p, 1, 0, cache=True)
            rec[key] = round(time.perf_counter() - t, 2)
            rec[key.replace("_s", "_state")] = c.cache_state
        rec["cache_equal"] = bool(np.array_equal(c.rowptr, g.rowptr) and np.array_equal(c.colind, g.colind))
        rec["cache_GB"] = round(os.path.getsize(gbo.mtx_cache_path(p, 1, 0)) / 1e9, 3)
    return rec


def keep_ppr(gbo, d_rp, d_ci, n):
    """The PPR SpMM runs at bs 8 (one 256-B row gather per set bit)."""
    return gbo.BsrMatrix.from_device_csr(d_rp, d_ci, n, n, 8, 0)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--bs", default="8,16,32,64")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-algos", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--variants", default="0,1,2,3", help="bs 8/16 kernel variants to time (best is kept)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched BFS / graph batch / PPR lines")
    ap.add_argument("--ingest", action="store_true",
                    help="only time .mtx ingest of the graph (write, both reader semantics, binary cache)")
    ap.add_argument("--ingest-dir", default=None, help="directory for the --ingest files (default: system temp)")
    ap.add_argument("--profile", action="store_true",
                    help="short run for rocprofv3: only the count sweeps and the copy kernel")
    ap.add_argument("--profile-sweeps", action="store_true",
                    help="short run for rocprofv3: BFS full steps and sum / min / min-plus sweeps at the "
                         "default plan of each --bs, --steps times each")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(the bit-BSR sweep would shard by tile-balanced block-row ranges, one range per GPU).",
              file=sys.stderr)
        sys.exit(2)

    import gbo

    t = time.time()
    g = gbo.gen_rmat(args.scale, 16, 0.57, 0.19, 0.19, seed=1)
    n, nnz = g.nrows, g.nnz
    log(f"[bench] R-MAT s{args.scale}: n={n} nnz={nnz} generated in {time.time() - t:.1f}s")
    if args.ingest:
        print(json.dumps({"ingest": {"graph": f"R-MAT s{args.scale}", "n": n, "nnz": nnz,
                                     **bench_ingest(gbo, g, args.ingest_dir)}}), flush=True)
        return
    d_rp = gbo.DeviceBuffer.from_numpy(g.rowptr)
    d_ci = gbo.DeviceBuffer.from_numpy(g.colind)
    outdeg = g.outdeg()
    timer = Timer(gbo)
    ceil = hbm_ceilings(gbo, timer)
    log(f"[bench] HBM ceilings: copy {ceil['copy']} GB/s, read-only {ceil['read']} GB/s")
    if args.profile_sweeps:
        # BFS full steps and value sweeps only, each block size's default plan
        w = gbo.edge_weights(g, 1, 64, 1)
        out = {}
        for bs in [int(b) for b in args.bs.split(",")]:
            A = gbo.BsrMatrix.from_device_csr(d_rp, d_ci, n, n, bs, 0)
            dxf = gbo.DeviceBuffer.from_numpy((np.random.default_rng(0).random(n) < 0.5).astype(np.float32))
            dx = gbo.DeviceBuffer((A.nblockcols * bs) // 8)
            gbo._check(gbo.lib().gbo_pack_vec(bs, dxf.ptr, dx.ptr, n, None))
            step_ms = time_bfs_step(gbo, A, dx, args.steps, args.warmup, timer)
            out[bs] = {"variant": A.s.plan_variant, "chunk": A.s.plan_chunk, "bfs_full_step_ms": round(step_ms, 4),
                       "bfs_full_step_bytes": sweep_bytes(A, n) - 4 * n + 3 * ((n + 7) // 8),
                       "value_sweeps": time_value_sweeps(gbo, A, g, w, args.steps, args.warmup, timer)}
            dx.free()
            dxf.free()
            A.free()
        print(json.dumps({"profile_sweeps": out}), flush=True)
        return
    if args.profile:
        args.no_algos = args.no_cpu = args.no_batch = True
    x = (np.random.default_rng(0).random(n) < 0.5).astype(np.float32)
    dxf = gbo.DeviceBuffer.from_numpy(x)
    dy = gbo.DeviceBuffer(4 * n)

    per_bs = {}
    keep = {}
    w_sssp = None
    for bs in [int(b) for b in args.bs.split(",")]:
        t = time.time()
        A = gbo.BsrMatrix.from_device_csr(d_rp, d_ci, n, n, bs, 0)
        gbo.synchronize()
        pack_s = time.time() - t
        dx = gbo.DeviceBuffer((A.nblockcols * bs) // 8)
        gbo._check(gbo.lib().gbo_pack_vec(bs, dxf.ptr, dx.ptr, n, None))
        # autotune: chunk size and (bs 8/16) kernel variant, best time kept
        tried = {}
        variants = [int(v) for v in args.variants.split(",")] if bs <= 16 else [0]
        for variant in variants:
            for chunk in ((0,) if args.profile else (0, 2 * gbo_default_chunk(bs), gbo_default_chunk(bs) // 2)):
                A.plan(chunk, variant)
                tried[(variant, A.s.plan_chunk)] = time_count(gbo, A, dx, dy, args.steps, args.warmup, timer)
        (bv, bc), ms = min(tried.items(), key=lambda kv: kv[1])
        A.plan(bc, bv)
        log(f"[bench] bs={bs} tuning (variant, chunk) -> ms: " + ", ".join(f"{k}: {v:.3f}" for k, v in tried.items()))
        byt = sweep_bytes(A, n)
        rec = {
            "bs": bs, "tiles": A.nblocks, "nnz_per_tile": round(nnz / max(A.nblocks, 1), 3),
            "tile_GB": round(A.tile_bytes() / 1e9, 3), "pack_s": round(pack_s, 3),
            "ms_per_sweep": round(ms, 4), "GTEPS": round(nnz / (ms * 1e-3) / 1e9, 3),
            "variant": bv, "chunk": bc,
            "bytes_per_sweep": byt, "achieved_GBs": round(byt / (ms * 1e-3) / 1e9, 1),
            "frac_hbm": round(byt / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        }
        log(f"[bench] bs={bs} T={A.nblocks} pack {pack_s:.2f}s sweep {ms:.3f} ms "
            f"{rec['GTEPS']} GTEPS {rec['achieved_GBs']} GB/s")
        if not args.no_algos:
            src = int(np.argmax(outdeg > 0))
            # every driver runs once untimed (device memory pools, plans), then
            # the median of `reps` timed runs is reported; inputs (outdeg) and
            # results stay in device memory, read back after the timing
            reps = 3
            d_res = gbo.DeviceBuffer(4 * n)
            d_od = gbo.DeviceBuffer.from_numpy(outdeg.astype(np.int32))

            def timed(fn):
                fn()
                gbo.synchronize()
                ts, r = [], None
                for _ in range(reps):
                    t = time.perf_counter()
                    r = fn()
                    gbo.synchronize()
                    ts.append(time.perf_counter() - t)
                return float(np.median(ts)), r

            bfs_s, (_, sizes) = timed(lambda: gbo.bfs(A, src, out=d_res))
            lv = d_res.to_numpy(np.int32, n)
            reached = lv >= 0
            edges = int(outdeg[reached].sum()) // 2
            # direction-optimizing: the graph is undirected, so A = AT
            do_s, _ = timed(lambda: gbo.bfs(A, src, A=A, out=d_res))
            lv_do = d_res.to_numpy(np.int32, n)
            # one BFS-step sweep with half the vertices in the frontier and none
            # visited: every tile is read (tiles, colind, rowptr, 3 bit vectors)
            step_ms = time_bfs_step(gbo, A, dx, args.steps, args.warmup, timer)
            step_bytes = sweep_bytes(A, n) - 4 * n + 3 * ((n + 7) // 8)
            rec["bfs"] = {"source": src, "levels": len(sizes), "reached": int(reached.sum()),
                          "ms": round(bfs_s * 1e3, 2), "graph500_GTEPS": round(edges / bfs_s / 1e9, 3),
                          "full_step_ms": round(step_ms, 4), "full_step_bytes": step_bytes,
                          "full_step_frac_hbm": round(step_bytes / (step_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                          "direction_optimizing_ms": round(do_s * 1e3, 2),
                          "direction_optimizing_graph500_GTEPS": round(edges / do_s / 1e9, 3),
                          "direction_optimizing_equal": bool(np.array_equal(lv_do, lv))}
            # PageRank, SSSP and CC each run on the plan fastest for their own
            # product (sum, weighted min-plus, min); BFS above used the count's
            if w_sssp is None:
                w_sssp = gbo.edge_weights(g, 1, 64, 1)
            tv = gbo.TileValues(A, g, w_sssp)
            L = gbo.lib()
            xs = gbo.DeviceBuffer.from_numpy(np.full(n, 1.0 / n, np.float32))
            xi = gbo.DeviceBuffer.from_numpy(np.arange(n, dtype=np.int32))
            xd = gbo.DeviceBuffer.from_numpy(np.where(np.arange(n) % 2 == 0, np.arange(n) % 1000,
                                                      0xFFFFFFFF).astype(np.uint32))
            dyv = gbo.DeviceBuffer(4 * max(n, 1))
            products = {
                "sum_f32": lambda: L.gbo_bsrbmv_sum_f32(A.ref, xs.ptr, dyv.ptr, None),
                "minplus_weighted": lambda: L.gbo_bsrbmv_minplus(A.ref, tv.tval.ptr, tv.woff.ptr, xd.ptr, dyv.ptr, None),
                "min_i32": lambda: L.gbo_bsrbmv_min_i32(A.ref, xi.ptr, dyv.ptr, None),
            }
            vplans, rec["value_plan"] = {}, {}
            for op, run in products.items():
                vplans[op], times = pick_plan(gbo, A, (bc, bv), run, args.steps, args.warmup, timer)
                rec["value_plan"][op] = {"chunk": vplans[op][0], "variant": vplans[op][1],
                                         "ms_by_plan": {f"{c},{v}": round(t, 4) for (c, v), t in times.items()}}
            for b in (xs, xi, xd, dyv):
                b.free()
            A.plan(*vplans["sum_f32"])
            pr_s, (_, it, delta) = timed(lambda: gbo.pagerank(A, d_od, 0.85, 1e-7, 1000, out=d_res))
            # per iteration: one sum sweep (tiles, colind, rowptr, read p, write p)
            pr_bytes = sweep_bytes(A, n) - 4 * n + 8 * n
            rec["pagerank"] = {"iters": it, "ms": round(pr_s * 1e3, 1), "ms_per_iter": round(pr_s * 1e3 / it, 3),
                               "last_l1_delta": delta, "bytes_per_iter": pr_bytes,
                               "frac_hbm_per_iter": round(pr_bytes / (pr_s / it) / 1e9 / HBM_PEAK_GBS, 4)}
            A.plan(*vplans["minplus_weighted"])
            sssp_s, (_, sit) = timed(lambda: gbo.sssp(A, src, tv, out=d_res))
            dist = d_res.to_numpy(np.uint32, n)
            tv.free()
            rec["sssp"] = {"source": src, "weights": "uint32 uniform [1,64], seed 1", "iters": sit,
                           "reached": int((dist != 0xFFFFFFFF).sum()), "ms": round(sssp_s * 1e3, 1)}
            A.plan(*vplans["min_i32"])
            cc_s, (_, cit) = timed(lambda: gbo.cc(A, out=d_res))
            lab = d_res.to_numpy(np.int32, n)
            d_res.free()
            d_od.free()
            rec["cc"] = {"iters": cit, "components": int(len(np.unique(lab))), "ms": round(cc_s * 1e3, 1)}
            rec["value_sweeps"] = time_value_sweeps(gbo, A, g, w_sssp, max(3, args.steps // 4), 1, timer, vplans)
            A.plan(bc, bv)  # back to the count-tuned plan
            log(f"[bench] bs={bs} bfs {rec['bfs']} pr {rec['pagerank']} sssp {rec['sssp']} cc {rec['cc']} "
                f"value plan {rec['value_plan']}")
        per_bs[bs] = rec
        keep[bs] = (A, dx)
        if len(keep) > 1:  # keep only the best so far resident
            best_bs = min(keep, key=lambda b: per_bs[b]["ms_per_sweep"])
            for b in list(keep):
                if b != best_bs:
                    keep[b][0].free()
                    del keep[b]

    best = min(per_bs.values(), key=lambda r: r["ms_per_sweep"])
    bs = best["bs"]
    A, dx = keep[bs]
    probes = gather_probes(gbo, A, dx, d_ci, nnz, n, timer)
    log(f"[bench] gather probes (bs {bs}): {probes}")
    extra = {}
    if not args.no_batch:
        for name, fn in (("multisource_bfs", lambda: bench_multisource(gbo, A, g, timer)),
                         ("graph_batch", lambda: bench_graph_batch(gbo)),
                         ("ppr", lambda: bench_ppr(gbo, keep_ppr(gbo, d_rp, d_ci, n), g, timer, d_ci))):
            try:
                t = time.time()
                extra[name] = fn()
                log(f"[bench] {name} ({time.time() - t:.1f}s): {extra[name]}")
            except Exception as e:  # a failing side line must not hide the headline
                extra[name] = {"error": str(e)}
                log(f"[bench] {name} failed: {e}")
    tc = None
    if not args.no_algos:
        try:
            tc = bench_tc(gbo, timer)
        except Exception as e:  # a failing side line must not hide the headline
            tc = {"error": str(e)}
            log(f"[bench] tc failed: {e}")
    cpu = None
    if not args.no_cpu:
        xb = gbo.pack_vec(bs, x)
        cpu = cpu_baseline(gbo, A, xb, n, args.cpu_seconds)
        log(f"[bench] cpu baseline {cpu}")
    out = {
        "metric": METRIC, "value": best["GTEPS"], "unit": "GTEPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": best["ms_per_sweep"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int32", "data": "synthetic",
        "config": {"workload": f"bsrbmv count sweep, R-MAT s{args.scale} (n={n}, ef16, a.57 b.19 c.19, "
                               f"symmetrized, dedup, seed 1), x=Bernoulli(0.5)",
                   "nnz": nnz, "bs": bs, "parallelism": "none (1 GPU)"},
        "roofline": {"bound": "hbm", "achieved": best["achieved_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": best["frac_hbm"], "traffic": measured_traffic(bs, best["variant"], args.scale),
                     "algorithmic_bytes": best["bytes_per_sweep"], "copy_ceiling_GBs": ceil["copy"],
                     "read_ceiling_GBs": ceil["read"],
                     "frac_of_read_ceiling": round(best["achieved_GBs"] / ceil["read"], 4),
                     "gather_probe": probes,
                     "sweep_over_gather_probe": round(best["ms_per_sweep"] / probes["count_x_gathers"]["ms"], 3)},
        "cpu_baseline": cpu,
        "per_block_size": per_bs,
        "triangle_counting": tc,
        "batched": extra,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
