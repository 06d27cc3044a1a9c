#!/usr/bin/env python3
"""Writes the committed golden fixtures under tests/golden/.

Inputs are small Matrix Market files written here (shaped like the
reference's own fixtures: a 64x64 symmetric pattern with diagonal and
off-diagonal entries, a 64x64 general matrix, a 1000-vertex lower-triangle
pattern, plus files with duplicates, self-loops, real/integer fields).
Expected outputs come from sources independent of the engine:

* readers_expected.npz -- CSR produced by the reference's own readers
  (algorithm/mmio_highlevel.h and algorithm/readMtx.hpp compiled in place
  into oracle/_ref by `make -C oracle ref`);
* c1_expected.npz -- config "ER n=1024, p=0.02, seed 0": dense numpy matvec
  y = A x for x = Bernoulli(0.1) (seed 0) and scipy BFS levels from vertex 0;
* counts.json -- triangle counts from networkx and closed forms, component
  counts from scipy, and the generators' edge-list checksums.

Run in the build container (needs /root/reference for the reader outputs):
    python tests/golden/make_golden.py
"""
import hashlib
import json
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "graph-bit-ops_amd"))

MTX = {
    "sym64.mtx": ("pattern symmetric", 64, 64,
                  [(1, 1), (2, 2), (5, 5), (33, 33), (40, 40), (64, 64), (2, 50), (10, 64), (33, 1)]),
    "gen64.mtx": ("pattern general", 64, 64,
                  [(17, 2), (18, 3), (19, 4), (40, 5), (2, 17), (63, 63), (64, 1), (1, 64), (33, 32), (32, 33)]),
    "dups.mtx": ("pattern general", 20, 20,
                 [(3, 7), (3, 1), (3, 7), (9, 9), (1, 20), (20, 1), (3, 1), (11, 12), (12, 11), (5, 5), (20, 19)]),
    "real.mtx": ("real symmetric", 30, 30,
                 [(2, 1, 0.5), (9, 3, -1.25), (30, 29, 3.0), (7, 7, 2.0), (15, 4, 1e-3)]),
    "int.mtx": ("integer general", 16, 24,
                [(1, 24, 3), (16, 1, -2), (8, 8, 7), (8, 9, 1), (2, 3, 5)]),
}


def write_mtx(path, kind, m, n, entries):
    with open(path, "w") as f:
        f.write(f"%%MatrixMarket matrix coordinate {kind}\n% golden fixture\n{m} {n} {len(entries)}\n")
        for e in entries:
            f.write(" ".join(str(v) for v in e) + "\n")


def lower1000(path):
    rng = np.random.default_rng(43)
    pairs = set()
    while len(pairs) < 5000:
        a, b = rng.integers(1, 1001, 2)
        if a > b:
            pairs.add((int(a), int(b)))
    write_mtx(path, "pattern symmetric", 1000, 1000, sorted(pairs, key=lambda p: (p[1], p[0])))


def main():
    import pyoracle as po
    import scipy.sparse as sp
    import scipy.sparse.csgraph as cg
    import networkx as nx
    import gbo

    for name, (kind, m, n, ent) in MTX.items():
        write_mtx(os.path.join(HERE, name), kind, m, n, ent)
    lower1000(os.path.join(HERE, "lower1000.mtx"))
    files = sorted(list(MTX) + ["lower1000.mtx"])

    # Reference readers (compiled from the reference tree).
    assert po.ref_available(), "build oracle/_ref first: make -C oracle ref"
    exp = {}
    for f in files:
        p = os.path.join(HERE, f)
        m, n, rp, ci = po.ref_mmio(p)
        exp[f"{f}|0|0|rowptr"], exp[f"{f}|0|0|colind"] = rp, ci
        for d in (0, 1, 2):
            if f == "int.mtx" and d == 2:
                continue  # non-square: readMtx would symmetrize out of range
            m, n, rp, ci = po.ref_readmtx(p, d)
            exp[f"{f}|1|{d}|rowptr"], exp[f"{f}|1|{d}|colind"] = rp, ci
    np.savez_compressed(os.path.join(HERE, "readers_expected.npz"), **exp)

    # Config C1 with independent expectations.
    g = gbo.gen_er(1024, 0.02, 0)
    A = sp.csr_matrix((np.ones(g.nnz), g.colind, g.rowptr), shape=(1024, 1024))
    x = (np.random.default_rng(0).random(1024) < 0.1).astype(np.float32)
    y = (A @ x.astype(np.float64)).astype(np.int32)
    lv = cg.shortest_path(A, unweighted=True, indices=0)
    lv = np.where(np.isinf(lv), -1, lv).astype(np.int32)
    np.savez_compressed(os.path.join(HERE, "c1_expected.npz"), rowptr=g.rowptr, colind=g.colind, x=x, y=y, level0=lv)

    counts = {}
    for m in (2, 3, 17, 64):
        counts[f"king{m}_triangles"] = 4 * (m - 1) ** 2
    for n in (3, 10, 65):
        counts[f"K{n}_triangles"] = math.comb(n, 3)
    for s in (10, 14):
        r = gbo.gen_rmat(s, 16, seed=2)
        G = nx.Graph()
        G.add_nodes_from(range(r.nrows))
        for u in range(r.nrows):
            for v in r.colind[r.rowptr[u]:r.rowptr[u + 1]]:
                if v > u:
                    G.add_edge(u, int(v))
        counts[f"rmat{s}_seed2_triangles_networkx"] = sum(nx.triangles(G).values()) // 3
        M = sp.csr_matrix((np.ones(r.nnz), r.colind, r.rowptr), shape=(r.nrows, r.nrows))
        counts[f"rmat{s}_seed2_components_scipy"] = int(cg.connected_components(M)[0])
        counts[f"rmat{s}_seed2_nnz"] = r.nnz
        counts[f"rmat{s}_seed2_colind_sha1"] = hashlib.sha1(r.colind.tobytes()).hexdigest()
    counts["er1024_p0.02_seed0_nnz"] = g.nnz
    counts["er1024_p0.02_seed0_colind_sha1"] = hashlib.sha1(g.colind.tobytes()).hexdigest()
    with open(os.path.join(HERE, "counts.json"), "w") as f:
        json.dump(counts, f, indent=1, sort_keys=True)
    print(json.dumps(counts, indent=1))


if __name__ == "__main__":
    main()
