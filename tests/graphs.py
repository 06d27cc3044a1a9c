"""Graph cases shared by the parity tests (CSR as int32 numpy arrays)."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def path_graph(n):
    u = np.arange(n - 1)
    return sym_csr(n, u, u + 1)


def star_graph(n):
    u = np.zeros(n - 1, np.int64)
    return sym_csr(n, u, np.arange(1, n))


def complete_graph(n):
    u, v = np.triu_indices(n, 1)
    return sym_csr(n, u, v)


def sym_csr(n, u, v):
    u = np.asarray(u, np.int64)
    v = np.asarray(v, np.int64)
    u, v = np.concatenate([u, v]), np.concatenate([v, u])
    keep = u != v
    key = np.unique(u[keep] * n + v[keep])
    r, c = key // n, key % n
    rp = np.zeros(n + 1, np.int64)
    np.add.at(rp, r + 1, 1)
    return np.cumsum(rp).astype(np.int32), c.astype(np.int32)


def random_sym(n, m, seed):
    rng = np.random.default_rng(seed)
    return sym_csr(n, rng.integers(0, n, m), rng.integers(0, n, m))


def block_diag_components(sizes, seed):
    """Disjoint random connected-ish pieces plus isolated vertices."""
    rng = np.random.default_rng(seed)
    us, vs, off = [], [], 0
    for s in sizes:
        if s > 1:
            perm = rng.permutation(s) + off
            us.append(perm[:-1])
            vs.append(perm[1:])  # a spanning path
            extra = rng.integers(0, s, (2, s)) + off
            us.append(extra[0])
            vs.append(extra[1])
        off += s
    return sym_csr(off, np.concatenate(us) if us else [], np.concatenate(vs) if vs else [])


def cases(gbo):
    """(name, Csr) pairs: ER C1, skewed R-MAT, grid, edge shapes."""
    out = [("er1024", gbo.gen_er(1024, 0.02, 0)),
           ("rmat12", gbo.gen_rmat(12, 16, seed=1)),
           ("king40", gbo.gen_king(40))]
    n = 1000
    rp, ci = random_sym(n, 3000, 7)
    out.append(("rand1000", gbo.Csr(n, n, rp, ci)))
    rp, ci = path_graph(333)
    out.append(("path333", gbo.Csr(333, 333, rp, ci)))
    rp, ci = star_graph(200)
    out.append(("star200", gbo.Csr(200, 200, rp, ci)))
    rp, ci = block_diag_components([50, 1, 1, 120, 3, 0, 77], 3)
    out.append(("comps", gbo.Csr(len(rp) - 1, len(rp) - 1, rp, ci)))
    out.append(("empty130", gbo.Csr(130, 130, np.zeros(131, np.int32), np.zeros(0, np.int32))))
    return out
