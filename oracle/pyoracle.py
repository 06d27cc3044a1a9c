"""ctypes/numpy wrapper of the oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the engine.  It also wraps the
reference's own .mtx readers compiled in oracle/_ref (when built).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import POINTER, c_double, c_int, c_int32, c_int64, c_uint64, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
WORD = {8: np.uint8, 16: np.uint16, 32: np.uint32, 64: np.uint64}
_L = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "-B", "all"], check=True)
    if os.path.isdir("/root/reference"):
        subprocess.run(["make", "-s", "-C", HERE, "-B", "ref"], check=True)


def lib():
    global _L
    if _L is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _L = ctypes.CDLL(path)
        P = c_void_p
        sig = {
            "or_rand64": ([c_uint64, c_uint64, c_uint64], c_uint64),
            "or_gen_er_edges": ([c_int32, c_double, c_uint64, P, P, c_int64], c_int64),
            "or_gen_rmat_edges": ([c_int32, c_int32, c_double, c_double, c_double, c_uint64, P, P], c_int64),
            "or_pack_count": ([c_int, c_int32, P, P, P], c_int64),
            "or_pack_fill": ([c_int, c_int, c_int32, P, P, P, P, P], None),
            "or_pack_vec": ([c_int, c_int32, P, P], None),
            "or_pack_values": ([c_int, c_int32, P, P, P, P, P, P, P, P], None),
            "or_bsrbmv_count": ([c_int, c_int32, c_int32, c_int32, P, P, P, P, P], None),
            "or_bsrbmv_bfs_step": ([c_int, c_int32, P, P, P, P, P, P], c_int64),
            "or_bsrbmv_minplus": ([c_int, c_int32, P, P, P, P, P, P, P], None),
            "or_bsrbmv_sum_f64": ([c_int, c_int32, P, P, P, P, P], None),
            "or_bsrbmv_min_i32": ([c_int, c_int32, P, P, P, P, P], None),
            "or_bsrbmm_masked": ([c_int, c_int32, P, P, P, P, P, P], c_uint64),
            "or_bsrbmm_sum": ([c_int, c_int32, P, P, P, P, P, P], c_uint64),
            "or_spgemm": ([c_int32, P, P, P, P, c_int32, P, P, P], c_int64),
            "or_bfs_bsr": ([c_int, c_int32, P, P, P, c_int32, P, P, c_int32], c_int32),
            "or_bfs_queue": ([c_int32, P, P, c_int32, P], c_int32),
            "or_dijkstra": ([c_int32, P, P, P, c_int32, P], None),
            "or_pagerank": ([c_int32, P, P, P, c_double, c_double, c_int32, P], c_int32),
            "or_ppr_solve": ([c_int32, P, P, P, c_double, c_int32, c_int, P, P, c_double, c_int32], c_int32),
            "or_ppr_check": ([c_int32, P, P, P, c_double, c_int32, c_int, P, P, P, P, P, P, P], None),
            "or_cc": ([c_int32, P, P, P], None),
            "or_tc": ([c_int32, P, P], c_uint64),
        }
        for k, (a, r) in sig.items():
            f = getattr(_L, k)
            f.argtypes = a
            f.restype = r
    return _L


def _p(a):
    return None if a is None else a.ctypes.data


# ------------------------------------------------------------- generators
def rand64(seed, stream, idx):
    return lib().or_rand64(seed, stream, idx)


def csr_from_edges(n, u, v, symmetric=True, drop_loops=True):
    u = np.asarray(u, np.int64)
    v = np.asarray(v, np.int64)
    if symmetric:
        u, v = np.concatenate([u, v]), np.concatenate([v, u])
    if drop_loops:
        keep = u != v
        u, v = u[keep], v[keep]
    key = np.unique(u * n + v)
    r, c = key // n, key % n
    rp = np.zeros(n + 1, np.int32)
    np.add.at(rp, r + 1, 1)
    return np.cumsum(rp).astype(np.int32), c.astype(np.int32)


def gen_er(n, p, seed):
    m = lib().or_gen_er_edges(n, p, seed, None, None, 0)
    u = np.empty(m, np.int32)
    v = np.empty(m, np.int32)
    lib().or_gen_er_edges(n, p, seed, _p(u), _p(v), m)
    return csr_from_edges(n, u, v)


def gen_rmat(scale, ef=16, a=0.57, b=0.19, c=0.19, seed=1):
    m = ef << scale
    u = np.empty(m, np.int32)
    v = np.empty(m, np.int32)
    lib().or_gen_rmat_edges(scale, ef, a, b, c, seed, _p(u), _p(v))
    return csr_from_edges(1 << scale, u, v)


# ---------------------------------------------------------------- packing
def pack(bs, nrows, rp, ci, direction=0):
    rp = np.ascontiguousarray(rp, np.int32)
    ci = np.ascontiguousarray(ci, np.int32)
    nbr = (nrows + bs - 1) // bs
    brp = np.zeros(nbr + 1, np.int32)
    T = lib().or_pack_count(bs, nrows, _p(rp), _p(ci), _p(brp))
    bci = np.zeros(max(T, 1), np.int32)
    tiles = np.zeros(max(T, 1) * bs, WORD[bs])
    lib().or_pack_fill(bs, direction, nrows, _p(rp), _p(ci), _p(brp), _p(bci), _p(tiles))
    return brp, bci[:T].copy(), tiles[: T * bs].copy()


def unpack(bs, nrows, ncols, brp, bci, tiles, direction=0):
    """Dense 0/1 matrix of a bit-BSR (for small cases)."""
    nbr = len(brp) - 1
    D = np.zeros((nbr * bs, ((ncols + bs - 1) // bs) * bs), np.uint8)
    for I in range(nbr):
        for t in range(brp[I], brp[I + 1]):
            J = bci[t]
            for w in range(bs):
                word = int(tiles[t * bs + w])
                for k in range(bs):
                    if (word >> (bs - 1 - k)) & 1:
                        r, c = (w, k) if direction == 0 else (k, w)
                        D[I * bs + r, J * bs + c] = 1
    return D[:nrows, :ncols]


def pack_vec(bs, x):
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros((len(x) + bs - 1) // bs, WORD[bs])
    lib().or_pack_vec(bs, len(x), _p(x), _p(out))
    return out


def pack_values(bs, nrows, rp, ci, val, brp, bci, tiles):
    T = len(bci)
    nnz_bits = int(sum(bin(int(w)).count("1") for w in tiles))
    tval = np.zeros(max(nnz_bits, 1), np.uint32)
    woff = np.zeros(T + 1, np.int64)
    lib().or_pack_values(bs, nrows, _p(np.ascontiguousarray(rp, np.int32)), _p(np.ascontiguousarray(ci, np.int32)),
                         _p(np.ascontiguousarray(val, np.uint32)), _p(brp), _p(bci), _p(tiles), _p(tval), _p(woff))
    return tval[:nnz_bits].copy(), woff


# ------------------------------------------------------------------ bsrbmv
def count(bs, nrows, brp, bci, tiles, xbits, r0=0, r1=None):
    nbr = len(brp) - 1
    y = np.zeros(nrows, np.int32)
    lib().or_bsrbmv_count(bs, nrows, r0, nbr if r1 is None else r1, _p(brp), _p(bci), _p(tiles),
                          _p(np.ascontiguousarray(xbits, WORD[bs])), _p(y))
    return y


def bfs_step(bs, nrows, brp, bci, tiles, frontier, visited):
    nx = np.zeros(len(brp) - 1, WORD[bs])
    c = lib().or_bsrbmv_bfs_step(bs, nrows, _p(brp), _p(bci), _p(tiles), _p(np.ascontiguousarray(frontier, WORD[bs])),
                                 _p(np.ascontiguousarray(visited, WORD[bs])), _p(nx))
    return nx, c


def minplus(bs, nrows, brp, bci, tiles, dist, tval=None, woff=None):
    out = np.zeros(nrows, np.uint32)
    lib().or_bsrbmv_minplus(bs, nrows, _p(brp), _p(bci), _p(tiles), _p(tval), _p(woff),
                            _p(np.ascontiguousarray(dist, np.uint32)), _p(out))
    return out


def sum_f64(bs, nrows, brp, bci, tiles, x):
    y = np.zeros(nrows, np.float64)
    lib().or_bsrbmv_sum_f64(bs, nrows, _p(brp), _p(bci), _p(tiles), _p(np.ascontiguousarray(x, np.float64)), _p(y))
    return y


def min_i32(bs, nrows, brp, bci, tiles, x):
    y = np.zeros(nrows, np.int32)
    lib().or_bsrbmv_min_i32(bs, nrows, _p(brp), _p(bci), _p(tiles), _p(np.ascontiguousarray(x, np.int32)), _p(y))
    return y


def bsrbmm_masked(bs, arp, aci, atiles, brp, bci, btiles):
    return lib().or_bsrbmm_masked(bs, len(arp) - 1, _p(arp), _p(aci), _p(atiles), _p(brp), _p(bci), _p(btiles))


def bsrbmm_sum(bs, arp, aci, atiles, brp, bci, btiles):
    return lib().or_bsrbmm_sum(bs, len(arp) - 1, _p(arp), _p(aci), _p(atiles), _p(brp), _p(bci), _p(btiles))


def spgemm(m, arp, aci, brp, bci, ncols):
    """C = A * B over the integers as element CSR (rowptr, colind, counts)."""
    arp, aci, brp, bci = (np.ascontiguousarray(a, np.int32) for a in (arp, aci, brp, bci))
    rp = np.zeros(m + 1, np.int32)
    nnz = lib().or_spgemm(m, _p(arp), _p(aci), _p(brp), _p(bci), ncols, _p(rp), None, None)
    ci = np.zeros(max(nnz, 1), np.int32)
    val = np.zeros(max(nnz, 1), np.int32)
    lib().or_spgemm(m, _p(arp), _p(aci), _p(brp), _p(bci), ncols, _p(rp), _p(ci), _p(val))
    return rp, ci[:nnz].copy(), val[:nnz].copy()


# ----------------------------------------------------------------- drivers
def bfs_bsr(bs, n, brp, bci, tiles, src, max_levels=4096):
    lv = np.zeros(n, np.int32)
    sizes = np.zeros(max_levels, np.int64)
    nl = lib().or_bfs_bsr(bs, n, _p(brp), _p(bci), _p(tiles), src, _p(lv), _p(sizes), max_levels)
    return lv, sizes[:nl].copy()


def bfs(n, rp, ci, src):
    lv = np.zeros(n, np.int32)
    lib().or_bfs_queue(n, _p(np.ascontiguousarray(rp, np.int32)), _p(np.ascontiguousarray(ci, np.int32)), src, _p(lv))
    return lv


def dijkstra(n, rp, ci, w, src):
    d = np.zeros(n, np.uint32)
    lib().or_dijkstra(n, _p(np.ascontiguousarray(rp, np.int32)), _p(np.ascontiguousarray(ci, np.int32)),
                      _p(None if w is None else np.ascontiguousarray(w, np.uint32)), src, _p(d))
    return d


def pagerank(n, rp, ci, outdeg, alpha=0.85, tol=1e-13, max_iter=10000):
    p = np.zeros(n, np.float64)
    it = lib().or_pagerank(n, _p(np.ascontiguousarray(rp, np.int32)), _p(np.ascontiguousarray(ci, np.int32)),
                           _p(np.ascontiguousarray(outdeg, np.int32)), alpha, tol, max_iter, _p(p))
    return p, it


def ppr_solve(n, rp, ci, outdeg, alpha, mode, R, X0=None, tol=1e-13, max_iter=10000):
    """FP64 fixed point of the PPR map (mode 0 forward over the pull CSR,
    mode 1 adjoint over the out-edge CSR); R and the result are [n][K]."""
    R = np.ascontiguousarray(R, np.float64)
    K = R.shape[1]
    X = np.ascontiguousarray(R if X0 is None else X0, np.float64).copy()
    it = lib().or_ppr_solve(n, _p(np.ascontiguousarray(rp, np.int32)), _p(np.ascontiguousarray(ci, np.int32)),
                            _p(np.ascontiguousarray(outdeg, np.int32)), alpha, K, mode, _p(R), _p(X), tol, max_iter)
    return X, it


def ppr_check(n, rp, ci, outdeg, alpha, mode, X, R, L=None):
    """Per column: (|X - T(X)|_1, |X - T(X)|_inf, dot, y_l1) in FP64 for FP32
    X, R (and L): see oracle.h."""
    X = np.ascontiguousarray(X, np.float32)
    R = np.ascontiguousarray(R, np.float32)
    K = X.shape[1]
    out = [np.zeros(K) for _ in range(4)]
    Lp = None if L is None else np.ascontiguousarray(L, np.float32)
    lib().or_ppr_check(n, _p(np.ascontiguousarray(rp, np.int32)), _p(np.ascontiguousarray(ci, np.int32)),
                       _p(np.ascontiguousarray(outdeg, np.int32)), alpha, K, mode, _p(X), _p(R), _p(Lp),
                       *[_p(o) for o in out])
    return tuple(out)


def cc(n, rp, ci):
    lab = np.zeros(n, np.int32)
    lib().or_cc(n, _p(np.ascontiguousarray(rp, np.int32)), _p(np.ascontiguousarray(ci, np.int32)), _p(lab))
    return lab


def tc(n, rp, ci):
    return int(lib().or_tc(n, _p(np.ascontiguousarray(rp, np.int32)), _p(np.ascontiguousarray(ci, np.int32))))


# ------------------------------------------- reference readers (oracle/_ref)
def ref_available():
    return (os.path.exists(os.path.join(HERE, "_ref", "libref_mmio.so"))
            and os.path.exists(os.path.join(HERE, "_ref", "libref_readmtx.so")))


def ref_mmio(path):
    """CSR from the reference's mmio_info/mmio_data (algorithm/mmio_highlevel.h)."""
    L = ctypes.CDLL(os.path.join(HERE, "_ref", "libref_mmio.so"))
    L.ref_mmio_info.argtypes = [ctypes.c_char_p] + [POINTER(c_int)] * 4
    L.ref_mmio_data.argtypes = [ctypes.c_char_p, c_void_p, c_void_p, c_void_p]
    m, n, nnz, sym = c_int(), c_int(), c_int(), c_int()
    rc = L.ref_mmio_info(os.fsencode(path), ctypes.byref(m), ctypes.byref(n), ctypes.byref(nnz), ctypes.byref(sym))
    if rc != 0:
        raise RuntimeError(f"mmio_info returned {rc}")
    rp = np.zeros(m.value + 1, np.int32)
    ci = np.zeros(max(nnz.value, 1), np.int32)
    val = np.zeros(max(nnz.value, 1), np.float32)
    rc = L.ref_mmio_data(os.fsencode(path), _p(rp), _p(ci), _p(val))
    if rc != 0:
        raise RuntimeError(f"mmio_data returned {rc}")
    return m.value, n.value, rp, ci[: nnz.value].copy()


def ref_readmtx(path, directed):
    """CSR from the reference's readMtx + coo2csr (algorithm/readMtx.hpp)."""
    L = ctypes.CDLL(os.path.join(HERE, "_ref", "libref_readmtx.so"))
    L.ref_readmtx.argtypes = [ctypes.c_char_p, c_int] + [POINTER(c_int)] * 3 + [POINTER(c_void_p)] * 2
    L.ref_free.argtypes = [c_void_p]
    nr, nc, nnz = c_int(), c_int(), c_int()
    rpp, cip = c_void_p(), c_void_p()
    L.ref_readmtx(os.fsencode(path), directed, ctypes.byref(nr), ctypes.byref(nc), ctypes.byref(nnz),
                  ctypes.byref(rpp), ctypes.byref(cip))
    rp = np.ctypeslib.as_array(ctypes.cast(rpp, POINTER(c_int32)), shape=(nr.value + 1,)).copy()
    ci = (np.ctypeslib.as_array(ctypes.cast(cip, POINTER(c_int32)), shape=(nnz.value,)).copy()
          if nnz.value else np.zeros(0, np.int32))
    L.ref_free(rpp)
    L.ref_free(cip)
    return nr.value, nc.value, rp, ci
