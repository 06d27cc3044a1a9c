"""gbo_bfs_noalloc: the pull BFS with every scratch buffer in the library's
static device array (no allocation in the call), against gbo_bfs.

Two calls in a row on one GPU (the second with a smaller graph, so stale
contents of the shared array from the first call would show), each at every
block size, levels and per-level sizes equal to gbo_bfs's; a graph too large
for the array and an unplanned matrix are refused before any launch.
"""
import ctypes
from ctypes import POINTER, byref, c_int, c_int32, c_int64, c_void_p

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(gbo):
    L = gbo.lib()
    fn = L.gbo_bfs_noalloc
    fn.argtypes = [POINTER(gbo.BsrStruct), c_int32, c_void_p, c_void_p, c_int32, POINTER(c_int32), c_void_p]
    fn.restype = c_int
    return fn


def _run_noalloc(gbo, A, src, maxl=4096):
    n = A.nrows
    d = gbo.DeviceBuffer(4 * n)
    sizes = np.zeros(maxl, np.int64)
    nl = c_int32()
    gbo._check(_noalloc(gbo)(A.ref, src, d.ptr, sizes.ctypes.data, maxl, byref(nl), None))
    lv = d.to_numpy(np.int32, n)
    d.free()
    return lv, sizes[:nl.value].copy()


@pytest.mark.parametrize("bs", (8, 16, 32, 64))
def test_noalloc_matches_gbo_bfs(gbo, bs):
    big = gbo.gen_rmat(14, 16, 0.57, 0.19, 0.19, seed=1)
    small = gbo.gen_rmat(10, 4, 0.57, 0.19, 0.19, seed=7)
    for g in (big, small):  # second call: smaller input over the same array
        A = gbo.BsrMatrix.from_csr(g, bs)
        src = int(np.argmax(g.outdeg() > 0))
        want, want_sizes = gbo.bfs(A, src)
        lv, sizes = _run_noalloc(gbo, A, src)
        assert np.array_equal(lv, want), (bs, g.nrows)
        assert np.array_equal(sizes, want_sizes), (bs, g.nrows)
        A.free()


@pytest.mark.parametrize("bs", (8, 16))
def test_noalloc_without_dense_counts(gbo, bs):
    # 8M vertices, one edge into vertex 0 from 8M-1 (symmetric): the frontier
    # words fit the array, the dense step's row counts (32 MiB) do not, so
    # every level takes the visited-row-skipping step; levels as gbo_bfs
    n = 1 << 23
    rowptr = np.zeros(n + 1, np.int32)
    rowptr[1:] = 1
    rowptr[n] = 2
    g = gbo.Csr(n, n, rowptr, np.array([n - 1, 0], np.int32))
    A = gbo.BsrMatrix.from_csr(g, bs)
    want, want_sizes = gbo.bfs(A, 0)
    lv, sizes = _run_noalloc(gbo, A, 0)
    assert np.array_equal(lv, want) and np.array_equal(sizes, want_sizes)
    assert lv[0] == 0 and lv[n - 1] == 1 and int((lv >= 0).sum()) == 2
    A.free()


def test_noalloc_refuses(gbo):
    # 100M vertices, no edges: the frontier / next / visited words (37.5 MB
    # at bs 64) exceed the array
    n = 100_000_000
    g = gbo.Csr(n, n, np.zeros(n + 1, np.int32), np.zeros(0, np.int32))
    A = gbo.BsrMatrix.from_csr(g, 64)
    with pytest.raises(gbo.GboError) as e:
        _run_noalloc(gbo, A, 0)
    assert e.value.code == -6 and "static array" in str(e.value)
    A.free()
    # an unplanned matrix would need a temporary plan (an allocation)
    B = gbo.BsrMatrix.from_csr(gbo.gen_er(200, 0.05, 1), 16, plan=False)
    with pytest.raises(gbo.GboError) as e:
        _run_noalloc(gbo, B, 0)
    assert e.value.code == -1 and "plan" in str(e.value)
    B.free()
