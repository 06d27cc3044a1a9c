"""C ABI checks that need no GPU: the library loads, exports every function
include/gbo.h declares, and its host-side entry points (generators, CSR
utilities, error reporting) behave; compute entry points are not called."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT


def test_exports_every_declared_symbol(gbo):
    declared = gbo.header_symbols(os.path.join(ROOT, "include", "gbo.h"))
    assert len(declared) > 40
    out = subprocess.run(["nm", "-D", "--defined-only", gbo.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [s for s in declared if s not in exported]
    assert not missing, missing
    L = gbo.lib()
    for s in declared:
        assert hasattr(L, s)


def test_no_oracle_dependency(gbo):
    """The engine never links or loads the oracle."""
    out = subprocess.run(["ldd", gbo.LIB_PATH], capture_output=True, text=True).stdout
    assert "oracle" not in out
    src = os.path.join(ROOT, "graph-bit-ops_amd")
    for dp, _, fs in os.walk(src):
        for f in fs:
            if f.endswith((".py", ".hip", ".cpp", ".hpp", ".h")):
                assert "pyoracle" not in open(os.path.join(dp, f), errors="ignore").read(), f


def test_version_and_errors(gbo):
    L = gbo.lib()
    assert L.gbo_version().startswith(b"gbo")
    nnz = ctypes.c_int64()
    rp, ci = ctypes.c_void_p(), ctypes.c_void_p()
    assert L.gbo_gen_er(-1, 0.1, 0, ctypes.byref(nnz), ctypes.byref(rp), ctypes.byref(ci)) == -1
    assert b"gbo_gen_er" in L.gbo_last_error()
    assert L.gbo_gen_rmat(12, 16, 0.9, 0.9, 0.1, 0, ctypes.byref(nnz), ctypes.byref(rp), ctypes.byref(ci)) == -1


def test_generator_properties(gbo):
    g = gbo.gen_er(1024, 0.02, 0)
    assert abs(g.nnz / 2 - 0.02 * 1024 * 1023 / 2) < 5 * np.sqrt(0.02 * 1024 * 1023 / 2)
    for c in (g, gbo.gen_rmat(12, 16, seed=1), gbo.gen_er_m(5000, 40000, 3)):
        r = np.repeat(np.arange(c.nrows), np.diff(c.rowptr))
        assert np.all(r != c.colind)                     # no self-loops
        for i in range(0, c.nrows, 97):                  # sorted, unique rows
            row = c.colind[c.rowptr[i]:c.rowptr[i + 1]]
            assert np.all(np.diff(row) > 0)
        key = np.sort(r.astype(np.int64) * c.nrows + c.colind)
        tkey = np.sort(c.colind.astype(np.int64) * c.nrows + r)
        assert np.array_equal(key, tkey)                 # symmetric
    # deterministic
    a, b = gbo.gen_rmat(12, 16, seed=5), gbo.gen_rmat(12, 16, seed=5)
    assert np.array_equal(a.colind, b.colind)


def test_tril_transpose(gbo):
    g = gbo.gen_er(500, 0.03, 1)
    L = g.tril(strict=True)
    r = np.repeat(np.arange(500), np.diff(L.rowptr))
    assert np.all(L.colind < r) and L.nnz * 2 == g.nnz
    T = L.transpose()
    rt = np.repeat(np.arange(500), np.diff(T.rowptr))
    assert np.array_equal(np.sort(rt.astype(np.int64) * 500 + T.colind), np.sort(L.colind.astype(np.int64) * 500 + r))


@pytest.mark.parametrize("bs", (3, 128))
def test_unsupported_block_size(gbo, bs):
    g = gbo.gen_er(100, 0.05, 0)
    nb = ctypes.c_int64()
    rc = gbo.lib().gbo_csr2bsr_nnzb(bs, g.rowptr.ctypes.data, g.colind.ctypes.data, 100, 100, ctypes.byref(nb))
    assert rc == -8
