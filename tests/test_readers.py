"""Matrix Market readers (host C++, gbo_mtx_read) against the reference's
own readers (golden CSR written by tests/golden/make_golden.py from the
reference sources compiled in oracle/_ref) and the Python restatement in
oracle/mtx_oracle.py; error handling of malformed input."""
import os

import numpy as np
import pytest

import mtx_oracle
from conftest import GOLDEN

FILES = ["sym64.mtx", "gen64.mtx", "dups.mtx", "real.mtx", "int.mtx", "lower1000.mtx"]


def reference_is_defined(path, directed):
    """readMtx.hpp:196-209 compacts marked entries with a scan that reads one
    element past the end when the sorted list ends in removed entries (a
    self-loop or repeat); the reference's output then depends on memory
    contents.  Those cases are compared with the restatement only."""
    m, n, symmetry, ent = mtx_oracle.parse(path)
    und = (symmetry == "symmetric" or directed == 2) and directed != 1
    pairs = list(ent) + ([(c, r) for r, c in ent if r != c] if und else [])
    pairs.sort()
    last = pairs[-1]
    return not (last[0] == last[1] or (len(pairs) > 1 and pairs[-2] == last))


@pytest.fixture(scope="module")
def expected():
    return np.load(os.path.join(GOLDEN, "readers_expected.npz"))


@pytest.mark.parametrize("name", FILES)
def test_mmio_semantics(gbo, expected, name):
    p = os.path.join(GOLDEN, name)
    c = gbo.mtx_read(p, 0, 0)
    assert np.array_equal(c.rowptr, expected[f"{name}|0|0|rowptr"])
    assert np.array_equal(c.colind, expected[f"{name}|0|0|colind"])
    m, n, rp, ci = mtx_oracle.read(p, 0)
    assert (c.nrows, c.ncols) == (m, n)
    assert np.array_equal(c.rowptr, rp) and np.array_equal(c.colind, ci)


@pytest.mark.parametrize("name", FILES)
@pytest.mark.parametrize("directed", (0, 1, 2))
def test_readmtx_semantics(gbo, expected, name, directed):
    if name == "int.mtx" and directed == 2:
        pytest.skip("non-square matrix cannot be symmetrized")
    p = os.path.join(GOLDEN, name)
    c = gbo.mtx_read(p, 1, directed)
    m, n, rp, ci = mtx_oracle.read(p, 1, directed)
    assert np.array_equal(c.rowptr, rp) and np.array_equal(c.colind, ci)
    if reference_is_defined(p, directed):
        assert np.array_equal(c.rowptr, expected[f"{name}|1|{directed}|rowptr"])
        assert np.array_equal(c.colind, expected[f"{name}|1|{directed}|colind"])


def test_selfloop_env(gbo, monkeypatch):
    p = os.path.join(GOLDEN, "dups.mtx")
    monkeypatch.setenv("GRB_UTIL_REMOVE_SELFLOOP", "0")
    c = gbo.mtx_read(p, 1, 0)
    m, n, rp, ci = mtx_oracle.read(p, 1, 0)
    assert np.array_equal(c.colind, ci)
    assert any(ci[k] == r for r in range(m) for k in range(rp[r], rp[r + 1]))


def test_values(gbo):
    c = gbo.mtx_read(os.path.join(GOLDEN, "real.mtx"), 0, 0, values=True)
    assert c.values is not None and len(c.values) == c.nnz
    np.testing.assert_allclose(sorted(set(c.values.tolist())), sorted([0.5, -1.25, 3.0, 2.0, 1e-3]), rtol=1e-6)


@pytest.mark.parametrize("text,code", [
    ("%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n", -5),
    ("not a banner\n", -5),
    ("%%MatrixMarket matrix coordinate pattern general\n3 3 4\n1 1\n2 2\n", -5),  # too few entries
    ("%%MatrixMarket matrix coordinate pattern general\n3 3 1\n4 1\n", -5),       # out of bounds
    ("%%MatrixMarket matrix coordinate pattern general\n3 x 1\n1 1\n", -5),       # bad size line
])
def test_malformed(gbo, tmp_path, text, code):
    p = tmp_path / "bad.mtx"
    p.write_text(text)
    with pytest.raises(gbo.GboError) as e:
        gbo.mtx_read(str(p), 1, 0)
    assert e.value.code == code


def test_missing_file(gbo):
    with pytest.raises(gbo.GboError) as e:
        gbo.mtx_read("/nonexistent/x.mtx", 0, 0)
    assert e.value.code == -4


def test_write_roundtrip(gbo, tmp_path):
    g = gbo.gen_er(300, 0.05, 2)
    p = str(tmp_path / "er.mtx")
    gbo.mtx_write(p, g)
    c = gbo.mtx_read(p, 0, 0)
    assert np.array_equal(c.rowptr, g.rowptr) and np.array_equal(c.colind, g.colind)
