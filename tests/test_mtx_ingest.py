"""Parallel Matrix Market ingest (gbo_mtx_read), its binary CSR cache
(gbo_mtx_read_cached) and the parallel writer.

- Randomized files with every layout a sequential fscanf reader accepts
  (entries split over lines, several entries on a line, tabs, CRLF, '+'
  signs, tokens past the declared count, comment blocks) and every field and
  symmetry, against the Python restatement of both reader semantics
  (oracle/mtx_oracle.py, pinned on the reference's own readers by
  test_readers.py) and, where oracle/_ref holds the reference readers built
  from their sources, against those as well.
- Cache: readMtx's own cache name, hit / rewrite / invalidation on source
  change, on a changed self-loop rule and on a corrupt cache, values kept.
- Writer: byte-identical to a plain per-entry formatter; read back equal.
"""
import os

import numpy as np
import pytest

import mtx_oracle

FIELDS = ("pattern", "real", "integer", "complex")
SYMS = ("general", "symmetric")


def _value_tok(rng, field):
    if field == "integer":
        return str(int(rng.integers(-50, 50)))
    # short decimals and dyadic values: a float32 conversion through double
    # and a direct one agree on them
    return rng.choice(["0.5", "-1.25", "3", "2.0", "1e-3", "7.75", "-0.125", "1.5e2"])


def _write_random(path, rng, m, n, nnz, field, sym, layout):
    """A random coordinate file; returns nothing (the file is the input)."""
    toks = []
    for _ in range(nnz):
        if sym == "symmetric":  # lower triangle and diagonal, duplicates allowed
            r = int(rng.integers(1, m + 1))
            c = int(rng.integers(1, r + 1))
        else:
            r, c = int(rng.integers(1, m + 1)), int(rng.integers(1, n + 1))
        e = [str(r), str(c)]
        if field in ("real", "integer"):
            e.append(_value_tok(rng, field))
        elif field == "complex":
            e += [_value_tok(rng, "real"), _value_tok(rng, "real")]
        toks.append(e)
    lines = []
    nl = "\r\n" if layout == "crlf" else "\n"
    if layout in ("plain", "crlf"):
        lines = [" ".join(e) for e in toks]
    elif layout == "messy":
        flat = [t for e in toks for t in e]
        i = 0
        while i < len(flat):
            k = int(rng.integers(1, 6))
            sep = rng.choice([" ", "\t", "  ", " \t "])
            words = [("+" + w if w.isdigit() and rng.random() < 0.1 else w) for w in flat[i:i + k]]
            lines.append(sep.join(words) + rng.choice(["", " ", "\t"]))
            i += k
        lines.append("99 99 99 extra tokens never read")
    with open(path, "w", newline="") as f:
        f.write(f"%%MatrixMarket matrix coordinate {field} {sym}{nl}")
        f.write(f"% a comment{nl}%{nl}   % indented comment{nl}")
        f.write(f"{m} {n} {nnz}{nl}")
        f.write(nl.join(lines) + nl)


def _ref_read(po, path, semantics, directed):
    """CSR from the reference's own reader (oracle/_ref), or None when the
    reference's result is undefined for this file (readMtx's compaction
    reading past its vector, see test_readers.reference_is_defined)."""
    from test_readers import reference_is_defined
    if semantics == 0:
        return po.ref_mmio(path)[2:]
    if not reference_is_defined(path, directed):
        return None
    return po.ref_readmtx(path, directed)[2:]


def _oracle_values(path, semantics, directed):
    """Values in CSR order per the restated semantics (first of a run of
    equal keys keeps its value under readMtx)."""
    with open(path) as f:
        banner = f.readline().split()
        field, sym = banner[3].lower(), banner[4].lower()
        line = f.readline()
        while line.lstrip().startswith("%") or not line.strip():
            line = f.readline()
        m, n, nnz = (int(t) for t in line.split())
        toks = f.read().split()
    per = {"pattern": 2, "real": 3, "integer": 3, "complex": 4}[field]
    ent = []
    for i in range(nnz):
        r, c = int(toks[i * per]) - 1, int(toks[i * per + 1]) - 1
        v = np.float32(float(toks[i * per + 2])) if per > 2 else np.float32(1.0)
        ent.append((r, c, v))
    rows = [[] for _ in range(m)]
    if semantics == 0:
        mirror = sym in ("symmetric", "hermitian")
        for r, c, v in ent:
            rows[r].append(v)
            if mirror and r != c:
                rows[c].append(v)
        return np.array([v for row in rows for v in row], np.float32)
    und = (sym == "symmetric" or directed == 2) and directed != 1
    seq = list(ent) + ([(c, r, v) for r, c, v in ent if r != c] if und else [])
    drop = int(os.environ.get("GRB_UTIL_REMOVE_SELFLOOP", "1") or 0) != 0
    first = {}
    for r, c, v in seq:
        if (r, c) not in first:
            first[(r, c)] = v
    keys = sorted(k for k in first if not (drop and k[0] == k[1]))
    return np.array([first[k] for k in keys], np.float32)


CASES = [(f, s, lay, seed) for seed, (f, s, lay) in enumerate(
    [(f, s, lay) for f in FIELDS for s in SYMS for lay in ("plain", "messy", "crlf")])]


@pytest.mark.parametrize("field,sym,layout,seed", CASES)
def test_random_files_match_restatement(gbo, po, tmp_path, field, sym, layout, seed):
    rng = np.random.default_rng(100 + seed)
    m = int(rng.integers(20, 300))
    n = m if sym == "symmetric" else int(rng.integers(20, 300))
    nnz = int(rng.integers(1, 3000))
    p = str(tmp_path / "r.mtx")
    _write_random(p, rng, m, n, nnz, field, sym, layout)
    ref = po.ref_available()
    for semantics, directed in ((0, 0), (1, 0), (1, 1), (1, 2)):
        if semantics == 1 and directed == 2 and m != n:
            continue
        c = gbo.mtx_read(p, semantics, directed, values=True)
        om, on, orp, oci = mtx_oracle.read(p, semantics, directed)
        assert (c.nrows, c.ncols) == (om, on)
        assert np.array_equal(c.rowptr, orp) and np.array_equal(c.colind, oci), (semantics, directed)
        if field != "complex":
            assert np.array_equal(c.values, _oracle_values(p, semantics, directed)), (semantics, directed)
        c2 = gbo.mtx_read(p, semantics, directed)  # no values: same structure
        assert np.array_equal(c2.rowptr, c.rowptr) and np.array_equal(c2.colind, c.colind)
        if ref and layout != "messy" and field != "complex":
            want = _ref_read(po, p, semantics, directed)
            if want is not None:
                assert np.array_equal(c.rowptr, want[0]) and np.array_equal(c.colind, want[1]), (semantics, directed)


def test_selfloop_rule_in_parallel_path(gbo, tmp_path, monkeypatch):
    rng = np.random.default_rng(7)
    p = str(tmp_path / "l.mtx")
    _write_random(p, rng, 200, 200, 5000, "real", "general", "plain")
    monkeypatch.setenv("GRB_UTIL_REMOVE_SELFLOOP", "0")
    c = gbo.mtx_read(p, 1, 2, values=True)
    om, on, orp, oci = mtx_oracle.read(p, 1, 2)
    assert np.array_equal(c.rowptr, orp) and np.array_equal(c.colind, oci)
    assert np.array_equal(c.values, _oracle_values(p, 1, 2))


@pytest.mark.parametrize("text,msg", [
    ("%%MatrixMarket matrix coordinate pattern general\n3 3 2\n1 1\n2 x\n", "malformed"),
    ("%%MatrixMarket matrix coordinate real general\n3 3 2\n1 1 0.5\n2 2 abc\n", "malformed"),
    ("%%MatrixMarket matrix coordinate pattern general\n3 3 2\n1 1\n2 4\n", "out of bounds"),
    ("%%MatrixMarket matrix coordinate pattern general\n3 3 2\n0 1\n2 2\n", "out of bounds"),
    ("%%MatrixMarket matrix coordinate pattern general\n3 3 3\n1 1 2\n2\n", "not enough"),
    ("%%MatrixMarket matrix coordinate pattern symmetric\n2 3 1\n1 3\n", "not square"),
])
def test_errors_name_the_first_bad_entry(gbo, tmp_path, text, msg):
    p = tmp_path / "bad.mtx"
    p.write_text(text)
    for semantics in (0, 1):
        with pytest.raises(gbo.GboError) as e:
            gbo.mtx_read(str(p), semantics, 0)
        assert e.value.code == -5 and msg in str(e.value)


def test_lowest_failing_entry_wins(gbo, tmp_path):
    """Two bad entries far apart (different parse ranges): the first is named."""
    lines = [f"{i % 1000 + 1} {(i * 7) % 1000 + 1}" for i in range(200000)]
    lines[150000] = "1 5000"
    lines[190000] = "zz 1"
    p = tmp_path / "two.mtx"
    p.write_text("%%MatrixMarket matrix coordinate pattern general\n1000 1000 200000\n" + "\n".join(lines) + "\n")
    with pytest.raises(gbo.GboError) as e:
        gbo.mtx_read(str(p), 0, 0)
    assert "entry 150000 out of bounds" in str(e.value)


@pytest.mark.parametrize("semantics", (0, 1))
def test_rmat_roundtrip(gbo, tmp_path, semantics):
    g = gbo.gen_rmat(14, 16, 0.57, 0.19, 0.19, seed=3)
    p = str(tmp_path / "rmat14.mtx")
    gbo.mtx_write(p, g)
    c = gbo.mtx_read(p, semantics, 0)
    assert np.array_equal(c.rowptr, g.rowptr) and np.array_equal(c.colind, g.colind)


def test_writer_bytes(gbo, tmp_path):
    g = gbo.gen_er(500, 0.03, 5)
    p = tmp_path / "w.mtx"
    gbo.mtx_write(str(p), g)
    want = [f"%%MatrixMarket matrix coordinate pattern general", f"{g.nrows} {g.ncols} {g.nnz}"]
    for r in range(g.nrows):
        for k in range(g.rowptr[r], g.rowptr[r + 1]):
            want.append(f"{r + 1} {g.colind[k] + 1}")
    assert p.read_text() == "\n".join(want) + "\n"


# ------------------------------------------------------------------ cache
def _sym_file(tmp_path, rng, name="g.mtx", field="real"):
    p = str(tmp_path / name)
    _write_random(p, rng, 400, 400, 6000, field, "symmetric", "plain")
    return p


def test_cache_name_is_readmtx_convention(gbo, tmp_path, monkeypatch):
    p = _sym_file(tmp_path, np.random.default_rng(1))
    d = str(tmp_path)
    assert gbo.mtx_cache_path(p, 1, 0) == f"{d}/.g.mtx.ud.nosl.bin"
    assert gbo.mtx_cache_path(p, 1, 1) == f"{d}/.g.mtx.d.nosl.bin"
    assert gbo.mtx_cache_path(p, 0, 0) == f"{d}/.g.mtx.mmio.bin"
    monkeypatch.setenv("GRB_UTIL_REMOVE_SELFLOOP", "0")
    assert gbo.mtx_cache_path(p, 1, 0) == f"{d}/.g.mtx.ud.sl.bin"


@pytest.mark.parametrize("semantics,directed", [(0, 0), (1, 0), (1, 1), (1, 2)])
def test_cache_roundtrip(gbo, tmp_path, semantics, directed):
    p = _sym_file(tmp_path, np.random.default_rng(2))
    want = gbo.mtx_read(p, semantics, directed, values=True)
    a = gbo.mtx_read(p, semantics, directed, values=True, cache=True)
    assert a.cache_state == 0 and os.path.exists(gbo.mtx_cache_path(p, semantics, directed))
    b = gbo.mtx_read(p, semantics, directed, values=True, cache=True)
    assert b.cache_state == 1
    for x in (a, b):
        assert (x.nrows, x.ncols) == (want.nrows, want.ncols)
        assert np.array_equal(x.rowptr, want.rowptr) and np.array_equal(x.colind, want.colind)
        assert np.array_equal(x.values, want.values)
    c = gbo.mtx_read(p, semantics, directed, cache=True)  # structure only, from the cache
    assert c.cache_state == 1 and c.values is None and np.array_equal(c.colind, want.colind)


def test_cache_pattern_values_are_ones(gbo, tmp_path):
    p = _sym_file(tmp_path, np.random.default_rng(3), field="pattern")
    gbo.mtx_read(p, 1, 0, cache=True)  # written without values
    b = gbo.mtx_read(p, 1, 0, values=True, cache=True)
    assert b.cache_state == 1 and np.all(b.values == 1.0) and len(b.values) == b.nnz


def test_cache_invalidation(gbo, tmp_path, monkeypatch):
    rng = np.random.default_rng(4)
    p = _sym_file(tmp_path, rng)
    assert gbo.mtx_read(p, 1, 0, cache=True).cache_state == 0
    # a changed source (new content, new size and mtime) is parsed again
    _write_random(p, rng, 300, 300, 4000, "real", "symmetric", "plain")
    os.utime(p, ns=(1, 1_000_000_123))
    c = gbo.mtx_read(p, 1, 0, cache=True)
    assert c.cache_state == 0
    want = gbo.mtx_read(p, 1, 0)
    assert np.array_equal(c.rowptr, want.rowptr) and np.array_equal(c.colind, want.colind)
    # same size, other mtime: parsed again
    os.utime(p, ns=(1, 2_000_000_456))
    assert gbo.mtx_read(p, 1, 0, cache=True).cache_state == 0
    assert gbo.mtx_read(p, 1, 0, cache=True).cache_state == 1
    # the self-loop rule selects another cache file
    monkeypatch.setenv("GRB_UTIL_REMOVE_SELFLOOP", "0")
    c = gbo.mtx_read(p, 1, 0, cache=True)
    assert c.cache_state == 0
    o = mtx_oracle.read(p, 1, 0)
    assert np.array_equal(c.colind, o[3])
    # a corrupt (truncated) cache is ignored and rewritten
    name = gbo.mtx_cache_path(p, 1, 0)
    with open(name, "r+b") as f:
        f.truncate(os.path.getsize(name) - 4)
    c = gbo.mtx_read(p, 1, 0, cache=True)
    assert c.cache_state == 0 and np.array_equal(c.colind, o[3])
    assert gbo.mtx_read(p, 1, 0, cache=True).cache_state == 1


@pytest.mark.parametrize("damage", ["column", "rowptr"])
def test_cache_damaged_content_is_parsed_again(gbo, tmp_path, damage):
    """A cache of the right key and size whose arrays do not form a CSR (a
    column index out of range, a decreasing rowptr) is not handed out."""
    p = _sym_file(tmp_path, np.random.default_rng(6))
    want = gbo.mtx_read(p, 1, 0)
    assert gbo.mtx_read(p, 1, 0, cache=True).cache_state == 0
    name = gbo.mtx_cache_path(p, 1, 0)
    hdr = 72  # magic, version, 3 flags, source size, mtime (s, ns), m, n, nnz, has_val, reserved
    raw = bytearray(open(name, "rb").read())
    has_val = int(np.frombuffer(bytes(raw[64:68]), np.uint32)[0])
    assert len(raw) == hdr + 4 * (want.nrows + 1) + 4 * want.nnz * (2 if has_val else 1)
    if damage == "column":
        at = hdr + 4 * (want.nrows + 1) + 4 * (want.nnz // 2)
        raw[at:at + 4] = np.int32(want.ncols).tobytes()
    else:
        r = int(np.argmax(np.diff(want.rowptr) > 0)) + 1  # rowptr[r] > rowptr[r - 1]
        raw[hdr + 4 * r:hdr + 4 * r + 4] = np.int32(want.nnz).tobytes()
    open(name, "r+b").write(bytes(raw))  # same size, same key
    c = gbo.mtx_read(p, 1, 0, cache=True)
    assert c.cache_state == 0
    assert np.array_equal(c.rowptr, want.rowptr) and np.array_equal(c.colind, want.colind)
    assert gbo.mtx_read(p, 1, 0, cache=True).cache_state == 1


def test_cache_not_writable(gbo, tmp_path):
    p = _sym_file(tmp_path, np.random.default_rng(5))
    os.mkdir(gbo.mtx_cache_path(p, 1, 0))  # a directory in the cache's place
    c = gbo.mtx_read(p, 1, 0, cache=True)
    assert c.cache_state == 2
    want = gbo.mtx_read(p, 1, 0)
    assert np.array_equal(c.colind, want.colind)
    assert not [f for f in os.listdir(tmp_path) if ".tmp" in f]  # no temporary left behind


def test_short_file_with_huge_count_is_refused_cheaply(gbo, tmp_path):
    """A header that declares far more entries than the file holds is refused
    from the token count, before any entry array is sized by it."""
    p = tmp_path / "short.mtx"
    p.write_text("%%MatrixMarket matrix coordinate pattern general\n5 5 2000000000\n1 1\n2 3\n")
    with pytest.raises(gbo.GboError) as e:
        gbo.mtx_read(str(p), 1, 0)
    assert e.value.code == -5 and "not enough entries" in str(e.value)
    p.write_text("%%MatrixMarket matrix coordinate pattern general\n5 5 2000000000\n1 1\n9 3\n")
    with pytest.raises(gbo.GboError) as e:
        gbo.mtx_read(str(p), 0, 0)
    assert "entry 1 out of bounds" in str(e.value)
