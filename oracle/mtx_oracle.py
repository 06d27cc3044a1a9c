"""Pure-Python restatement of the reference's two Matrix Market readers.

TEST INFRASTRUCTURE ONLY (small files).  Pinned against the reference's own
readers compiled in oracle/_ref (tests/golden/readers_expected.npz).

semantics 0 -- mmio_info/mmio_data (algorithm/mmio_highlevel.h:7-309):
  1-based -> 0-based; symmetric or hermitian banners add the mirrored entry
  (r != c) right after its source; duplicates and self-loops kept; entries
  of a row in file order.
semantics 1 -- readMtx + coo2csr (algorithm/readMtx.hpp:352-421, 151-216,
  247-304): undirected iff (banner symmetric and directed != 1) or directed
  == 2; undirected adds (c, r) for every r != c; sort by (row, col); drop
  repeats; drop self-loops unless GRB_UTIL_REMOVE_SELFLOOP is "0".
"""
import os


def parse(path):
    with open(path) as f:
        banner = f.readline().split()
        field, symmetry = banner[3].lower(), banner[4].lower()
        line = f.readline()
        while line.lstrip().startswith("%") or not line.strip():
            line = f.readline()
        m, n, nnz = (int(t) for t in line.split())
        toks = f.read().split()
    per = {"pattern": 2, "real": 3, "integer": 3, "complex": 4}[field]
    ent = [(int(toks[i * per]) - 1, int(toks[i * per + 1]) - 1) for i in range(nnz)]
    return m, n, symmetry, ent


def to_csr(m, rows):
    rp = [0]
    ci = []
    for r in range(m):
        ci.extend(rows[r])
        rp.append(len(ci))
    return rp, ci


def read(path, semantics, directed=0):
    m, n, symmetry, ent = parse(path)
    rows = [[] for _ in range(m)]
    if semantics == 0:
        sym = symmetry in ("symmetric", "hermitian")
        for r, c in ent:
            rows[r].append(c)
            if sym and r != c:
                rows[c].append(r)
        return (m, n) + to_csr(m, rows)
    undirected = symmetry == "symmetric" or directed == 2
    if directed == 1:
        undirected = False
    pairs = list(ent)
    if undirected:
        pairs += [(c, r) for r, c in ent if r != c]
    drop_loops = int(os.environ.get("GRB_UTIL_REMOVE_SELFLOOP", "1") or 0) != 0
    keep = sorted(set(p for p in pairs if not (drop_loops and p[0] == p[1])))
    for r, c in keep:
        rows[r].append(c)
    return (m, n) + to_csr(m, rows)
