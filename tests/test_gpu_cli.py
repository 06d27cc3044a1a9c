"""The .mtx-in command lines (graph-bit-ops_amd/bin/<name>): reference-named
binaries, reference stdout keys, results equal to the oracle."""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

pytestmark = pytest.mark.gpu
BIN = os.path.join(ROOT, "graph-bit-ops_amd", "bin")


def run(name, *args):
    p = subprocess.run([os.path.join(BIN, name), *args], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    return p.stdout


@pytest.mark.parametrize("bs", (8, 16, 32, 64))
def test_bsrbmv_cli(bs):
    out = run(f"bsrbmv{bs}", os.path.join(GOLDEN, "lower1000.mtx"))
    assert f"blocksize: {bs}," in out
    assert f"BSR BMV-{bs} success: 1" in out
    assert re.search(rf"BSR BMV-{bs}: [0-9.]+", out)


def test_bfs_cli_levels(gbo, po, tmp_path):
    path = os.path.join(GOLDEN, "lower1000.mtx")
    outf = str(tmp_path / "lv.txt")
    out = run("bfs32", path, "0", "--out", outf)
    g = gbo.mtx_read(path, 1, 0)
    want = po.bfs(g.nrows, g.rowptr, g.colind, 0)
    assert np.array_equal(np.loadtxt(outf, dtype=np.int64), want)
    succ = [int(v) for v in re.findall(r"succ: (\d+)", out)]
    assert succ[:-1] == list(np.bincount(want[want >= 0])[1:]) and succ[-1] == 0


def test_cc_tc_bsrbmm_cli(gbo, po):
    path = os.path.join(GOLDEN, "lower1000.mtx")
    g = gbo.mtx_read(path, 1, 0)
    out = run("cc16", path, "0")
    assert "CORRECT" in out
    lab = po.cc(g.nrows, g.rowptr, g.colind)
    assert f"found with {len(set(lab.tolist()))} components" in out
    out = run("tc8", path, "0")
    assert f"ntris_bmm: {po.tc(g.nrows, g.rowptr, g.colind)}" in out
    out = run("bsrbmm64", path)
    a, b = re.search(r"ntris_bmm: (\d+)", out).group(1), re.search(r"ntris_csr: (\d+)", out).group(1)
    assert a == b


def test_sssp_pr_cli(gbo, po, tmp_path):
    path = os.path.join(GOLDEN, "lower1000.mtx")
    g = gbo.mtx_read(path, 0, 0)
    outf = str(tmp_path / "d.txt")
    run("sssp8", path, "--out", outf)
    lv = po.bfs(g.nrows, g.rowptr, g.colind, 0)
    want = np.where(lv < 0, 0xFFFFFFFF, lv)
    assert np.array_equal(np.loadtxt(outf, dtype=np.int64), want)
    outf = str(tmp_path / "p.txt")
    out = run("pr32", path, "--max-iter", "500", "--tol", "1e-9", "--out", outf)
    p = np.loadtxt(outf)
    ref, _ = po.pagerank(g.nrows, g.rowptr, g.colind, g.outdeg(), 0.85, 1e-14, 100000)
    assert np.abs(p - ref).sum() <= 1e-6
    assert "niter:" in out


def test_cli_usage_errors():
    p = subprocess.run([os.path.join(BIN, "bsrbmv32")], capture_output=True, text=True)
    assert p.returncode == 1 and "./exe [xxx.mtx]" in p.stdout
    p = subprocess.run([os.path.join(BIN, "bfs32"), "/nonexistent.mtx", "0"], capture_output=True, text=True)
    assert p.returncode != 0 and "not found" in p.stderr


def test_cli_csr_cache(gbo, po, tmp_path):
    """--cache: the first run parses and writes readMtx's cache file, the
    second loads it; both give the oracle's BFS levels."""
    import shutil
    path = str(tmp_path / "g.mtx")
    shutil.copy(os.path.join(GOLDEN, "lower1000.mtx"), path)
    g = gbo.mtx_read(path, 1, 0)
    want = po.bfs(g.nrows, g.rowptr, g.colind, 0)
    for state in ("written", "loaded"):
        outf = str(tmp_path / f"lv_{state}.txt")
        out = run("bfs16", path, "0", "--cache", "--out", outf)
        assert f"csr cache: {tmp_path}/.g.mtx.ud.nosl.bin ({state})" in out
        assert np.array_equal(np.loadtxt(outf, dtype=np.int64), want)
