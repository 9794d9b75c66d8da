"""Edge-file converter (mgpm_graph_read_edges, bin/convert): MatrixMarket and
GraphChallenge-style edge lists into the reference's CSR files, checked
against a CSR built directly from the same edges (CPU only)."""
import random
import subprocess
from pathlib import Path

import numpy as np
import pytest

import mercury_amd as ma
from conftest import csr_from_edges

BIN = Path(__file__).resolve().parent.parent / "mercury_amd" / "bin"


def _edges(seed, n=60, m=300):
    rng = random.Random(seed)
    e = [(rng.randrange(n), rng.randrange(n)) for _ in range(m)]
    e += [(3, 3), (5, 7), (7, 5), (5, 7)]  # self-loop and duplicates
    return n, e


def _write_mtx(path, n, edges, base=1, symmetric=True):
    kind = "symmetric" if symmetric else "general"
    lines = [f"%%MatrixMarket matrix coordinate pattern {kind}",
             "% a comment", f"{n} {n} {len(edges)}"]
    lines += [f"{a + base} {b + base}" for a, b in edges]
    path.write_text("\n".join(lines) + "\n")


def _write_tsv(path, edges, base=1):
    path.write_text("".join(f"{a + base}\t{b + base}\t1\n" for a, b in edges))


def _csr(g):
    rp, ci = g.csr()
    return rp.copy(), ci.copy()


def test_mtx_and_tsv_equal_direct_csr(tmp_path):
    n, e = _edges(1)
    want_rp, want_ci = csr_from_edges(n, e)
    _write_mtx(tmp_path / "g.mtx", n, e)
    rp, ci = _csr(ma.Graph.read_edges(str(tmp_path / "g.mtx")))
    assert np.array_equal(rp, want_rp) and np.array_equal(ci, want_ci)
    # the TSV route names vertices by their largest id: drop isolated tail
    _write_tsv(tmp_path / "g.tsv", e)
    g = ma.Graph.read_edges(str(tmp_path / "g.tsv"))
    rp2, ci2 = _csr(g)
    k = g.n
    assert np.array_equal(rp2, want_rp[:k + 1]) and np.array_equal(ci2, want_ci)


def test_zero_based_lists_and_general_mtx(tmp_path):
    n, e = _edges(2)
    want_rp, want_ci = csr_from_edges(n, e)
    _write_mtx(tmp_path / "g0.mtx", n, e, base=0, symmetric=False)
    rp, ci = _csr(ma.Graph.read_edges(str(tmp_path / "g0.mtx"), base=0))
    assert np.array_equal(rp, want_rp) and np.array_equal(ci, want_ci)
    (tmp_path / "g0.txt").write_text(
        "# comment\n" + "".join(f"{a} {b}\n" for a, b in e))
    rp, ci = _csr(ma.Graph.read_edges(str(tmp_path / "g0.txt"), base=0))
    assert np.array_equal(ci, want_ci)


def test_cli_writes_loadable_reference_files(tmp_path):
    n, e = _edges(3)
    _write_tsv(tmp_path / "in.tsv", e)
    out = tmp_path / "graph"
    r = subprocess.run([str(BIN / "convert"), str(tmp_path / "in.tsv"),
                        str(out)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert "vertices" in r.stdout
    g = ma.Graph.load(str(out))
    ref = ma.Graph.read_edges(str(tmp_path / "in.tsv"))
    a, b = _csr(g), _csr(ref)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    meta = (tmp_path / "graph.meta.txt").read_text().split()
    assert meta[2:6] == ["4", "8", "1", "2"]


def test_errors(tmp_path):
    with pytest.raises(ma.MgpmError) as ei:
        ma.Graph.read_edges(str(tmp_path / "missing.mtx"))
    assert ei.value.code == -1
    (tmp_path / "bad.mtx").write_text("not a banner\n1 1 0\n")
    with pytest.raises(ma.MgpmError) as ei:
        ma.Graph.read_edges(str(tmp_path / "bad.mtx"))
    assert ei.value.code == -2
    (tmp_path / "range.mtx").write_text(
        "%%MatrixMarket matrix coordinate pattern symmetric\n3 3 1\n1 9\n")
    with pytest.raises(ma.MgpmError) as ei:
        ma.Graph.read_edges(str(tmp_path / "range.mtx"))
    assert ei.value.code == -2
    (tmp_path / "count.mtx").write_text(
        "%%MatrixMarket matrix coordinate pattern symmetric\n3 3 2\n1 2\n")
    with pytest.raises(ma.MgpmError) as ei:
        ma.Graph.read_edges(str(tmp_path / "count.mtx"))
    assert ei.value.code == -2
    (tmp_path / "zero.tsv").write_text("0\t1\n")
    with pytest.raises(ma.MgpmError) as ei:  # 1-based by default
        ma.Graph.read_edges(str(tmp_path / "zero.tsv"))
    assert ei.value.code == -2
    (tmp_path / "junk.tsv").write_text("1\tx\n")
    with pytest.raises(ma.MgpmError) as ei:
        ma.Graph.read_edges(str(tmp_path / "junk.tsv"))
    assert ei.value.code == -2
    r = subprocess.run([str(BIN / "convert")], capture_output=True, text=True)
    assert r.returncode == 1 and "Usage:" in r.stdout
