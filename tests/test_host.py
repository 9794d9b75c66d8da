"""Host-side tests (CPU only): the C ABI library loads and exports every
declared symbol, the reference-format loader/writer and its error paths,
the generators, the product's DAG preprocessing versus the oracle and the
reference, and the CLIs' argument handling.  No GPU calls."""
import hashlib
import json
import os
import re
import shutil
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

import mercury_amd as ma
import oracle
from conftest import csr_from_edges

ROOT = Path(__file__).resolve().parent.parent
GOLDEN = ROOT / "tests" / "golden"
BIN = ROOT / "mercury_amd" / "bin"


def header_functions():
    text = (ROOT / "include" / "mgpm.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mgpm_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    names = header_functions()
    assert len(names) >= 19
    lib = ma.native.lib
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python mirror covers them all
    assert set(names) <= set(ma.native.EXPORTED)


def test_abi_version_and_strerror():
    assert ma.native.lib.mgpm_abi_version() == 1
    assert ma.native.strerror(-8) == "Not support."
    assert ma.native.strerror(0) == "ok"


def test_device_count_without_gpu_is_not_an_error():
    assert ma.device_count() >= 0


@pytest.mark.skipif(ma.device_count() > 0, reason="checks the no-GPU path")
def test_device_calls_without_gpu_fail_cleanly():
    g = ma.Graph.rmat(6, 4, seed=1)
    for make in (lambda: ma.DeviceGraph(g, 0), lambda: ma.Batch(0),
                 lambda: ma.Server(0)):
        with pytest.raises(ma.MgpmError) as ei:
            make()
        assert ei.value.code == -6  # MGPM_ERR_NODEVICE


# ----------------------------------------------------------------- loader

def _copy_fixture(tmp_path, name="gnp40"):
    d = tmp_path / name
    shutil.copytree(GOLDEN / name, d)
    return d


def test_load_golden_fixtures_roundtrip(tmp_path):
    golden = json.loads((GOLDEN / "golden.json").read_text())
    for name, meta in golden.items():
        g = ma.Graph.load(str(GOLDEN / name / "graph"))
        assert g.n == meta["n"] and g.nnz == meta["nnz"]
        rp, ci = g.csr()
        out = tmp_path / name
        out.mkdir()
        g.write(str(out / "graph"))
        for f in ("graph.vertex.bin", "graph.edge.bin"):
            assert (out / f).read_bytes() == (GOLDEN / name / f).read_bytes()
        g2 = ma.Graph.load(str(out / "graph"))
        rp2, ci2 = g2.csr()
        assert np.array_equal(rp, rp2) and np.array_equal(ci, ci2)


def _expect_error(prefix, code):
    with pytest.raises(ma.MgpmError) as ei:
        ma.Graph.load(str(prefix))
    assert ei.value.code == code, str(ei.value)


def test_load_missing_file_is_io_error(tmp_path):
    _expect_error(tmp_path / "nope" / "graph", -1)


def test_load_truncated_edge_file_is_io_error(tmp_path):
    d = _copy_fixture(tmp_path)
    data = (d / "graph.edge.bin").read_bytes()
    (d / "graph.edge.bin").write_bytes(data[:-8])
    _expect_error(d / "graph", -1)


def test_load_wrong_vid_size_is_format_error(tmp_path):
    d = _copy_fixture(tmp_path)
    f = (d / "graph.meta.txt").read_text().split()
    f[2] = "8"
    (d / "graph.meta.txt").write_text(" ".join(f))
    _expect_error(d / "graph", -2)


def test_load_bad_max_degree_is_format_error(tmp_path):
    # graph.cc:28 asserts 0 < max_degree < n
    d = _copy_fixture(tmp_path)
    f = (d / "graph.meta.txt").read_text().split()
    f[6] = "0"
    (d / "graph.meta.txt").write_text(" ".join(f))
    _expect_error(d / "graph", -2)


def test_load_short_meta_is_format_error(tmp_path):
    d = _copy_fixture(tmp_path)
    (d / "graph.meta.txt").write_text("40 390 4 8")
    _expect_error(d / "graph", -2)


def test_load_rejects_asymmetric_and_unsorted(tmp_path):
    d = _copy_fixture(tmp_path)
    ci = np.fromfile(d / "graph.edge.bin", dtype="<i4")
    rp = np.fromfile(d / "graph.vertex.bin", dtype="<i8")
    bad = ci.copy()
    row = slice(rp[0], rp[1])
    bad[row] = bad[row][::-1]  # unsorted row
    bad.tofile(d / "graph.edge.bin")
    _expect_error(d / "graph", -7)
    bad = ci.copy()
    bad[rp[1] - 1] = (bad[rp[1] - 1] + 1) % 40  # breaks symmetry / order
    bad.tofile(d / "graph.edge.bin")
    _expect_error(d / "graph", -7)


def test_from_csr_rejects_self_loops_and_duplicates():
    with pytest.raises(ma.MgpmError) as ei:
        ma.Graph.from_csr([0, 2, 3], [0, 1, 0])  # self-loop at 0
    assert ei.value.code == -7
    with pytest.raises(ma.MgpmError) as ei:
        ma.Graph.from_csr([0, 2, 4], [1, 1, 0, 0])  # duplicates
    assert ei.value.code == -7


def test_empty_graph_is_accepted_from_csr():
    g = ma.Graph.from_csr(np.zeros(4, dtype=np.int64), [])
    assert g.n == 3 and g.nnz == 0


# ------------------------------------------------------------- generators

def _digest(g):
    rp, ci = g.csr()
    h = hashlib.sha256()
    h.update(rp.tobytes())
    h.update(ci.tobytes())
    return h.hexdigest()


def test_rmat_deterministic_and_seeded():
    a = ma.Graph.rmat(12, 16, seed=1)
    b = ma.Graph.rmat(12, 16, seed=1)
    c = ma.Graph.rmat(12, 16, seed=2)
    assert _digest(a) == _digest(b) != _digest(c)
    rp, ci = a.csr()
    deg = np.diff(rp)
    assert rp[-1] == len(ci) and deg.max() < a.n
    # skew: RMAT without permutation puts the hub at vertex 0
    assert deg[0] == deg.max()


def test_rmat_independent_of_thread_count():
    code = ("import sys; sys.path.insert(0, %r); import mercury_amd as ma, "
            "hashlib; g = ma.Graph.rmat(13, 8, seed=5); rp, ci = g.csr(); "
            "print(hashlib.sha256(rp.tobytes() + ci.tobytes()).hexdigest())"
            % str(ROOT))
    outs = set()
    for t in ("1", "3", "8"):
        env = dict(os.environ, OMP_NUM_THREADS=t)
        outs.add(subprocess.run([sys.executable, "-c", code], env=env,
                                capture_output=True, text=True,
                                check=True).stdout.strip())
    assert len(outs) == 1


def test_rmat16_seed1_shape():
    """Config: RMAT scale 16, ef 16, seed 1 -- record the cleaned size."""
    g = ma.Graph.rmat(16, 16, seed=1)
    assert g.n == 65536
    assert g.nnz == 1819108  # 2m after symmetrize/dedup/no self-loops


def test_er_shape():
    g = ma.Graph.er(5000, 40000, seed=3)
    rp, ci = g.csr()
    assert g.n == 5000 and 70000 < g.nnz <= 80000
    assert np.all(np.diff(rp) >= 0)


# --------------------------------------------------- DAG preprocessing

@pytest.mark.parametrize("scale,ef,seed,thd", [(10, 16, 7, 128), (12, 8, 3, 128),
                                               (11, 32, 2, 16)])
def test_product_orientation_equals_oracle(scale, ef, seed, thd):
    g = ma.Graph.rmat(scale, ef, seed=seed)
    rp, ci = g.csr()
    mine = g.orientation(thd)
    ref = oracle.orientation(rp, ci, thd)
    for k in ("dag_rowptr", "dag_colidx", "small_rowptr", "small_colidx",
              "large_rowptr", "large_colidx"):
        assert np.array_equal(mine[k], ref[k]), k


@pytest.mark.skipif(not oracle.ref_available(),
                    reason="oracle/_ref not built (reference sources absent)")
def test_product_orientation_equals_reference_graph_cc(tmp_path):
    g = ma.Graph.rmat(11, 16, seed=9)
    g.write(str(tmp_path / "graph"))
    mine = g.orientation(128)
    ref = oracle.ref_orientation(str(tmp_path / "graph"), g.n, 128)
    for k in ("dag_rowptr", "dag_colidx", "small_rowptr", "small_colidx",
              "large_rowptr", "large_colidx"):
        assert np.array_equal(mine[k], ref[k]), k


# -------------------------------------------------------------------- CLIs

def test_tc_challenge_usage():
    r = subprocess.run([str(BIN / "tc_challenge")], capture_output=True,
                       text=True)
    assert r.returncode == 1 and "Usage:" in r.stdout


def test_sm_usage_and_unknown_pattern():
    r = subprocess.run([str(BIN / "sm")], capture_output=True, text=True)
    assert r.returncode == 1 and "Usage:" in r.stdout
    r = subprocess.run([str(BIN / "sm"), str(GOLDEN / "k6" / "graph"), "P9"],
                       capture_output=True, text=True)
    assert r.returncode == 255 and "Not support." in r.stdout


def test_cli_load_error_is_reported_not_crashed(tmp_path):
    r = subprocess.run([str(BIN / "tc_challenge"), str(tmp_path / "graph")],
                       capture_output=True, text=True)
    assert r.returncode == 1 and "Failed to open file" in r.stderr
