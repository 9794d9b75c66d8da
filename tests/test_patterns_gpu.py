"""P1..P8 parity on the GPU: the HIP kernels (through the C ABI) against
the committed golden fixtures and the CPU oracle, bit-exact uint64."""
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

import mercury_amd as ma
import oracle
from conftest import csr_from_edges

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
GOLDEN = ROOT / "tests" / "golden"
PATTERNS = list(range(1, 9))


def _golden():
    return json.loads((GOLDEN / "golden.json").read_text())


@pytest.mark.parametrize("name", list(_golden()))
def test_patterns_golden_fixtures(gpu_available, name):
    want = _golden()[name]["counts"]
    g = ma.Graph.load(str(GOLDEN / name / "graph"))
    dg = ma.DeviceGraph(g, 0)
    got = {"TC": dg.tc_count()[0]}
    for p in PATTERNS:
        got[f"P{p}"] = dg.pattern_count(p)[0]
    print(name, got, flush=True)
    assert got == want


@pytest.mark.parametrize("scale,ef,seed", [(8, 8, 1), (9, 4, 2), (8, 16, 3)])
def test_patterns_rmat_vs_oracle(gpu_available, scale, ef, seed):
    g = ma.Graph.rmat(scale, ef, seed=seed)
    rp, ci = g.csr()
    dg = ma.DeviceGraph(g, 0)
    for p in PATTERNS:
        want = oracle.pattern(rp, ci, p)[0]
        got = dg.pattern_count(p)[0]
        print(f"rmat{scale} ef{ef} seed{seed} P{p}: {got} vs {want}",
              flush=True)
        assert got == want, p


def test_patterns_er_vs_oracle(gpu_available):
    g = ma.Graph.er(700, 700 * 6, seed=5)
    rp, ci = g.csr()
    dg = ma.DeviceGraph(g, 0)
    for p in PATTERNS:
        assert dg.pattern_count(p)[0] == oracle.pattern(rp, ci, p)[0], p


def test_patterns_edge_cases(gpu_available):
    # empty graph, single edge, a star, a complete graph
    cases = [
        (np.zeros(5, dtype=np.int64), np.zeros(0, dtype=np.int32)),
        csr_from_edges(2, [(0, 1)]),
        csr_from_edges(9, [(0, i) for i in range(1, 9)]),
        csr_from_edges(12, [(a, b) for a in range(12)
                            for b in range(a + 1, 12)]),
    ]
    for rp, ci in cases:
        g = ma.Graph.from_csr(rp, ci)
        dg = ma.DeviceGraph(g, 0)
        for p in PATTERNS:
            assert dg.pattern_count(p)[0] == oracle.pattern(rp, ci, p)[0]


def test_patterns_repeatable(gpu_available):
    g = ma.Graph.rmat(9, 8, seed=7)
    dg = ma.DeviceGraph(g, 0)
    for p in PATTERNS:
        assert len({dg.pattern_count(p)[0] for _ in range(5)}) == 1


def test_unknown_pattern_rejected(gpu_available):
    g = ma.Graph.rmat(6, 4, seed=1)
    dg = ma.DeviceGraph(g, 0)
    with pytest.raises(ma.MgpmError) as ei:
        dg.pattern_count(9)
    assert ei.value.code == -8


@pytest.mark.parametrize("algebra", [[], ["--set-algebra"]])
def test_sm_cli_matches_golden(gpu_available, algebra):
    """Default (P1, P4, P7, P8 by census, the rest per edge) and
    --set-algebra (every pattern per edge)."""
    want = _golden()["gnp40"]["counts"]
    for p in range(1, 9):
        # the reference's form: <graph> <pattern> <devices per node>
        # (main_mpi.cpp:20-25); 4 devices per node is device 0 on rank 0
        r = subprocess.run([str(ROOT / "mercury_amd" / "bin" / "sm"),
                            str(GOLDEN / "gnp40" / "graph"), f"P{p}"]
                           + (["4"] if p % 2 else ["--device", "0"])
                           + algebra,
                           capture_output=True, text=True, timeout=60)
        assert r.returncode == 0, r.stderr
        line = [l for l in r.stdout.splitlines()
                if l.startswith("total_num_triangles = ")][0]
        assert int(line.split("=")[1]) == want[f"P{p}"]


@pytest.mark.parametrize("plan", [[], ["--host-plan"]])
def test_tc_challenge_cli_matches_golden(gpu_available, plan):
    """Default (plan built on the device) and --host-plan."""
    want = _golden()["rmat9"]["counts"]["TC"]
    r = subprocess.run([str(ROOT / "mercury_amd" / "bin" / "tc_challenge"),
                        str(GOLDEN / "rmat9" / "graph")] + plan,
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert "Runtime  = " in r.stdout
    line = [l for l in r.stdout.splitlines()
            if l.startswith("total_num_triangles = ")][0]
    assert int(line.split("=")[1]) == want


@pytest.mark.parametrize("split", [1, 6, 12, -1])
def test_patterns_split_shares_vs_oracle(gpu_available, split):
    """pattern_split forces heavy tasks into shares: at 2^1 nearly every edge
    and vertex is cut into min(d0, d1) (or d) shares, at -1 none is.  Every
    count must stay equal to the oracle's."""
    for scale, ef, seed in [(8, 8, 31), (7, 16, 32)]:
        g = ma.Graph.rmat(scale, ef, seed=seed)
        rp, ci = g.csr()
        dg = ma.DeviceGraph(g, 0)
        t = ma.make_tune(pattern_split=split, pattern_census=-1)
        for p in PATTERNS:
            want = oracle.pattern(rp, ci, p)[0]
            got = dg.pattern_count(p, t)[0]
            print(f"split {split} rmat{scale} P{p}: {got} vs {want}",
                  flush=True)
            assert got == want, (split, p)


def test_patterns_split_hub_star_and_clique(gpu_available):
    """A hub joined to a clique: the hub's edges are the heaviest tasks and
    are split at the default threshold."""
    n = 300
    edges = [(0, i) for i in range(1, n)]
    edges += [(a, b) for a in range(1, 60) for b in range(a + 1, 60)]
    edges += [(i, i + 1) for i in range(60, n - 1)]
    rp, ci = csr_from_edges(n, edges)
    g = ma.Graph.from_csr(rp, ci)
    dg = ma.DeviceGraph(g, 0)
    for split in (0, 8, -1):
        t = ma.make_tune(pattern_split=split, pattern_census=-1) if split else None
        for p in PATTERNS:
            assert dg.pattern_count(p, t)[0] == oracle.pattern(rp, ci, p)[0], \
                (split, p)
