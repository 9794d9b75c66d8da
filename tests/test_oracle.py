"""Pins the CPU oracle (CPU only, no GPU).

The reference ships no tests or golden vectors (it has no test directory),
so every oracle function is pinned here by something other than its own
formula: exhaustive brute-force enumeration, closed forms, mathematical
identities, and -- for the DAG preprocessing -- the reference's own
graph.cc compiled from its sources into oracle/_ref.
"""
import itertools
import json
import random
from math import comb
from pathlib import Path

import numpy as np
import pytest

import oracle
from conftest import csr_from_edges

GOLDEN = Path(__file__).resolve().parent / "golden"


def rand_graph(rng, n, p):
    return csr_from_edges(n, [(a, b) for a in range(n)
                              for b in range(a + 1, n) if rng.random() < p])


# ------------------------------------------------------------ brute force

def test_brute_closed_forms():
    # K_n: only cliques; C(n,3) triangles, C(n,4) 4-cliques, C(n,5) 5-cliques
    rp, ci = csr_from_edges(7, itertools.combinations(range(7), 2))
    c = oracle.brute(rp, ci)
    assert c["triangle"] == comb(7, 3) and c["k4"] == comb(7, 4)
    assert c["k5"] == comb(7, 5)
    assert all(c[k] == 0 for k in ("claw", "path4", "paw", "k14", "fork",
                                   "bull", "diamond_pendant", "k113",
                                   "k4_pendant"))
    # star K_{1,k}: claws C(k,3), K_{1,4}'s C(k,4)
    k = 9
    rp, ci = csr_from_edges(k + 1, [(0, i) for i in range(1, k + 1)])
    c = oracle.brute(rp, ci)
    assert c["claw"] == comb(k, 3) and c["k14"] == comb(k, 4)
    assert c["triangle"] == 0 and c["path4"] == 0
    # cycle C_k (k >= 6): k induced P4, no claws
    rp, ci = csr_from_edges(11, [(i, (i + 1) % 11) for i in range(11)])
    c = oracle.brute(rp, ci)
    assert c["path4"] == 11 and c["claw"] == 0 and c["fork"] == 0
    # bull graph itself: one bull, paws = 2
    rp, ci = csr_from_edges(5, [(0, 1), (1, 2), (2, 0), (0, 3), (1, 4)])
    c = oracle.brute(rp, ci)
    assert c["bull"] == 1 and c["paw"] == 2 and c["triangle"] == 1


# --------------------------------------------------------------- patterns

@pytest.mark.parametrize("seed", range(6))
def test_pattern_restatement_matches_brute_force(seed):
    """The literal restatement of pattern_kernels.cuh (set algebra over
    directed edges) equals brute-force induced counts times the
    multiplicities of oracle.expected_patterns, for TC and P1..P8."""
    rng = random.Random(seed)
    for _ in range(12):
        n = rng.randint(4, 22)
        rp, ci = rand_graph(rng, n, rng.uniform(0.1, 0.8))
        exp = oracle.expected_patterns(oracle.brute(rp, ci))
        assert oracle.tc(rp, ci) == exp[0]
        for p in range(1, 9):
            assert oracle.pattern(rp, ci, p)[0] == exp[p], (n, p)


def test_pattern_planted_mistake_is_caught():
    """A plausible transcription slip (dropping the 3*claw term of P3, i.e.
    excluding v1 from L0) is detected by the brute-force pin."""
    rng = random.Random(3)
    rp, ci = rand_graph(rng, 18, 0.3)
    census = oracle.brute(rp, ci)
    wrong_p3 = 4 * census["k14"]
    assert census["claw"] > 0
    assert oracle.pattern(rp, ci, 3)[0] != wrong_p3


def test_pattern_counters_combination():
    """sm_multigpu.cu:209-232: P4/P6 report counters[1]; the others
    counters[1] - counters[2]."""
    rng = random.Random(9)
    rp, ci = rand_graph(rng, 20, 0.35)
    for p in range(1, 9):
        r, c1, c2 = oracle.pattern(rp, ci, p)
        if p in (4, 6):
            assert r == c1 and c2 == 0
        else:
            assert r == c1 - c2 and c1 >= c2


def test_pattern_range_partition_sums():
    rng = random.Random(5)
    rp, ci = rand_graph(rng, 30, 0.3)
    n = len(rp) - 1
    for p in range(1, 9):
        whole = oracle.pattern(rp, ci, p)
        parts = [oracle.pattern(rp, ci, p, v_lo=a, v_hi=b)
                 for a, b in ((0, 7), (7, 19), (19, n))]
        assert sum(x[1] for x in parts) == whole[1]
        assert sum(x[2] for x in parts) == whole[2]


# --------------------------------------------------------------------- TC

def test_tc_closed_forms():
    for n in (3, 4, 10, 25):
        rp, ci = csr_from_edges(n, itertools.combinations(range(n), 2))
        assert oracle.tc(rp, ci) == comb(n, 3)
    for k in (4, 5, 12):  # wheel with k rim vertices: k triangles
        e = [(0, i) for i in range(1, k + 1)] + \
            [(i, i % k + 1) for i in range(1, k + 1)]
        rp, ci = csr_from_edges(k + 1, e)
        assert oracle.tc(rp, ci) == k
    rp, ci = csr_from_edges(9, [(i, 4 + j) for i in range(4) for j in range(5)])
    assert oracle.tc(rp, ci) == 0  # K_{4,5}
    rp, ci = csr_from_edges(30, [(i, (i - 1) // 2) for i in range(1, 30)])
    assert oracle.tc(rp, ci) == 0  # tree


def test_tc_identity_dag_sum():
    """TC over small+large division == plain sum over DAG arcs == brute."""
    rng = random.Random(2)
    for _ in range(10):
        rp, ci = rand_graph(rng, rng.randint(5, 50), rng.uniform(0.05, 0.6))
        o = oracle.orientation(rp, ci, 128)
        t = oracle.tc(rp, ci)
        assert t == oracle.tc_dag(o["dag_rowptr"], o["dag_colidx"])
        if len(rp) - 1 <= 64:
            assert t == oracle.brute(rp, ci)["triangle"]


def test_tc_division_threshold_invariant():
    """The small/large split only routes work; a hub-heavy graph has large
    edges at DEG_THD=128 and the count does not depend on it."""
    rng = np.random.default_rng(4)
    n = 320
    edges = [(a, b) for a in range(n) for b in range(a + 1, n)
             if rng.random() < 0.9]
    rp, ci = csr_from_edges(n, edges)
    o = oracle.orientation(rp, ci, 128)
    assert len(o["large_colidx"]) > 0 and len(o["small_colidx"]) > 0
    assert oracle.tc(rp, ci) == oracle.tc_dag(o["dag_rowptr"],
                                              o["dag_colidx"])


# ------------------------------------------------------------ orientation

def _dag_invariants(rp, ci, o):
    n = len(rp) - 1
    deg = np.diff(rp)
    drp, dci = o["dag_rowptr"], o["dag_colidx"]
    assert drp[-1] == len(ci) // 2  # m+ = m
    seen = set()
    for v in range(n):
        row = dci[drp[v]:drp[v + 1]]
        assert np.all(np.diff(row) > 0)  # sorted
        for u in row:
            # tie rule of graph.cc:207-213
            assert deg[u] > deg[v] or (deg[u] == deg[v] and u > v)
            key = (min(v, u), max(v, u))
            assert key not in seen
            seen.add(key)
    assert len(seen) == len(ci) // 2


def test_orientation_invariants():
    rng = random.Random(8)
    for _ in range(5):
        rp, ci = rand_graph(rng, 60, 0.2)
        _dag_invariants(rp, ci, oracle.orientation(rp, ci, 128))


@pytest.mark.skipif(not oracle.ref_available(),
                    reason="oracle/_ref not built (reference sources absent)")
@pytest.mark.parametrize("name", ["k6", "wheel8", "petersen", "gnp40",
                                  "gnp60", "rmat7", "rmat9"])
@pytest.mark.parametrize("thd", [128, 3, 8])
def test_orientation_matches_reference_graph_cc(name, thd):
    """The restatement equals the reference's own orientation_with_division
    (graph.cc compiled from its sources) on every fixture, including small
    thresholds that put edges on both sides of the division."""
    meta = json.loads((GOLDEN / "golden.json").read_text())[name]
    prefix = GOLDEN / name / "graph"
    rp = np.fromfile(GOLDEN / name / "graph.vertex.bin", dtype="<i8")
    ci = np.fromfile(GOLDEN / name / "graph.edge.bin", dtype="<i4")
    assert len(rp) == meta["n"] + 1
    mine = oracle.orientation(rp, ci, thd)
    ref = oracle.ref_orientation(str(prefix), meta["n"], thd)
    for k in ("dag_rowptr", "dag_colidx", "small_rowptr", "small_colidx",
              "large_rowptr", "large_colidx", "max_out"):
        assert np.array_equal(mine[k], ref[k]), k


@pytest.mark.skipif(not oracle.ref_available(),
                    reason="oracle/_ref not built (reference sources absent)")
def test_edgelist_order_matches_reference():
    """graph.cc:336-365: the task edge list is every directed CSR entry in
    CSR order; the oracle walks edges in exactly that order."""
    prefix = GOLDEN / "rmat9" / "graph"
    rp = np.fromfile(GOLDEN / "rmat9" / "graph.vertex.bin", dtype="<i8")
    ci = np.fromfile(GOLDEN / "rmat9" / "graph.edge.bin", dtype="<i4")
    src, dst = oracle.ref_edgelist(str(prefix))
    assert np.array_equal(dst, ci)
    assert np.array_equal(src, np.repeat(np.arange(len(rp) - 1), np.diff(rp)))


# ------------------------------------------------------------------ golden

def test_golden_counts_stable():
    golden = json.loads((GOLDEN / "golden.json").read_text())
    for name, g in golden.items():
        rp = np.fromfile(GOLDEN / name / "graph.vertex.bin", dtype="<i8")
        ci = np.fromfile(GOLDEN / name / "graph.edge.bin", dtype="<i4")
        assert oracle.tc(rp, ci) == g["counts"]["TC"]
        for p in range(1, 9):
            assert oracle.pattern(rp, ci, p)[0] == g["counts"][f"P{p}"], \
                (name, p)


def test_pattern_work_known_answer():
    """Set-op work of the reference's P4 algebra on the path 0-1-2, by hand:
    edge (1,0): two differences of sizes 2+1 and 1+2, then membership tests
    of L1 = {1} in N(0) and N(2) (1+1 each); edge (2,1): differences 1+2
    and 2+1, then tests of {0,2} in N(1) (1+2 each): 22 in all."""
    rp, ci = csr_from_edges(3, [(0, 1), (1, 2)])
    count, work = oracle.pattern_work(rp, ci, 4)
    assert count == oracle.pattern(rp, ci, 4)[0] == 0
    assert work == 22
