"""P2 (induced forks), P3 (4 x induced K1,4 + 3 x induced claws), P5
(2 x induced bulls) and P6 (2 x induced diamond + pendant + 2 x induced
paws) in closed form from the device 5-vertex census (census.cc,
census_kernels.hip, census5_kernels.hip) against the oracle's literal
restatement of the reference's set algebra, closed forms, and the
reference-form per-edge kernels (pattern_census = -1), bit-exact.  The
graphs cover every anchor class of the pair passes (DAG out-degree up to
64, 128, 256, 512 and above)."""
import random
from math import comb

import pytest

import mercury_amd as ma
import oracle
from conftest import csr_from_edges

pytestmark = pytest.mark.gpu

SET_ALGEBRA = dict(pattern_census=-1)
PATTERNS = (2, 3, 5, 6)


def _census(dg, p):
    c, st = dg.pattern_count(p)
    assert st.path == 30 + p  # the census ran
    return c


@pytest.mark.parametrize("scale,ef,seed",
                         [(7, 8, 1), (9, 8, 3), (10, 8, 7), (11, 4, 2)])
def test_pairs_rmat_vs_oracle(gpu_available, scale, ef, seed):
    g = ma.Graph.rmat(scale, ef, seed=seed)
    rp, ci = g.csr()
    dg = ma.DeviceGraph(g, 0)
    for p in PATTERNS:
        want = oracle.pattern(rp, ci, p)[0]
        got = _census(dg, p)
        print(f"P{p} rmat{scale} ef{ef} seed{seed}: census {got} oracle "
              f"{want}", flush=True)
        assert got == want, p


def test_pairs_golden_and_random(gpu_available):
    import json
    from pathlib import Path
    gold = Path(__file__).resolve().parent / "golden"
    for name, rec in json.loads((gold / "golden.json").read_text()).items():
        dg = ma.DeviceGraph(ma.Graph.load(str(gold / name / "graph")), 0)
        for p in PATTERNS:
            assert _census(dg, p) == rec["counts"][f"P{p}"], (name, p)
    rng = random.Random(11)
    for i in range(40):
        n = rng.randint(5, 60)
        pr = rng.choice([0.05, 0.2, 0.5, 0.8])
        edges = [(a, b) for a in range(n) for b in range(a + 1, n)
                 if rng.random() < pr]
        rp, ci = csr_from_edges(n, edges)
        dg = ma.DeviceGraph(ma.Graph.from_csr(rp, ci), 0)
        for p in PATTERNS:
            assert _census(dg, p) == oracle.pattern(rp, ci, p)[0], (i, p)


def test_pairs_er_and_hub(gpu_available):
    g = ma.Graph.er(3000, 30000, seed=5)
    rp, ci = g.csr()
    dg = ma.DeviceGraph(g, 0)
    for p in PATTERNS:
        assert _census(dg, p) == oracle.pattern(rp, ci, p)[0], p
    # a hub joined to a 40-clique and 100 leaves
    edges = [(a, b) for a in range(1, 41) for b in range(a + 1, 41)]
    edges += [(0, v) for v in range(1, 141)]
    edges += [(v, v + 100) for v in range(41, 60)]
    rp, ci = csr_from_edges(200, edges)
    dg = ma.DeviceGraph(ma.Graph.from_csr(rp, ci), 0)
    for p in PATTERNS:
        assert _census(dg, p) == oracle.pattern(rp, ci, p)[0], p


def test_pairs_closed_forms(gpu_available):
    def counts(n, edges):
        rp, ci = csr_from_edges(n, edges)
        dg = ma.DeviceGraph(ma.Graph.from_csr(rp, ci), 0)
        return {p: _census(dg, p) for p in PATTERNS}
    for k in (3, 4, 10, 200):  # star K1,k: P3 = 4 C(k,4) + 3 C(k,3)
        c = counts(k + 1, [(0, i) for i in range(1, k + 1)])
        assert c == {2: 0, 3: 4 * comb(k, 4) + 3 * comb(k, 3), 5: 0, 6: 0}
    for n in (5, 9, 40):  # complete graphs hold no induced pattern
        c = counts(n, [(a, b) for a in range(n) for b in range(a + 1, n)])
        assert c == {2: 0, 3: 0, 5: 0, 6: 0}
    for k in (6, 50):  # cycles: no vertex of degree 3
        c = counts(k, [(i, (i + 1) % k) for i in range(k)])
        assert c == {2: 0, 3: 0, 5: 0, 6: 0}
    # the bull itself: one induced bull, P5 = 2; its forks: none
    c = counts(5, [(0, 1), (0, 2), (1, 2), (0, 3), (1, 4)])
    assert c[5] == 2
    # the wheel W5 (hub + 5-cycle) and the empty graph
    c = counts(6, [(0, i) for i in range(1, 6)] +
               [(i, i % 5 + 1) for i in range(1, 6)])
    rp, ci = csr_from_edges(6, [(0, i) for i in range(1, 6)] +
                            [(i, i % 5 + 1) for i in range(1, 6)])
    assert c == {p: oracle.pattern(rp, ci, p)[0] for p in PATTERNS}
    assert counts(8, []) == {2: 0, 3: 0, 5: 0, 6: 0}


_GNP = {}


def _gnp(n, pr, seed):
    if (n, pr, seed) not in _GNP:
        rng = random.Random(seed)
        edges = [(a, b) for a in range(n) for b in range(a + 1, n)
                 if rng.random() < pr]
        _GNP[(n, pr, seed)] = csr_from_edges(n, edges)
    return _GNP[(n, pr, seed)]


@pytest.mark.parametrize("n,pr,seed", [(300, 0.45, 1), (700, 0.45, 2),
                                       (1400, 0.42, 3)])
@pytest.mark.parametrize("p", PATTERNS)
def test_pairs_every_anchor_class_vs_kernels(gpu_available, n, pr, seed, p):
    """Dense G(n, p): DAG out-degrees reach past 128, 256 and 512, so every
    anchor class (LDS rows of 1..8 words, and the global-row class) runs;
    the per-edge kernels give the reference's count (one pattern per test:
    the per-edge kernels take up to a minute on n = 1400)."""
    rp, ci = _gnp(n, pr, seed)
    dg = ma.DeviceGraph(ma.Graph.from_csr(rp, ci), 0)
    a = _census(dg, p)
    b = dg.pattern_count(p, ma.make_tune(**SET_ALGEBRA))[0]
    print(f"P{p} gnp n={n} p={pr}: census {a} kernels {b}", flush=True)
    assert a == b, p
    dg.close()


@pytest.mark.parametrize("scale,seed", [(12, 4), (13, 1)])
def test_pairs_rmat_vs_kernels(gpu_available, scale, seed):
    dg = ma.DeviceGraph(ma.Graph.rmat(scale, 16, seed=seed), 0)
    for p in PATTERNS:
        a = _census(dg, p)
        b = dg.pattern_count(p, ma.make_tune(**SET_ALGEBRA))[0]
        print(f"P{p} rmat{scale} seed{seed}: census {a} kernels {b}",
              flush=True)
        assert a == b, p


def test_pairs_repeat_and_interleave(gpu_available):
    """The workspace is reused across calls and patterns: counts repeat."""
    g = ma.Graph.rmat(9, 16, seed=9)
    rp, ci = g.csr()
    dg = ma.DeviceGraph(g, 0)
    want = {p: oracle.pattern(rp, ci, p)[0] for p in range(1, 9)}
    for _ in range(2):
        for p in (2, 1, 3, 4, 5, 7, 6, 8):
            assert dg.pattern_count(p)[0] == want[p], p
