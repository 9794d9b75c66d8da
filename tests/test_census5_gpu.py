"""P7 (2 x induced K1,1,3) and P8 (3 x induced K4 + pendant) in closed form
from the device 4-clique and 5-clique census (census.cc,
census_kernels.hip) against the oracle, closed forms, and the
reference-form per-edge kernels (pattern_census = -1), bit-exact."""
from math import comb

import pytest

import mercury_amd as ma
import oracle
from conftest import csr_from_edges

pytestmark = pytest.mark.gpu

SET_ALGEBRA = dict(pattern_census=-1)


def _both(g, p):
    """(census count, reference-form kernel count) of pattern p."""
    dg = ma.DeviceGraph(g, 0)
    a, st = dg.pattern_count(p)
    b = dg.pattern_count(p, ma.make_tune(**SET_ALGEBRA))[0]
    dg.close()
    assert st.path == 30 + p  # the census path ran
    return a, b


@pytest.mark.parametrize("p", [7, 8])
@pytest.mark.parametrize("scale,ef,seed",
                         [(8, 8, 1), (10, 16, 7), (11, 4, 3), (12, 16, 5)])
def test_census5_rmat_vs_oracle(gpu_available, p, scale, ef, seed):
    g = ma.Graph.rmat(scale, ef, seed=seed)
    rp, ci = g.csr()
    want = oracle.pattern(rp, ci, p)[0]
    a, b = _both(g, p)
    print(f"P{p} rmat{scale} ef{ef} seed{seed}: census {a} kernels {b} "
          f"oracle {want}", flush=True)
    assert a == want and b == want


@pytest.mark.parametrize("p", [7, 8])
@pytest.mark.parametrize("scale,seed", [(13, 4), (14, 2)])
def test_census5_matches_kernels(gpu_available, p, scale, seed):
    a, b = _both(ma.Graph.rmat(scale, 16, seed=seed), p)
    print(f"P{p} rmat{scale} seed{seed}: census {a} kernels {b}", flush=True)
    assert a == b


@pytest.mark.parametrize("p", [7, 8])
def test_census5_er(gpu_available, p):
    g = ma.Graph.er(1 << 12, (1 << 12) * 12, seed=9)
    rp, ci = g.csr()
    a, b = _both(g, p)
    assert a == b == oracle.pattern(rp, ci, p)[0]


def _closed(n, edges, want7, want8):
    rp, ci = csr_from_edges(n, edges)
    g = ma.Graph.from_csr(rp, ci)
    for p, want in ((7, want7), (8, want8)):
        a, b = _both(g, p)
        assert a == want and b == want, (p, n, len(edges), a, b, want)


def _clique(vs):
    return [(a, b) for i, a in enumerate(vs) for b in vs[i + 1:]]


def test_census5_closed_forms(gpu_available):
    for n in (3, 4, 5, 9, 30):                # complete graphs: none
        _closed(n, _clique(list(range(n))), 0, 0)
    for r in (2, 3, 5, 40):                   # K_{1,1,r}: C(r,3) induced K113
        edges = [(0, 1)] + [(h, 2 + j) for h in (0, 1) for j in range(r)]
        _closed(r + 2, edges, 2 * comb(r, 3), 0)
    for p in (1, 3):                          # K4, p pendants per vertex
        edges = _clique([0, 1, 2, 3])
        edges += [(v, 4 + v * p + j) for v in range(4) for j in range(p)]
        _closed(4 + 4 * p, edges, 0, 3 * 4 * p)
    # K_n with one pendant per vertex: 3 n C(n-1,3).  n = 150 gives common
    # neighbourhoods over 128 (one wave per triangle) and triangle sets over
    # 64 (multi-word adjacency rows).
    for n in (6, 70, 150):
        edges = _clique(list(range(n))) + [(v, n + v) for v in range(n)]
        _closed(2 * n, edges, 0, 3 * n * comb(n - 1, 3))
    _closed(5, [], 0, 0)


def test_census5_hub_and_clique(gpu_available):
    # a hub joined to a 30-clique and to 200 leaves
    edges = _clique(list(range(1, 31)))
    edges += [(0, v) for v in range(1, 231)]
    rp, ci = csr_from_edges(231, edges)
    g = ma.Graph.from_csr(rp, ci)
    for p, closed in ((7, 0), (8, 3 * 200 * comb(30, 3))):
        want = oracle.pattern(rp, ci, p)[0]
        a, b = _both(g, p)
        assert a == want == closed and b == want, p


def test_census5_dense_mix(gpu_available):
    # two overlapping 90-cliques plus a sparse random part: triangle sets
    # over 64 with mixed adjacency
    import random
    rng = random.Random(11)
    edges = set(_clique(list(range(90))) + _clique(list(range(50, 140))))
    for _ in range(3000):
        a, b = rng.randrange(400), rng.randrange(400)
        if a != b:
            edges.add((min(a, b), max(a, b)))
    for _ in range(4000):                     # thin out the overlap
        a, b = rng.randrange(50, 90), rng.randrange(50, 90)
        edges.discard((min(a, b), max(a, b)))
    rp, ci = csr_from_edges(400, sorted(edges))
    g = ma.Graph.from_csr(rp, ci)
    for p in (7, 8):
        want = oracle.pattern(rp, ci, p)[0]
        a, b = _both(g, p)
        print(f"P{p} dense mix: census {a} kernels {b} oracle {want}",
              flush=True)
        assert a == want and b == want, p


def test_census_dense_random_golden(gpu_available):
    # dense G(n, p) graphs whose oracle counts are committed
    # (tests/golden/make_dense_counts.py); P1 and P4 by the census too
    import json
    import sys
    from pathlib import Path
    golden = Path(__file__).resolve().parent / "golden"
    sys.path.insert(0, str(golden))
    from make_dense_counts import dense_edges
    for row in json.loads((golden / "dense_counts.json").read_text()):
        n = row["n"]
        rp, ci = csr_from_edges(n, dense_edges(n, row["p"], row["seed"]))
        assert len(ci) // 2 == row["edges"]
        g = ma.Graph.from_csr(rp, ci)
        for p in (1, 4, 7, 8):
            a, b = _both(g, p)
            print(f"P{p} G({n}, {row['p']}): census {a} kernels {b} "
                  f"oracle {row[f'P{p}']}", flush=True)
            assert a == b == row[f"P{p}"], (n, p)
