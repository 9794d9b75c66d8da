"""P1 (induced claws) in closed form from the device triangle census
(census.cc, census_kernels.hip) against the oracle, closed forms, and the
reference-form per-edge kernels (pattern_census = -1), bit-exact."""
import pytest

import mercury_amd as ma
import oracle
from conftest import csr_from_edges

pytestmark = pytest.mark.gpu

SET_ALGEBRA = dict(pattern_census=-1)


def _p1(g):
    dg = ma.DeviceGraph(g, 0)
    a, st = dg.pattern_count(1)
    b = dg.pattern_count(1, ma.make_tune(**SET_ALGEBRA))[0]
    dg.close()
    assert st.path == 31  # the census path ran
    return a, b


@pytest.mark.parametrize("scale,ef,seed", [(8, 8, 1), (10, 16, 7), (11, 4, 3)])
def test_census_p1_rmat_vs_oracle(gpu_available, scale, ef, seed):
    g = ma.Graph.rmat(scale, ef, seed=seed)
    rp, ci = g.csr()
    want = oracle.pattern(rp, ci, 1)[0]
    a, b = _p1(g)
    print(f"rmat{scale} ef{ef} seed{seed}: census {a} kernels {b} oracle {want}",
          flush=True)
    assert a == want and b == want


@pytest.mark.parametrize("scale,seed", [(13, 4), (14, 2), (16, 1)])
def test_census_p1_matches_kernels(gpu_available, scale, seed):
    a, b = _p1(ma.Graph.rmat(scale, 16, seed=seed))
    print(f"rmat{scale} seed{seed}: census {a} kernels {b}", flush=True)
    assert a == b


def test_census_p1_er(gpu_available):
    g = ma.Graph.er(1 << 12, (1 << 12) * 12, seed=9)
    rp, ci = g.csr()
    a, b = _p1(g)
    assert a == b == oracle.pattern(rp, ci, 1)[0]


def _closed(n, edges, want):
    rp, ci = csr_from_edges(n, edges)
    a, b = _p1(ma.Graph.from_csr(rp, ci))
    assert a == want and b == want, (n, len(edges), a, b, want)


def test_census_p1_closed_forms(gpu_available):
    from math import comb
    for k in (0, 1, 2, 3, 7, 40):            # star K1,k: C(k,3) claws
        _closed(k + 1, [(0, j) for j in range(1, k + 1)], comb(k, 3))
    for n in (3, 4, 9, 30):                   # complete graphs: none
        _closed(n, [(a, b) for a in range(n) for b in range(a + 1, n)], 0)
    for a, b in ((2, 5), (3, 3), (4, 7)):     # K_{a,b}: a C(b,3) + b C(a,3)
        _closed(a + b, [(i, a + j) for i in range(a) for j in range(b)],
                a * comb(b, 3) + b * comb(a, 3))
    _closed(5, [], 0)


def test_census_p1_hub_and_clique(gpu_available):
    # a hub joined to a 30-clique and to 200 leaves
    edges = [(a, b) for a in range(1, 31) for b in range(a + 1, 31)]
    edges += [(0, v) for v in range(1, 231)]
    rp, ci = csr_from_edges(231, edges)
    g = ma.Graph.from_csr(rp, ci)
    want = oracle.pattern(rp, ci, 1)[0]
    a, b = _p1(g)
    assert a == want and b == want
