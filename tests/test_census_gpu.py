"""P1 (induced claws) and P4 (induced 4-paths) in closed form from the
device triangle and 4-cycle census (census.cc, census_kernels.hip) against
the oracle, closed forms, and the reference-form per-edge kernels
(pattern_census = -1), bit-exact."""
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


@pytest.mark.parametrize("p", [1, 4])
@pytest.mark.parametrize("scale,ef,seed", [(8, 8, 1), (10, 16, 7), (11, 4, 3)])
def test_census_rmat_vs_oracle(gpu_available, p, scale, ef, seed):
    g = ma.Graph.rmat(scale, ef, seed=seed)
    rp, ci = g.csr()
    want = oracle.pattern(rp, ci, p)[0]
    a, b = _both(g, p)
    print(f"P{p} rmat{scale} ef{ef} seed{seed}: census {a} kernels {b} "
          f"oracle {want}", flush=True)
    assert a == want and b == want


@pytest.mark.parametrize("p", [1, 4])
@pytest.mark.parametrize("scale,seed", [(13, 4), (14, 2), (16, 1)])
def test_census_matches_kernels(gpu_available, p, scale, seed):
    a, b = _both(ma.Graph.rmat(scale, 16, seed=seed), p)
    print(f"P{p} rmat{scale} seed{seed}: census {a} kernels {b}", flush=True)
    assert a == b


@pytest.mark.parametrize("p", [1, 4])
def test_census_er(gpu_available, p):
    g = ma.Graph.er(1 << 12, (1 << 12) * 12, seed=9)
    rp, ci = g.csr()
    a, b = _both(g, p)
    assert a == b == oracle.pattern(rp, ci, p)[0]


def _closed(n, edges, want1, want4):
    rp, ci = csr_from_edges(n, edges)
    g = ma.Graph.from_csr(rp, ci)
    for p, want in ((1, want1), (4, want4)):
        a, b = _both(g, p)
        assert a == want and b == want, (p, n, len(edges), a, b, want)


def test_census_closed_forms(gpu_available):
    from math import comb
    for k in (0, 1, 2, 3, 7, 40):            # star K1,k: C(k,3) claws
        _closed(k + 1, [(0, j) for j in range(1, k + 1)], comb(k, 3), 0)
    for n in (3, 4, 9, 30):                   # complete graphs: none
        _closed(n, [(a, b) for a in range(n) for b in range(a + 1, n)], 0, 0)
    for a, b in ((2, 5), (3, 3), (4, 7)):     # K_{a,b}: a C(b,3) + b C(a,3)
        _closed(a + b, [(i, a + j) for i in range(a) for j in range(b)],
                a * comb(b, 3) + b * comb(a, 3), 0)
    for k in (4, 5, 6, 11):                   # cycle C_k: k induced P4 (k>4)
        _closed(k, [(i, (i + 1) % k) for i in range(k)], 0, k if k > 4 else 0)
    for k in (3, 4, 9):                       # path on k vertices: k - 3
        _closed(k, [(i, i + 1) for i in range(k - 1)], 0, max(k - 3, 0))
    _closed(5, [], 0, 0)


def test_census_hub_and_clique(gpu_available):
    # a hub joined to a 30-clique and to 200 leaves
    edges = [(a, b) for a in range(1, 31) for b in range(a + 1, 31)]
    edges += [(0, v) for v in range(1, 231)]
    rp, ci = csr_from_edges(231, edges)
    g = ma.Graph.from_csr(rp, ci)
    for p in (1, 4):
        want = oracle.pattern(rp, ci, p)[0]
        a, b = _both(g, p)
        assert a == want and b == want, p
