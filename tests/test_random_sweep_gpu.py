"""A randomized parity sweep: small graphs of varied shape (Erdos-Renyi of
varied density, RMAT of varied edge factor, a clique with random extra
edges and pendants), each counted through the C ABI -- TC with the host-
and the device-built plan, P1..P8 by the default path (the census) and by
the per-edge kernels -- and compared with the oracle, bit-exact."""
import random

import pytest

import mercury_amd as ma
import oracle
from conftest import csr_from_edges

pytestmark = pytest.mark.gpu

SET_ALGEBRA = ma.make_tune(pattern_census=-1)


def _case(seed):
    rng = random.Random(seed)
    kind = seed % 3
    if kind == 0:  # G(n, p)
        n = rng.randint(5, 140)
        p = rng.choice([0.02, 0.08, 0.2, 0.45])
        edges = [(a, b) for a in range(n) for b in range(a + 1, n)
                 if rng.random() < p]
        return f"gnp n={n} p={p}", csr_from_edges(n, edges)
    if kind == 1:  # RMAT
        scale = rng.randint(5, 9)
        ef = rng.choice([1, 2, 4, 8, 16])
        g = ma.Graph.rmat(scale, ef, seed=seed)
        return f"rmat scale={scale} ef={ef}", g.csr()
    k = rng.randint(4, 40)  # a clique, random extra edges, pendants
    n = k + rng.randint(5, 60)
    edges = [(a, b) for a in range(k) for b in range(a + 1, k)]
    edges += [(rng.randrange(n), rng.randrange(n)) for _ in range(2 * n)]
    edges += [(rng.randrange(k), v) for v in range(k, n) if rng.random() < 0.3]
    return f"clique k={k} n={n}", csr_from_edges(n, edges)


@pytest.mark.parametrize("seed", range(96))
def test_random_graph_all_counts(gpu_available, seed):
    name, (rp, ci) = _case(seed)
    print(f"seed {seed}: {name}, {len(ci) // 2} edges", flush=True)
    g = ma.Graph.from_csr(rp, ci)
    want_tc = oracle.tc(rp, ci)
    for device_plan in (False, True):
        dg = ma.DeviceGraph(g, 0, device_plan=device_plan)
        assert dg.tc_count()[0] == want_tc, (name, device_plan)
        if device_plan:
            dg.close()
            continue
        for p in range(1, 9):
            want = oracle.pattern(rp, ci, p)[0]
            assert dg.pattern_count(p)[0] == want, (name, p)
            assert dg.pattern_count(p, SET_ALGEBRA)[0] == want, (name, p)
        dg.close()
