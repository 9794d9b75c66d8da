"""The closed forms behind the P7 and P8 census (census.cc), evaluated
term by term in plain Python on small random graphs and checked against
brute-force induced-subgraph counts (oracle.brute) and the oracle's
literal restatement of the reference's set algebra: P7 = 2 x induced
K1,1,3, P8 = 3 x induced (K4 + pendant)."""
import random

import pytest

import oracle
from conftest import csr_from_edges


def _terms(adj):
    """Census terms of census.cc from first principles (sets, no DAG)."""
    n = len(adj)
    d = [len(s) for s in adj]
    edges = [(a, b) for a in range(n) for b in adj[a] if a < b]
    t = {e: len(adj[e[0]] & adj[e[1]]) for e in edges}
    tris = [(a, b, c) for (a, b) in edges for c in adj[a] & adj[b] if c > b]
    kt = {T: len(adj[T[0]] & adj[T[1]] & adj[T[2]]) for T in tris}

    def te(a, b):
        return t[(min(a, b), max(a, b))]

    k5 = 0
    for a, b, c in tris:
        common = sorted(x for x in adj[a] & adj[b] & adj[c] if x > c)
        k5 += sum(1 for i, x in enumerate(common) for y in common[i + 1:]
                  if y in adj[x])
    return {
        "edge_triples": sum(x * (x - 1) * (x - 2) // 6 for x in t.values()),
        "tri_pairs": sum(k * (k - 1) // 2 for k in kt.values()),
        "tri_edge_k": sum(kt[T] * (te(T[0], T[1]) + te(T[0], T[2]) +
                                   te(T[1], T[2]) - 6) for T in tris),
        "tri_deg_k": sum(kt[T] * (d[T[0]] + d[T[1]] + d[T[2]] - 9)
                         for T in tris),
        "k5": k5,
    }


@pytest.mark.parametrize("seed", range(16))
def test_p7_p8_closed_forms(seed):
    rng = random.Random(seed)
    n = rng.randint(8, 36)
    p = rng.choice([0.2, 0.4, 0.6, 0.85])
    edges = [(a, b) for a in range(n) for b in range(a + 1, n)
             if rng.random() < p]
    adj = [set() for _ in range(n)]
    for a, b in edges:
        adj[a].add(b)
        adj[b].add(a)
    c = _terms(adj)
    assert c["tri_edge_k"] % 2 == 0 and c["tri_deg_k"] % 3 == 0
    p7 = 2 * (c["edge_triples"] - c["tri_edge_k"] // 2 + 3 * c["tri_pairs"]
              - 10 * c["k5"])
    p8 = 3 * (c["tri_deg_k"] // 3 - c["tri_edge_k"] + 6 * c["tri_pairs"]
              - 20 * c["k5"])
    rp, ci = csr_from_edges(n, edges)
    brute = oracle.brute(rp, ci)
    assert p7 == 2 * brute["k113"] == oracle.pattern(rp, ci, 7)[0]
    assert p8 == 3 * brute["k4_pendant"] == oracle.pattern(rp, ci, 8)[0]
