"""The census runs on the graph relabelled in DAG order (census_relabel in
census5_kernels.hip: vertex v becomes its rank by (degree, id)).  Its sums
count subgraphs, so the counts must not depend on the labelling: the
oracle's TC and P1-P8 on RMAT graphs and on their relabellings, built here
the same way (a stable sort of the vertices by degree, then the arcs by
(rank of the row, rank of the head)), agree, and the relabelled rows have
the layout the census kernels rely on."""
import numpy as np
import pytest

import mercury_amd as ma
import oracle


def relabel_dag_order(rp, ci):
    n = len(rp) - 1
    deg = np.diff(rp)
    perm = np.argsort(deg, kind="stable")  # rank -> old id
    rank = np.empty(n, np.int64)
    rank[perm] = np.arange(n)
    src = np.repeat(np.arange(n), deg)
    key = rank[src] * n + rank[ci]
    ci2 = (np.sort(key) % n).astype(np.int32)
    rp2 = np.zeros(n + 1, np.int64)
    rp2[1:] = np.cumsum(deg[perm])
    return rp2, ci2


@pytest.mark.parametrize("scale,ef,seed", [(8, 4, 1), (9, 8, 3)])
def test_counts_invariant_under_dag_order_relabelling(scale, ef, seed):
    g = ma.Graph.rmat(scale, ef, seed=seed)
    rp, ci = g.csr()
    rp2, ci2 = relabel_dag_order(rp, ci)
    n = len(rp) - 1
    d2 = np.diff(rp2)
    # degrees non-decreasing in the new id: DAG order (degree, then id) is
    # id order, so each row lists the vertices below its own first
    assert (np.diff(d2) >= 0).all()
    for v in range(n):
        row = ci2[rp2[v]:rp2[v + 1]]
        assert (np.diff(row) > 0).all()
        below = int((row < v).sum())
        above = [u for u in row if d2[u] > d2[v] or (d2[u] == d2[v] and u > v)]
        assert len(above) == len(row) - below
        assert (row[below:] > v).all()
    assert oracle.tc(rp, ci) == oracle.tc(rp2, ci2)
    for p in range(1, 9):
        assert oracle.pattern(rp, ci, p)[0] == oracle.pattern(rp2, ci2, p)[0], p
