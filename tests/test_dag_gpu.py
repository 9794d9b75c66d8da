"""Device-built degree-oriented DAG and small/large division
(mgpm_dgraph_orientation) against the oracle's restatement of graph.cc:196-321
(oracle.orientation, itself bit-equal to the reference's own graph.cc where
that builds) and against the host build, array for array."""
import json
from pathlib import Path

import numpy as np
import pytest

import mercury_amd as ma
import oracle
from conftest import csr_from_edges

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"
KEYS = ("dag_rowptr", "dag_colidx", "small_rowptr", "small_colidx",
        "large_rowptr", "large_colidx")


def _same(g, thd):
    rp, ci = g.csr()
    want = oracle.orientation(rp, ci, thd)
    host = g.orientation(thd)
    got, st = ma.DeviceGraph(g, 0).orientation(thd)
    for k in KEYS:
        assert np.array_equal(got[k], want[k]), (k, thd)
        assert np.array_equal(host[k], want[k]), (k, thd)
    assert st.m_plus == len(want["dag_colidx"]) and st.kernel_ms > 0


@pytest.mark.parametrize("scale,ef,seed", [(8, 8, 1), (12, 16, 2), (16, 16, 1)])
@pytest.mark.parametrize("thd", [128, 3, 0])
def test_dag_rmat_equals_host(gpu_available, scale, ef, seed, thd):
    _same(ma.Graph.rmat(scale, ef, seed=seed), thd)


def test_dag_er_and_fixtures(gpu_available):
    _same(ma.Graph.er(20000, 200000, seed=3), 128)
    for name in json.loads((GOLDEN / "golden.json").read_text()):
        _same(ma.Graph.load(str(GOLDEN / name / "graph")), 8)


def test_dag_edge_cases(gpu_available):
    cases = [
        (np.zeros(5, dtype=np.int64), np.zeros(0, dtype=np.int32)),
        csr_from_edges(2, [(0, 1)]),
        csr_from_edges(9, [(0, i) for i in range(1, 9)]),
        csr_from_edges(300, [(a, b) for a in range(300) for b in range(a + 1, 300)
                             if (a * 7 + b) % 5 == 0]),
    ]
    for rp, ci in cases:
        _same(ma.Graph.from_csr(rp, ci), 128)


def test_dag_repeat_and_threshold_change(gpu_available):
    g = ma.Graph.rmat(11, 16, seed=5)
    dg = ma.DeviceGraph(g, 0)
    rp, ci = g.csr()
    for thd in (128, 16, 128):
        got, _ = dg.orientation(thd)
        want = oracle.orientation(rp, ci, thd)
        for k in KEYS:
            assert np.array_equal(got[k], want[k]), (k, thd)
    # the triangle count still works on the same handle afterwards
    assert dg.tc_count()[0] == ma.DeviceGraph(g, 0).tc_count()[0]
