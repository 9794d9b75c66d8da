"""Batches replayed from a recorded HIP graph (mgpm_batch_opts.use_graph):
every replay equals the single-query counts, with single-query work and
synchronisation between replays, and the graph is re-recorded when a plan
changes."""
import random

import pytest

import mercury_amd as ma

pytestmark = pytest.mark.gpu


def _graphs(n, seed, scales=(6, 9)):
    rng = random.Random(seed)
    return [ma.Graph.rmat(rng.randint(*scales), rng.choice((4, 8, 16)),
                          seed=seed * 1000 + i) for i in range(n)]


def test_graph_replays_match_single_queries(gpu_available):
    dgs = [ma.DeviceGraph(g, 0) for g in _graphs(12, 41)]
    pats = [0, 2, 4, 1, 8, 5, 6, 7, 3]
    queries = [(dgs[i], pats[i % len(pats)]) for i in range(len(dgs))]
    queries += [(dg, 0) for dg in dgs[:4]]
    single = [dg.pattern_count(p)[0] for dg, p in queries]
    b = ma.Batch(0, lanes=4, use_graph=True)
    counts, st = b.run(queries)
    assert counts == single and st.path == 101
    for r in range(4):
        counts, st = b.run(queries)
        assert counts == single, r
        assert st.prep_ms == 0.0 and st.path == 101
        assert dgs[r].tc_count()[0] == single[len(dgs) + r]  # eager work between
    b.close()


def test_graph_rerecorded_after_plan_change(gpu_available):
    dgs = [ma.DeviceGraph(g, 0) for g in _graphs(6, 42, scales=(8, 10))]
    want = [dg.tc_count()[0] for dg in dgs]
    b = ma.Batch(0, lanes=2, use_graph=True)
    assert b.run([(dg, 0) for dg in dgs])[0] == want
    assert dgs[0].tc_count(ma.make_tune(chunk_work=256))[0] == want[0]
    counts, st = b.run([(dg, 0) for dg in dgs])
    assert counts == want and st.prep_ms > 0
