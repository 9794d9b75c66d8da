"""Batched multi-graph counting (mgpm_batch_run): every query's count equals
its single-query count and the CPU oracle, whatever the number of lanes,
repeated runs, repeated graphs and patterns in one batch, and plan changes
between runs."""
import random

import pytest

import mercury_amd as ma
import oracle

pytestmark = pytest.mark.gpu


def _graphs(n, seed, scales=(6, 10), efs=(4, 8, 16)):
    rng = random.Random(seed)
    return [ma.Graph.rmat(rng.randint(*scales), rng.choice(efs),
                          seed=seed * 1000 + i) for i in range(n)]


@pytest.mark.parametrize("lanes", [1, 4])
def test_batch_mixed_matches_single_and_oracle(gpu_available, lanes):
    graphs = _graphs(18, 1, scales=(6, 9))
    dgs = [ma.DeviceGraph(g, 0) for g in graphs]
    pats = [0, 2, 4, 1, 8, 5, 6, 7, 3]
    queries = [(dgs[i], pats[i % len(pats)]) for i in range(len(dgs))]
    queries += [(dgs[i], 0) for i in range(0, len(dgs), 3)]  # same graph twice
    queries += [(dgs[1], 2), (dgs[1], 4)]                    # and again
    single = [dg.pattern_count(p)[0] for dg, p in queries]
    single_tc = [dg.tc_count()[0] for dg in dgs[:3]]
    b = ma.Batch(0, lanes=lanes)
    counts, st = b.run(queries)
    assert counts == single
    for (dg, p), c, g in zip(queries[:4], counts, graphs):
        rp, ci = g.csr()
        want = oracle.tc(rp, ci) if p == 0 else oracle.pattern(rp, ci, p)[0]
        assert c == want, p
    for r in range(3):  # reruns, with single-query work between them
        counts2, st2 = b.run(queries)
        assert counts2 == single, r
        assert st2.prep_ms == 0.0
        assert dgs[r].tc_count()[0] == single_tc[r]
        assert dgs[r].pattern_count(4)[0] == dgs[r].pattern_count(4)[0]
    b.close()


def test_batch_tc_many_small(gpu_available):
    graphs = _graphs(200, 2, scales=(6, 11))
    dgs = [ma.DeviceGraph(g, 0) for g in graphs]
    single = [dg.tc_count()[0] for dg in dgs]
    b = ma.Batch(0)
    for _ in range(3):
        counts, _ = b.run([(dg, 0) for dg in dgs])
        assert counts == single


def test_batch_tc_large_and_small_together(gpu_available):
    big = ma.Graph.rmat(16, 16, seed=1)
    rp, ci = big.csr()
    want_big = oracle.tc(rp, ci)
    small = _graphs(40, 3, scales=(6, 10))
    dgs = [ma.DeviceGraph(big, 0)] + [ma.DeviceGraph(g, 0) for g in small]
    single = [dg.tc_count()[0] for dg in dgs]
    assert single[0] == want_big
    counts, _ = ma.Batch(0, lanes=3).run([(dg, 0) for dg in dgs])
    assert counts == single


def test_batch_replans_when_plan_changes(gpu_available):
    g = ma.Graph.rmat(10, 16, seed=3)
    dg = ma.DeviceGraph(g, 0)
    want = dg.tc_count()[0]
    b = ma.Batch(0)
    assert b.run([(dg, 0)])[0] == [want]
    # a different tune rebuilds the graph's TC plan; the batch notices
    assert dg.tc_count(ma.make_tune(chunk_work=64, group_work=64))[0] == want
    counts, st = b.run([(dg, 0)])
    assert counts == [want] and st.prep_ms > 0


def test_batch_empty_and_bad_pattern(gpu_available):
    g = ma.Graph.rmat(6, 4, seed=1)
    dg = ma.DeviceGraph(g, 0)
    b = ma.Batch(0)
    assert b.run([])[0] == []
    with pytest.raises(ma.MgpmError) as ei:
        b.run([(dg, 11)])
    assert ei.value.code == -8
    # still usable afterwards
    assert b.run([(dg, 0)])[0] == [dg.tc_count()[0]]


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_after_graph_freed_and_replaced(gpu_available, use_graph):
    """A graph freed and replaced by another (glibc often hands the new
    handle the freed one's address) must not match the batch's recorded
    plan: the rerun counts the new graph."""
    g1 = ma.Graph.rmat(9, 8, seed=71)
    g2 = ma.Graph.rmat(10, 16, seed=72)
    b = ma.Batch(0, lanes=2, use_graph=use_graph)
    for _ in range(3):
        dg = ma.DeviceGraph(g1, 0)
        c1, _ = b.run([(dg, 0), (dg, 4)])
        assert c1 == [dg.tc_count()[0], dg.pattern_count(4)[0]]
        dg.close()
        dg = ma.DeviceGraph(g2, 0)
        c2, _ = b.run([(dg, 0), (dg, 4)])
        rp, ci = g2.csr()
        assert c2 == [oracle.tc(rp, ci), oracle.pattern(rp, ci, 4)[0]]
        dg.close()
