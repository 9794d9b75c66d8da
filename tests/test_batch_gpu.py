"""Batched multi-graph launch: every query's count equals its per-graph
(eager) count and the CPU oracle; the captured HIP graph is replayed several
times with eager work and synchronisation between replays."""
import random

import pytest

import mercury_amd as ma
import oracle

pytestmark = [
    pytest.mark.gpu,
    pytest.mark.skip(reason=(
        "the persistent multi-graph batch kernel these tests exercised "
        "timed out on its first GPU run (inside mgpm_batch_run, first "
        "replay of the captured HIP graph) and no cause was found in the "
        "code; that kernel and its graph-capture path are switched off. "
        "The multi-stream batch engine is tested in test_batch_streams_gpu.py")),
]


def _graphs(n, seed, scales=(6, 10), efs=(4, 8, 16)):
    rng = random.Random(seed)
    out = []
    for i in range(n):
        s = rng.randint(*scales)
        ef = rng.choice(efs)
        out.append(ma.Graph.rmat(s, ef, seed=1000 + i))
    return out


def test_batch_mixed_matches_eager_and_oracle(gpu_available):
    graphs = _graphs(24, 1)
    dgs = [ma.DeviceGraph(g, 0) for g in graphs]
    pats = [0, 2, 4, 1, 8, 5, 6, 7, 3]
    queries = [(dgs[i], pats[i % len(pats)]) for i in range(len(dgs))]
    queries += [(dgs[i], 0) for i in range(0, len(dgs), 3)]  # repeats
    eager = [dg.pattern_count(p)[0] for dg, p in queries]
    eager_tc = [dg.tc_count()[0] for dg in dgs]
    b = ma.Batch(0)
    counts, st = b.run(queries)
    assert counts == eager
    for (dg, p), c, g in zip(queries[:6], counts, graphs):
        rp, ci = g.csr()
        want = oracle.tc(rp, ci) if p == 0 else oracle.pattern(rp, ci, p)[0]
        assert c == want
    # replays of the captured graph, eager work and syncs between them
    for r in range(4):
        counts2, st2 = b.run(queries)
        assert counts2 == eager, r
        assert st2.prep_ms == 0.0  # replay, no re-prepare
        assert dgs[r].tc_count()[0] == eager_tc[r]
        assert dgs[r].pattern_count(4)[0] == dgs[r].pattern_count(4)[0]


def test_batch_tc_many_small(gpu_available):
    graphs = _graphs(200, 2, scales=(6, 11))
    dgs = [ma.DeviceGraph(g, 0) for g in graphs]
    eager = [dg.tc_count()[0] for dg in dgs]
    b = ma.Batch(0)
    for _ in range(3):
        counts, _ = b.run([(dg, 0) for dg in dgs])
        assert counts == eager


def test_batch_reprepares_when_plan_changes(gpu_available):
    g = ma.Graph.rmat(10, 16, seed=3)
    dg = ma.DeviceGraph(g, 0)
    want = dg.tc_count()[0]
    b = ma.Batch(0)
    assert b.run([(dg, 0)])[0] == [want]
    # a different tune rebuilds the graph's TC plan; the batch notices
    assert dg.tc_count(ma.make_tune(chunk_work=64, group_work=64))[0] == want
    counts, st = b.run([(dg, 0)])
    assert counts == [want] and st.prep_ms > 0


def test_batch_rejects_bad_pattern(gpu_available):
    g = ma.Graph.rmat(6, 4, seed=1)
    dg = ma.DeviceGraph(g, 0)
    b = ma.Batch(0)
    with pytest.raises(ma.MgpmError) as ei:
        b.run([(dg, 11)])
    assert ei.value.code == -8
