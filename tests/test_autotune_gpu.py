"""Autotuner: every candidate it times must give the same count (it fails
otherwise), the chosen knobs stay resident and keep the count bit-exact
against the CPU oracle, and batches with per-size-class tuning match the
untuned single-query counts."""
import random

import pytest

import mercury_amd as ma
import oracle

pytestmark = pytest.mark.gpu


def test_autotune_tc_keeps_count_and_plan(gpu_available):
    g = ma.Graph.rmat(13, 16, seed=21)
    rp, ci = g.csr()
    want = oracle.tc(rp, ci)
    dg = ma.DeviceGraph(g, 0)
    t, ms = dg.autotune(0, reps=2)
    assert ms > 0
    assert t.tc_items in (4, 8)
    # NULL tune now runs the tuned plan
    c, st = dg.tc_count()
    assert c == want
    # an explicit tune still overrides it, and the count is unchanged
    assert dg.tc_count(ma.make_tune(small_max_deg=-1, wave_max_deg=-1,
                                    block_max_deg=-1))[0] == want


def test_autotune_pattern_grid(gpu_available):
    g = ma.Graph.rmat(9, 8, seed=22)
    rp, ci = g.csr()
    dg = ma.DeviceGraph(g, 0)
    for p in (1, 2, 4):
        want = oracle.pattern(rp, ci, p)[0]
        t, ms = dg.autotune(p, reps=2)
        assert t.grid_blocks >= 0 and ms > 0
        # the tuned knobs drive the per-edge kernels: count them that way
        assert dg.pattern_count(p, ma.make_tune(pattern_census=-1))[0] == want
        assert dg.pattern_count(p)[0] == want


def test_autotune_rejects_bad_pattern(gpu_available):
    dg = ma.DeviceGraph(ma.Graph.rmat(6, 4, seed=1), 0)
    with pytest.raises(ma.MgpmError) as ei:
        dg.autotune(9)
    assert ei.value.code == -8


def test_batch_autotune_size_classes(gpu_available):
    rng = random.Random(23)
    graphs = [ma.Graph.rmat(rng.randint(7, 12), rng.choice((4, 16)),
                            seed=2300 + i) for i in range(24)]
    plain = [ma.DeviceGraph(g, 0) for g in graphs]
    single = [dg.tc_count()[0] for dg in plain]
    dgs = [ma.DeviceGraph(g, 0) for g in graphs]
    b = ma.Batch(0, autotune=True)
    counts, st = b.run([(dg, 0) for dg in dgs] + [(dgs[0], 4)])
    assert counts[:-1] == single
    assert counts[-1] == plain[0].pattern_count(4)[0]
    # second run: no re-tune, no re-plan
    counts2, st2 = b.run([(dg, 0) for dg in dgs] + [(dgs[0], 4)])
    assert counts2 == counts and st2.prep_ms == 0.0
