"""Continuous-batching server: every request's count is the same alone and
in the mix; admission by HBM footprint; open-loop replay; the one-launch-
per-request baseline; error isolation."""
import random

import numpy as np
import pytest

import mercury_amd as ma

pytestmark = pytest.mark.gpu

PATTERNS = list(range(0, 9))


def _pool(n, seed, scales=(6, 10)):
    rng = random.Random(seed)
    return [ma.Graph.rmat(rng.randint(*scales), rng.choice((4, 8, 16)),
                          seed=seed * 1000 + i) for i in range(n)]


def _alone(graphs):
    dgs = [ma.DeviceGraph(g, 0) for g in graphs]
    return {(i, p): dgs[i].pattern_count(p)[0]
            for i in range(len(graphs)) for p in PATTERNS}


def test_server_results_same_alone_and_in_mix(gpu_available):
    graphs = _pool(10, 3)
    alone = _alone(graphs)
    srv = ma.Server(0, slots=3)
    ids = [srv.add_graph(g) for g in graphs]
    rng = random.Random(4)
    reqs = [(rng.randrange(len(ids)), rng.choice(PATTERNS))
            for _ in range(300)]
    tickets = [srv.submit(ids[g], p) for g, p in reqs]
    batched = 0
    for t, (g, p) in zip(tickets, reqs):
        r = srv.wait(t)
        assert r.status == 0
        assert r.count == alone[(g, p)], (g, p)
        batched = max(batched, r.batch_size)
    assert batched > 1  # requests really were served together
    st = srv.stats()
    assert st.completed == 300 and st.failed == 0
    assert st.launches < 300
    srv.close()


def test_server_replay_poisson(gpu_available):
    graphs = _pool(8, 5)
    alone = _alone(graphs)
    srv = ma.Server(0, slots=3)
    ids = [srv.add_graph(g) for g in graphs]
    rng = np.random.default_rng(6)
    t = np.cumsum(rng.exponential(1e6 / 2000.0, size=400))  # 2000 req/s
    sched = [(int(rng.integers(len(ids))), int(rng.integers(0, 9)), int(x))
             for x in t]
    out = srv.replay(sched)
    for (g, p, _), r in zip(sched, out):
        assert r.status == 0
        assert r.count == alone[(g, p)]
        assert r.latency_ms > 0
    srv.close()


def test_server_one_launch_per_request_baseline(gpu_available):
    graphs = _pool(4, 7)
    alone = _alone(graphs)
    srv = ma.Server(0, slots=1, max_batch=1)
    ids = [srv.add_graph(g) for g in graphs]
    sched = [(i % 4, (3 * i) % 9, 200 * i) for i in range(60)]
    out = srv.replay(sched)
    for (g, p, _), r in zip(sched, out):
        assert r.status == 0 and r.batch_size == 1
        assert r.count == alone[(g, p)]
    assert srv.stats().launches == 60
    srv.close()


def test_server_admission_by_footprint(gpu_available):
    small = ma.Graph.rmat(6, 4, seed=1)
    big = ma.Graph.rmat(16, 16, seed=2)  # tens of MiB resident
    srv = ma.Server(0, hbm_budget=8 << 20)  # 8 MiB
    gid = srv.add_graph(small)
    with pytest.raises(ma.MgpmError) as ei:
        srv.add_graph(big)
    assert ei.value.code == -9  # MGPM_ERR_CAPACITY
    # the server keeps serving what it admitted
    r = srv.wait(srv.submit(gid, 0))
    assert r.status == 0
    assert r.count == ma.DeviceGraph(small, 0).tc_count()[0]
    st = srv.stats()
    assert st.resident_bytes <= st.budget_bytes
    srv.close()


def test_server_rejects_bad_requests(gpu_available):
    srv = ma.Server(0)
    gid = srv.add_graph(ma.Graph.rmat(6, 4, seed=1))
    with pytest.raises(ma.MgpmError) as ei:
        srv.submit(gid + 5, 0)
    assert ei.value.code == -11
    with pytest.raises(ma.MgpmError) as ei:
        srv.submit(gid, 12)
    assert ei.value.code == -8
    with pytest.raises(ma.MgpmError) as ei:
        srv.wait(123456)
    assert ei.value.code == -11
    srv.close()


def _lane_bytes(g, p):
    """Workspace one census query takes on one lane (a one-lane batch)."""
    b = ma.Batch(0, lanes=1)
    dg = ma.DeviceGraph(g, 0)
    b.run([(dg, p)])
    return b.footprint()


def test_server_waits_for_memory_and_fails_oversized_alone(gpu_available):
    """A budget that holds one small census workspace at a time: requests
    wait for memory to free instead of failing (the stats count the
    waits), every one of them completes with its single-query count, and a
    request whose workspace cannot fit even with every idle workspace
    released fails alone with MGPM_ERR_CAPACITY."""
    small = [ma.Graph.rmat(8, 8, seed=90 + i) for i in range(4)]
    big = ma.Graph.rmat(14, 16, seed=95)
    w_small = max(_lane_bytes(g, 2) for g in small)
    w_big = _lane_bytes(big, 2)
    assert w_big > 3 * w_small
    probe = ma.Server(0, slots=2, lanes=1)
    for g in small + [big]:
        probe.add_graph(g)
    resident = probe.stats().resident_bytes
    probe.close()
    srv = ma.Server(0, slots=2, lanes=1,
                    hbm_budget=resident + (3 * w_small) // 2)
    ids = [srv.add_graph(g) for g in small + [big]]
    alone = {(i, p): ma.DeviceGraph(g, 0).pattern_count(p)[0]
             for i, g in enumerate(small) for p in (0, 2, 5)}
    reqs = [(i % 4, (0, 2, 5)[i % 3]) for i in range(24)]
    tickets = [srv.submit(ids[g], p) for g, p in reqs[:12]]
    big_ticket = srv.submit(ids[4], 2)
    tickets += [srv.submit(ids[g], p) for g, p in reqs[12:]]
    for t, (g, p) in zip(tickets, reqs):
        r = srv.wait(t)
        assert r.status == 0, (g, p)
        assert r.count == alone[(g, p)], (g, p)
    with pytest.raises(ma.MgpmError) as ei:
        srv.wait(big_ticket)
    assert ei.value.code == -9
    st = srv.stats()
    assert st.completed == 24 and st.failed == 1
    assert st.deferred_failed == 1 and st.deferred >= 1
    assert st.resident_bytes + st.workspace_bytes <= st.budget_bytes
    srv.close()
