"""Triangle-count parity on the GPU: HIP kernels (through the C ABI) versus
the CPU oracle, bit-exact uint64.  Every kernel class runs both with the
partner lists gathered from the DAG (tc_layout = -1) and streamed from
their task-by-task layout (tc_layout = 1)."""
import numpy as np
import pytest

import mercury_amd as ma
import oracle
from conftest import csr_from_edges

pytestmark = pytest.mark.gpu

# Tunes that route every task through one kernel class, so each kernel is
# exercised on every graph: wave-small, wave-medium, workgroup, global.
FORCE = {
    "default": None,
    "small_only": dict(small_max_deg=128, wave_max_deg=-1, block_max_deg=-1),
    "medium_only": dict(small_max_deg=-1, wave_max_deg=1024,
                        block_max_deg=-1),
    "block_only": dict(small_max_deg=-1, wave_max_deg=-1,
                       block_max_deg=16384),
    "global_only": dict(small_max_deg=-1, wave_max_deg=-1, block_max_deg=-1),
    "tiny_chunks": dict(chunk_work=7),
    "items4": dict(tc_items=4),
    "cursor_always": dict(tc_cursor_min=1),
    "cursor_never": dict(tc_cursor_min=-1),
    "items8": dict(tc_items=8),
    "items8_global": dict(tc_items=8, small_max_deg=-1, wave_max_deg=-1,
                          block_max_deg=-1),
    "half": dict(tc_half_waves=1),
    "half_items8": dict(tc_half_waves=1, tc_items=8),
    "half_small_only": dict(tc_half_waves=1, small_max_deg=128,
                            wave_max_deg=-1, block_max_deg=-1),
    "half_tiny_chunks": dict(tc_half_waves=1, chunk_work=7),
    "merge16": dict(tc_merge_max=16),
    "merge64": dict(tc_merge_max=64),
    "merge64_items4": dict(tc_merge_max=64, tc_items=4),
}


def _tune(name, layout=0):
    kw = dict(FORCE[name] or {})
    if layout:
        kw["tc_layout"] = layout
    return ma.make_tune(**kw) if kw else None


def _gpu_tc(g, tune=None):
    dg = ma.DeviceGraph(g, 0)
    c, st = dg.tc_count(tune)
    dg.close()
    return c, st


@pytest.mark.parametrize("scale,ef,seed", [(8, 4, 1), (10, 16, 7),
                                           (12, 8, 3)])
@pytest.mark.parametrize("mode", list(FORCE))
@pytest.mark.parametrize("layout", [-1, 1])
def test_tc_rmat_small(gpu_available, scale, ef, seed, mode, layout):
    g = ma.Graph.rmat(scale, ef, seed=seed)
    rp, ci = g.csr()
    want = oracle.tc(rp, ci)
    dg = ma.DeviceGraph(g, 0)
    got, st = dg.tc_count(_tune(mode, layout))
    print(f"rmat{scale} ef{ef} seed{seed} {mode} layout {layout}: {got} vs "
          f"{want}", flush=True)
    assert got == want
    assert st.path == (2 if layout > 0 else 1)


def test_tc_layout_kept_and_rebuilt(gpu_available):
    """The laid-out lists are built once and reused; a re-plan (new task
    cut) drops them and the next streamed count rebuilds them; the gather
    path still runs beside them; the device-built plan streams too."""
    g = ma.Graph.rmat(13, 16, seed=5)
    rp, ci = g.csr()
    want = oracle.tc(rp, ci)
    dg = ma.DeviceGraph(g, 0)
    lay = ma.make_tune(tc_layout=1)
    c1, s1 = dg.tc_count(lay)
    c2, s2 = dg.tc_count(lay)
    assert c1 == c2 == want and s2.prep_ms <= s1.prep_ms
    assert dg.tc_count(ma.make_tune(tc_layout=-1))[0] == want
    c3, s3 = dg.tc_count(ma.make_tune(tc_layout=1, chunk_work=9))
    assert c3 == want and s3.path == 2
    dg.close()
    dg = ma.DeviceGraph(g, 0, device_plan=True)
    for t in (lay, ma.make_tune(tc_layout=-1), lay):
        assert dg.tc_count(t)[0] == want
    dg.close()


def test_tc_rmat16_seed1(gpu_available):
    """Config: RMAT scale 16, ef 16, seed 1 (the reference's CPU-sized case)."""
    g = ma.Graph.rmat(16, 16, seed=1)
    rp, ci = g.csr()
    want = oracle.tc(rp, ci)
    got, st = _gpu_tc(g)
    assert got == want
    assert st.m == len(ci) // 2


def test_tc_er(gpu_available):
    g = ma.Graph.er(1 << 14, (1 << 14) * 16, seed=3)
    rp, ci = g.csr()
    want = oracle.tc(rp, ci)
    assert _gpu_tc(g)[0] == want
    # every owner small: all tasks merged, from both list sources; and the
    # small class two tasks per wave
    for layout in (-1, 1):
        assert _gpu_tc(g, ma.make_tune(tc_merge_max=64,
                                       tc_layout=layout))[0] == want
    for items in (4, 8):
        assert _gpu_tc(g, ma.make_tune(tc_half_waves=1, tc_layout=1,
                                       tc_items=items))[0] == want


@pytest.mark.parametrize("n", [1, 2, 3, 5, 40])
def test_tc_complete_graph(gpu_available, n):
    rp, ci = csr_from_edges(n, [(a, b) for a in range(n)
                                for b in range(a + 1, n)])
    g = ma.Graph.from_csr(rp, ci)
    want = n * (n - 1) * (n - 2) // 6
    for mode in FORCE:
        assert _gpu_tc(g, _tune(mode))[0] == want


def test_tc_empty_and_edgeless(gpu_available):
    g = ma.Graph.from_csr(np.zeros(6, dtype=np.int64),
                          np.zeros(0, dtype=np.int32))
    assert _gpu_tc(g)[0] == 0


def test_tc_dense_hub_uses_big_tables(gpu_available):
    """A clique of 1100 vertices gives d+ up to 1099 (> 1024): the
    workgroup-table class and, when forced, the global binary search."""
    n = 1100
    rp, ci = csr_from_edges(n, [(a, b) for a in range(n)
                                for b in range(a + 1, n)])
    g = ma.Graph.from_csr(rp, ci)
    want = n * (n - 1) * (n - 2) // 6
    for mode in ("default", "global_only", "tiny_chunks", "merge64", "half"):
        for layout in (-1, 1):
            assert _gpu_tc(g, _tune(mode, layout))[0] == want, (mode, layout)


def test_tc_repeatable(gpu_available):
    g = ma.Graph.rmat(14, 16, seed=11)
    dg = ma.DeviceGraph(g, 0)
    counts = {dg.tc_count()[0] for _ in range(10)}
    assert len(counts) == 1


@pytest.mark.parametrize("items", [4, 8])
def test_tc_long_lists_owner_cursor(gpu_available, items):
    """Dense G(n, p): partner lists average well over 32 items, so the
    streaming loop takes its per-lane owner-cursor path (not the start-mark
    path) in every class it is forced into."""
    g = ma.Graph.er(3000, 450000, seed=13)
    rp, ci = g.csr()
    want = oracle.tc(rp, ci)
    for mode in ("default", "small_only", "medium_only", "block_only",
                 "global_only", "tiny_chunks"):
        for layout in (-1, 1):
            kw = dict(FORCE[mode] or {})
            kw["tc_items"] = items
            kw["tc_layout"] = layout
            assert _gpu_tc(g, ma.make_tune(**kw))[0] == want, (mode, layout)
