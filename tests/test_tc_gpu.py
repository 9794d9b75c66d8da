"""Triangle-count parity on the GPU: HIP kernels (through the C ABI) versus
the CPU oracle, bit-exact uint64."""
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
}


def _tune(name):
    kw = FORCE[name]
    return ma.make_tune(**kw) if kw else None


def _gpu_tc(g, tune=None):
    dg = ma.DeviceGraph(g, 0)
    c, st = dg.tc_count(tune)
    dg.close()
    return c, st


@pytest.mark.parametrize("scale,ef,seed", [(8, 4, 1), (10, 16, 7),
                                           (12, 8, 3)])
@pytest.mark.parametrize("mode", list(FORCE))
def test_tc_rmat_small(gpu_available, scale, ef, seed, mode):
    g = ma.Graph.rmat(scale, ef, seed=seed)
    rp, ci = g.csr()
    want = oracle.tc(rp, ci)
    dg = ma.DeviceGraph(g, 0)
    got, st = dg.tc_count(_tune(mode))
    print(f"rmat{scale} ef{ef} seed{seed} {mode}: {got} vs {want}", flush=True)
    assert got == want


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
    assert _gpu_tc(g)[0] == oracle.tc(rp, ci)


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
    for mode in ("default", "global_only", "tiny_chunks"):
        assert _gpu_tc(g, _tune(mode))[0] == want


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
        kw = dict(FORCE[mode] or {})
        kw["tc_items"] = items
        assert _gpu_tc(g, ma.make_tune(**kw))[0] == want, mode
