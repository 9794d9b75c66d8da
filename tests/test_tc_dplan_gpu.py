"""Triangle counts with the plan built on the device
(``mgpm_dgraph_upload_ex(..., MGPM_UPLOAD_DEVICE_PLAN, ...)``): the same
counts as the host-built plan and the CPU oracle, bit-exact, and the same
plan totals (DAG arcs, wedge checks S+, probes)."""
import json
from pathlib import Path

import pytest

import mercury_amd as ma
import oracle
from conftest import csr_from_edges

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"

TUNES = {
    "default": None,
    "small_only": dict(small_max_deg=128, wave_max_deg=-1, block_max_deg=-1),
    "medium_only": dict(small_max_deg=-1, wave_max_deg=1024,
                        block_max_deg=-1),
    "block_only": dict(small_max_deg=-1, wave_max_deg=-1,
                       block_max_deg=16384),
    "global_only": dict(small_max_deg=-1, wave_max_deg=-1, block_max_deg=-1),
    "tiny_chunks": dict(chunk_work=7),
    "tiny_groups": dict(group_work=1),
}


def _tune(name):
    kw = TUNES[name]
    return ma.make_tune(**kw) if kw else None


def _both(g, tune=None):
    """(device-plan count, stats), (host-plan count, stats)."""
    dd = ma.DeviceGraph(g, 0, device_plan=True)
    dh = ma.DeviceGraph(g, 0)
    a = dd.tc_count(tune)
    b = dh.tc_count(tune)
    dd.close()
    dh.close()
    return a, b


def _check(g, want, tune=None, label=""):
    (cd, sd), (ch, sh) = _both(g, tune)
    print(f"{label}: device plan {cd}, host plan {ch}, want {want}", flush=True)
    assert cd == want and ch == want
    assert (sd.m_plus, sd.s_plus, sd.probes) == (sh.m_plus, sh.s_plus,
                                                 sh.probes)


@pytest.mark.parametrize("scale,ef,seed", [(8, 4, 1), (10, 16, 7),
                                           (12, 8, 3), (14, 16, 2)])
@pytest.mark.parametrize("mode", list(TUNES))
def test_dplan_rmat(gpu_available, scale, ef, seed, mode):
    g = ma.Graph.rmat(scale, ef, seed=seed)
    rp, ci = g.csr()
    _check(g, oracle.tc(rp, ci), _tune(mode),
           f"rmat{scale} ef{ef} seed{seed} {mode}")


def test_dplan_rmat16_seed1(gpu_available):
    g = ma.Graph.rmat(16, 16, seed=1)
    rp, ci = g.csr()
    _check(g, oracle.tc(rp, ci), None, "rmat16 seed1")


def test_dplan_er(gpu_available):
    g = ma.Graph.er(1 << 14, (1 << 14) * 16, seed=3)
    rp, ci = g.csr()
    _check(g, oracle.tc(rp, ci), None, "er14")


def test_dplan_golden(gpu_available):
    golden = json.loads((GOLDEN / "golden.json").read_text())
    for name, rec in golden.items():
        g = ma.Graph.load(str(GOLDEN / name / "graph"))
        _check(g, rec["counts"]["TC"], None, name)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 40])
def test_dplan_complete_graph(gpu_available, n):
    rp, ci = csr_from_edges(n, [(a, b) for a in range(n)
                                for b in range(a + 1, n)])
    g = ma.Graph.from_csr(rp, ci)
    for mode in TUNES:
        _check(g, n * (n - 1) * (n - 2) // 6, _tune(mode), f"K{n} {mode}")


def test_dplan_edgeless_and_star(gpu_available):
    for n, edges in ((4, []), (9, [(0, k) for k in range(1, 9)]),
                     (2, [(0, 1)])):
        rp, ci = csr_from_edges(n, edges)
        _check(ma.Graph.from_csr(rp, ci), 0, None, f"n{n} e{len(edges)}")


def test_dplan_replan_and_batch(gpu_available):
    """A tune on a device-planned graph re-plans on the device; batches over
    device-planned graphs give the single-query counts."""
    gs = [ma.Graph.rmat(s, 8, seed=s) for s in (9, 10, 11)]
    dgs = [ma.DeviceGraph(g, 0, device_plan=True) for g in gs]
    want = [oracle.tc(*g.csr()) for g in gs]
    for dg, w in zip(dgs, want):
        assert dg.tc_count(ma.make_tune(chunk_work=64))[0] == w
        assert dg.tc_count(None)[0] == w
    b = ma.Batch(0, lanes=2)
    counts, _ = b.run([(dg, 0) for dg in dgs])
    assert list(counts) == want
    for dg in dgs:
        dg.close()
