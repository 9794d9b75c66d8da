"""Full-size workloads against oracle counts recorded in
tests/golden/large_counts.json (made by tests/golden/make_large_counts.py):
RMAT-22 and ER-22 triangle counts with the host- and the device-built plan,
and RMAT-18 P1 / P4 by census and by the per-edge kernels; the pattern
configuration (RMAT-20, seed 4) for all eight patterns against the counts
recorded in tests/golden/rmat20_counts.json, with P1 and P4 recounted by
the per-edge kernels; every pattern by census against the per-edge
kernels on RMAT-16."""
import json
from pathlib import Path

import pytest

import mercury_amd as ma

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"
LARGE = json.loads((GOLDEN / "large_counts.json").read_text())
RMAT20 = json.loads((GOLDEN / "rmat20_counts.json").read_text())
KERN = dict(pattern_census=-1)


@pytest.mark.parametrize("key,make", [
    ("tc_rmat22", lambda: ma.Graph.rmat(22, 16, seed=2)),
    ("tc_er22", lambda: ma.Graph.er(1 << 22, 16 << 22, 3)),
])
def test_tc_full_size(gpu_available, key, make):
    g = make()
    for device_plan in (False, True):
        dg = ma.DeviceGraph(g, 0, device_plan=device_plan)
        got = dg.tc_count()[0]
        dg.close()
        print(f"{key} device_plan={device_plan}: {got}", flush=True)
        assert got == LARGE[key]


def test_p1_p4_rmat18(gpu_available):
    dg = ma.DeviceGraph(ma.Graph.rmat(18, 16, seed=4), 0)
    kern = ma.make_tune(pattern_census=-1)
    p1 = dg.pattern_count(1)[0]
    print(f"rmat18 P1 census {p1}", flush=True)
    assert p1 == LARGE["p1_rmat18"]
    assert dg.pattern_count(1, kern)[0] == p1
    p4 = dg.pattern_count(4)[0]
    p4k = dg.pattern_count(4, kern)[0]
    print(f"rmat18 P4 census {p4} kernels {p4k}", flush=True)
    assert p4 == p4k
    if "p4_rmat18" in LARGE:
        assert p4 == LARGE["p4_rmat18"]
    dg.close()


def test_p7_p8_rmat16_census_matches_kernels(gpu_available):
    # the census closed forms against the reference-form per-edge kernels
    # (about 15 s of per-edge kernels) on RMAT-16, ef 16, seed 1
    dg = ma.DeviceGraph(ma.Graph.rmat(16, 16, seed=1), 0)
    kern = ma.make_tune(pattern_census=-1)
    for p in (7, 8):
        a, st = dg.pattern_count(p)
        b = dg.pattern_count(p, kern)[0]
        print(f"P{p} rmat16: census {a} kernels {b}", flush=True)
        assert st.path == 30 + p and a == b
    dg.close()


def test_p7_p8_rmat18_census(gpu_available):
    # P7 on RMAT-18, ef 16, seed 4 by the census; the value is the one the
    # reference-form per-edge kernels return for this graph (77 s of
    # per-edge kernels, too long for this suite)
    dg = ma.DeviceGraph(ma.Graph.rmat(18, 16, seed=4), 0)
    p7, st = dg.pattern_count(7)
    print(f"rmat18 P7 census {p7} in {st.kernel_ms:.1f} ms", flush=True)
    assert st.path == 37 and p7 == 17468419210578
    p8, st = dg.pattern_count(8)
    print(f"rmat18 P8 census {p8} in {st.kernel_ms:.1f} ms", flush=True)
    assert st.path == 38 and p8 == RMAT20["rmat18_seed4"]["P8"]
    dg.close()


@pytest.fixture(scope="module")
def rmat20():
    dg = ma.DeviceGraph(ma.Graph.rmat(20, 16, seed=4), 0)
    yield dg
    dg.close()


@pytest.mark.parametrize("p", range(1, 9))
def test_config3_rmat20_census(gpu_available, rmat20, p):
    """The pattern configuration, every pattern by the census."""
    got, st = rmat20.pattern_count(p)
    print(f"rmat20 P{p} census {got} in {st.kernel_ms:.0f} ms", flush=True)
    assert st.path == 30 + p
    assert got == RMAT20["counts"][f"P{p}"]


@pytest.mark.parametrize("p", [1, 4])
def test_config3_rmat20_kernels(gpu_available, rmat20, p):
    """P1 and P4 of the pattern configuration recounted by the reference's
    per-edge set algebra (about 30 s and 80 s)."""
    got, st = rmat20.pattern_count(p, ma.make_tune(**KERN))
    print(f"rmat20 P{p} kernels {got} in {st.kernel_ms:.0f} ms", flush=True)
    assert got == RMAT20["counts"][f"P{p}"]


@pytest.mark.parametrize("p", [2, 3, 5, 6])
def test_pairs_rmat16_census_matches_kernels(gpu_available, p):
    """P2, P3, P5, P6 by the census against the per-edge kernels on
    RMAT-16, ef 16, seed 1 (17-80 s of per-edge kernels each)."""
    dg = ma.DeviceGraph(ma.Graph.rmat(16, 16, seed=1), 0)
    a, st = dg.pattern_count(p)
    b = dg.pattern_count(p, ma.make_tune(**KERN))[0]
    print(f"P{p} rmat16: census {a} kernels {b}", flush=True)
    assert st.path == 30 + p and a == b
    dg.close()
