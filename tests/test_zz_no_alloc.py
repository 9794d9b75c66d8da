"""mgpm_pattern_count_noalloc (the allocation-free count, scratch in the
library's fixed per-GPU array) against the entry points it varies:
mgpm_tc_count for MGPM_TC and mgpm_pattern_count's per-edge kernels for
P1..P8, on two calls in a row per pattern, the second on a smaller graph
(the array keeps the first call's contents), plus the oracle on the small
graph and the error paths."""
import pytest

import mercury_amd as ma
import oracle

pytestmark = pytest.mark.gpu

PER_EDGE = dict(pattern_census=-1)


def _reference(dg, p):
    if p == 0:
        return dg.tc_count()[0]
    return dg.pattern_count(p, ma.make_tune(**PER_EDGE))[0]


def test_noalloc_matches_entry_points_big_then_small(gpu_available):
    big = ma.DeviceGraph(ma.Graph.rmat(11, 16, seed=81), 0)
    g_small = ma.Graph.rmat(8, 8, seed=82)
    small = ma.DeviceGraph(g_small, 0)
    rp, ci = g_small.csr()
    for p in range(0, 9):
        want_big = _reference(big, p)      # also makes the plan resident
        want_small = _reference(small, p)
        print(f"noalloc P{p}: big then small", flush=True)
        assert big.pattern_count_noalloc(p) == want_big, p
        assert small.pattern_count_noalloc(p) == want_small, p
        want = oracle.tc(rp, ci) if p == 0 else oracle.pattern(rp, ci, p)[0]
        assert want_small == want, p


def test_noalloc_repeat_and_alternate(gpu_available):
    g1 = ma.DeviceGraph(ma.Graph.rmat(10, 8, seed=83), 0)
    g2 = ma.DeviceGraph(ma.Graph.er(300, 3000, seed=84), 0)
    want = {(i, p): _reference(g, p) for i, g in enumerate((g1, g2))
            for p in (0, 2, 6)}
    for _ in range(3):
        for i, g in enumerate((g1, g2)):
            for p in (0, 2, 6):
                assert g.pattern_count_noalloc(p) == want[(i, p)]


def test_noalloc_errors(gpu_available):
    dg = ma.DeviceGraph(ma.Graph.rmat(7, 4, seed=85), 0)
    with pytest.raises(ma.MgpmError) as ei:
        dg.pattern_count_noalloc(3)        # plan not resident yet
    assert ei.value.code == -3
    with pytest.raises(ma.MgpmError) as ei:
        dg.pattern_count_noalloc(9)
    assert ei.value.code == -8
    # a star whose hub needs more slab per workgroup than the array holds:
    # rejected before any launch
    n = 600_000
    star = ma.Graph.from_csr(*_star(n))
    sdg = ma.DeviceGraph(star, 0)
    with pytest.raises(ma.MgpmError) as ei:
        sdg.pattern_count_noalloc(4)
    assert ei.value.code == -9


def _star(n):
    import numpy as np
    rp = np.zeros(n + 1, dtype=np.int64)
    rp[1] = n - 1
    rp[2:] = n - 1 + np.arange(1, n, dtype=np.int64)
    ci = np.zeros(2 * (n - 1), dtype=np.int32)
    ci[:n - 1] = np.arange(1, n, dtype=np.int32)
    return rp, ci
