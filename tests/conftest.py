import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers",
                            "gpu: needs an MI355X (runs through the C ABI)")


def csr_from_edges(n, edges):
    """Simple undirected CSR from an edge iterable (dedup, no self-loops)."""
    adj = [set() for _ in range(n)]
    for a, b in edges:
        if a != b:
            adj[a].add(b)
            adj[b].add(a)
    rp = [0]
    ci = []
    for v in range(n):
        ci.extend(sorted(adj[v]))
        rp.append(len(ci))
    return np.array(rp, dtype=np.int64), np.array(ci, dtype=np.int32)


@pytest.fixture(scope="session")
def gpu_available():
    import mercury_amd as ma
    if ma.device_count() < 1:
        pytest.skip("no GPU visible")
    return True
