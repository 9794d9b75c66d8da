"""Writes tests/golden/large_counts.json: counts of the full-size synthetic
workloads computed by the CPU oracle (oracle/oracle.c), for the GPU tests
that run those workloads (tests/test_large_gpu.py).

  tc_rmat22  triangles, RMAT scale 22, edge factor 16, seed 2
  tc_er22    triangles, Erdos-Renyi n = 2^22, 16 * 2^22 draws, seed 3
  p1_rmat18  P1, RMAT scale 18, edge factor 16, seed 4
  p4_rmat18  P4, RMAT scale 18, edge factor 16, seed 4

Slow: the P1 count takes about 33 minutes on 8 host threads, P4 longer.
Run from the repository root:  python tests/golden/make_large_counts.py
[key ...]  (only the keys named; the others keep their recorded values).
"""
from __future__ import annotations

import json
import sys
import time
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))

import mercury_amd as ma  # noqa: E402
import oracle  # noqa: E402


def main() -> None:
    path = HERE / "large_counts.json"
    out = json.loads(path.read_text()) if path.exists() else {}
    want = set(sys.argv[1:])
    for key, make, what in (
            ("tc_rmat22", lambda: ma.Graph.rmat(22, 16, seed=2), "tc"),
            ("tc_er22", lambda: ma.Graph.er(1 << 22, 16 << 22, 3), "tc"),
            ("p1_rmat18", lambda: ma.Graph.rmat(18, 16, seed=4), 1),
            ("p4_rmat18", lambda: ma.Graph.rmat(18, 16, seed=4), 4)):
        if want and key not in want:
            continue
        rp, ci = make().csr()
        t0 = time.perf_counter()
        c = oracle.tc(rp, ci) if what == "tc" else oracle.pattern(rp, ci, what)[0]
        print(f"{key}: {c} ({time.perf_counter() - t0:.0f} s)", flush=True)
        out[key] = int(c)
        path.write_text(json.dumps(out, indent=1) + "\n")


if __name__ == "__main__":
    main()
