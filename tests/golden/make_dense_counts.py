"""Writes tests/golden/dense_counts.json: P1, P4, P7 and P8 of dense random
graphs, computed by the CPU oracle (oracle/oracle.c), for the GPU census tests
(tests/test_census5_gpu.py).  An edge of G(n, p) lies on about n p^2
triangles (49 to 128 here), so the census's triangle sets cross 64 (its
multi-word adjacency rows; for P1 and P4 its K4 step intersects per
triangle once a set exceeds 16) and, at p = 0.8, about half the common
neighbourhoods exceed 128 (its wave-per-triangle path), on random rather
than clique-shaped neighbourhoods.

Each graph is G(n, p) drawn by dense_edges() below (Python's random with
the given seed, pairs a < b in order).  Takes about a minute on 8 host
threads.  Run from the repository root:
    python tests/golden/make_dense_counts.py
"""
from __future__ import annotations

import json
import random
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(HERE.parent))

import oracle  # noqa: E402
from conftest import csr_from_edges  # noqa: E402

CASES = [(300, 0.5, 1), (200, 0.8, 2), (400, 0.35, 3)]


def dense_edges(n: int, p: float, seed: int) -> list[tuple[int, int]]:
    rng = random.Random(seed)
    return [(a, b) for a in range(n) for b in range(a + 1, n)
            if rng.random() < p]


def main() -> None:
    out = []
    for n, p, seed in CASES:
        rp, ci = csr_from_edges(n, dense_edges(n, p, seed))
        row = {"n": n, "p": p, "seed": seed, "edges": int(len(ci) // 2)}
        for pat in (1, 4, 7, 8):
            row[f"P{pat}"] = oracle.pattern(rp, ci, pat)[0]
        print(row, flush=True)
        out.append(row)
    (HERE / "dense_counts.json").write_text(json.dumps(out, indent=1) + "\n")


if __name__ == "__main__":
    main()
