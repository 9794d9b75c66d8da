"""Writes the golden fixtures under tests/golden/.

Each fixture is a small graph in the reference's on-disk format
(<name>/graph.meta.txt, graph.vertex.bin, graph.edge.bin; graph.cc:18-35)
plus its counts in golden.json: triangles and P1..P8 from the CPU oracle
(oracle/oracle.c), each cross-checked here against exhaustive brute force
(n <= 64) through the multiplicities of oracle.expected_patterns, and
against the closed forms recorded with the named graphs.

The reference ships no fixtures or golden vectors of its own, so these are
drift fixtures: they pin the oracle and the GPU path to fixed numbers.

Run from the repository root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import itertools
import json
import sys
from math import comb
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402


def csr(n, edges):
    adj = [set() for _ in range(n)]
    for a, b in edges:
        if a != b:
            adj[a].add(b)
            adj[b].add(a)
    rp, ci = [0], []
    for v in range(n):
        ci.extend(sorted(adj[v]))
        rp.append(len(ci))
    return np.array(rp, dtype=np.int64), np.array(ci, dtype=np.int32)


def complete(n):
    return n, list(itertools.combinations(range(n), 2))


def star(k):
    return k + 1, [(0, i) for i in range(1, k + 1)]


def cycle(k):
    return k, [(i, (i + 1) % k) for i in range(k)]


def path(k):
    return k, [(i, i + 1) for i in range(k - 1)]


def wheel(k):  # hub 0, rim 1..k
    return k + 1, [(0, i) for i in range(1, k + 1)] + \
        [(i, i % k + 1) for i in range(1, k + 1)]


def kab(a, b):
    return a + b, [(i, a + j) for i in range(a) for j in range(b)]


def petersen():
    outer = [(i, (i + 1) % 5) for i in range(5)]
    spokes = [(i, i + 5) for i in range(5)]
    inner = [(5 + i, 5 + (i + 2) % 5) for i in range(5)]
    return 10, outer + spokes + inner


def gnp(n, p, seed):
    rng = np.random.default_rng(seed)
    return n, [(a, b) for a in range(n) for b in range(a + 1, n)
               if rng.random() < p]


def rmat_edges(scale, ef, seed):
    """Small RMAT (a,b,c,d = 0.57,0.19,0.19,0.05) with numpy's PCG64; only
    used to shape fixtures, not the library's generator."""
    rng = np.random.default_rng(seed)
    n = 1 << scale
    e = []
    for _ in range(ef * n):
        s = d = 0
        for lvl in range(scale):
            r = rng.random()
            bit = 1 << (scale - 1 - lvl)
            if r < 0.57:
                pass
            elif r < 0.76:
                d |= bit
            elif r < 0.95:
                s |= bit
            else:
                s |= bit
                d |= bit
        e.append((s, d))
    return n, e


# name -> (graph, closed-form expectations {pattern id: count})
FIXTURES = {
    "k6": (complete(6), {0: comb(6, 3), **{p: 0 for p in range(1, 9)}}),
    "star7": (star(7), {0: 0, 1: comb(7, 3), 3: 4 * comb(7, 4) + 3 * comb(7, 3),
                        2: 0, 4: 0, 5: 0, 6: 0, 7: 0, 8: 0}),
    "cycle9": (cycle(9), {0: 0, 1: 0, 4: 9}),
    "path8": (path(8), {0: 0, 1: 0, 4: 5}),
    "wheel8": (wheel(8), {0: 8}),
    "k34": (kab(3, 4), {0: 0}),
    "petersen": (petersen(), {0: 0}),
    "gnp40": (gnp(40, 0.25, 7), {}),
    "gnp60": (gnp(60, 0.12, 11), {}),
    "rmat7": (rmat_edges(7, 8, 5), {}),
    "rmat9": (rmat_edges(9, 8, 6), {}),
}


def main() -> None:
    golden = {}
    for name, ((n, edges), closed) in FIXTURES.items():
        rp, ci = csr(n, edges)
        d = HERE / name
        d.mkdir(exist_ok=True)
        md = int(np.diff(rp).max()) if n else 0
        (d / "graph.meta.txt").write_text(
            f"{n} {len(ci)} 4 8 1 2 {md} 0 0 0\n")
        rp.astype("<i8").tofile(d / "graph.vertex.bin")
        ci.astype("<i4").tofile(d / "graph.edge.bin")
        counts = {"TC": oracle.tc(rp, ci)}
        for p in range(1, 9):
            counts[f"P{p}"] = oracle.pattern(rp, ci, p)[0]
        if n <= 64:
            exp = oracle.expected_patterns(oracle.brute(rp, ci))
            for p, v in exp.items():
                key = "TC" if p == 0 else f"P{p}"
                assert counts[key] == v, (name, key, counts[key], v)
        for p, v in closed.items():
            key = "TC" if p == 0 else f"P{p}"
            assert counts[key] == v, (name, key, counts[key], v)
        golden[name] = {"n": n, "nnz": int(len(ci)), "counts": counts}
        print(name, golden[name])
    (HERE / "golden.json").write_text(json.dumps(golden, indent=1) + "\n")


if __name__ == "__main__":
    main()
