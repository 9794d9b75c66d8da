"""Summaries of rocprofv3 CSV output for the counting kernels.

  python -m mercury_amd.profsum kernels <rocprof_dir>
      per-kernel calls / average / total device time (kernel_stats.csv)

  python -m mercury_amd.profsum traffic --copy-fetch D --copy-write D \\
         --fetch D --write D --copy-bytes B --passes K [--out FILE]
      HBM-side bytes per counting pass: FETCH_SIZE and WRITE_SIZE (KiB per
      dispatch) of the counting kernels, summed per pass and scaled by the
      factor that makes a plain copy of B bytes read / write exactly B.

  python -m mercury_amd.profsum counters <rocprof_dir>
      sum of each PMC counter per kernel over all dispatches
"""
from __future__ import annotations

import argparse
import collections
import csv
import json
from pathlib import Path

COUNTING = ("tc_wave_tables", "tc_half_tables", "tc_block_tables",
            "tc_global_search", "pattern_", "batch_kernel")


def _short(name: str) -> str:
    name = name.replace("void ", "").replace("(anonymous namespace)::", "")
    name = name.replace("mgpm::", "")
    return name.split("(")[0]


def _find(d: str, suffix: str) -> Path:
    hits = sorted(Path(d).rglob(f"*{suffix}"))
    if not hits:
        raise FileNotFoundError(f"no *{suffix} under {d}")
    return hits[0]


def kernels(d: str) -> list[dict]:
    rows = list(csv.DictReader(open(_find(d, "kernel_stats.csv"))))
    out = []
    for r in rows:
        out.append({"kernel": _short(r["Name"]), "calls": int(r["Calls"]),
                    "avg_ms": float(r["AverageNs"]) / 1e6,
                    "total_ms": float(r["TotalDurationNs"]) / 1e6,
                    "pct": float(r["Percentage"])})
    return out


def counter_per_dispatch(d: str) -> dict:
    """{(dispatch, kernel): {counter: value}} summed over dimensions."""
    agg = collections.defaultdict(lambda: collections.defaultdict(float))
    for r in csv.DictReader(open(_find(d, "counter_collection.csv"))):
        key = (int(r["Dispatch_Id"]), _short(r["Kernel_Name"]))
        agg[key][r["Counter_Name"]] += float(r["Counter_Value"])
    return agg


def _copy_kib(d: str, counter: str) -> float:
    vals = [v[counter] for (_, k), v in counter_per_dispatch(d).items()
            if "copyBuffer" in k and v.get(counter, 0) > 1024]
    return sum(vals) / len(vals)


def _counting_kib(d: str, counter: str) -> float:
    return sum(v.get(counter, 0.0)
               for (_, k), v in counter_per_dispatch(d).items()
               if any(c in k for c in COUNTING))


def traffic(a) -> dict:
    cf = _copy_kib(a.copy_fetch, "FETCH_SIZE")
    cw = _copy_kib(a.copy_write, "WRITE_SIZE")
    f_scale = a.copy_bytes / (cf * 1024)
    w_scale = a.copy_bytes / (cw * 1024)
    fetch = _counting_kib(a.fetch, "FETCH_SIZE") * 1024 * f_scale / a.passes
    write = _counting_kib(a.write, "WRITE_SIZE") * 1024 * w_scale / a.passes
    return {"copy_bytes": a.copy_bytes,
            "copy_fetch_kib": cf, "copy_write_kib": cw,
            "fetch_scale": f_scale, "write_scale": w_scale,
            "fetch_bytes_per_pass": fetch, "write_bytes_per_pass": write,
            "hbm_bytes_per_pass": fetch + write}


def counters(d: str) -> dict:
    per = collections.defaultdict(lambda: collections.defaultdict(float))
    for (_, k), v in counter_per_dispatch(d).items():
        for c, x in v.items():
            per[k][c] += x
    return {k: dict(v) for k, v in per.items()}


def main() -> None:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    k = sub.add_parser("kernels")
    k.add_argument("dir")
    c = sub.add_parser("counters")
    c.add_argument("dir")
    t = sub.add_parser("traffic")
    for f in ("copy_fetch", "copy_write", "fetch", "write"):
        t.add_argument("--" + f.replace("_", "-"), dest=f, required=True)
    t.add_argument("--copy-bytes", type=int, required=True)
    t.add_argument("--passes", type=int, required=True)
    t.add_argument("--out")
    a = ap.parse_args()
    if a.cmd == "kernels":
        for r in kernels(a.dir):
            print(f"{r['kernel']:<40} calls {r['calls']:>5}  avg "
                  f"{r['avg_ms']:10.4f} ms  total {r['total_ms']:10.3f} ms  "
                  f"{r['pct']:6.2f}%")
    elif a.cmd == "counters":
        print(json.dumps(counters(a.dir), indent=1))
    else:
        res = traffic(a)
        text = json.dumps(res, indent=1)
        if a.out:
            Path(a.out).write_text(text + "\n")
        print(text)


if __name__ == "__main__":
    main()
