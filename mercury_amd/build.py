"""Builds the native library ``mercury_amd/libmgpm.so`` and the CLIs.

HIP sources (``*.hip``) are compiled by hipcc for gfx950; host C++ sources
(``*.cc``) by g++ with OpenMP; everything is linked into one shared library
with a C ABI (``include/mgpm.h``).  The CLIs ``tc_challenge`` and ``sm`` are
linked against it into ``mercury_amd/bin/``.  Outputs stay in-tree.

Usage: ``python -m mercury_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "libmgpm.so"
BIN = PKG / "bin"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
ARCH = os.environ.get("MGPM_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # Packed FP32 ops are kept out of every code object.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-fopenmp", "-Wall",
             "-D__HIP_PLATFORM_AMD__"]

# Translation units of the library (CLI mains are separate).
LIB_SOURCES = [
    "host_graph.cc", "tc_plan.cc", "pattern_plan.cc", "capi.cc",
    "patterns.cc", "batch.cc", "autotune.cc", "server.cc", "dag.cc",
    "tc_dplan.cc", "census.cc", "tc_kernels.hip", "pattern_kernels.hip",
    "dag_kernels.hip", "tc_dplan_kernels.hip", "census_kernels.hip",
    "census5_kernels.hip", "noalloc.hip",
]
CLI_SOURCES = {"tc_challenge": "cli_tc_challenge.cc", "sm": "cli_sm.cc",
               "convert": "cli_convert.cc"}


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                          text=True)
    out = "\n".join(l for l in proc.stdout.splitlines()
                    if "packed-fp32-ops" not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"build step failed ({proc.returncode}): "
                           f"{' '.join(cmd)}\n{out}")
    if out.strip():
        print(out, file=sys.stderr)


def _compile(src: str) -> Path:
    path = CSRC / src
    obj = OBJ / (path.name + ".o")
    inc = [f"-I{ROOT / 'include'}", f"-I{CSRC}"]
    if path.suffix == ".hip":
        cmd = [str(ROCM / "bin" / "hipcc"), *HIPCC_FLAGS, *inc, "-c",
               str(path), "-o", str(obj)]
    else:
        cmd = ["g++", *CXX_FLAGS, *inc, f"-I{ROCM / 'include'}", "-c",
               str(path), "-o", str(obj)]
    _run(cmd)
    return obj


def build(verbose: bool = True) -> Path:
    """Compiles every source (always) and links lib + CLIs; returns LIB."""
    OBJ.mkdir(exist_ok=True)
    BIN.mkdir(exist_ok=True)
    srcs = LIB_SOURCES + [s for s in CLI_SOURCES.values()
                          if (CSRC / s).exists()]
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        objs = dict(zip(srcs, ex.map(_compile, srcs)))
    link = ["-L" + str(ROCM / "lib"), "-lamdhip64", "-fopenmp", "-pthread",
            f"-Wl,-rpath,{ROCM / 'lib'}"]
    tmp = LIB.with_suffix(".so.tmp")
    _run(["g++", "-shared", "-o", str(tmp),
          *[str(objs[s]) for s in LIB_SOURCES], *link])
    os.replace(tmp, LIB)
    for name, src in CLI_SOURCES.items():
        if src not in objs:
            continue
        _run(["g++", "-o", str(BIN / name), str(objs[src]), str(LIB), *link,
              f"-Wl,-rpath,$ORIGIN/.."])
    if verbose:
        print(f"built {LIB.relative_to(ROOT)}", file=sys.stderr)
    return LIB


def clean() -> None:
    shutil.rmtree(OBJ, ignore_errors=True)


if __name__ == "__main__":
    build()
