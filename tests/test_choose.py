"""The census's binomial helpers (mercury_amd/csrc/choose.hpp) are exact
modulo 2^64 far beyond where the plain x(x-1)(x-2)/6 wraps (x ~ 2.6 M):
compiled for the host and compared with Python's exact integers."""
import subprocess
from math import comb
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent

DRIVER = r"""
#include <cstdio>
#include <cstdlib>
#include "choose.hpp"
int main(int argc, char **argv) {
  for (int i = 1; i < argc; ++i) {
    unsigned long long x = std::strtoull(argv[i], nullptr, 10);
    std::printf("%llu %llu %llu\n",
                (unsigned long long)mgpm::choose2(x),
                (unsigned long long)mgpm::choose3(x),
                (unsigned long long)mgpm::choose4(x));
  }
  return 0;
}
"""


@pytest.fixture(scope="module")
def choose_bin(tmp_path_factory):
    d = tmp_path_factory.mktemp("choose")
    src = d / "drv.cc"
    src.write_text(DRIVER)
    exe = d / "drv"
    subprocess.run(["g++", "-O2", "-std=c++17",
                    f"-I{ROOT / 'mercury_amd' / 'csrc'}", str(src), "-o",
                    str(exe)], check=True)
    return exe


def test_choose_exact_mod_2_64(choose_bin):
    xs = list(range(0, 40)) + [2_642_245, 2_642_246, 3_000_000, 65_536,
                               1 << 21, (1 << 32) - 1, 1 << 32,
                               12_345_678_901, (1 << 63) + 5, (1 << 64) - 1]
    out = subprocess.run([str(choose_bin)] + [str(x) for x in xs],
                         capture_output=True, text=True, check=True).stdout
    M = 1 << 64
    for x, line in zip(xs, out.splitlines()):
        c2, c3, c4 = (int(v) for v in line.split())
        assert (c2, c3, c4) == (comb(x, 2) % M, comb(x, 3) % M,
                                comb(x, 4) % M), x
    # the plain form does wrap at d = 3,000,000; the helper does not
    d = 3_000_000
    assert (d * (d - 1) * (d - 2) % M) // 6 != comb(d, 3)
