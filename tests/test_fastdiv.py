"""The multiplicative-inverse division the line passes and the spreading use to decode block indices
(FastDiv in csrc/tr_internal.hpp): its host code path, compiled from the header itself, against
plain unsigned division -- every divisor below 3000 with the dividends 0..19999, random dividends and
2^31 - 1, and 300 000 random (divisor, dividend) pairs below 2^31.  CPU only (hipcc host build)."""

import os
import shutil
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

DRIVER = r"""
#include "tr_internal.hpp"
#include <cstdio>
#include <random>
int main() {
  std::mt19937_64 g(1);
  long bad = 0, n = 0;
  for (unsigned d = 1; d < 3000; ++d) {
    const tr::FastDiv f = tr::make_fastdiv(d);
    for (unsigned x = 0; x < 20000; ++x, ++n) bad += f.div(x) != x / d;
    for (int i = 0; i < 2000; ++i, ++n) { const unsigned x = g() & 0x7fffffffu; bad += f.div(x) != x / d; }
    bad += f.div(0x7fffffffu) != 0x7fffffffu / d;
    ++n;
  }
  for (int i = 0; i < 300000; ++i, ++n) {
    const unsigned d = (unsigned)(g() % 0x7fffffffu) + 1u, x = g() & 0x7fffffffu;
    bad += tr::make_fastdiv(d).div(x) != x / d;
  }
  printf("%ld %ld\n", n, bad);
  return bad != 0;
}
"""


def test_fastdiv_matches_division(tmp_path):
    if not (os.path.exists(HIPCC) or shutil.which("hipcc")):
        pytest.skip("hipcc not available")
    src = tmp_path / "fastdiv.cpp"
    src.write_text(DRIVER)
    exe = tmp_path / "fastdiv"
    subprocess.run([HIPCC, "-O2", "-std=c++17", "-I", os.path.join(REPO, "include"),
                    "-I", os.path.join(REPO, "multinodemri-miccai26_amd", "csrc"), str(src), "-o", str(exe)],
                   check=True, capture_output=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    checked, bad = (int(v) for v in out.stdout.split())
    assert checked > 66_000_000 and bad == 0, out.stdout
