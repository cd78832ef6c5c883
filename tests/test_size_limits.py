"""The 2D blocked kernels' plane-size predicate (ceep-v1_amd/csrc/limits2d.h),
compiled with the host compiler and checked at its edges: a plane whose FP64
byte offsets reach 2^31 -- where the kernels' out-of-window marker begins --
must take the one-step kernels."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "ceep-v1_amd", "csrc")

PROGRAM = r"""
#include <cstdio>
#include "limits2d.h"
int main() {
  const long long cases[] = {1, 600LL * 640, (1LL << 28) - 1, 1LL << 28, (1LL << 28) + 1,
                             23170LL * 23232, 0, -5};
  for (long long c : cases) std::printf("%lld %d\n", c, (int)ceep::tb2d_fits(c));
  static_assert(ceep::tb2d_fits(571LL * 576), "the 600^2 workload's box must fit");
  return 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs the host C++ compiler")
def test_blocked_2d_plane_limit(tmp_path):
    src = tmp_path / "limits.cpp"
    exe = tmp_path / "limits"
    src.write_text(PROGRAM)
    subprocess.run(["g++", "-std=c++17", "-I", CSRC, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    got = {int(a): bool(int(b)) for a, b in (line.split() for line in out.split("\n") if line)}
    for cells, fits in got.items():
        # every byte offset of an FP64 plane (8 B per cell) below the 2^31 marker
        assert fits == (cells > 0 and cells * 8 <= (1 << 31) - 8), cells
    assert got[(1 << 28) - 1] and not got[1 << 28]
    assert not got[23170 * 23232]
