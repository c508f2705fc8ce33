"""The FP32 propagation-blocked PageRank sweep sums contributions as 64-bit
fixed point (2^-62 resolution).  Phase 1 writes each contribution as a
32-bit word (encode_fixed) that phase 2 widens with one shift
(decode_fixed); the pair must reproduce to_fixed, the direct rounding, for
every float the normalised iteration can produce (all of [0, 2)).  Host-side
build of the header's functions; no GPU."""
from __future__ import annotations

import shutil
import subprocess
import tempfile
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

SRC = r'''
#include <cstdio>
#include "pr_blocking.hpp"
using gx::pb::to_fixed; using gx::pb::encode_fixed; using gx::pb::decode_fixed;
int main() {
  unsigned long long bad = 0;
  for (uint32_t b = 0; b < 0x40000000u; ++b) {  // every float in [0, 2)
    const float v = __builtin_bit_cast(float, b);
    if (decode_fixed(encode_fixed(v)) != to_fixed(v)) ++bad;
  }
  // closed forms: 1 -> 2^62, 2^-62 -> 1, 1.5 * 2^-62 -> 2 (tie rounds up),
  // 2^-63 -> 1 (tie), 2^-64 -> 0, 0.75 -> 3 * 2^60
  const bool ok = to_fixed(1.0f) == (1ull << 62) && to_fixed(0x1p-62f) == 1 && to_fixed(0x1.8p-62f) == 2 &&
                  to_fixed(0x1p-63f) == 1 && to_fixed(0x1p-64f) == 0 && to_fixed(0.75f) == (3ull << 60);
  std::printf("%llu %d\n", bad, ok ? 1 : 0);
  return 0;
}
'''


@pytest.mark.skipif(not Path(HIPCC).exists(), reason="hipcc not available")
def test_encode_decode_matches_to_fixed_for_every_float():
    with tempfile.TemporaryDirectory() as td:
        src = Path(td) / "fx.hip"
        exe = Path(td) / "fx"
        src.write_text(SRC)
        subprocess.run([HIPCC, "-O2", "-std=c++17", "-w", "--offload-arch=gfx950", "-Xclang", "-target-feature",
                        "-Xclang", "-packed-fp32-ops", "-I", str(ROOT / "include"), "-I",
                        str(ROOT / "galois_amd" / "csrc"), str(src), "-o", str(exe)],
                       check=True, capture_output=True, timeout=300)
        out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120, check=True).stdout.split()
    assert out == ["0", "1"], out
