"""The FMA divisions of pangu_amd/csrc/exact_div.h against the IEEE quotient
(CPU): the header is compiled for the host with the ROCm clang, the same
compiler front end the device code goes through, and every result is
compared bit for bit with the hardware division."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "pangu_amd", "csrc")
CLANG = "/opt/rocm/llvm/bin/clang++"

HARNESS = r"""
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include "exact_div.h"
using pangu::div12;
using pangu::div_by;
static uint64_t s = 0x9e3779b97f4a7c15ull;
static inline uint64_t rnd() { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return s; }
static inline double bits(uint64_t b) { double d; std::memcpy(&d, &b, 8); return d; }
static inline bool same(double a, double b) { return std::memcmp(&a, &b, 8) == 0; }
static const uint64_t M52 = (1ull << 52) - 1;
static double operand(uint64_t mant, int spread) {
  const uint64_t e = (uint64_t)(1023 + (int)(rnd() % (2 * spread + 1)) - spread);
  return bits(((rnd() & 1ull) << 63) | (e << 52) | (mant & M52));
}
static uint64_t mantissa(int mode, uint64_t k) {
  switch (mode) {
    case 0: return rnd();
    case 1: return k;                       // 1 + k ulp
    case 2: return M52 - k;                 // just below 2
    case 3: return (1ull << 51) + k - 2048; // around 1.5
    case 4: return rnd() | rnd();           // dense ones
    case 5: return rnd() & rnd();           // sparse ones
    default: return 0x5555555555555ull ^ (k & 0xff);  // near 4/3
  }
}
int main(int argc, char **argv) {
  const long n = std::atol(argv[1]);
  long bad12 = 0, badby = 0;
  for (long t = 0; t < n; ++t) {
    const uint64_t k = rnd() % 4096;
    const double x = operand(mantissa((int)(rnd() % 7), rnd() % 4096), 300);
    const double y = operand(mantissa((int)(rnd() % 7), k), 300);
    volatile double q12 = x / 12.0, qy = x / y, ry = 1.0 / y;
    if (!same(div12(x), q12)) {
      if (bad12 < 5) std::printf("div12 %a: %a vs %a\n", x, div12(x), (double)q12);
      ++bad12;
    }
    if (!same(div_by(x, y, ry), qy)) {
      if (badby < 5) std::printf("div_by %a / %a: %a vs %a\n", x, y, div_by(x, y, ry), (double)qy);
      ++badby;
    }
  }
  // signed zeros and exact quotients
  volatile double z[6] = {0.0, -0.0, 24.0, -36.0, 12.0, -1.5};
  for (int i = 0; i < 6; ++i) {
    const double x = z[i];
    volatile double q = x / 12.0, r3 = 1.0 / 3.0, qq = x / 3.0;
    if (!same(div12(x), q)) { std::printf("special %a\n", x); ++bad12; }
    if (!same(div_by(x, 3.0, r3), qq)) { std::printf("special/3 %a\n", x); ++badby; }
  }
  std::printf("%ld %ld\n", bad12, badby);
  return (bad12 || badby) ? 1 : 0;
}
"""


@pytest.mark.skipif(not os.path.exists(CLANG), reason="ROCm clang++ not present")
def test_fma_divisions_return_the_ieee_quotient(tmp_path):
    src = tmp_path / "div.cpp"
    src.write_text(HARNESS)
    exe = tmp_path / "div"
    subprocess.run([CLANG, "-O2", "-std=c++17", "-march=x86-64-v3", "-ffp-contract=off",
                    "-I", CSRC, str(src), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe), "20000000"], capture_output=True, text=True)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert r.stdout.strip().splitlines()[-1] == "0 0"


def test_planted_division_error_is_caught(tmp_path):
    """The harness notices a quotient that is one ulp off: a version of
    div12 without its correction step (q = x * RN(1/12)) fails it."""
    if not os.path.exists(CLANG):
        pytest.skip("ROCm clang++ not present")
    hdr = tmp_path / "exact_div.h"
    with open(os.path.join(CSRC, "exact_div.h")) as f:
        text = f.read()
    planted = text.replace("return fma(-e, ry, q);\n}\n\n// x / y", "return q + 0.0 * e;\n}\n\n// x / y", 1)
    assert planted != text
    hdr.write_text(planted)
    src = tmp_path / "div.cpp"
    src.write_text(HARNESS)
    exe = tmp_path / "div"
    subprocess.run([CLANG, "-O2", "-std=c++17", "-march=x86-64-v3", "-ffp-contract=off",
                    "-I", str(tmp_path), str(src), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe), "200000"], capture_output=True, text=True)
    assert r.returncode == 1
    assert int(r.stdout.strip().splitlines()[-1].split()[0]) > 0
