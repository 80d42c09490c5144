"""Division by a literal via multiply-high (oxidelake_amd/csrc/divlit.h),
checked on the CPU against the hardware divide for every divisor shape
(1, powers of two, general, negative) and edge dividends."""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]

PROGRAM = r"""
#include <cstdio>
#include <cstdint>
#include <random>
#include "divlit.h"
using namespace oxr;
int main() {
  std::mt19937_64 rng(12345);
  const int64_t divisors[] = {1, -1, 2, -2, 3, 5, 6, 7, -7, 10, 11, 12, 16, 25, 100, 127, 1000,
                              641, 1 << 20, 1000000007, (1ll << 32) + 1, INT64_MAX, INT64_MIN + 1,
                              -3, -1000, 1ll << 62, 6700417};
  const int64_t edges[] = {0, 1, -1, 2, -2, INT64_MAX, INT64_MIN, INT64_MIN + 1, INT64_MAX - 1,
                           1ll << 40, -(1ll << 40), 7, -7, 6, -6};
  long bad = 0, checked = 0;
  auto check = [&](int64_t n, int64_t d) {
    if (n == INT64_MIN && d == -1) return;  // overflow, reported by the caller
    DivMagic g = div_magic(d);
    int64_t q, r;
    sdivmod_magic(n, g, &q, &r);
    checked++;
    if (q != n / d || r != n % d) {
      if (bad++ < 5) printf("n=%lld d=%lld got %lld %lld want %lld %lld\n", (long long)n,
                            (long long)d, (long long)q, (long long)r, (long long)(n / d),
                            (long long)(n % d));
    }
  };
  for (int64_t d : divisors) {
    for (int64_t n : edges) check(n, d);
    for (int i = 0; i < 200000; i++) check((int64_t)rng(), d);
    for (int i = 0; i < 20000; i++) check((int64_t)(rng() >> (rng() % 63)), d);
  }
  for (int i = 0; i < 300000; i++) {
    int64_t d = (int64_t)(rng() >> (rng() % 64));
    if (d == 0) continue;
    check((int64_t)rng(), d);
  }
  printf("checked %ld bad %ld\n", checked, bad);
  return bad ? 1 : 0;
}
"""


def test_division_by_literal_matches_hardware_divide(tmp_path):
    src = tmp_path / "divlit_test.cpp"
    exe = tmp_path / "divlit_test"
    src.write_text(PROGRAM)
    res = subprocess.run(["g++", "-O2", "-std=c++17", f"-I{ROOT / 'oxidelake_amd' / 'csrc'}", str(src),
                          "-o", str(exe)], capture_output=True, text=True)
    if res.returncode != 0:
        pytest.fail(res.stderr)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout
    assert "bad 0" in out.stdout
