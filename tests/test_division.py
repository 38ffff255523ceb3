"""The U sweep of the step kernel divides by the diagonal as
q = b*r, e = fma(-d, q, b), x = fma(e, r, q) with r = 1/d correctly rounded
(Markstein's correction).  This pins that the form equals the IEEE quotient
b/d bit for bit, on random operands over a wide exponent range and on
divisors whose significand is near all-zeros or all-ones; a dropped
correction term or a wrong sign fails it.  Runs on the CPU (gcc, FMA
through fma())."""
import os
import shutil
import subprocess
import tempfile

import pytest

SRC = r"""
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
static uint64_t s = 0x9E3779B97F4A7C15ull;
static uint64_t nx(void) { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return s; }
static double mk(uint64_t mant, int e, int neg) {
  uint64_t b = ((uint64_t)(e + 1023) << 52) | (mant & ((1ull << 52) - 1));
  double d; memcpy(&d, &b, 8); return neg ? -d : d;
}
static double corr(double b, double d, double r) { double q = b * r; double e = fma(-d, q, b); return fma(e, r, q); }
int main(int argc, char** argv) {
  long n = atol(argv[1]), bad = 0, tested = 0;
  for (long i = 0; i < n; ++i) {
    double b, d;
    if (i % 2) { b = mk(nx(), (int)(nx() % 600) - 300, nx() & 1); d = mk(nx(), (int)(nx() % 600) - 300, nx() & 1); }
    else {
      uint64_t m = (nx() & 1) ? ((1ull << 52) - 1 - (nx() & 0xFFF)) : (nx() & 0xFFF);
      d = mk(m, (int)(nx() % 60) - 30, nx() & 1); b = mk(nx(), (int)(nx() % 60) - 30, nx() & 1);
    }
    double ref = b / d, r = 1.0 / d;
    if (!isnormal(ref) || !isnormal(r)) continue;
    ++tested;
    if (corr(b, d, r) != ref) ++bad;
  }
  /* a plausible mistake: no correction step -> must disagree somewhere */
  long naive_bad = 0;
  for (long i = 0; i < 100000; ++i) { double b = mk(nx(), 0, 0), d = mk(nx(), 0, 0); if (b * (1.0 / d) != b / d) ++naive_bad; }
  printf("%ld %ld %ld\n", tested, bad, naive_bad);
  return 0;
}
"""


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_reciprocal_correction_is_ieee_division():
    with tempfile.TemporaryDirectory() as td:
        c = os.path.join(td, "div.c")
        exe = os.path.join(td, "div")
        open(c, "w").write(SRC.replace("#include <math.h>", "#include <math.h>\n#include <stdlib.h>"))
        subprocess.run(["gcc", "-O2", "-mfma", "-ffp-contract=off", c, "-o", exe, "-lm"], check=True)
        out = subprocess.run([exe, "3000000"], check=True, capture_output=True, text=True).stdout.split()
    tested, bad, naive_bad = map(int, out)
    assert tested > 2_900_000
    assert bad == 0
    assert naive_bad > 0
