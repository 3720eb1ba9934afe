"""The binomial inversion's bottom-up stage divides integer counts by 3 with
Markstein's three operations (q = a rb, r = fma(-b, q, a), fma(r, rb, q) with
rb = fl(1/3)) instead of the IEEE division sequence.  This checks, on the CPU
with correctly rounded float32 fmaf and division (the operations gfx950's
v_fma_f32 and IEEE division perform), that the two agree on every
integer-valued float 0 .. 2^24."""

import shutil
import subprocess

import pytest

SRC = r"""
#include <math.h>
#include <stdio.h>
#include <string.h>
int main(void) {
  const volatile float b = 3.0f;
  const float rb = 1.0f / b;
  long bad = 0;
  for (long i = 0; i <= (1L << 24); ++i) {
    const float a = (float)i;
    const float q = a * rb;
    const float r = fmaf(-b, q, a);
    const float m = fmaf(r, rb, q);
    const float e = a / b;
    if (memcmp(&m, &e, 4) != 0) ++bad;
  }
  printf("%ld\n", bad);
  return 0;
}
"""


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_markstein_division_by_three_is_exact_on_integers(tmp_path):
    src = tmp_path / "div3.c"
    src.write_text(SRC)
    exe = tmp_path / "div3"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", str(exe), str(src), "-lm"], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    assert int(out.strip()) == 0
