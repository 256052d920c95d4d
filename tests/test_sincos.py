"""The Ewald phase sine/cosine (cm3d_amd/csrc/nb_sincos.h: Cody-Waite
reduction + Taylor series) against the C library's long-double sinl/cosl:
absolute error below 4e-16 over the argument range of the Ewald phases
(|k.r| <= 3 pi kmax, kmax <= 64 here) and at the quadrant boundaries."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SRC = r'''
#include "nb_sincos.h"
#include <stdio.h>
int main() {
  double me = 0; unsigned long long st = 2024;
  const double pio2 = 1.5707963267948966;
  for (int i = 0; i < 400000; i++) {
    st = st * 6364136223846793005ULL + 1442695040888963407ULL;
    double u = (st >> 11) * (1.0 / 9007199254740992.0) * 2 - 1;
    double x = i % 3 == 0 ? u * 610.0 : (i % 3 == 1 ? u * 4.0 : (i % 1000) * pio2 + u * 1e-9);
    double s, c;
    nb_sincos(x, &s, &c);
    long double es = sinl((long double)x), ec = cosl((long double)x);
    double e1 = fabsl(s - es), e2 = fabsl(c - ec);
    if (e1 > me) me = e1;
    if (e2 > me) me = e2;
  }
  printf("%.6e\n", me);
  return 0;
}
'''


def test_sincos_error_bound(tmp_path):
    src = tmp_path / "sc.cpp"
    src.write_text(SRC)
    exe = tmp_path / "sc"
    r = subprocess.run(["g++", "-O2", "-I", os.path.join(ROOT, "cm3d_amd", "csrc"), str(src),
                        "-o", str(exe), "-lm"], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail(r.stderr)
    err = float(subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout)
    assert err < 4e-16, err
