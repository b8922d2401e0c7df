"""exp_cr (bem3_aw_qd2_amd/csrc/hip/exp_cr.h), the correctly rounded exponential
the GPU uses for the two factors of the z == 0 stopping residue, pinned on the
CPU: the same header compiled for the host (the double-double arithmetic is
fma-based and rounds the same way on both), against mpmath at 200 bits rounded
once to double.  Zero mismatches over arguments spanning the reduction table,
tiny arguments and the ends of the range."""
import os
import subprocess

import mpmath
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
HDR = os.path.join(os.path.dirname(HERE), "bem3_aw_qd2_amd", "csrc", "hip", "exp_cr.h")
PROG = r'''
#include <stdio.h>
#include <stdlib.h>
#include "%s"
int main(int argc, char **argv) {
  int n = atoi(argv[1]);
  double *x = (double *)malloc(8 * (size_t)n);
  if (fread(x, 8, n, stdin) != (size_t)n) return 1;
  for (int i = 0; i < n; i++) { double y = exp_cr(x[i]); fwrite(&y, 8, 1, stdout); }
  return 0;
}
'''


def test_exp_cr_correctly_rounded(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text(PROG % HDR)
    exe = tmp_path / "t"
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-Wno-unknown-pragmas", str(src), "-o", str(exe)], check=True)
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.uniform(-30.0, 30.0, 12000), rng.uniform(-700.0, 709.0, 4000),
                        rng.uniform(-1e-3, 1e-3, 2000), rng.uniform(-1e-15, 1e-15, 500),
                        np.array([0.0, 1.0, -1.0, 709.0, -700.0, 0.5 * np.log(2.0) / 64])])
    out = subprocess.run([str(exe), str(len(x))], input=x.tobytes(), capture_output=True, check=True).stdout
    y = np.frombuffer(out, dtype=np.float64)
    mpmath.mp.prec = 200
    bad = [xi for xi, yi in zip(x, y) if float(mpmath.exp(mpmath.mpf(float(xi)))) != yi]
    assert not bad, bad[:5]
