/*
 * faddeeva_cpu.c -- oracle Faddeeva function w(z) = exp(-z^2) erfc(-i z).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The reference calls Faddeeva_w(z, 1e-15) (qgf2_src/Faddeeva.cc:679-972) from
 * its Ewald sums (qgf2_src/d3_qpgf_d2.c:1831, :1982).  This restatement uses
 * Weideman's published rational approximation (SIAM J. Numer. Anal. 31 (1994)
 * 1497-1518) with N = 40 terms on Im z >= 0 and the reflection
 * w(z) = 2 exp(-z^2) - w(-z) below the real axis.
 *
 * Pins: tests/test_oracle_faddeeva.py checks it against the reference's own
 * known-answer table (Faddeeva.cc:1918-2075, committed as
 * tests/golden/faddeeva_kat.json) and against scipy.special.wofz, plus the
 * identities w(-z) = 2 exp(-z^2) - w(z), w(-conj z) = conj w(z) and
 * Re w(x) = exp(-x^2).  The Green's function tests compare whole lattice sums
 * with the reference's, which calls its own Faddeeva_w.
 */
#include <math.h>
#include <complex.h>
#include "oracle.h"
#include "weideman_coef.h"

static const double weid_a[WEID_N] = { WEID_COEFS };

/* w(z) for Im z >= 0 */
static double complex w_upper(double complex z)
{
  const double L = WEID_L;
  double complex den = L - I * z;
  double complex Z = (L + I * z) / den;
  double complex p = weid_a[WEID_N - 1];
  for (int n = WEID_N - 2; n >= 0; n--) p = p * Z + weid_a[n];
  double complex iden = 1.0 / den;
  return 2.0 * p * iden * iden + M_2_SQRTPI * 0.5 * iden;
}

double complex orc_faddeeva_w(double complex z)
{
  /* on the real axis Re w(x) = exp(-x^2) exactly as the reference returns it
   * (Faddeeva.cc:684-686); the Ewald spectral sum at z = 0 subtracts it from
   * cos(0) and its stopping rule then sees that difference */
  if (cimag(z) == 0.0) return orc_exp_z0(-creal(z) * creal(z)) + I * cimag(w_upper(z));
  if (cimag(z) > 0.0) return w_upper(z);
  /* w(z) = 2 exp(-z^2) - w(-z) */
  return 2.0 * cexp(-z * z) - w_upper(-z);
}

double complex orc_scaled_w(double x, double y)
{
  if (y >= 0.0) return exp(x * x - y * y) * w_upper(x + I * y);
  /* exp(x^2-y^2) * 2 exp(-z^2) = 2 exp(-2ixy) */
  double a = -2.0 * x * y;
  return 2.0 * (cos(a) + I * sin(a)) - exp(x * x - y * y) * w_upper(-x - I * y);
}

/* w at n points z (interleaved re, im) into w (interleaved), for callers
 * without a C99 complex ABI */
void orc_faddeeva_w_batch(int n, const double *z, double *w)
{
  for (int i = 0; i < n; i++) {
    double complex v = orc_faddeeva_w(z[2 * i] + I * z[2 * i + 1]);
    w[2 * i] = creal(v);
    w[2 * i + 1] = cimag(v);
  }
}
