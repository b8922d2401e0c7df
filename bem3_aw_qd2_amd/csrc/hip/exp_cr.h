// exp_cr.h -- correctly rounded exp(x) in double precision.
//
// The Ewald spectral sum of the lattice Green's function decides where to stop
// at z == 0 from the rounding residue of exp(bn^2 E^2) * exp(-(bn E)^2) - 1
// (qgf2_src/d3_qpgf_d2.c:1976-1989 with the stopping rule of :471-482), so the
// two exponentials there must round the same way on every platform.  exp_cr()
// returns the correctly rounded value, which libm's of the glibc line before 2.28
// (IBM's accurate library) returned always and glibc 2.35's returns for all but
// about 0.08 % of arguments.
//
// Method: Tang's table-driven reduction (ACM TOMS 15 (1989) 144-157) in
// double-double arithmetic (Dekker, Numer. Math. 18 (1971) 224-242):
//   x = k ln2/64 + r, k = 64 m + j, |r| <= ln2/128,
//   exp(x) = 2^m 2^(j/64) exp(r), exp(r) by its Taylor series to degree 11.
// The double-double result is within about 2^-100 of exp(x); one rounding of
// hi + lo then gives the correctly rounded value unless exp(x) lies within
// 2^-100 of a rounding midpoint (binary64's hardest cases of exp need about
// 2^-113, Lefevre and Muller, ARITH 15, 2001).  Constants: gen_exp_cr.py.
//
// Valid for -700 <= x <= 709 (normal results); outside, the library exp is
// returned.  Header-only and host-compilable so that the same arithmetic can be
// checked on the CPU.
#pragma once
#include <math.h>
#include "../../../include/exp_cr_table.h"

#if defined(__HIPCC__)
#define EXPCR_HD __host__ __device__
#else
#define EXPCR_HD
#endif

#pragma clang fp contract(off)

struct dd_t {
  double h, l;
};

// error-free transformations (no contraction: each operation rounds once)
EXPCR_HD inline dd_t dd_two_sum(double a, double b) {
  const double s = a + b;
  const double bb = s - a;
  return dd_t{s, (a - (s - bb)) + (b - bb)};
}
EXPCR_HD inline dd_t dd_fast_two_sum(double a, double b) {  // |a| >= |b|
  const double s = a + b;
  return dd_t{s, b - (s - a)};
}
EXPCR_HD inline dd_t dd_two_prod(double a, double b) {
  const double p = a * b;
  return dd_t{p, fma(a, b, -p)};
}
EXPCR_HD inline dd_t dd_mul(dd_t a, dd_t b) {
  dd_t p = dd_two_prod(a.h, b.h);
  const double c = (a.h * b.l + a.l * b.h) + p.l;
  return dd_fast_two_sum(p.h, c);
}
EXPCR_HD inline dd_t dd_add(dd_t a, dd_t b) {
  dd_t s = dd_two_sum(a.h, b.h);
  dd_t t = dd_two_sum(a.l, b.l);
  s = dd_fast_two_sum(s.h, s.l + t.h);
  return dd_fast_two_sum(s.h, s.l + t.l);
}

// exp(x) as a double-double, -700 <= x <= 709
EXPCR_HD inline dd_t exp_dd(double x) {
  const dd_t T[64] = EXPCR_T_INIT;
  const dd_t C[12] = EXPCR_C_INIT;
  const double kd = rint(x * EXPCR_INV);
  const int k = (int)kd;
  // r = x - k ln2/64: x - k L1 is exact (k L1 has at most 48 significant bits
  // and lies within a factor two of x when k != 0)
  const double t1 = x - kd * EXPCR_L1;
  const dd_t k2 = dd_two_prod(kd, EXPCR_L2);
  dd_t r = dd_add(dd_t{t1, 0.0}, dd_t{-k2.h, -k2.l});
  r = dd_add(r, dd_t{-kd * EXPCR_L3, 0.0});
  // Taylor series of exp(r), Horner in double-double
  dd_t p = C[11];
  for (int n = 10; n >= 0; n--) p = dd_add(dd_mul(p, r), C[n]);
  const int j = k & 63, m = (k - j) / 64;
  const dd_t q = dd_mul(T[j], p);
  return dd_t{ldexp(q.h, m), ldexp(q.l, m)};
}

EXPCR_HD inline double exp_cr(double x) {
  if (!(x >= -700.0 && x <= 709.0)) return exp(x);
  // the high part is hi + lo rounded to nearest (fast two-sum normalisation)
  return exp_dd(x).h;
}
