// gf_dev.h -- device quasi-periodic Helmholtz Green's function (2-D lattice in
// the xy-plane) and the Faddeeva function it needs.
//
// Restates qgf2_src/d3_qpgf_d2.c of the reference: method choice between the
// Ewald split and the pure spectral sum (:124-176), unit-cell reduction with the
// Bloch phase (:409-438), Ewald parameter (:1680-1693), spatial image sum
// (:1714-1838) and spectral sum (:1840-2009) with the same row-by-row adaptive
// stopping rules (:471-482), and the Hessian variants (:2011-2398).  The
// reference evaluates one point per call; here one lane evaluates one point and
// the instance constants (lattice, reciprocal lattice, Ewald parameter, method
// probe) are computed once on the host into GfParams.  The reference's probe is
// read at |z| / fd_lift: between its switch height and fd_lift times that, where
// the pure spectral sum needs up to ~900 orders, the Ewald sum is taken instead
// (the same value to ~1e-15; fd_lift = 1 restores the reference's choice).
//
// Faddeeva w(z): Weideman's rational approximation, N = 40 (SIAM J. Numer.
// Anal. 31 (1994) 1497-1518), branch-free on Im z >= 0, reflection below.
#pragma once
#include "cplx.h"
#include "../../../include/weideman_coef.h"
#include "../../../include/erfcx_coef.h"
#include "../../../include/bqd2.h"
#include "exp_cr.h"

// Products and sums are rounded one at a time, as the reference's C does on
// x86-64 (no contraction into FMA): the Ewald spectral sum at z == 0 decides
// where to stop from the rounding residue of exp(bn^2 E^2) * exp(-(bn E)^2) - 1,
// so the inputs to that residue have to be rounded identically.  Explicit fma()
// calls (complex products, the Weideman recurrence) are unaffected.
#pragma clang fp contract(off)

#define BQD2_EFS 0.1
#define BQD2_EW_LIMIT 100
#define BQD2_FD_LIMIT 100
#define EW_EXPM_D 3.0
// |z| / E below which the Ewald spectral terms are formed exactly as the
// reference forms them (its stopping rule reads rounding residue at z == 0)
#define BQD2_ZFAST 2.0e-8

enum { GF_QG = 0, GF_DQG = 1, GF_V1 = 2, GF_V2 = 3 };

// ---------------------------------------------------------------- Faddeeva w
// p(Z) = sum a_k Z^k with real a_k at complex Z, by division through the real
// quadratic (x - Z)(x - conj Z) = x^2 - t x + s (Goertzel's method; Knuth,
// TAOCP vol. 2, 4.6.4): b_k = a_k + t b_{k+1} - s b_{k+2}, p(Z) = b_1 Z + a_0 - s b_2.
// Two real FMAs per coefficient where Horner's rule takes a complex multiply-add.
template <int N>
__device__ __forceinline__ cd poly_real_at(const double (&a)[N], cd Z) {
  const double t = 2.0 * Z.x, s = fma(Z.x, Z.x, Z.y * Z.y);
  double b2 = 0.0, b1 = a[N - 1];
#pragma unroll
  for (int k = N - 2; k >= 1; k--) {
    const double b = fma(t, b1, fma(-s, b2, a[k]));
    b2 = b1;
    b1 = b;
  }
  return cd{fma(b1, Z.x, fma(-s, b2, a[0])), b1 * Z.y};
}

__device__ __forceinline__ cd weideman_upper(double x, double y) {
  // w(z) ~ 2 p(Z)/(L - iz)^2 + (1/sqrt(pi))/(L - iz), Z = (L + iz)/(L - iz)
  const double L = WEID_L;
  const double a[WEID_N] = {WEID_COEFS};
  cd iden = crecip_nz(cd{L + y, -x});  // 1/(L - i z)
  cd Z = cd{L - y, x} * iden;           // (L + i z)/(L - i z)
  cd p = poly_real_at(a, Z);
  const double isp = 0.56418958354775628695;  // 1/sqrt(pi)
  return 2.0 * (p * iden * iden) + isp * iden;
}

// the same with N = 24, for 1.5 <= Im z <= 12 and Re z <= 3, where its relative
// error (< 4e-15) is that of the N = 40 table
__device__ __forceinline__ cd weideman24_upper(double x, double y) {
  const double L = WEID24_L;
  const double a[WEID24_N] = {WEID24_COEFS};
  cd iden = crecip_nz(cd{L + y, -x});
  cd Z = cd{L - y, x} * iden;
  cd p = poly_real_at(a, Z);
  const double isp = 0.56418958354775628695;
  return 2.0 * (p * iden * iden) + isp * iden;
}

// erfcx(y) = exp(y^2) erfc(y) = w(i y): on the imaginary axis Weideman's Z is
// real and covers (-1, 1] once, so a real polynomial in Z fitted to erfcx's
// relative error (degree 20, < 7e-16 on t >= 0; coefficients from
// gen_erfcx.py) replaces the N = 40 Fourier table at half the multiply-adds;
// y < 0 by erfcx(y) = 2 exp(y^2) - erfcx(-y).
__device__ __forceinline__ double erfcx_body(double y) {
  const double L = ERFCX_L;
  const double c[ERFCX_N] = {ERFCX_COEFS};
  const double t = fabs(y);
  const double iden = rcp_pos(L + t), Z = (L - t) * iden;
  double p = c[ERFCX_N - 1];
#pragma unroll
  for (int n = ERFCX_N - 2; n >= 0; n--) p = fma(p, Z, c[n]);
  const double isp = 0.56418958354775628695;
  const double v = 2.0 * p * iden * iden + isp * iden;
  return y >= 0.0 ? v : 2.0 * exp(y * y) - v;
}
__device__ __noinline__ double erfcx_call(double y) { return erfcx_body(y); }

// Inlined in the value-and-gradient sums; in the Hessian sums (GF_V2) a call,
// which keeps their register pressure lower, unless the translation unit asks
// for it inline everywhere (BQD2_ERFCX_INLINE: the assembly kernels).
template <int M>
__device__ __forceinline__ double erfcx_w(double y) {
#if defined(BQD2_ERFCX_INLINE)
  return erfcx_body(y);
#else
  return M == GF_V2 ? erfcx_call(y) : erfcx_body(y);
#endif
}

// w(x + i y) with the reference's conventions: real axis returns Re w = exp(-x^2)
// exactly (Faddeeva.cc:684-686), correctly rounded (exp_cr.h), lower half plane
// by reflection.
__device__ __forceinline__ cd faddeeva_w(double x, double y) {
  if (y > 0.0) return weideman_upper(x, y);
  if (y == 0.0) {
    cd w = weideman_upper(x, 0.0);
    return cd{exp_cr(-__dmul_rn(x, x)), w.y};
  }
  // w(z) = 2 exp(-z^2) - w(-z)
  cd wm = weideman_upper(-x, -y);
  double e = 2.0 * exp(y * y - x * x);
  cd ph = cis(-2.0 * x * y);
  return cd{e * ph.x - wm.x, e * ph.y - wm.y};
}

// ---------------------------------------------------------------- sums
struct Sum10 {
  cd g, dg[3], d2g[6];
};

__device__ __forceinline__ bool conv_same(cd a, cd b, double eps) {
  return 2.0 * fabs(a.x - b.x) <= eps * (fabs(a.x) + fabs(b.x)) &&
         2.0 * fabs(a.y - b.y) <= eps * (fabs(a.y) + fabs(b.y));
}
__device__ __forceinline__ bool conv_small(cd t, cd s, double eps) {
  return fabs(t.x) <= fabs(s.x) * BQD2_EFS * eps && fabs(t.y) <= fabs(s.y) * BQD2_EFS * eps;
}

template <int M>
__device__ __forceinline__ void sum_add(Sum10& s, const Sum10& t) {
  if (M != GF_DQG) s.g += t.g;
  if (M != GF_QG) {
    s.dg[0] += t.dg[0];
    s.dg[1] += t.dg[1];
    s.dg[2] += t.dg[2];
  }
  if (M == GF_V2) {
#pragma unroll
    for (int i = 0; i < 6; i++) s.d2g[i] += t.d2g[i];
  }
}

// the components a row-convergence test reads (per sum and mode); the sum as it
// stood before the row is kept as these only
struct Key2 {
  cd a, b;
};
template <int M>
__device__ __forceinline__ Key2 key_spatial(const Sum10& s) {
  if (M == GF_QG || M == GF_V1) return Key2{s.g, cd{0, 0}};
  if (M == GF_DQG) return Key2{s.dg[0], s.dg[1]};
  return Key2{s.d2g[0], s.d2g[1]};
}
template <int M>
__device__ __forceinline__ Key2 key_spectral(const Sum10& s) {
  if (M == GF_QG) return Key2{s.g, cd{0, 0}};
  if (M == GF_DQG || M == GF_V1) return Key2{s.dg[2], cd{0, 0}};
  return Key2{s.d2g[0], s.d2g[1]};
}
template <int M>
__device__ __forceinline__ Key2 key_fd(const Sum10& s) {
  if (M == GF_QG) return Key2{s.g, cd{0, 0}};
  if (M == GF_DQG || M == GF_V1) return Key2{s.dg[2], cd{0, 0}};
  return Key2{s.d2g[2], cd{0, 0}};
}
// stopping quantities of each mode (d3_qpgf_d2.c, per chain)
template <int M>
__device__ __forceinline__ bool rows_conv_spatial(const Sum10& s, const Key2& o, double eps) {
  const Key2 k = key_spatial<M>(s);
  if (M == GF_QG || M == GF_V1) return conv_same(k.a, o.a, eps);
  return conv_same(k.a, o.a, eps) && conv_same(k.b, o.b, eps);
}
template <int M>
__device__ __forceinline__ bool term_small_spatial(const Sum10& t, const Sum10& s, double eps) {
  if (M == GF_QG || M == GF_V1) return conv_small(t.g, s.g, eps);
  if (M == GF_DQG) return conv_small(t.dg[0], s.dg[0], eps) && conv_small(t.dg[1], s.dg[1], eps);
  return conv_small(t.d2g[0], s.d2g[0], eps) && conv_small(t.d2g[1], s.d2g[1], eps);
}
template <int M>
__device__ __forceinline__ bool rows_conv_spectral(const Sum10& s, const Key2& o, double eps) {
  const Key2 k = key_spectral<M>(s);
  if (M == GF_V2) return conv_same(k.a, o.a, eps) && conv_same(k.b, o.b, eps);
  return conv_same(k.a, o.a, eps);
}
template <int M>
__device__ __forceinline__ bool term_small_spectral(const Sum10& t, const Sum10& s, double eps) {
  if (M == GF_QG) return conv_small(t.g, s.g, eps);
  if (M == GF_DQG || M == GF_V1) return conv_small(t.dg[2], s.dg[2], eps);
  return conv_small(t.d2g[0], s.d2g[0], eps) && conv_small(t.d2g[1], s.d2g[1], eps);
}
template <int M>
__device__ __forceinline__ bool rows_conv_fd(const Sum10& s, const Key2& o, double eps) {
  return conv_same(key_fd<M>(s).a, o.a, eps);
}
template <int M>
__device__ __forceinline__ bool term_small_fd(const Sum10& t, const Sum10& s, double eps) {
  if (M == GF_QG) return conv_small(t.g, s.g, eps);
  if (M == GF_DQG || M == GF_V1) return conv_small(t.dg[2], s.dg[2], eps);
  return conv_small(t.d2g[2], s.d2g[2], eps);
}

// ---------------------------------------------------------------- terms
// spatial image (l1, l2): d3_qpgf_d2.c:1814-1838 (v1) and :2131-2170 (v2)
// ce is the Bloch phase exp(i k_t . R_l) of the image, carried along the row by
// the caller (one sincos per row instead of per image)
// ef: exp(k^2 E^2 - |r + R_l|^2 / 4E^2), also carried along the row by the caller
template <int M>
__device__ __forceinline__ void spatial_term(const GfParams& P, double rx, double ry, double rz, double E,
                                             int l1, int l2, cd ce, double ef, Sum10& t) {
  // no z == 0 residue is read in the image sum: products may be fused
#pragma clang fp contract(fast)
  double Rl0 = l1 * P.d1x + l2 * P.d2x;
  double Rl1 = l1 * P.d1y + l2 * P.d2y;
  double ra0 = rx + Rl0, ra1 = ry + Rl1;
  const double rl2 = ra0 * ra0 + ra1 * ra1 + rz * rz;
  double irl = rsqrt_pos(rl2);
  double rl = rl2 * irl;
  double iE = 1.0 / E;
  double ke = P.k * E, y = rl * 0.5 * iE;
  cd w = (y >= 1.5 && y <= 12.0 && ke <= 3.0) ? weideman24_upper(ke, y) : weideman_upper(ke, y);
  const double isp = 0.56418958354775628695;
  cd cer = ce * (ef * w.x);
  t.g = cer * irl;
  if (M != GF_QG) {
    cd cf1 = (irl * irl) * (cer * irl - ce * (P.k * ef * w.y) + ce * (ef * iE * isp));
    t.dg[0] = ra0 * cf1;
    t.dg[1] = ra1 * cf1;
    t.dg[2] = cf1;
    if (M == GF_V2) {
      cd cf2 = (irl * irl) * (cer * (P.k * P.k * irl * 2.0) - ce * (ef * iE * iE * iE * isp));
      cd cf3 = cf1 * (-6.0 * irl * irl) + cf2;
      t.d2g[0] = 2.0 * cf1 + (ra0 * ra0) * cf3;
      t.d2g[1] = 2.0 * cf1 + (ra1 * ra1) * cf3;
      t.d2g[2] = 2.0 * cf1 + (rz * rz) * cf3;
      t.d2g[3] = (ra0 * ra1) * cf3;
      t.d2g[4] = ra1 * cf3;
      t.d2g[5] = ra0 * cf3;
    }
  }
}

// spectral order (n1, n2): d3_qpgf_d2.c:1965-2009 (v1) and :2349-2398 (v2).
// returns 0 propagating, 1 evanescent, -3 Wood anomaly.
// The z-derivative of an order is what the stopping rules read.  At z == 0 it
// is formed exactly as the reference forms it (the evanescent difference is 0
// and the propagating one the rounding residue of two correctly rounded
// exponentials); at 0 < |z| < BQD2_ZFAST E, where it is cancellation noise, it
// is taken in exact arithmetic to first order in z, so the sum is continuous in
// z down to z -> 0.
// fast: ce is the planar phase exp(i a_n . r) carried along the row by the
// caller, and the evanescent pair e^{+-b z} erfc(b E +- z/2E) is formed as
// exp(-(b^2 E^2 + z^2/4E^2)) erfcx(b E +- z/2E) (the same value, one exp and two
// scaled complementary error functions).  Used away from z == 0 only, where the
// stopping rule reads well-conditioned terms; at z == 0 the terms are formed as
// the reference forms them.
template <int M>
__device__ __forceinline__ int spectral_term(const GfParams& P, double rx, double ry, double rz, double E,
                                             int n1, int n2, bool fast, cd cef, double eff, Sum10& t) {
  double an0 = -P.vkx + n1 * P.g1x + n2 * P.g2x;
  double an1 = -P.vky + n1 * P.g1y + n2 * P.g2y;
  double bn2 = P.k * P.k - an0 * an0 - an1 * an1;
  cd ce = fast ? cef : cis(an0 * rx + an1 * ry);
  double azve = rz / (2.0 * E);
  const double isp = 0.56418958354775628695;
  if (bn2 > 0.0) {
    double bn = sqrt(bn2);
    cd ff = cis(bn * rz);
    // ef and w kept apart as the reference does: at z == 0, ef*Re(w) - ff is
    // pure rounding and the row stopping rule reads it (products not fused;
    // both exponentials correctly rounded, so the residue does not depend on
    // which libm the reference was built with)
    double ef = fast ? eff : exp_cr(__dmul_rn(__dmul_rn(bn2, E), E) - __dmul_rn(azve, azve));
    cd w = faddeeva_w(bn * E, azve);
    cd F = cd{2.0 * ff.x, 2.0 * (__dmul_rn(ef, w.y) + ff.y)};
    double ib = 1.0 / bn;
    t.g = ce * F * ib;
    if (M != GF_QG) {
      cd Gz = cd{2.0 * (__dmul_rn(ef, w.x) - ff.x), -2.0 * ff.y};
      if (!fast && rz != 0.0) {
        // level to within rounding, not exactly: ef Re w - cos(bn z) is pure
        // cancellation in floating point (the reference's stops there follow
        // its libm); taken in exact arithmetic to first order in z:
        // -y e^{x^2} (2/sqrt(pi) - 2 x Im w(x)), x = bn E, y = z/2E
        Gz.x = -2.0 * azve * ef * (2.0 * isp - 2.0 * (bn * E) * w.y);
      }
      t.dg[0] = t.g * an0;
      t.dg[1] = t.g * an1;
      t.dg[2] = ce * Gz;
      if (M == GF_V2) {
        t.d2g[0] = t.dg[0] * an0;
        t.d2g[1] = t.dg[1] * an1;
        t.d2g[2] = ce * (mul_i(F) * bn + cd{2.0 * isp / E * ef, 0.0});
        t.d2g[3] = t.dg[0] * an1;
        t.d2g[4] = an1 * t.dg[2];
        t.d2g[5] = an0 * t.dg[2];
      }
    }
    return 0;
  } else if (bn2 < 0.0) {
    double bn, ib, f1, f2, efs = 0.0;
    if (fast) {
#pragma clang fp contract(fast)
      ib = rsqrt_pos(-bn2);
      bn = -bn2 * ib;
      efs = eff;
      f1 = efs * erfcx_w<M>(bn * E + azve);
      f2 = efs * erfcx_w<M>(bn * E - azve);
    } else {
      bn = sqrt(-bn2);
      ib = 1.0 / bn;
      double ebz = exp(bn * rz);
      f1 = ebz * erfc(bn * E + azve);
      f2 = 1.0 / ebz * erfc(bn * E - azve);
    }
    t.g = mul_i(ce) * (-(f1 + f2) * ib);
    if (M != GF_QG) {
      t.dg[0] = t.g * an0;
      t.dg[1] = t.g * an1;
      t.dg[2] = ce * (f1 - f2);
      if (!fast && rz != 0.0) {
        // level to within rounding, not exactly: f1 - f2 is pure cancellation;
        // in exact arithmetic to first order in z, with g(d) = e^{2xd} erfc(x + d),
        // x = bn E, d = z/2E: f1 - f2 = 2 d g'(0) = 2 d (2 x erfc(x) - (2/sqrt(pi)) e^{-x^2})
        const double x = bn * E;
        t.dg[2] = ce * (2.0 * azve * (2.0 * x * erfc(x) - 2.0 * isp * exp(-x * x)));
      }
      if (M == GF_V2) {
        double ef = fast ? efs : exp(bn2 * E * E - azve * azve);
        t.d2g[0] = t.dg[0] * an0;
        t.d2g[1] = t.dg[1] * an1;
        t.d2g[2] = ce * (-bn * (f1 + f2) + 2.0 * isp / E * ef);
        t.d2g[3] = t.dg[0] * an1;
        t.d2g[4] = an1 * t.dg[2];
        t.d2g[5] = an0 * t.dg[2];
      }
    }
    return 1;
  }
  return -3;
}

// pure spectral order: d3_qpgf_d2.c:581-604, :1013-1043
// ce: the planar phase exp(i a_n . r), carried along the row by the caller
template <int M>
__device__ __forceinline__ int fd_term(const GfParams& P, double rx, double ry, double rz, int n1, int n2, cd ce,
                                       Sum10& t) {
  // the pure spectral sum is never taken at z == 0: products may be fused
#pragma clang fp contract(fast)
  double an0 = -P.vkx + n1 * P.g1x + n2 * P.g2x;
  double an1 = -P.vky + n1 * P.g1y + n2 * P.g2y;
  double bn2 = P.k * P.k - an0 * an0 - an1 * an1;
  int kind;
  cd e;
  double bn;
  if (bn2 > 0.0) {
    bn = sqrt(bn2);
    e = cis(an0 * rx + an1 * ry + bn * fabs(rz));
    t.g = e / bn;
    kind = 0;
  } else if (bn2 < 0.0) {
    bn = sqrt(-bn2);
    e = ce * exp(-bn * fabs(rz));
    t.g = mul_i(e) * (-1.0 / bn);
    kind = 1;
  } else {
    return -2;
  }
  t.dg[2] = e;
  t.dg[0] = t.g * an0;
  t.dg[1] = t.g * an1;
  if (M == GF_V2) {
    t.d2g[0] = an0 * t.dg[0];
    t.d2g[1] = an1 * t.dg[1];
    t.d2g[2] = bn2 * t.g;
    t.d2g[3] = an0 * t.dg[1];
    t.d2g[4] = an1 * t.dg[2];
    t.d2g[5] = an0 * t.dg[2];
  }
  return kind;
}

// ---------------------------------------------------------------- rows
// terms summed by one evaluation, split by kind (spatial images, propagating and
// evanescent spectral orders of the Ewald sum, orders of the pure spectral sum)
struct GfTally {
  int sp, pr, ev, fd;
};

// Each row (and each sum of rows below) runs as one loop: the centre term, then
// outward in +1 direction until the stopping rule fires, then in -1 direction
// from the centre.  The order of the terms and the stopping rules are those of
// the reference's three loops (centre, up, down); one loop keeps a single copy
// of the term code and lets lanes in different phases run together.

// cs: exp(i k_t . d2), the Bloch phase step from image l2 to l2 + 1
template <int M>
__device__ int spatial_row(const GfParams& P, double rx, double ry, double rz, double E, double eps, int l1,
                           cd cs, Sum10& S) {
  const int lc = -(int)round(((rx + l1 * P.d1x) * P.d2x + (ry + l1 * P.d1y) * P.d2y) / P.d2n);
  const cd c0 = cis(P.vkx * (l1 * P.d1x + lc * P.d2x) + P.vky * (l1 * P.d1y + lc * P.d2y));
  const cd csb = conj(cs);
  // exp(k^2 E^2 - rl^2 / 4E^2) along the row: rl^2 is quadratic in l2, so the
  // factor from one image to the next changes by the constant q each step
  const double i4e = 0.25 / (E * E);
  const double ax = rx + l1 * P.d1x + lc * P.d2x, ay = ry + l1 * P.d1y + lc * P.d2y;
  const double bb = ax * P.d2x + ay * P.d2y;
  const double e0 = exp(P.k * P.k * E * E - (ax * ax + ay * ay + rz * rz) * i4e);
  const double q = exp(-2.0 * P.d2n * i4e);
  const double fu0 = exp(-(2.0 * bb + P.d2n) * i4e), fd0 = exp((2.0 * bb - P.d2n) * i4e);
  Sum10 t;
  int l2 = lc, dir = 0, cc = 0;
  cd ce = c0;
  double ef = e0, fs = fu0;
  for (;;) {
    spatial_term<M>(P, rx, ry, rz, E, l1, l2, ce, ef, t);
    sum_add<M>(S, t);
    cc++;
    if (dir == 0) {
      dir = 1;
    } else if (term_small_spatial<M>(t, S, eps)) {
      if (dir < 0) break;
      dir = -1;
      l2 = lc;
      ce = c0;
      ef = e0;
      fs = fd0;
    } else if (dir > 0 ? l2 + 1 == lc + BQD2_EW_LIMIT : l2 - 1 == lc - BQD2_EW_LIMIT) {
      return -1;
    }
    l2 += dir;
    ce = ce * (dir > 0 ? cs : csb);
    ef = ef * fs;
    fs = fs * q;
  }
  return cc;
}

// the row flag is 1 when no order other than the centre one was propagating:
// the reference counts propagating orders only in its outward loops.
// fast: planar phases carried along the row (cs = exp(i g2 . r) per step)
template <int M>
__device__ int spectral_row(const GfParams& P, double rx, double ry, double rz, double E, double eps, int n1,
                            bool fast, cd cs, Sum10& S, int& flag, int& nev) {
  const int nc = (int)round(-((-P.vkx + n1 * P.g1x) * P.g2x + (-P.vky + n1 * P.g1y) * P.g2y) / P.g2n);
  cd c0 = cd{1.0, 0.0};
  // fast: exp((k^2 - |a_n|^2) E^2 - z^2/4E^2) along the row, |a_n|^2 quadratic in n2
  double e0 = 0.0, q = 0.0, fu0 = 0.0, fd0 = 0.0;
  if (fast) {
    const double ax = -P.vkx + n1 * P.g1x + nc * P.g2x, ay = -P.vky + n1 * P.g1y + nc * P.g2y;
    c0 = cis(ax * rx + ay * ry);
    const double e2 = E * E, azve = rz / (2.0 * E), bb = ax * P.g2x + ay * P.g2y;
    e0 = exp((P.k * P.k - ax * ax - ay * ay) * e2 - azve * azve);
    q = exp(-2.0 * P.g2n * e2);
    fu0 = exp(-(2.0 * bb + P.g2n) * e2);
    fd0 = exp((2.0 * bb - P.g2n) * e2);
  }
  const cd csb = conj(cs);
  Sum10 t;
  int n2 = nc, dir = 0, cc = 0, prop = 0;
  cd ce = c0;
  double ef = e0, fs = fu0;
  for (;;) {
    const int k = spectral_term<M>(P, rx, ry, rz, E, n1, n2, fast, ce, ef, t);
    if (k < 0) return k;
    sum_add<M>(S, t);
    nev += k;
    cc++;
    if (dir == 0) {
      dir = 1;
    } else if (k > 0 && term_small_spectral<M>(t, S, eps)) {
      if (dir < 0) break;
      dir = -1;
      n2 = nc;
      ce = c0;
      ef = e0;
      fs = fd0;
    } else {
      if (k == 0) prop++;
      if (dir > 0 ? n2 + 1 == nc + BQD2_EW_LIMIT : n2 - 1 == nc - BQD2_EW_LIMIT) return -2;
    }
    n2 += dir;
    ce = ce * (dir > 0 ? cs : csb);
    ef = ef * fs;
    fs = fs * q;
  }
  flag = (prop == 0);
  return cc;
}

template <int M>
__device__ int fd_row(const GfParams& P, double rx, double ry, double rz, double eps, int n1, cd cs, Sum10& S) {
  const int nc = (int)round((P.vkx * P.g2x + P.vky * P.g2y - n1 * P.g12) / P.g2n);
  const cd c0 = cis((-P.vkx + n1 * P.g1x + nc * P.g2x) * rx + (-P.vky + n1 * P.g1y + nc * P.g2y) * ry);
  const cd csb = conj(cs);
  Sum10 t;
  int n2 = nc, dir = 0, cc = 0;
  cd ce = c0;
  for (;;) {
    const int k = fd_term<M>(P, rx, ry, rz, n1, n2, ce, t);
    if (k < 0) return k;
    sum_add<M>(S, t);
    cc++;
    if (dir == 0) {
      dir = 1;
    } else if (k > 0 && term_small_fd<M>(t, S, eps)) {
      if (dir < 0) break;
      dir = -1;
      n2 = nc;
      ce = c0;
    } else if (dir > 0 ? n2 + 1 == nc + BQD2_FD_LIMIT : n2 - 1 == nc - BQD2_FD_LIMIT) {
      return -1;
    }
    n2 += dir;
    ce = ce * (dir > 0 ? cs : csb);
  }
  return cc;
}

__device__ __forceinline__ void sum_zero(Sum10& s) {
  s.g = cd{0, 0};
#pragma unroll
  for (int i = 0; i < 3; i++) s.dg[i] = cd{0, 0};
#pragma unroll
  for (int i = 0; i < 6; i++) s.d2g[i] = cd{0, 0};
}

// Ewald evaluation in the unit cell (d3_qpgf_d2.c:1671-1712, :2011-2052).
// parts selects the shares summed (1: images, 2: spectral orders, 3: both):
// the result is linear in the two, so a caller may take them in two calls (two
// kernels) and hold neither across the other's sum.
template <int M, int parts = 3>
__device__ int ewald_cell(const GfParams& P, double rx, double ry, double rz, double eps, Sum10& out, GfTally& tl) {
  double E = P.E0;
  double r2 = rx * rx + ry * ry + rz * rz;
  double k2 = P.k * P.k;
  if (k2 * E * E - r2 / (4.0 * E * E) > EW_EXPM_D) E = sqrt(EW_EXPM_D + sqrt(k2 * r2 + EW_EXPM_D * EW_EXPM_D)) / (sqrt(2.0) * P.k);

  // Bloch phase step of the images and planar phase step of the spectral
  // orders along a row; the spectral recurrence is used away from z == 0
  const cd css = cis(P.vkx * P.d2x + P.vky * P.d2y);
  const cd csg = cis(P.g2x * rx + P.g2y * ry);
  const bool fast = fabs(rz) >= BQD2_ZFAST * E;
  const double ifp = 0.079577471545947667884;  // 1/(4 pi)
  const double iep = 0.039788735772973833942;  // 1/(8 pi)
  Key2 old;
  int c1 = 0;
  if (parts & 1) {
  Sum10 S1;
  sum_zero(S1);
  for (int l1 = 0, dir = 0;;) {
    old = key_spatial<M>(S1);
    const int c = spatial_row<M>(P, rx, ry, rz, E, eps, l1, css, S1);
    if (c < 0) return -1;
    c1 += c;
    if (dir == 0) {
      dir = 1;
    } else if (rows_conv_spatial<M>(S1, old, eps)) {
      if (dir < 0) break;
      dir = -1;
      l1 = 0;
    } else if (dir > 0 ? l1 + 1 == BQD2_EW_LIMIT : l1 - 1 == -BQD2_EW_LIMIT) {
      return -1;
    }
    l1 += dir;
  }
  // the image sum's share of the result (S1 is not needed past this point)
  if (M != GF_DQG) out.g = S1.g * ifp;
  if (M != GF_QG) {
    out.dg[0] = S1.dg[0] * (-ifp);
    out.dg[1] = S1.dg[1] * (-ifp);
    out.dg[2] = S1.dg[2] * (-rz * ifp);
  }
  if (M == GF_V2) {
    out.d2g[0] = S1.d2g[0] * (-iep);
    out.d2g[1] = S1.d2g[1] * (-iep);
    out.d2g[2] = S1.d2g[2] * (-iep);
    out.d2g[3] = S1.d2g[3] * (-iep);
    out.d2g[4] = S1.d2g[4] * (-rz * iep);
    out.d2g[5] = S1.d2g[5] * (-rz * iep);
  }
  }

  if (!(parts & 2)) {
    tl.sp += c1;
    return c1;
  }
  Sum10 S2;
  sum_zero(S2);
  int flag = 0, nev = 0, c2 = 0;
  for (int n1 = P.nc, dir = 0;;) {
    old = key_spectral<M>(S2);
    const int c = spectral_row<M>(P, rx, ry, rz, E, eps, n1, fast, csg, S2, flag, nev);
    if (c < 0) return c;
    c2 += c;
    if (dir == 0) {
      dir = 1;
    } else if (flag && rows_conv_spectral<M>(S2, old, eps)) {
      if (dir < 0) break;
      dir = -1;
      n1 = P.nc;
    } else if (dir > 0 ? n1 + 1 == P.nc + BQD2_EW_LIMIT : n1 - 1 == P.nc - BQD2_EW_LIMIT) {
      return -2;
    }
    n1 += dir;
  }

  tl.sp += c1;
  tl.ev += nev;
  tl.pr += c2 - nev;
  double cdd = 0.25 * P.aidd;
  if (M != GF_DQG) out.g = out.g + mul_i(S2.g) * cdd;
  if (M != GF_QG) {
    out.dg[0] = out.dg[0] - S2.dg[0] * cdd;
    out.dg[1] = out.dg[1] - S2.dg[1] * cdd;
    out.dg[2] = out.dg[2] + S2.dg[2] * cdd;
  }
  if (M == GF_V2) {
    out.d2g[0] = out.d2g[0] - mul_i(S2.d2g[0]) * cdd;
    out.d2g[1] = out.d2g[1] - mul_i(S2.d2g[1]) * cdd;
    out.d2g[2] = out.d2g[2] - S2.d2g[2] * cdd;
    out.d2g[3] = out.d2g[3] - mul_i(S2.d2g[3]) * cdd;
    out.d2g[4] = out.d2g[4] + mul_i(S2.d2g[4]) * cdd;
    out.d2g[5] = out.d2g[5] + mul_i(S2.d2g[5]) * cdd;
  }
  return c1 + c2;
}

// pure spectral evaluation in the unit cell (d3_qpgf_d2.c:484-533, :887-950)
template <int M>
__device__ int fd_cell(const GfParams& P, double rx, double ry, double rz, double eps, Sum10& out, GfTally& tl) {
  if (rz == 0.0) return -4;
  const cd csg = cis(P.g2x * rx + P.g2y * ry);
  Sum10 S;
  Key2 old;
  sum_zero(S);
  int cc = 0;
  for (int n1 = P.nc, dir = 0;;) {
    old = key_fd<M>(S);
    const int c = fd_row<M>(P, rx, ry, rz, eps, n1, csg, S);
    if (c < 0) return c;
    cc += c;
    if (dir == 0) {
      dir = 1;
    } else if (rows_conv_fd<M>(S, old, eps)) {
      if (dir < 0) break;
      dir = -1;
      n1 = P.nc;
    } else if (dir > 0 ? n1 + 1 == P.nc + BQD2_FD_LIMIT : n1 - 1 == P.nc - BQD2_FD_LIMIT) {
      return -1;
    }
    n1 += dir;
  }
  tl.fd += cc;
  double a = P.aidd, sz = rz / fabs(rz);
  out.g = mul_i(S.g) * (0.5 * a);
  out.dg[0] = S.dg[0] * (-0.5 * a);
  out.dg[1] = S.dg[1] * (-0.5 * a);
  out.dg[2] = S.dg[2] * (-0.5 * a * sz);
  if (M == GF_V2) {
#pragma unroll
    for (int i = 0; i < 4; i++) out.d2g[i] = mul_i(S.d2g[i]) * (-0.5 * a);
    out.d2g[4] = mul_i(S.d2g[4]) * (-0.5 * a * sz);
    out.d2g[5] = mul_i(S.d2g[5]) * (-0.5 * a * sz);
  }
  return cc;
}

// whether the evaluation at offset z from the lattice plane takes the pure
// spectral sum (its truncation error estimate below eps; d3_qpgf_d2.c's choice)
__device__ __forceinline__ bool qpgf_uses_fd(const GfParams& P, double rz, double eps) {
  if (P.ewald_only) return false;
  const double zs = fabs(rz) / P.fd_lift;
  return (P.bN / P.bNp * exp(-P.bNp * zs) < eps) && (P.bN / P.bNm * exp(-P.bNm * zs) < eps);
}

// full evaluation at offset r (source minus target); returns the term count or
// the reference's negative code (-1x pure spectral, -2x Ewald).  PARTS: the
// shares to sum, 1 Ewald images, 2 Ewald spectral orders, 4 pure spectral sum
// (the other kind of evaluation then returns zero); the result is linear in
// them, so a caller may take them in separate kernels, each of which holds
// only its own sum's registers.
template <int M, int PARTS = 7>
__device__ int qpgf_eval(const GfParams& P, double rx, double ry, double rz, double eps, Sum10& out, GfTally& tl) {
  const bool use_fd = qpgf_uses_fd(P, rz, eps);
  int l1 = (int)floor((rx * P.d2y - ry * P.d2x) * P.idd + 0.5);
  int l2 = (int)floor(-(rx * P.d1y - ry * P.d1x) * P.idd + 0.5);
  double tx = rx - l1 * P.d1x - l2 * P.d2x;
  double ty = ry - l1 * P.d1y - l2 * P.d2y;
  sum_zero(out);
  int err;
  if (use_fd) {
    if (!(PARTS & 4)) return 0;
    err = fd_cell<M>(P, tx, ty, rz, eps, out, tl);
  } else {
    if (!(PARTS & 3)) return 0;
    err = ewald_cell<M, PARTS & 3>(P, tx, ty, rz, eps, out, tl);
  }
  if (l1 != 0 || l2 != 0) {
    double arg = -(P.vkx * (l1 * P.d1x + l2 * P.d2x) + P.vky * (l1 * P.d1y + l2 * P.d2y));
    cd cf = cis(arg);
    out.g *= cf;
#pragma unroll
    for (int i = 0; i < 3; i++) out.dg[i] *= cf;
    if (M == GF_V2) {
#pragma unroll
      for (int i = 0; i < 6; i++) out.d2g[i] *= cf;
    }
  }
  if (err < 0) err -= use_fd ? 10 : 20;
  return err;
}

// Value and gradient as the reference's singular (polar) integrands take them:
// the value from the value-only sum and the gradient from the gradient-only sum
// (d3hm_qpgf_d2_qG / _dqG, d3_qpgf_d2.c:10-122; called by q0_bil_sqG_r /
// q0_bil_sqH_r, q0_bil_coef.c:428-497, and q0_lit_sqG_r / _sqH_r).  Their
// stopping rules read different partial sums, which matters only where the
// offset is level with the source (z == 0: the Ewald spectral stop reads a
// rounding residue, ewald_cell's non-fast range); elsewhere one value-and-
// gradient sum gives both to rounding and is taken instead.
__device__ __forceinline__ int qpgf_eval_g_dg(const GfParams& P, double rx, double ry, double rz, double eps,
                                              Sum10& out, GfTally& tl) {
  if (fabs(rz) >= BQD2_ZFAST * P.E0) return qpgf_eval<GF_V1>(P, rx, ry, rz, eps, out, tl);
  const int c0 = qpgf_eval<GF_QG>(P, rx, ry, rz, eps, out, tl);
  if (c0 < 0) return c0;
  Sum10 d;
  const int c1 = qpgf_eval<GF_DQG>(P, rx, ry, rz, eps, d, tl);
  if (c1 < 0) return c1;
#pragma unroll
  for (int i = 0; i < 3; i++) out.dg[i] = d.dg[i];
  return c0 + c1;
}

template <int M>
__device__ __forceinline__ int qpgf_eval(const GfParams& P, double rx, double ry, double rz, double eps, Sum10& out) {
  GfTally tl{0, 0, 0, 0};
  return qpgf_eval<M>(P, rx, ry, rz, eps, out, tl);
}

