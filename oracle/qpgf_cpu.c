/*
 * qpgf_cpu.c -- oracle quasi-periodic Helmholtz Green's function (2-D lattice).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates qgf2_src/d3_qpgf_d2.c: the method choice of d3hm_qpgf_d2_v1/_v2/_qG/
 * _dqG (:10-230), the unit-cell reduction and Bloch phase (:232-468), the pure
 * spectral ("fd") sums (:470-1043), and the Ewald split into a spatial sum over
 * lattice images (:1714-1838, :2054-2170) and a spectral sum over reciprocal
 * vectors (:1840-2009, :2172-2398), with the same row-by-row adaptive stopping
 * rules (d3hm_sf_comp_z / d3hm_sf_diff_z, :471-482) so that the term counts agree
 * with the reference's return values.  One routine serves all four modes; the
 * mode selects which partial sums drive the stopping tests, exactly as the four
 * reference chains do.
 *
 * Pins: tests/test_oracle_gf.py compares every mode, value and term count
 * with the reference's own d3_qpgf_d2.c compiled into oracle/_ref, and checks
 * quasi-periodicity G(r+R) = exp(-i k.R) G(r), the Helmholtz equation
 * trace(Hessian) = -k^2 G and the smoothness of G across the switch between
 * the Ewald split and the pure spectral sum.
 */
#include <math.h>
#include <complex.h>
#include <stdlib.h>
#include <quadmath.h>
#include "oracle.h"

/* Where the offset lies level with the source the stopping rules of the Ewald
 * spectral sum read cancellation: the z-derivative of a propagating order,
 * exp(bn^2 E^2) Re w(bn E + i z/2E) - cos(bn z) (d3_qpgf_d2.c:1976-1989,
 * Faddeeva.cc:684-686 on the real axis), and of an evanescent one,
 * e^{bz} erfc(bE + z/2E) - e^{-bz} erfc(bE - z/2E) (:1991-2004).
 *   z == 0 exactly: the evanescent difference is exactly 0 on every platform
 *     and the propagating one is the rounding residue of the two exponentials.
 *   0 < |z| < 2e-8 E: both are floating-point noise, so where the reference
 *     stops follows its libm's rounding (glibc's values there are up to 1e-2
 *     off the limit z -> 0 of its own values).
 * Mode 0: this libm's exp and erfc throughout, as the reference compiled here.
 * Mode 1, the platform-independent statement that the GPU evaluates: at z == 0
 * the two exponentials correctly rounded (libquadmath's 113-bit expq rounded
 * once to double; glibc 2.35's exp misrounds about 0.08 % of arguments); at
 * 0 < |z| < 2e-8 E the two differences in exact arithmetic, to first order in
 * z (relative error below 1e-15):
 *   propagating  -y e^{x^2} (2/sqrt(pi) - 2 x Im w(x)),
 *   evanescent   2 d (2 x erfc(x) - (2/sqrt(pi)) e^{-x^2}),
 * x = bn E, y = d = z/2E.  Mode 1 is continuous in z down to z -> 0. */
static int cr_exp_mode = 0;
void orc_set_cr_exp(int on) { cr_exp_mode = on; }
int orc_cr_mode(void) { return cr_exp_mode; }
double orc_exp_z0(double x) { return cr_exp_mode ? (double)expq((__float128)x) : exp(x); }
#define ZLEVEL 2.0e-8

enum { M_QG = 0, M_DQG = 1, M_V1 = 2, M_V2 = 3 };

/* d3_qpgf_d2.c:471-476 : two totals agree to eps in both parts */
static int conv_same(double complex a, double complex b, double eps)
{
  return 2.0 * fabs(creal(a) - creal(b)) <= eps * (fabs(creal(a)) + fabs(creal(b))) &&
         2.0 * fabs(cimag(a) - cimag(b)) <= eps * (fabs(cimag(a)) + fabs(cimag(b)));
}

/* d3_qpgf_d2.c:478-482 : a term is negligible against the running total */
static int conv_small(double complex t, double complex s, double eps)
{
  return fabs(creal(t)) <= fabs(creal(s)) * EFS * eps &&
         fabs(cimag(t)) <= fabs(cimag(s)) * EFS * eps;
}

typedef struct {
  double complex g, dg[3], d2g[6];
} sum10;

static void sum_add(sum10 *s, const sum10 *t, int mode)
{
  s->g += t->g;
  for (int i = 0; i < 3; i++) s->dg[i] += t->dg[i];
  if (mode == M_V2) for (int i = 0; i < 6; i++) s->d2g[i] += t->d2g[i];
}

/* totals that decide when a whole row of terms has stopped changing */
static int rows_converged_spatial(int mode, const sum10 *s, const sum10 *o, double eps)
{
  switch (mode) {
  case M_QG: case M_V1: return conv_same(s->g, o->g, eps);
  case M_DQG: return conv_same(s->dg[0], o->dg[0], eps) && conv_same(s->dg[1], o->dg[1], eps);
  default: return conv_same(s->d2g[0], o->d2g[0], eps) && conv_same(s->d2g[1], o->d2g[1], eps);
  }
}
static int term_small_spatial(int mode, const sum10 *t, const sum10 *s, double eps)
{
  switch (mode) {
  case M_QG: case M_V1: return conv_small(t->g, s->g, eps);
  case M_DQG: return conv_small(t->dg[0], s->dg[0], eps) && conv_small(t->dg[1], s->dg[1], eps);
  default: return conv_small(t->d2g[0], s->d2g[0], eps) && conv_small(t->d2g[1], s->d2g[1], eps);
  }
}
static int rows_converged_spectral(int mode, const sum10 *s, const sum10 *o, double eps)
{
  switch (mode) {
  case M_QG: return conv_same(s->g, o->g, eps);
  case M_DQG: case M_V1: return conv_same(s->dg[2], o->dg[2], eps);
  default: return conv_same(s->d2g[0], o->d2g[0], eps) && conv_same(s->d2g[1], o->d2g[1], eps);
  }
}
static int term_small_spectral(int mode, const sum10 *t, const sum10 *s, double eps)
{
  switch (mode) {
  case M_QG: return conv_small(t->g, s->g, eps);
  case M_DQG: case M_V1: return conv_small(t->dg[2], s->dg[2], eps);
  default: return conv_small(t->d2g[0], s->d2g[0], eps) && conv_small(t->d2g[1], s->d2g[1], eps);
  }
}
/* the pure spectral sum watches dG/dz (value mode: G; Hessian mode: d2G/dz2) */
static int rows_converged_fd(int mode, const sum10 *s, const sum10 *o, double eps)
{
  switch (mode) {
  case M_QG: return conv_same(s->g, o->g, eps);
  case M_DQG: case M_V1: return conv_same(s->dg[2], o->dg[2], eps);
  default: return conv_same(s->d2g[2], o->d2g[2], eps);
  }
}
static int term_small_fd(int mode, const sum10 *t, const sum10 *s, double eps)
{
  switch (mode) {
  case M_QG: return conv_small(t->g, s->g, eps);
  case M_DQG: case M_V1: return conv_small(t->dg[2], s->dg[2], eps);
  default: return conv_small(t->d2g[2], s->d2g[2], eps);
  }
}

static inline double complex cis(double a) { return cos(a) + I * sin(a); }

/* ---------------- Ewald spatial term (d3_qpgf_d2.c:1814-1838, :2131-2170) ---------- */
static void spatial_term(int mode, const QPD2 *qd, const double *r, double E, int l1, int l2, sum10 *t)
{
  double Rl0 = l1 * qd->vd1[0] + l2 * qd->vd2[0];
  double Rl1 = l1 * qd->vd1[1] + l2 * qd->vd2[1];
  double ra0 = r[0] + Rl0, ra1 = r[1] + Rl1, ra2 = r[2];
  double rl2 = ra0 * ra0 + ra1 * ra1 + ra2 * ra2;
  double rl = sqrt(rl2), irl = 1.0 / rl;
  double ke = qd->k * E, iE = 1.0 / E, y = rl * 0.5 * iE;
  double complex ce = cis(qd->vk[0] * Rl0 + qd->vk[1] * Rl1);
  double complex efw = orc_scaled_w(ke, y);          /* exp(ke^2 - y^2) w(ke + i y) */
  double complex cef_re = ce * creal(efw);
  double complex cef_im = ce * cimag(efw);
  double ef = exp(ke * ke - y * y);
  double isp = 1.0 / sqrt(M_PI);
  t->g = cef_re * irl;
  double complex cf1 = irl * irl * (cef_re * irl - qd->k * cef_im + ce * ef * iE * isp);
  t->dg[0] = ra0 * cf1;
  t->dg[1] = ra1 * cf1;
  t->dg[2] = cf1;
  if (mode == M_V2) {
    double complex cf2 = irl * irl * (qd->k * qd->k * irl * 2.0 * cef_re - ce * ef * iE * iE * iE * isp);
    double complex cf3 = -3.0 * irl * irl * 2.0 * cf1 + cf2;
    t->d2g[0] = 2.0 * cf1 + ra0 * ra0 * cf3;
    t->d2g[1] = 2.0 * cf1 + ra1 * ra1 * cf3;
    t->d2g[2] = 2.0 * cf1 + ra2 * ra2 * cf3;
    t->d2g[3] = ra0 * ra1 * cf3;
    t->d2g[4] = ra1 * cf3;
    t->d2g[5] = ra0 * cf3;
  }
}

/* ---------------- Ewald spectral term (d3_qpgf_d2.c:1965-2009, :2349-2398) -------- */
static int spectral_term(int mode, const QPD2 *qd, const double *r, double E,
                         const double *g1, const double *g2, int n1, int n2, sum10 *t)
{
  double an0 = -qd->vk[0] + n1 * g1[0] + n2 * g2[0];
  double an1 = -qd->vk[1] + n1 * g1[1] + n2 * g2[1];
  double bn2 = qd->k * qd->k - an0 * an0 - an1 * an1;
  double complex ce = cis(an0 * r[0] + an1 * r[1]);
  double azve = r[2] / (2.0 * E);
  double isp = 1.0 / sqrt(M_PI);
  if (bn2 > 0.0) {
    double bn = sqrt(bn2);
    double complex ff = cis(bn * r[2]);
    /* exp(bn^2 E^2 - azve^2) and w(bn E + i azve) kept apart as in :1976-1985:
     * at z == 0 the difference ef*Re(w) - ff is pure rounding and the row
     * stopping rule depends on it */
    const int level = fabs(r[2]) < ZLEVEL * E;
    double ef = r[2] == 0.0 ? orc_exp_z0(bn2 * E * E - azve * azve) : exp(bn2 * E * E - azve * azve);
    double complex w = orc_faddeeva_w(bn * E + I * azve);
    double complex F = 2.0 * (I * ef * cimag(w) + ff);
    double complex Gz = 2.0 * (ef * creal(w) - ff);
    if (cr_exp_mode && level && r[2] != 0.0) {
      /* level but not exactly (0 < |z| < 2e-8 E): ef Re w - cos(bn z) in
       * exact arithmetic, -y e^{x^2} (2/sqrt(pi) - 2 x Im w(x)) + O(y^3) with
       * x = bn E, y = z/2E, instead of the rounding residue it is in floating
       * point (whose stopping decisions follow the libm) */
      Gz = 2.0 * (-azve * ef * (1.1283791670955126 - 2.0 * (bn * E) * cimag(w))) - 2.0 * I * cimag(ff);
    }
    t->g = ce / bn * F;
    t->dg[0] = t->g * an0;
    t->dg[1] = t->g * an1;
    t->dg[2] = ce * Gz;
    if (mode == M_V2) {
      t->d2g[0] = t->dg[0] * an0;
      t->d2g[1] = t->dg[1] * an1;
      t->d2g[2] = ce * (I * bn * F + 2.0 * isp / E * ef);
      t->d2g[3] = t->dg[0] * an1;
      t->d2g[4] = an1 * t->dg[2];
      t->d2g[5] = an0 * t->dg[2];
    }
    return 0;
  } else if (bn2 < 0.0) {
    double bn = sqrt(-bn2);
    const int level = fabs(r[2]) < ZLEVEL * E;
    double ebz = exp(bn * r[2]);
    double f1 = ebz * erfc(bn * E + azve);
    double f2 = 1.0 / ebz * erfc(bn * E - azve);
    t->g = -I * ce / bn * (f1 + f2);
    t->dg[0] = t->g * an0;
    t->dg[1] = t->g * an1;
    t->dg[2] = ce * (f1 - f2);
    if (cr_exp_mode && level && r[2] != 0.0) {
      /* level but not exactly: f1 - f2 = g(d) - g(-d), g(d) = e^{2xd} erfc(x + d),
       * x = bn E, d = z/2E, in exact arithmetic 2 d g'(0) + O(d^3) with
       * g'(0) = 2 x erfc(x) - (2/sqrt(pi)) e^{-x^2} */
      const double x = bn * E;
      t->dg[2] = ce * (2.0 * azve * (2.0 * x * erfc(x) - 1.1283791670955126 * exp(-x * x)));
    }
    if (mode == M_V2) {
      double ef = exp(bn2 * E * E - azve * azve);
      t->d2g[0] = t->dg[0] * an0;
      t->d2g[1] = t->dg[1] * an1;
      t->d2g[2] = ce * (-bn * (f1 + f2) + 2.0 * isp / E * ef);
      t->d2g[3] = t->dg[0] * an1;
      t->d2g[4] = an1 * t->dg[2];
      t->d2g[5] = an0 * t->dg[2];
    }
    return 1;
  }
  return -3; /* Wood anomaly, beta_n == 0 */
}

/* ---------------- pure spectral term (d3_qpgf_d2.c:581-604, :1013-1043) ---------- */
static int fd_term(int mode, const QPD2 *qd, const double *r, const double *g1, const double *g2,
                   int n1, int n2, sum10 *t)
{
  double an0 = -qd->vk[0] + n1 * g1[0] + n2 * g2[0];
  double an1 = -qd->vk[1] + n1 * g1[1] + n2 * g2[1];
  double bn2 = qd->k * qd->k - an0 * an0 - an1 * an1;
  int kind;
  if (bn2 > 0.0) {
    double bn = sqrt(bn2);
    t->dg[2] = cis(an0 * r[0] + an1 * r[1] + bn * fabs(r[2]));
    t->g = t->dg[2] / bn;
    kind = 0;
  } else if (bn2 < 0.0) {
    double bn = sqrt(-bn2);
    t->dg[2] = cis(an0 * r[0] + an1 * r[1]) * exp(-bn * fabs(r[2]));
    t->g = -I * t->dg[2] / bn;
    kind = 1;
  } else return -2;
  t->dg[0] = t->g * an0;
  t->dg[1] = t->g * an1;
  if (mode == M_V2) {
    t->d2g[0] = an0 * t->dg[0];
    t->d2g[1] = an1 * t->dg[1];
    t->d2g[2] = bn2 * t->g;
    t->d2g[3] = an0 * t->dg[1];
    t->d2g[4] = an1 * t->dg[2];
    t->d2g[5] = an0 * t->dg[2];
  }
  return kind;
}

static void reciprocal(const QPD2 *qd, double idd, double *g1, double *g2)
{
  g1[0] = 2.0 * M_PI * idd * qd->vd2[1];
  g1[1] = -2.0 * M_PI * idd * qd->vd2[0];
  g2[0] = -2.0 * M_PI * idd * qd->vd1[1];
  g2[1] = 2.0 * M_PI * idd * qd->vd1[0];
}

/* one row of the spatial sum: l2 outward from the image nearest to -r (:1772-1812) */
static int spatial_row(int mode, const QPD2 *qd, const double *r, double E, double eps, int l1, sum10 *S)
{
  double d2n = qd->vd2[0] * qd->vd2[0] + qd->vd2[1] * qd->vd2[1];
  int lc = -(int)round(((r[0] + l1 * qd->vd1[0]) * qd->vd2[0] + (r[1] + l1 * qd->vd1[1]) * qd->vd2[1]) / d2n);
  sum10 t;
  int cc = 1, l2;
  spatial_term(mode, qd, r, E, l1, lc, &t);
  sum_add(S, &t, mode);
  for (l2 = lc + 1; l2 < lc + EW_LIMIT; l2++) {
    cc++;
    spatial_term(mode, qd, r, E, l1, l2, &t);
    sum_add(S, &t, mode);
    if (term_small_spatial(mode, &t, S, eps)) break;
  }
  if (l2 == lc + EW_LIMIT) return -1;
  for (l2 = lc - 1; l2 > lc - EW_LIMIT; l2--) {
    cc++;
    spatial_term(mode, qd, r, E, l1, l2, &t);
    sum_add(S, &t, mode);
    if (term_small_spatial(mode, &t, S, eps)) break;
  }
  if (l2 == lc - EW_LIMIT) return -1;
  return cc;
}

/* spatial sum over image rows l1 = 0, 1, 2, ... then -1, -2, ... (:1740-1770) */
static int spatial_sum(int mode, const QPD2 *qd, const double *r, double E, double eps, sum10 *S)
{
  sum10 old;
  int cc = 0, c, l1;
  c = spatial_row(mode, qd, r, E, eps, 0, S);
  if (c < 0) return c;
  cc += c;
  for (l1 = 1; l1 < EW_LIMIT; l1++) {
    old = *S;
    c = spatial_row(mode, qd, r, E, eps, l1, S);
    if (c < 0) return c;
    cc += c;
    if (rows_converged_spatial(mode, S, &old, eps)) break;
  }
  if (l1 == EW_LIMIT) return -1;
  for (l1 = -1; l1 > -EW_LIMIT; l1--) {
    old = *S;
    c = spatial_row(mode, qd, r, E, eps, l1, S);
    if (c < 0) return c;
    cc += c;
    if (rows_converged_spatial(mode, S, &old, eps)) break;
  }
  if (l1 == -EW_LIMIT) return -1;
  return cc;
}

/* one row of the Ewald spectral sum (:1916-1963). *flag: 1 if no order other
 * than the row's centre term was propagating (the reference counts propagating
 * orders only inside its two outward loops, :1938-1960); the row total may then
 * be tested for convergence. */
static int spectral_row(int mode, const QPD2 *qd, const double *r, double E, double eps,
                        const double *g1, const double *g2, int n1, sum10 *S, int *flag)
{
  double g2n = g2[0] * g2[0] + g2[1] * g2[1];
  int nc = (int)round(-((-qd->vk[0] + n1 * g1[0]) * g2[0] + (-qd->vk[1] + n1 * g1[1]) * g2[1]) / g2n);
  sum10 t;
  int cc = 1, n2, k, prop = 0;
  k = spectral_term(mode, qd, r, E, g1, g2, n1, nc, &t);
  if (k < 0) return k;
  sum_add(S, &t, mode);
  for (n2 = nc + 1; n2 < nc + EW_LIMIT; n2++) {
    cc++;
    k = spectral_term(mode, qd, r, E, g1, g2, n1, n2, &t);
    if (k < 0) return k;
    sum_add(S, &t, mode);
    if (k > 0) { if (term_small_spectral(mode, &t, S, eps)) break; }
    else prop++;
  }
  if (n2 == nc + EW_LIMIT) return -2;
  for (n2 = nc - 1; n2 > nc - EW_LIMIT; n2--) {
    cc++;
    k = spectral_term(mode, qd, r, E, g1, g2, n1, n2, &t);
    if (k < 0) return k;
    sum_add(S, &t, mode);
    if (k > 0) { if (term_small_spectral(mode, &t, S, eps)) break; }
    else prop++;
  }
  if (n2 == nc - EW_LIMIT) return -2;
  *flag = (prop == 0);
  return cc;
}

static int spectral_sum(int mode, const QPD2 *qd, const double *r, double E, double eps, sum10 *S)
{
  double det = qd->vd1[0] * qd->vd2[1] - qd->vd2[0] * qd->vd1[1];
  if (det == 0.0) return -4;
  double idd = 1.0 / det, g1[2], g2[2];
  int nc = (int)round((qd->vk[0] * qd->vd1[0] + qd->vk[1] * qd->vd1[1]) / (2.0 * M_PI));
  reciprocal(qd, idd, g1, g2);
  sum10 old;
  int cc = 0, c, n1, flag = 0;
  c = spectral_row(mode, qd, r, E, eps, g1, g2, nc, S, &flag);
  if (c < 0) return c;
  cc += c;
  for (n1 = nc + 1; n1 < nc + EW_LIMIT; n1++) {
    old = *S;
    c = spectral_row(mode, qd, r, E, eps, g1, g2, n1, S, &flag);
    if (c < 0) return c;
    cc += c;
    if (flag && rows_converged_spectral(mode, S, &old, eps)) break;
  }
  if (n1 == nc + EW_LIMIT) return -2;
  for (n1 = nc - 1; n1 > nc - EW_LIMIT; n1--) {
    old = *S;
    c = spectral_row(mode, qd, r, E, eps, g1, g2, n1, S, &flag);
    if (c < 0) return c;
    cc += c;
    if (flag && rows_converged_spectral(mode, S, &old, eps)) break;
  }
  if (n1 == nc - EW_LIMIT) return -2;
  return cc;
}

/* Ewald evaluation in the unit cell (:1671-1712, :2011-2052) */
static int ewald_cell(int mode, const QPD2 *qd, const double *r, double eps, sum10 *out)
{
  double det = qd->vd1[0] * qd->vd2[1] - qd->vd2[0] * qd->vd1[1];
  if (det == 0.0) return -4;
  double sx = qd->vd1[0] + qd->vd2[0], sy = qd->vd1[1] + qd->vd2[1];
  double dx = qd->vd1[0] - qd->vd2[0], dy = qd->vd1[1] - qd->vd2[1];
  double adp = sqrt(sx * sx + sy * sy), adm = sqrt(dx * dx + dy * dy);
  double E = sqrt(fabs(det) / (4.0 * M_PI) * adp / adm);
  double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
  double k2 = qd->k * qd->k;
  if (k2 * E * E - r2 / (4.0 * E * E) > EW_EXPM)
    E = sqrt(EW_EXPM + sqrt(k2 * r2 + EW_EXPM * EW_EXPM)) / (sqrt(2.0) * qd->k);

  sum10 S1 = {0}, S2 = {0};
  int c1 = spatial_sum(mode, qd, r, E, eps, &S1);
  if (c1 < 0) return c1;
  int c2 = spectral_sum(mode, qd, r, E, eps, &S2);
  if (c2 < 0) return c2;

  double ifp = 1.0 / (4.0 * M_PI);
  double cd = 0.25 * fabs(1.0 / det);
  out->g = S1.g * ifp + S2.g * (I * cd);
  out->dg[0] = -S1.dg[0] * ifp - S2.dg[0] * cd;
  out->dg[1] = -S1.dg[1] * ifp - S2.dg[1] * cd;
  out->dg[2] = -S1.dg[2] * r[2] * ifp + S2.dg[2] * cd;
  if (mode == M_V2) {
    double iep = 1.0 / (8.0 * M_PI);
    double complex ci = I * cd;
    for (int i = 0; i < 4; i++) out->d2g[i] = -S1.d2g[i] * iep;
    out->d2g[4] = -S1.d2g[4] * r[2] * iep;
    out->d2g[5] = -S1.d2g[5] * r[2] * iep;
    out->d2g[0] += -S2.d2g[0] * ci;
    out->d2g[1] += -S2.d2g[1] * ci;
    out->d2g[2] += -S2.d2g[2] * cd;
    out->d2g[3] += -S2.d2g[3] * ci;
    out->d2g[4] += S2.d2g[4] * ci;
    out->d2g[5] += S2.d2g[5] * ci;
  }
  return c1 + c2;
}

/* one row of the pure spectral sum (:535-579) */
static int fd_row(int mode, const QPD2 *qd, const double *r, double eps,
                  const double *g1, const double *g2, int n1, sum10 *S)
{
  double g2n = g2[0] * g2[0] + g2[1] * g2[1];
  int nc = (int)round((qd->vk[0] * g2[0] + qd->vk[1] * g2[1] - n1 * (g1[0] * g2[0] + g1[1] * g2[1])) / g2n);
  sum10 t;
  int cc = 1, n2, k;
  k = fd_term(mode, qd, r, g1, g2, n1, nc, &t);
  if (k < 0) return k;
  sum_add(S, &t, mode);
  for (n2 = nc + 1; n2 < nc + FD_LIMIT; n2++) {
    k = fd_term(mode, qd, r, g1, g2, n1, n2, &t);
    if (k < 0) return k;
    sum_add(S, &t, mode);
    cc++;
    if (k > 0 && term_small_fd(mode, &t, S, eps)) break;
  }
  if (n2 == nc + FD_LIMIT) return -1;
  for (n2 = nc - 1; n2 > nc - FD_LIMIT; n2--) {
    k = fd_term(mode, qd, r, g1, g2, n1, n2, &t);
    if (k < 0) return k;
    sum_add(S, &t, mode);
    cc++;
    if (k > 0 && term_small_fd(mode, &t, S, eps)) break;
  }
  if (n2 == nc - FD_LIMIT) return -1;
  return cc;
}

/* pure spectral evaluation in the unit cell (:484-533, :887-950) */
static int fd_cell(int mode, const QPD2 *qd, const double *r, double eps, sum10 *out)
{
  if (r[2] == 0.0) return -4;
  double det = qd->vd1[0] * qd->vd2[1] - qd->vd2[0] * qd->vd1[1];
  if (det == 0.0) return -3;
  double idd = 1.0 / det, g1[2], g2[2];
  reciprocal(qd, idd, g1, g2);
  int nc = (int)round((qd->vk[0] * qd->vd1[0] + qd->vk[1] * qd->vd1[1]) / (2.0 * M_PI));
  sum10 S = {0}, old;
  int cc = 0, c, n1;
  c = fd_row(mode, qd, r, eps, g1, g2, nc, &S);
  if (c < 0) return c;
  cc += c;
  for (n1 = nc + 1; n1 < nc + FD_LIMIT; n1++) {
    old = S;
    c = fd_row(mode, qd, r, eps, g1, g2, n1, &S);
    if (c < 0) return c;
    cc += c;
    if (rows_converged_fd(mode, &S, &old, eps)) break;
  }
  if (n1 == nc + FD_LIMIT) return -1;
  for (n1 = nc - 1; n1 > nc - FD_LIMIT; n1--) {
    old = S;
    c = fd_row(mode, qd, r, eps, g1, g2, n1, &S);
    if (c < 0) return c;
    cc += c;
    if (rows_converged_fd(mode, &S, &old, eps)) break;
  }
  if (n1 == nc - FD_LIMIT) return -1;
  double a = fabs(idd), sz = r[2] / fabs(r[2]);
  out->g = S.g * (0.5 * I * a);
  out->dg[0] = S.dg[0] * (-0.5 * a);
  out->dg[1] = S.dg[1] * (-0.5 * a);
  out->dg[2] = S.dg[2] * (-0.5 * a * sz);
  if (mode == M_V2) {
    double complex c5 = -0.5 * I * a;
    for (int i = 0; i < 4; i++) out->d2g[i] = S.d2g[i] * c5;
    out->d2g[4] = S.d2g[4] * c5 * sz;
    out->d2g[5] = S.d2g[5] * c5 * sz;
  }
  return cc;
}

int orc_qpgf(int mode, const QPD2 *qd, const double *r, double eps,
             double complex *G, double complex *dG, double complex *d2G)
{
  /* method choice (:10-65 and its three siblings) */
  double k2 = qd->k * qd->k;
  double det = qd->vd1[0] * qd->vd2[1] - qd->vd2[0] * qd->vd1[1];
  double idd = 1.0 / det, cpd = 2.0 * M_PI * idd;
  double ad12 = qd->vd1[0] * qd->vd1[0] + qd->vd1[1] * qd->vd1[1];
  double ad22 = qd->vd2[0] * qd->vd2[0] + qd->vd2[1] * qd->vd2[1];
  int n1 = (int)round((qd->vk[0] * qd->vd1[0] + qd->vk[1] * qd->vd1[1]) / (2.0 * M_PI));
  int n2 = (int)round((qd->vk[0] * qd->vd2[0] + qd->vk[1] * qd->vd2[1]) / (2.0 * M_PI));
  double aNx = -qd->vk[0] + n1 * cpd * qd->vd2[1] - n2 * cpd * qd->vd1[1];
  double aNy = -qd->vk[1] - n1 * cpd * qd->vd2[0] + n2 * cpd * qd->vd1[0];
  double bN = sqrt(k2 - aNx * aNx - aNy * aNy);
  int p1 = n1, m1 = n1, p2 = n2, m2 = n2;
  if (ad22 < ad12) { p1 = n1 + S_LIMIT; m1 = n1 - S_LIMIT; }
  else { p2 = n2 + S_LIMIT; m2 = n2 - S_LIMIT; }
  double apx = -qd->vk[0] + p1 * cpd * qd->vd2[1] - p2 * cpd * qd->vd1[1];
  double apy = -qd->vk[1] - p1 * cpd * qd->vd2[0] + p2 * cpd * qd->vd1[0];
  double amx = -qd->vk[0] + m1 * cpd * qd->vd2[1] - m2 * cpd * qd->vd1[1];
  double amy = -qd->vk[1] - m1 * cpd * qd->vd2[0] + m2 * cpd * qd->vd1[0];
  double bNp = k2 - apx * apx - apy * apy;
  double bNm = k2 - amx * amx - amy * amy;
  int use_fd = 0;
  if (!(bNp > 0.0 || bNm > 0.0)) {
    bNp = sqrt(-bNp);
    bNm = sqrt(-bNm);
    if (bN / bNp * exp(-bNp * fabs(r[2])) < eps && bN / bNm * exp(-bNm * fabs(r[2])) < eps) use_fd = 1;
  }

  /* unit-cell reduction and Bloch phase (:232-468) */
  int l1 = (int)floor((r[0] * qd->vd2[1] - r[1] * qd->vd2[0]) * idd + 0.5);
  int l2 = (int)floor(-(r[0] * qd->vd1[1] - r[1] * qd->vd1[0]) * idd + 0.5);
  double rt[3] = { r[0] - l1 * qd->vd1[0] - l2 * qd->vd2[0], r[1] - l1 * qd->vd1[1] - l2 * qd->vd2[1], r[2] };
  sum10 o = {0};
  int err = use_fd ? fd_cell(mode, qd, rt, eps, &o) : ewald_cell(mode, qd, rt, eps, &o);
  if (l1 != 0 || l2 != 0) {
    double arg = -(qd->vk[0] * (l1 * qd->vd1[0] + l2 * qd->vd2[0]) + qd->vk[1] * (l1 * qd->vd1[1] + l2 * qd->vd2[1]));
    double complex cf = cis(arg);
    o.g *= cf;
    for (int i = 0; i < 3; i++) o.dg[i] *= cf;
    if (mode == M_V2) for (int i = 0; i < 6; i++) o.d2g[i] *= cf;
  }
  if (G) *G = o.g;
  if (dG) for (int i = 0; i < 3; i++) dG[i] = o.dg[i];
  if (d2G && mode == M_V2) for (int i = 0; i < 6; i++) d2G[i] = o.d2g[i];
  if (err < 0) err -= use_fd ? 10 : 20;
  return err;
}

void orc_qpgf_batch(int mode, const QPD2 *qd, int n, const double *r, double eps,
                    double complex *G, double complex *dG, double complex *d2G, int *ret)
{
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < n; i++) {
    double complex g, dg[3], d2g[6];
    int e = orc_qpgf(mode, qd, r + 3 * i, eps, &g, dg, d2g);
    if (G) G[i] = g;
    if (dG) for (int j = 0; j < 3; j++) dG[3 * i + j] = dg[j];
    if (d2G) for (int j = 0; j < 6; j++) d2G[6 * i + j] = d2g[j];
    if (ret) ret[i] = e;
  }
}
