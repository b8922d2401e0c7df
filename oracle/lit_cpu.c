/*
 * lit_cpu.c -- oracle element-coefficient quadrature for linear triangles.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Element conventions of the reference (d3qd2_elem.c:1035-1163): the reference
 * triangle is the equilateral one with corners (-1/2,-sqrt3/2), (1,0),
 * (-1/2,sqrt3/2); r(z,e) = cr0 + cr1 z + cr2 e, constant normal W = cw0;
 * interpolation functions N_n on the nodes at 2/5 of the way to the corners.
 *
 * Restated reference routines:
 *   periodic pair rule   q0_lit_dcoef_v2_cc_GL/GH   (q0_lit_dcoef_v2_cc.c:156-321, point :345-381)
 *   free-space pair      lit_dcoef_v2_GL/GH         (lit_dcoef_v2.c:185-357, point :381-413)
 *   periodic self        q0_lit_coef_bd             (q0_lit_coef.c:297-372, :522-602); dC = 0
 *                        (q0_lit_dcoef_v2_cc.c:331-343)
 *   free-space self      lit_coef_bd                (lit_coef.c:290-350, :543-578)
 *                        + principal-value dG       (lit_dcoef.c:461-637, :845-952)
 *   rule switch          GL dF/dv > CBD_SDF -> GH  (d3qd2_elem.c:552-555, :866-869)
 * The pair rules split the triangle into three quadrilaterals around its
 * centroid, each the image of [-1,1]^2 with Jacobian factor 1 + a/4 + b/4.  The
 * GL dF/dv that decides the switch follows the reference's operation order
 * (three parts summed per j, times the weight, times w_j, summed, times w_i,
 * then the constant) so that the choice is the same pair by pair.
 *
 * Pins: tests/test_oracle_lit.py (solid angle of closed triangulated surfaces,
 * brute-force quadrature of the pair integrals on subdivided triangles,
 * the k -> 0 potential of a flat triangle in closed form, symmetry of the three
 * interpolation functions, and agreement of the triangle and quadrilateral
 * discretisations of one sphere).
 */
#include <math.h>
#include <complex.h>
#include <string.h>
#include "oracle.h"
#include "oracle_elem.h"

#define I_FP 7.957747154594766788444e-2    /* 1/(4 pi) */
#define SQ3 1.732050807568877293528e+0     /* sqrt(3) */
#define I_SPSQ3 3.062938307898844719507e-2 /* 1/(6 pi sqrt(3)) */
#define P34_N0 0.2

static inline double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline double abs3(const double *a) { return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }

/* d3qd2_elem.c:1047-1058 */
double orc_lit_N(int n, double z, double e)
{
  double ia = 0.5 / P34_N0;
  switch (n) {
  case 0: return 1.0 / 3.0 * (1.0 - ia * (z + sqrt(3.0) * e));
  case 1: return 1.0 / 3.0 * (1.0 + 2.0 * ia * z);
  default: return 1.0 / 3.0 * (1.0 - ia * (z - sqrt(3.0) * e));
  }
}

/* d/dr of N_n along (cos, sin) (lit_dcoef.c:845-856) */
static double lit_dN(int n, double c, double s)
{
  double ia = 0.5 / P34_N0;
  switch (n) {
  case 0: return -2.0 / 3.0 * ia * (c + SQ3 * s);
  case 1: return 4.0 / 3.0 * ia * c;
  default: return -2.0 / 3.0 * ia * (c - SQ3 * s);
  }
}

/* the three quadrilateral parts of the triangle as images of (a, b) */
void orc_lit_part(int part, double a, double b, double *z, double *e)
{
  if (part == 0) {
    *z = 0.125 * (3.0 + 2.0 * a + 2.0 * b + a * b);
    *e = 0.125 * SQ3 * (-a + b);
  } else if (part == 1) {
    *z = 0.0625 * (-3.0 + a - 5.0 * b - a * b);
    *e = 0.0625 * SQ3 * (3.0 + 3.0 * a + b + a * b);
  } else {
    *z = 0.0625 * (-3.0 - 5.0 * a + b - a * b);
    *e = 0.0625 * SQ3 * (-3.0 - a - 3.0 * b - a * b);
  }
}

typedef struct {
  double cr[3][3], V[3], aV;
} litc;

static void lit_load(const BOUD *bd, int s, litc *E)
{
  for (int l = 0; l < 3; l++) {
    for (int m = 0; m < 3; m++) E->cr[l][m] = bd->cr[s][l][m];
    E->V[l] = bd->cw[s][l][0];
  }
  E->aV = abs3(E->V);
}

static inline void lit_r(const litc *E, double z, double e, double *r)
{
  for (int i = 0; i < 3; i++) r[i] = E->cr[i][0] + E->cr[i][1] * z + E->cr[i][2] * e;
}

/* dF/dv integrand of one point: i_aR^3 (3 (R.v)(R.V)/R^2 - V.v) */
static inline void lit_dFpt(const litc *E, double z, double e, const double *rt, const double *v0,
                            const double *v1, double *d0, double *d1, double *f, double *vR, double *aRo)
{
  lit_r(E, z, e, vR);
  for (int l = 0; l < 3; l++) vR[l] -= rt[l];
  double aR = abs3(vR), iR = 1.0 / aR;
  double rdV = dot3(vR, E->V) * iR;
  double rdv0 = dot3(vR, v0) * iR, rdv1 = dot3(vR, v1) * iR;
  double Vdv0 = dot3(E->V, v0), Vdv1 = dot3(E->V, v1);
  *d0 = iR * iR * iR * (3.0 * rdv0 * rdV - Vdv0);
  *d1 = iR * iR * iR * (3.0 * rdv1 * rdV - Vdv1);
  *f = (vR[0] * E->V[0] + vR[1] * E->V[1] + vR[2] * E->V[2]) * iR * iR * iR;
  *aRo = aR;
}

/* GL12 dF/dv0, dF/dv1 of a triangle pair in the reference's operation order */
void orc_lit_dF_GL(const BOUD *bd, const double *rt, const double *v0, const double *v1, int s, double *dF0,
                   double *dF1)
{
  litc E;
  lit_load(bd, abs(s), &E);
  double a0 = 0.0, a1 = 0.0;
  for (int i = 0; i < GLN; i++) {
    double t0 = 0.0, t1 = 0.0;
    double a = bd->xli[i];
    for (int j = 0; j < GLN; j++) {
      double b = bd->xli[j];
      double wt = 1.0 + 0.25 * a + 0.25 * b;
      double p0 = 0.0, p1 = 0.0;
      for (int part = 0; part < 3; part++) {
        double z, e, d0, d1, f, vR[3], aR;
        orc_lit_part(part, a, b, &z, &e);
        lit_dFpt(&E, z, e, rt, v0, v1, &d0, &d1, &f, vR, &aR);
        p0 += d0;
        p1 += d1;
      }
      t0 += p0 * wt * bd->wli[j];
      t1 += p1 * wt * bd->wli[j];
    }
    a0 += t0 * bd->wli[i];
    a1 += t1 * bd->wli[i];
  }
  *dF0 = a0 * (0.0625 * SQ3 * I_SPSQ3);
  *dF1 = a1 * (0.0625 * SQ3 * I_SPSQ3);
}

/* periodic pair on an n x n rule per part; dC0/dC1 carry only dF ([8]) */
long orc_q0_lit_pair(const BOUD *bd, const double *rt, const double *v0, const double *v1, int s, int n,
                     const double *x, const double *w, double complex *CC, double complex *dC0,
                     double complex *dC1)
{
  litc E;
  lit_load(bd, s, &E);
  for (int l = 0; l < 9; l++) CC[l] = dC0[l] = dC1[l] = 0.0;
  long terms = 0;
  double F = 0.0, F0 = 0.0, F1 = 0.0;
  for (int i = 0; i < n; i++) {
    double complex tg[3] = {0}, th[3] = {0};
    double tf = 0.0, tf0 = 0.0, tf1 = 0.0;
    double a = x[i];
    for (int j = 0; j < n; j++) {
      double b = x[j], wt = (1.0 + 0.25 * a + 0.25 * b) * w[j];
      for (int part = 0; part < 3; part++) {
        double z, e, d0, d1, f, vR[3], aR;
        orc_lit_part(part, a, b, &z, &e);
        lit_dFpt(&E, z, e, rt, v0, v1, &d0, &d1, &f, vR, &aR);
        double complex qG, dqG[3];
        int err = orc_qpgf(2, &bd->qd, vR, MEPS, &qG, dqG, NULL);
        if (err < 0) return err;
        terms += err;
        double complex hn = dqG[0] * E.V[0] + dqG[1] * E.V[1] + dqG[2] * E.V[2];
        for (int ep = 0; ep < 3; ep++) {
          double Np = orc_lit_N(ep, z, e);
          tg[ep] += qG * Np * wt;
          th[ep] += hn * Np * wt;
        }
        tf += f * wt;
        tf0 += d0 * wt;
        tf1 += d1 * wt;
      }
    }
    for (int ep = 0; ep < 3; ep++) {
      CC[ep] += tg[ep] * w[i];
      CC[ep + 4] += th[ep] * w[i];
    }
    F += tf * w[i];
    F0 += tf0 * w[i];
    F1 += tf1 * w[i];
  }
  for (int ep = 0; ep < 3; ep++) {
    CC[ep] *= 1.0 / 24.0 * E.aV;
    CC[ep + 4] *= 1.0 / 24.0;
  }
  const double cF = 0.0625 * SQ3 * I_SPSQ3;
  CC[8] = F * cF;
  dC0[8] = F0 * cF;
  dC1[8] = F1 * cF;
  return terms;
}

/* free-space pair with tangential derivatives; s < 0 reverses the normal:
 * H, dH, F, dF change sign (lit_convert_CC / _dCC, d3qd2_elem.c:1075-1093) */
void orc_lit_pair(const BOUD *bd, const double *rt, const double *v0, const double *v1, int s, double k, int n,
                  const double *x, const double *w, double complex *CC, double complex *dC0, double complex *dC1)
{
  int sig = s < 0;
  litc E;
  lit_load(bd, abs(s), &E);
  for (int l = 0; l < 9; l++) CC[l] = dC0[l] = dC1[l] = 0.0;
  double F = 0.0, F0 = 0.0, F1 = 0.0;
  for (int i = 0; i < n; i++) {
    double complex tg[3] = {0}, th[3] = {0}, tdg0[3] = {0}, tdg1[3] = {0}, tdh0[3] = {0}, tdh1[3] = {0};
    double tf = 0.0, tf0 = 0.0, tf1 = 0.0;
    double a = x[i];
    for (int j = 0; j < n; j++) {
      double b = x[j], wt = (1.0 + 0.25 * a + 0.25 * b) * w[j];
      for (int part = 0; part < 3; part++) {
        double z, e, d0, d1, f, vR[3], aR;
        orc_lit_part(part, a, b, &z, &e);
        lit_dFpt(&E, z, e, rt, v0, v1, &d0, &d1, &f, vR, &aR);
        double iR = 1.0 / aR;
        double rdW = dot3(vR, E.V) * iR, rdv0 = dot3(vR, v0) * iR, rdv1 = dot3(vR, v1) * iR;
        double Wdv0 = dot3(E.V, v0), Wdv1 = dot3(E.V, v1);
        double complex kR = k * aR, ca = I * kR, cb = ca - 1.0;
        double complex ce = cos(k * aR) + I * sin(k * aR);
        for (int ep = 0; ep < 3; ep++) {
          double Np = orc_lit_N(ep, z, e);
          tg[ep] += ce * iR * Np * wt;
          th[ep] += cb * ce * rdW * iR * iR * Np * wt;
          tdg0[ep] += cb * ce * iR * iR * rdv0 * Np * wt;
          tdg1[ep] += cb * ce * iR * iR * rdv1 * Np * wt;
          tdh0[ep] += ce * iR * iR * iR * ((kR * kR + 3.0 * cb) * rdv0 * rdW - cb * Wdv0) * Np * wt;
          tdh1[ep] += ce * iR * iR * iR * ((kR * kR + 3.0 * cb) * rdv1 * rdW - cb * Wdv1) * Np * wt;
        }
        tf += f * wt;
        tf0 += d0 * wt;
        tf1 += d1 * wt;
      }
    }
    for (int ep = 0; ep < 3; ep++) {
      CC[ep] += tg[ep] * w[i];
      CC[ep + 4] += th[ep] * w[i];
      dC0[ep] += tdg0[ep] * w[i];
      dC1[ep] += tdg1[ep] * w[i];
      dC0[ep + 4] += tdh0[ep] * w[i];
      dC1[ep + 4] += tdh1[ep] * w[i];
    }
    F += tf * w[i];
    F0 += tf0 * w[i];
    F1 += tf1 * w[i];
  }
  const double c = 0.0625 * SQ3 * I_SPSQ3;
  for (int ep = 0; ep < 3; ep++) {
    CC[ep] *= c * E.aV;
    CC[ep + 4] *= c;
    dC0[ep] *= -c * E.aV;
    dC1[ep] *= -c * E.aV;
    dC0[ep + 4] *= c;
    dC1[ep + 4] *= c;
  }
  CC[8] = F * c;
  dC0[8] = F0 * c;
  dC1[8] = F1 * c;
  if (sig) {
    for (int l = 4; l < 9; l++) {
      if (l == 7) continue;
      CC[l] = -CC[l];
      dC0[l] = -dC0[l];
      dC1[l] = -dC1[l];
    }
  }
}

/* ------------------------------------------------------------ polar self rules */
/* corner distances and angles seen from (zt, et) (d3qd2_elem.c:1149-1163) */
void orc_lit_corners(double zt, double et, double *r, double *th)
{
  double ztp = 1.0 - zt, ztm = -0.5 - zt, etp = 0.5 * SQ3 - et, etm = -0.5 * SQ3 - et;
  r[0] = sqrt(ztm * ztm + etm * etm);
  r[1] = sqrt(ztp * ztp + et * et);
  r[2] = sqrt(ztm * ztm + etp * etp);
  r[3] = r[0];
  th[0] = atan2(etm, ztm);
  th[1] = atan2(-et, ztp);
  th[2] = atan2(etp, ztm);
  th[3] = th[0] + 2.0 * M_PI;
}

static inline double theta_node(const double *th, int j, int half, double xm)
{
  return half == 0 ? 0.25 * (th[j + 1] - th[j]) * xm + 0.25 * (th[j + 1] + 3.0 * th[j])
                   : 0.25 * (th[j + 1] - th[j]) * xm + 0.25 * (3.0 * th[j + 1] + th[j]);
}

static inline double polar_rm(const double *r, const double *th, int j, double t)
{
  return r[j] * r[j + 1] * sin(th[j + 1] - th[j]) / (r[j] * sin(t - th[j]) - r[j + 1] * sin(t - th[j + 1]));
}

/* periodic self element: G with the value chain, H with the gradient chain */
long orc_q0_lit_self(const BOUD *bd, int s, double zt, double et, double complex *CC)
{
  litc E;
  lit_load(bd, s, &E);
  double r[4], th[4];
  orc_lit_corners(zt, et, r, th);
  for (int l = 0; l < 9; l++) CC[l] = 0.0;
  long terms = 0;
  for (int j = 0; j < 3; j++) {
    double complex tG[3] = {0}, tH[3] = {0};
    for (int m = 0; m < GHN; m++) {
      double complex sG[3] = {0}, sH[3] = {0};
      for (int h = 0; h < 2; h++) {
        double t = theta_node(th, j, h, bd->xhi[m]);
        double rm = polar_rm(r, th, j, t);
        double c = cos(t), sn = sin(t);
        double complex aG[3] = {0}, aH[3] = {0};
        for (int i = 0; i < GLN; i++) {
          double rr = 0.5 * rm * (bd->xli[i] + 1.0);
          double z = rr * c + zt, e = rr * sn + et, vR[3];
          for (int l = 0; l < 3; l++) vR[l] = rr * (E.cr[l][1] * c + E.cr[l][2] * sn);
          double complex qG, dqG[3];
          int err = orc_qpgf(0, &bd->qd, vR, MEPS, &qG, NULL, NULL);
          if (err < 0) return err;
          terms += err;
          err = orc_qpgf(1, &bd->qd, vR, MEPS, NULL, dqG, NULL);
          if (err < 0) return err;
          terms += err;
          double complex hn = dqG[0] * E.V[0] + dqG[1] * E.V[1] + dqG[2] * E.V[2];
          for (int n = 0; n < 3; n++) {
            double Nn = orc_lit_N(n, z, e);
            aG[n] += qG * Nn * rr * bd->wli[i];
            aH[n] += hn * Nn * rr * bd->wli[i];
          }
        }
        for (int n = 0; n < 3; n++) {
          sG[n] += aG[n] * 0.5 * rm;
          sH[n] += aH[n] * 0.5 * rm;
        }
      }
      for (int n = 0; n < 3; n++) {
        tG[n] += sG[n] * bd->whi[m];
        tH[n] += sH[n] * bd->whi[m];
      }
    }
    for (int n = 0; n < 3; n++) {
      CC[n] += tG[n] * 0.25 * (th[j + 1] - th[j]);
      CC[n + 4] += tH[n] * 0.25 * (th[j + 1] - th[j]);
    }
  }
  for (int n = 0; n < 3; n++) {
    CC[n] *= 2.0 / (3.0 * SQ3) * E.aV;
    CC[n + 4] *= 2.0 / (3.0 * SQ3);
  }
  return terms;
}

/* free-space self element: G only (H and F vanish on a flat element) */
void orc_lit_self(const BOUD *bd, int s_signed, double zt, double et, double k, double complex *CC)
{
  litc E;
  lit_load(bd, abs(s_signed), &E);
  double r[4], th[4];
  orc_lit_corners(zt, et, r, th);
  for (int l = 0; l < 9; l++) CC[l] = 0.0;
  for (int n = 0; n < 3; n++) {
    for (int j = 0; j < 3; j++) {
      double complex tC = 0.0;
      for (int m = 0; m < GHN; m++) {
        double complex s2 = 0.0;
        for (int h = 0; h < 2; h++) {
          double t = theta_node(th, j, h, bd->xhi[m]);
          double rm = polar_rm(r, th, j, t);
          double c = cos(t), sn = sin(t), K[3];
          for (int l = 0; l < 3; l++) K[l] = E.cr[l][1] * c + E.cr[l][2] * sn;
          double aK = abs3(K);
          double complex acc = 0.0;
          for (int i = 0; i < GLN; i++) {
            double rr = 0.5 * rm * (bd->xli[i] + 1.0);
            double complex ce = cos(k * aK * rr) + I * sin(k * aK * rr);
            acc += ce / aK * orc_lit_N(n, rr * c + zt, rr * sn + et) * bd->wli[i];
          }
          s2 += acc * 0.5 * rm;
        }
        tC += s2 * bd->whi[m];
      }
      CC[n] += tC * 0.25 * (th[j + 1] - th[j]);
    }
    CC[n] *= I_SPSQ3 * E.aV;
  }
}

/* principal-value tangential derivative of the self element's G along the
 * element parameter dir (0: zeta, 1: eta) (lit_dcoef_bd_tz0_PV / _te0_PV) */
void orc_lit_self_dPV(const BOUD *bd, int s_signed, double zt, double et, double k, int dir, double complex *dCC)
{
  litc E;
  lit_load(bd, abs(s_signed), &E);
  double r[4], th[4];
  orc_lit_corners(zt, et, r, th);
  double vt[3];
  for (int l = 0; l < 3; l++) vt[l] = E.cr[l][dir == 0 ? 1 : 2];
  double i_aT = 1.0 / abs3(vt);
  for (int l = 0; l < 9; l++) dCC[l] = 0.0;
  for (int n = 0; n < 3; n++) {
    double complex acc = 0.0;
    /* D_bar(eps): sub-triangles, r in [DEPS, rm], GHN in r */
    for (int j = 0; j < 3; j++) {
      double complex tC = 0.0;
      for (int m = 0; m < GHN; m++) {
        double complex s2 = 0.0;
        for (int h = 0; h < 2; h++) {
          double t = theta_node(th, j, h, bd->xhi[m]);
          double rm = polar_rm(r, th, j, t);
          double c = cos(t), sn = sin(t), K[3];
          for (int l = 0; l < 3; l++) K[l] = E.cr[l][1] * c + E.cr[l][2] * sn;
          double aK = abs3(K), iK = 1.0 / aK, Kdv = dot3(K, vt) * iK;
          double complex sum = 0.0;
          for (int i = 0; i < GHN; i++) {
            double rr = 0.5 * (rm - DEPS) * bd->xhi[i] + 0.5 * (rm + DEPS);
            double complex kR = k * aK * rr, ca = I * kR, cb = ca - 1.0;
            double complex ce = cos(k * aK * rr) + I * sin(k * aK * rr);
            sum += cb * ce / rr * iK * iK * Kdv * orc_lit_N(n, rr * c + zt, rr * sn + et) * bd->whi[i];
          }
          s2 += sum * 0.5 * (rm - DEPS);
        }
        tC += s2 * bd->whi[m];
      }
      acc += tC * 0.25 * (th[j + 1] - th[j]);
    }
    /* D_eps: four quarters of the circle */
    {
      double complex tC = 0.0;
      for (int m = 0; m < GHN; m++) {
        double complex s4 = 0.0;
        const double off[4] = {-0.75 * M_PI, -0.25 * M_PI, 0.25 * M_PI, 0.75 * M_PI};
        for (int q = 0; q < 4; q++) {
          double t = 0.25 * M_PI * bd->xhi[m] + off[q];
          double c = cos(t), sn = sin(t), K[3];
          for (int l = 0; l < 3; l++) K[l] = E.cr[l][1] * c + E.cr[l][2] * sn;
          double aK = abs3(K), iK = 1.0 / aK, Kdv = dot3(K, vt) * iK;
          double complex keK = k * DEPS * aK;
          double complex cc = cos(k * DEPS * aK), cs = sin(k * DEPS * aK);
          s4 += lit_dN(n, c, sn) * iK * iK * iK * Kdv * ((keK + I) * cs + (2.0 - I * keK) * cc - 2.0);
        }
        tC += s4 * bd->whi[m];
      }
      tC *= 0.25 * M_PI;
      acc += I * tC / k;
    }
    /* D_eps principal value: two quarters */
    {
      double complex tC = 0.0;
      for (int m = 0; m < GHN; m++) {
        double complex s2 = 0.0;
        const double off[2] = {0.25 * M_PI, 0.75 * M_PI};
        for (int q = 0; q < 2; q++) {
          double t = 0.25 * M_PI * bd->xhi[m] + off[q];
          double c = cos(t), sn = sin(t), K[3];
          for (int l = 0; l < 3; l++) K[l] = E.cr[l][1] * c + E.cr[l][2] * sn;
          double aK = abs3(K), iK = 1.0 / aK, Kdv = dot3(K, vt) * iK;
          double cs = sin(k * aK * DEPS);
          s2 += lit_dN(n, c, sn) * cs * iK * iK * iK * Kdv;
        }
        tC += s2 * bd->whi[m];
      }
      tC *= 0.25 * M_PI;
      acc -= tC / k;
    }
    dCC[n] = acc * (-I_SPSQ3 * E.aV * i_aT);
  }
  /* s < 0: dG keeps its sign (only dH and dF change, and they are zero here) */
}
