/*
 * elem_cpu.c -- oracle element-coefficient quadrature for bilinear elements.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Output convention of the reference (d3qd2_elem.h:113): CC[0..3] = G, CC[4..7]
 * = H, CC[8] = F; dC*[0..8] their derivatives along the node tangents.
 *
 * Restated reference routines:
 *   periodic pair rule   q0_bil_dcoef_v2_cc_GL/GH   (q0_bil_dcoef_v2_cc.c:129-281)
 *   periodic self        q0_bil_coef_bd            (q0_bil_coef.c:246-338, :408-497)
 *                        + central-difference dF   (q0_bil_dcoef_v2_cc.c:336-454)
 *   free-space pair      bil_dcoef_v2_GL/GH        (bil_dcoef_v2.c:141-325)
 *   free-space self      bil_coef_bd               (bil_coef.c:244-342, :449-536)
 *                        + principal-value dG,dH,dF (bil_dcoef.c:369-740, :916-1630)
 *   rule switch          dcoef_bd_node_t2 / q0_dcoef_bd_node_t2_cc (d3qd2_elem.c:515-572, :830-886)
 * The per-point arithmetic follows the reference operation by operation (no FMA
 * contraction on the host compiler) so that the GL dF/dv that decides the
 * GL12 -> GH64 switch is bit-identical; the Green's function is evaluated once
 * per quadrature point instead of once per shape function, which changes no
 * sum because each shape function keeps its own accumulator in the same order.
 *
 * Pins: tests/test_oracle_sample.py checks rows of the assembled system against
 * the reference's own solution in ex.dat (residual of A x = b), the particle
 * velocity recovery against ex.dat's pv (pins dG, dH, dF including the
 * principal-value self terms) and the dpv system residual; tests/test_elem_math.py
 * checks the static F row sums against the solid angle of a closed surface
 * (Gauss) and its vanishing tangential derivatives.
 */
#include <math.h>
#include <complex.h>
#include <string.h>
#include "oracle.h"
#include "oracle_elem.h"

#define I_FP 7.957747154594766788444e-2   /* 1/(4 pi) */
#define P44 0.5773502691896257645091      /* 1/sqrt(3) */

static inline double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline double abs3(const double *a) { return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }

/* interpolation functions on the collocation nodes (d3qd2_elem.c:903-915) */
double orc_bil_N(int n, double z, double e)
{
  double ia = 1.0 / P44;
  switch (n) {
  case 0: return 0.25 * (1.0 - z * ia) * (1.0 - e * ia);
  case 1: return 0.25 * (1.0 + z * ia) * (1.0 - e * ia);
  case 2: return 0.25 * (1.0 + z * ia) * (1.0 + e * ia);
  default: return 0.25 * (1.0 - z * ia) * (1.0 + e * ia);
  }
}

void orc_elem_copy(const BOUD *bd, int s, double cr[3][4], double cw[3][3])
{
  for (int l = 0; l < 3; l++) {
    for (int m = 0; m < 4; m++) cr[l][m] = bd->cr[s][l][m];
    for (int m = 0; m < 3; m++) cw[l][m] = bd->cw[s][l][m];
  }
}

static inline void bil_rw(double cr[3][4], double cw[3][3], double z, double e, double *r, double *w)
{
  for (int i = 0; i < 3; i++) {
    r[i] = cr[i][0] + cr[i][1] * z + cr[i][2] * e + cr[i][3] * z * e;
    w[i] = cw[i][0] + cw[i][1] * z + cw[i][2] * e;
  }
}

/* unit tangents at a collocation node (d3qd2_elem.c:122-143) */
void orc_node_tangents(const BOUD *bd, int s, int nid, double *vtz, double *vte)
{
  s = abs(s);
  if (bd->ed[s][3] != 0) {
    for (int i = 0; i < 3; i++) {
      vtz[i] = bd->cr[s][i][1] + bd->cr[s][i][3] * bd->et_44[nid];
      vte[i] = bd->cr[s][i][2] + bd->cr[s][i][3] * bd->zt_44[nid];
    }
  } else {
    for (int i = 0; i < 3; i++) {
      vtz[i] = bd->cr[s][i][1];
      vte[i] = bd->cr[s][i][2];
    }
  }
  orc_unit(vtz);
  orc_unit(vte);
}

/* F-side geometry of one point, shared by both kernels */
typedef struct {
  double vR[3], vW[3], aR, iR, aW, rdW, rdv0, rdv1, Wdv0, Wdv1;
} ptgeo;

static inline void point_geo(double cr[3][4], double cw[3][3], double z, double e, const double *rt,
                             const double *v0, const double *v1, ptgeo *g)
{
  bil_rw(cr, cw, z, e, g->vR, g->vW);
  for (int l = 0; l < 3; l++) g->vR[l] -= rt[l];
  g->aR = abs3(g->vR);
  g->iR = 1.0 / g->aR;
  g->aW = abs3(g->vW);
  g->rdW = dot3(g->vR, g->vW) * g->iR;
  g->rdv0 = dot3(g->vR, v0) * g->iR;
  g->rdv1 = dot3(g->vR, v1) * g->iR;
  g->Wdv0 = dot3(g->vW, v0);
  g->Wdv1 = dot3(g->vW, v1);
}

/*
 * GL12 estimate of dF/dv0, dF/dv1 for the rule switch, with the reference's
 * summation order (inner j sum times w_i, then 1/(4 pi)): the values that
 * d3qd2_elem.c:536-537 and :851-852 compare against CBD_SDF.
 */
void orc_bil_dF_GL(const BOUD *bd, const double *rt, const double *v0, const double *v1, int s, double *dF0,
                   double *dF1)
{
  double cr[3][4], cw[3][3];
  orc_elem_copy(bd, abs(s), cr, cw);
  double a0 = 0.0, a1 = 0.0;
  for (int i = 0; i < GLN; i++) {
    double t0 = 0.0, t1 = 0.0;
    for (int j = 0; j < GLN; j++) {
      ptgeo g;
      point_geo(cr, cw, bd->xli[i], bd->xli[j], rt, v0, v1, &g);
      t0 += bd->wli[j] * g.iR * g.iR * g.iR * (3.0 * g.rdv0 * g.rdW - g.Wdv0);
      t1 += bd->wli[j] * g.iR * g.iR * g.iR * (3.0 * g.rdv1 * g.rdW - g.Wdv1);
    }
    a0 += bd->wli[i] * t0;
    a1 += bd->wli[i] * t1;
  }
  *dF0 = a0 * I_FP;
  *dF1 = a1 * I_FP;
}

/* periodic pair, n x n Gauss-Legendre rule (q0_bil_dcoef_v2_cc.c:129-281).
 * dC0/dC1 carry only the F derivative ([8]); [0..7] stay zero as in the
 * reference's "cc" variant. Returns the summed GF term count or a GF error. */
long orc_q0_bil_pair(const BOUD *bd, const double *rt, const double *v0, const double *v1, int s, int n,
                     const double *x, const double *w, double complex *CC, double complex *dC0,
                     double complex *dC1)
{
  double cr[3][4], cw[3][3];
  orc_elem_copy(bd, s, cr, cw);
  for (int l = 0; l < 9; l++) CC[l] = dC0[l] = dC1[l] = 0.0;
  long terms = 0;
  double F = 0.0, F0 = 0.0, F1 = 0.0;
  for (int i = 0; i < n; i++) {
    double complex tg[4] = {0}, th[4] = {0};
    double tf = 0.0, tf0 = 0.0, tf1 = 0.0;
    for (int j = 0; j < n; j++) {
      double z = x[i], e = x[j];
      ptgeo g;
      point_geo(cr, cw, z, e, rt, v0, v1, &g);
      double complex qG, dqG[3];
      int err = orc_qpgf(2, &bd->qd, g.vR, MEPS, &qG, dqG, NULL);
      if (err < 0) return err;
      terms += err;
      double complex hn = dqG[0] * g.vW[0] + dqG[1] * g.vW[1] + dqG[2] * g.vW[2];
      for (int ep = 0; ep < 4; ep++) {
        double Np = orc_bil_N(ep, z, e);
        tg[ep] += w[j] * qG * Np * g.aW;
        th[ep] += w[j] * hn * Np;
      }
      tf += w[j] * g.rdW * g.iR * g.iR;
      tf0 += w[j] * g.iR * g.iR * g.iR * (3.0 * g.rdv0 * g.rdW - g.Wdv0);
      tf1 += w[j] * g.iR * g.iR * g.iR * (3.0 * g.rdv1 * g.rdW - g.Wdv1);
    }
    for (int ep = 0; ep < 4; ep++) {
      CC[ep] += w[i] * tg[ep];
      CC[ep + 4] += w[i] * th[ep];
    }
    F += w[i] * tf;
    F0 += w[i] * tf0;
    F1 += w[i] * tf1;
  }
  CC[8] = F * I_FP;
  dC0[8] = F0 * I_FP;
  dC1[8] = F1 * I_FP;
  return terms;
}

/* free-space pair, n x n rule (bil_dcoef_v2.c:141-325); s signed: for s < 0
 * H, dH, F, dF change sign (the sub side sees the reversed normal). */
void orc_bil_pair(const BOUD *bd, const double *rt, const double *v0, const double *v1, int s, double k,
                  int n, const double *x, const double *w, double complex *CC, double complex *dC0,
                  double complex *dC1)
{
  int sig = s < 0;
  double cr[3][4], cw[3][3];
  orc_elem_copy(bd, abs(s), cr, cw);
  for (int l = 0; l < 9; l++) CC[l] = dC0[l] = dC1[l] = 0.0;
  double F = 0.0, F0 = 0.0, F1 = 0.0;
  for (int i = 0; i < n; i++) {
    double complex tg[4] = {0}, th[4] = {0}, tdg0[4] = {0}, tdg1[4] = {0}, tdh0[4] = {0}, tdh1[4] = {0};
    double tf = 0.0, tf0 = 0.0, tf1 = 0.0;
    for (int j = 0; j < n; j++) {
      double z = x[i], e = x[j];
      ptgeo g;
      point_geo(cr, cw, z, e, rt, v0, v1, &g);
      double complex ca = I * k * g.aR, cb = ca - 1.0;
      double complex ce = cos(k * g.aR) + I * sin(k * g.aR);
      double tD = g.rdW * g.iR * g.iR;
      for (int ep = 0; ep < 4; ep++) {
        double Np = orc_bil_N(ep, z, e);
        tg[ep] += w[j] * ce * g.iR * g.aW * Np;
        th[ep] += w[j] * (ca - 1.0) * ce * tD * Np;
        tdg0[ep] += w[j] * cb * ce * g.iR * g.iR * g.rdv0 * g.aW * Np;
        tdg1[ep] += w[j] * cb * ce * g.iR * g.iR * g.rdv1 * g.aW * Np;
        tdh0[ep] += w[j] * ce * g.iR * g.iR * g.iR * ((k * k * g.aR * g.aR + 3.0 * cb) * g.rdv0 * g.rdW - cb * g.Wdv0) * Np;
        tdh1[ep] += w[j] * ce * g.iR * g.iR * g.iR * ((k * k * g.aR * g.aR + 3.0 * cb) * g.rdv1 * g.rdW - cb * g.Wdv1) * Np;
      }
      tf += w[j] * tD;
      tf0 += w[j] * g.iR * g.iR * g.iR * (3.0 * g.rdv0 * g.rdW - g.Wdv0);
      tf1 += w[j] * g.iR * g.iR * g.iR * (3.0 * g.rdv1 * g.rdW - g.Wdv1);
    }
    for (int ep = 0; ep < 4; ep++) {
      CC[ep] += w[i] * tg[ep];
      CC[ep + 4] += w[i] * th[ep];
      dC0[ep] += w[i] * tdg0[ep];
      dC1[ep] += w[i] * tdg1[ep];
      dC0[ep + 4] += w[i] * tdh0[ep];
      dC1[ep + 4] += w[i] * tdh1[ep];
    }
    F += w[i] * tf;
    F0 += w[i] * tf0;
    F1 += w[i] * tf1;
  }
  for (int ep = 0; ep < 4; ep++) {
    CC[ep] *= I_FP;
    CC[ep + 4] *= I_FP;
    dC0[ep] *= -I_FP;
    dC1[ep] *= -I_FP;
    dC0[ep + 4] *= I_FP;
    dC1[ep + 4] *= I_FP;
  }
  CC[8] = F * I_FP;
  dC0[8] = F0 * I_FP;
  dC1[8] = F1 * I_FP;
  if (sig) {
    for (int l = 4; l < 9; l++) {
      CC[l] = -CC[l];
      dC0[l] = -dC0[l];
      dC1[l] = -dC1[l];
    }
  }
}

/* ------------------------------------------------------------ polar self rules */
/* corner distances and angles seen from (zt, et) (d3qd2_elem.c:1017-1034) */
static void bil_corners(double zt, double et, double *r, double *th)
{
  double zp = 1.0 - zt, zm = -1.0 - zt, ep = 1.0 - et, em = -1.0 - et;
  r[0] = sqrt(zm * zm + em * em);
  r[1] = sqrt(zp * zp + em * em);
  r[2] = sqrt(zp * zp + ep * ep);
  r[3] = sqrt(zm * zm + ep * ep);
  r[4] = r[0];
  th[0] = atan2(em, zm);
  th[1] = atan2(em, zp);
  th[2] = atan2(ep, zp);
  th[3] = atan2(ep, zm);
  th[4] = th[0] + 2.0 * M_PI;
}

int orc_bil_is_plane(double cr[3][4], double cw[3][3])
{
  double bdc = cr[0][1] * cw[0][2] + cr[1][1] * cw[1][2] + cr[2][1] * cw[2][2];
  double ab = sqrt(cr[0][1] * cr[0][1] + cr[1][1] * cr[1][1] + cr[2][1] * cr[2][1]);
  double aC = sqrt(cw[0][2] * cw[0][2] + cw[1][2] * cw[1][2] + cw[2][2] * cw[2][2]);
  return fabs(bdc / (ab * aC)) < GEPS;
}

/* radial direction vector K(r) with r(zt + r cos, et + r sin) - r(zt, et) = r K */
static inline void polar_K(double cr[3][4], double c, double s, double r, double zt, double et, double *K)
{
  for (int i = 0; i < 3; i++) K[i] = cr[i][1] * c + cr[i][2] * s + cr[i][3] * (r * c * s + zt * s + et * c);
}

/* theta abscissae of the sub-triangle j rule: two GHN halves (q0_bil_coef.c:286-300) */
static inline double theta_node(const double *th, int j, int half, double xm)
{
  return half == 0 ? 0.25 * (th[j + 1] - th[j]) * xm + 0.25 * (th[j + 1] + 3.0 * th[j])
                   : 0.25 * (th[j + 1] - th[j]) * xm + 0.25 * (3.0 * th[j + 1] + th[j]);
}

static inline double polar_rm(const double *r, const double *th, int j, double t)
{
  return r[j] * r[j + 1] * sin(th[j + 1] - th[j]) / (r[j] * sin(t - th[j]) - r[j + 1] * sin(t - th[j + 1]));
}

/* static F of the self element (bil_coef.c:525-536, q0_bil_coef.c:315-337) */
static double bil_self_F(const BOUD *bd, double cr[3][4], double cw[3][3], double zt, double et)
{
  if (orc_bil_is_plane(cr, cw)) return 0.0;
  double r[5], th[5];
  bil_corners(zt, et, r, th);
  double bdC = cr[0][1] * cw[0][2] + cr[1][1] * cw[1][2] + cr[2][1] * cw[2][2];
  double F = 0.0;
  for (int j = 0; j < 4; j++) {
    double tD = 0.0;
    for (int m = 0; m < GHN; m++) {
      double sum2 = 0.0;
      for (int h = 0; h < 2; h++) {
        double t = theta_node(th, j, h, bd->xhi[m]);
        double rm = polar_rm(r, th, j, t);
        double c = cos(t), s = sin(t), acc = 0.0;
        for (int i = 0; i < GLN; i++) {
          double rr = 0.5 * rm * (bd->xli[i] + 1.0), K[3];
          polar_K(cr, c, s, rr, zt, et, K);
          double aK = abs3(K);
          acc += 1.0 / (aK * aK * aK) * c * s * bd->wli[i];
        }
        sum2 += acc * 0.5 * rm;
      }
      tD += sum2 * bd->whi[m];
    }
    F += tD * 0.25 * (th[j + 1] - th[j]);
  }
  return F * I_FP * bdC;
}

/* periodic self element (q0_bil_coef_bd, q0_bil_coef.c:246-338 with the
 * integrands of :408-497), node at (zt, et) of element s > 0 */
long orc_q0_bil_self(const BOUD *bd, int s, double zt, double et, double complex *CC)
{
  double cr[3][4], cw[3][3];
  orc_elem_copy(bd, s, cr, cw);
  double r[5], th[5];
  bil_corners(zt, et, r, th);
  for (int l = 0; l < 9; l++) CC[l] = 0.0;
  long terms = 0;
  for (int j = 0; j < 4; j++) {
    double complex tG[4] = {0}, tH[4] = {0};
    for (int m = 0; m < GHN; m++) {
      double complex sG[4] = {0}, sH[4] = {0};
      for (int h = 0; h < 2; h++) {
        double t = theta_node(th, j, h, bd->xhi[m]);
        double rm = polar_rm(r, th, j, t);
        double c = cos(t), sn = sin(t);
        double complex aG[4] = {0}, aH[4] = {0};
        for (int i = 0; i < GLN; i++) {
          double rr = 0.5 * rm * (bd->xli[i] + 1.0);
          double z = rr * c + zt, e = rr * sn + et, vR[3], vW[3];
          for (int l = 0; l < 3; l++) {
            vR[l] = rr * (cr[l][1] * c + cr[l][2] * sn + cr[l][3] * (rr * c * sn + zt * sn + et * c));
            vW[l] = cw[l][0] + cw[l][1] * z + cw[l][2] * e;
          }
          double aW = abs3(vW);
          /* value and gradient from their separate chains, as q0_bil_sqG_r and
           * q0_bil_sqH_r call d3hm_qpgf_d2_qG and _dqG (q0_bil_coef.c:440, :488) */
          double complex qG, dqG[3];
          int err = orc_qpgf(0, &bd->qd, vR, MEPS, &qG, NULL, NULL);
          if (err < 0) return err;
          terms += err;
          err = orc_qpgf(1, &bd->qd, vR, MEPS, NULL, dqG, NULL);
          if (err < 0) return err;
          terms += err;
          double complex hn = dqG[0] * vW[0] + dqG[1] * vW[1] + dqG[2] * vW[2];
          for (int n = 0; n < 4; n++) {
            double Nn = orc_bil_N(n, z, e);
            aG[n] += qG * Nn * aW * rr * bd->wli[i];
            aH[n] += hn * Nn * rr * bd->wli[i];
          }
        }
        for (int n = 0; n < 4; n++) {
          sG[n] += aG[n] * 0.5 * rm;
          sH[n] += aH[n] * 0.5 * rm;
        }
      }
      for (int n = 0; n < 4; n++) {
        tG[n] += sG[n] * bd->whi[m];
        tH[n] += sH[n] * bd->whi[m];
      }
    }
    for (int n = 0; n < 4; n++) {
      CC[n] += tG[n] * 0.25 * (th[j + 1] - th[j]);
      CC[n + 4] += tH[n] * 0.25 * (th[j + 1] - th[j]);
    }
  }
  CC[8] = bil_self_F(bd, cr, cw, zt, et);
  return terms;
}

/* dF/dv of the periodic self element by central difference of F along the
 * element parameter, scaled by 1/|T| (q0_bil_dcoef_v2_cc.c:336-390) */
void orc_q0_bil_self_dF(const BOUD *bd, int s, double zt, double et, double *dFz, double *dFe)
{
  double cr[3][4], cw[3][3];
  orc_elem_copy(bd, s, cr, cw);
  double Tz[3], Te[3];
  for (int i = 0; i < 3; i++) {
    Tz[i] = cr[i][1] + cr[i][3] * et;
    Te[i] = cr[i][2] + cr[i][3] * zt;
  }
  double Fp = bil_self_F(bd, cr, cw, zt + CDFH, et), Fm = bil_self_F(bd, cr, cw, zt - CDFH, et);
  *dFz = 1.0 / abs3(Tz) * (Fp - Fm) / (2.0 * CDFH);
  Fp = bil_self_F(bd, cr, cw, zt, et + CDFH);
  Fm = bil_self_F(bd, cr, cw, zt, et - CDFH);
  *dFe = 1.0 / abs3(Te) * (Fp - Fm) / (2.0 * CDFH);
}

/* free-space self element G, H, F (bil_coef_bd, bil_coef.c:244-342) */
void orc_bil_self(const BOUD *bd, int s_signed, double zt, double et, double k, double complex *CC)
{
  int sig = s_signed < 0, s = abs(s_signed);
  double cr[3][4], cw[3][3];
  orc_elem_copy(bd, s, cr, cw);
  double r[5], th[5];
  bil_corners(zt, et, r, th);
  int plane = orc_bil_is_plane(cr, cw);
  double bdC = cr[0][1] * cw[0][2] + cr[1][1] * cw[1][2] + cr[2][1] * cw[2][2];
  for (int l = 0; l < 9; l++) CC[l] = 0.0;
  for (int j = 0; j < 4; j++) {
    double complex tG[4] = {0}, tH[4] = {0};
    for (int m = 0; m < GHN; m++) {
      double complex sG[4] = {0}, sH[4] = {0};
      for (int h = 0; h < 2; h++) {
        double t = theta_node(th, j, h, bd->xhi[m]);
        double rm = polar_rm(r, th, j, t);
        double c = cos(t), sn = sin(t);
        double complex aG[4] = {0}, aH[4] = {0};
        for (int i = 0; i < GLN; i++) {
          double rr = 0.5 * rm * (bd->xli[i] + 1.0);
          double z = rr * c + zt, e = rr * sn + et, K[3], W[3];
          polar_K(cr, c, sn, rr, zt, et, K);
          for (int l = 0; l < 3; l++) W[l] = cw[l][0] + cw[l][1] * z + cw[l][2] * e;
          double aK = abs3(K), aW = abs3(W);
          double complex ce = cos(k * aK * rr) + I * sin(k * aK * rr);
          double complex kR = k * rr * aK;
          for (int n = 0; n < 4; n++) {
            double Nn = orc_bil_N(n, z, e);
            aG[n] += ce / aK * aW * Nn * bd->wli[i];
            aH[n] += (I * kR - 1.0) * ce / (aK * aK * aK) * c * sn * Nn * bd->wli[i];
          }
        }
        for (int n = 0; n < 4; n++) {
          sG[n] += aG[n] * 0.5 * rm;
          sH[n] += aH[n] * 0.5 * rm;
        }
      }
      for (int n = 0; n < 4; n++) {
        tG[n] += sG[n] * bd->whi[m];
        tH[n] += sH[n] * bd->whi[m];
      }
    }
    for (int n = 0; n < 4; n++) {
      CC[n] += tG[n] * 0.25 * (th[j + 1] - th[j]);
      CC[n + 4] += tH[n] * 0.25 * (th[j + 1] - th[j]);
    }
  }
  for (int n = 0; n < 4; n++) {
    CC[n] *= I_FP;
    CC[n + 4] = plane ? 0.0 : CC[n + 4] * (I_FP * bdC);
  }
  CC[8] = bil_self_F(bd, cr, cw, zt, et);
  if (sig) for (int l = 4; l < 9; l++) CC[l] = -CC[l];
}

/*
 * Principal-value tangential derivative of the free-space self coefficients
 * along the element's unit tangent vt = T/|T| (T = dr/dzeta for dir 0, dr/deta for
 * dir 1): bil_dcoef_bd_tz0_PV / te0_PV (bil_dcoef.c:369-740).  The element is
 * split into the disc D_eps of parameter radius DEPS around the node and the
 * rest; D_eps is integrated in a principal-value form (terms "De" and "DP").
 */
typedef struct {
  double cr[3][4], cw[3][3], zt, et, k, vt[3];
  int dir; /* 0: zeta, 1: eta */
} pvctx;

/* the geometric factor that picks the tangent component: sin for zeta, cos for eta */
static inline double pv_sel(const pvctx *P, double c, double s) { return P->dir == 0 ? s : c; }

static void pv_terms_Db(const pvctx *P, double r, double c, double s, double complex *dG, double complex *dH,
                        double *dF)
{
  /* D_bar integrands at radius r (bil_dcoef.c:942-970, :1175-1199, :1300-1312, and the eta versions) */
  double z = r * c + P->zt, e = r * s + P->et, K[3], W[3];
  polar_K((double (*)[4])P->cr, c, s, r, P->zt, P->et, K);
  for (int l = 0; l < 3; l++) W[l] = P->cw[l][0] + P->cw[l][1] * z + P->cw[l][2] * e;
  double aK = abs3(K), iK = 1.0 / aK, aW = abs3(W);
  double k = P->k;
  double complex ca = I * k * aK * r, cb = ca - 1.0;
  double complex ce = cos(k * aK * r) + I * sin(k * aK * r);
  double KdTn = dot3(K, P->vt) * iK;
  double KdT = dot3(K, P->vt);
  double sel = P->dir == 0 ? c : s; /* cth for zeta, sth for eta inside the H and F kernels */
  double complex kR = k * r * aK;
  for (int n = 0; n < 4; n++) {
    double Nn = orc_bil_N(n, z, e);
    dG[n] = cb * ce * iK * iK / r * KdTn * Nn * aW;
    dH[n] = ce / (r * aK * aK * aK * aK * aK) * ((kR * kR + 3.0 * (I * kR - 1.0)) * KdT * sel - (I * kR - 1.0) * aK * aK) * Nn;
  }
  *dF = 1.0 / (r * aK * aK * aK * aK * aK) * (3.0 * KdT * sel - aK * aK);
}

static void pv_terms_De(const pvctx *P, double r, double c, double s, double complex *dG, double complex *dH0,
                        double complex *dH1)
{
  /* D_eps integrands (bil_dcoef.c:990-1022, :1237-1284) */
  double z = r * c + P->zt, e = r * s + P->et, K[3], W[3];
  polar_K((double (*)[4])P->cr, c, s, r, P->zt, P->et, K);
  for (int l = 0; l < 3; l++) W[l] = P->cw[l][0] + P->cw[l][1] * z + P->cw[l][2] * e;
  double aK = abs3(K), iK = 1.0 / aK, aW = abs3(W);
  double k = P->k;
  double kR = k * aK * r;
  double complex ca = I * kR, cb = ca - 1.0;
  double cc = cos(kR), cs = sin(kR);
  double KdTn = dot3(K, P->vt) * iK, KdT = dot3(K, P->vt);
  double sel = P->dir == 0 ? c : s;
  double complex ce = cc + I * cs, cbk = I * k;
  for (int n = 0; n < 4; n++) {
    double Nn = orc_bil_N(n, z, e);
    dG[n] = (cb * cs + kR * cc) * iK * iK / r * KdTn * aW * Nn;
    dH0[n] = ce / (aK * aK * aK * aK) * ((k * kR + 3.0 * cbk) * KdT * sel - cbk * aK * aK) * Nn;
    dH1[n] = cs / (aK * aK * aK * aK * aK * r) * (3.0 * KdT * sel - aK * aK) * Nn;
  }
}

static void pv_terms_DP(const pvctx *P, double r, double c, double s, double complex *dG, double complex *dH,
                        double *dF)
{
  /* symmetric PV pair at +r and -r (bil_dcoef.c:1042-1087, :1149-1173, :1338-1358) */
  double zp = r * c + P->zt, ep = r * s + P->et, zm = -r * c + P->zt, em = -r * s + P->et;
  double Kp[3], Km[3], Wp[3], Wm[3];
  for (int i = 0; i < 3; i++) {
    Kp[i] = P->cr[i][1] * c + P->cr[i][2] * s + P->cr[i][3] * (r * c * s + P->zt * s + P->et * c);
    Km[i] = P->cr[i][1] * c + P->cr[i][2] * s + P->cr[i][3] * (-r * c * s + P->zt * s + P->et * c);
    Wp[i] = P->cw[i][0] + P->cw[i][1] * zp + P->cw[i][2] * ep;
    Wm[i] = P->cw[i][0] + P->cw[i][1] * zm + P->cw[i][2] * em;
  }
  double aKp = abs3(Kp), aKm = abs3(Km), iKp = 1.0 / aKp, iKm = 1.0 / aKm, aWp = abs3(Wp), aWm = abs3(Wm);
  double k = P->k;
  double ccp = cos(k * aKp * r), ccm = cos(-k * aKm * r);
  double KpT = dot3(Kp, P->vt), KmT = dot3(Km, P->vt);
  double sel = P->dir == 0 ? c : s;
  for (int n = 0; n < 4; n++) {
    double Np = orc_bil_N(n, zp, ep), Nm = orc_bil_N(n, zm, em);
    dG[n] = 1.0 / r * (ccp * iKp * iKp * (KpT * iKp) * Np * aWp - ccm * iKm * iKm * (KmT * iKm) * Nm * aWm);
    dH[n] = 1.0 / r * (ccp / (aKp * aKp * aKp * aKp * aKp) * (3.0 * KpT * sel - aKp * aKp) * Np -
                       ccm / (aKm * aKm * aKm * aKm * aKm) * (3.0 * KmT * sel - aKm * aKm) * Nm);
  }
  *dF = 1.0 / r * (1.0 / (aKp * aKp * aKp * aKp * aKp) * (3.0 * KpT * sel - aKp * aKp) -
                   1.0 / (aKm * aKm * aKm * aKm * aKm) * (3.0 * KmT * sel - aKm * aKm));
}

void orc_bil_self_dPV(const BOUD *bd, int s_signed, double zt, double et, double k, int dir, double complex *dCC)
{
  int sig = s_signed < 0, s = abs(s_signed);
  pvctx P;
  orc_elem_copy(bd, s, P.cr, P.cw);
  P.zt = zt;
  P.et = et;
  P.k = k;
  P.dir = dir;
  for (int i = 0; i < 3; i++) P.vt[i] = dir == 0 ? P.cr[i][1] + P.cr[i][3] * et : P.cr[i][2] + P.cr[i][3] * zt;
  double iT = 1.0 / abs3(P.vt);
  double r[5], th[5];
  bil_corners(zt, et, r, th);
  int plane = orc_bil_is_plane(P.cr, P.cw);
  double bdC = P.cr[0][1] * P.cw[0][2] + P.cr[1][1] * P.cw[1][2] + P.cr[2][1] * P.cw[2][2];
  double complex G[4] = {0}, H[4] = {0};
  double F = 0.0;

  /* D_bar: sub-triangles, r in [DEPS, rm], GHN in r */
  for (int j = 0; j < 4; j++) {
    double complex tG[4] = {0}, tH[4] = {0};
    double tF = 0.0;
    for (int m = 0; m < GHN; m++) {
      double complex sG[4] = {0}, sH[4] = {0};
      double sF = 0.0;
      for (int h = 0; h < 2; h++) {
        double t = theta_node(th, j, h, bd->xhi[m]);
        double rm = polar_rm(r, th, j, t);
        double c = cos(t), sn = sin(t), sel = pv_sel(&P, c, sn);
        double complex aG[4] = {0}, aH[4] = {0};
        double aF = 0.0;
        for (int i = 0; i < GHN; i++) {
          double rr = 0.5 * (rm - DEPS) * bd->xhi[i] + 0.5 * (rm + DEPS);
          double complex g[4], hh[4];
          double f;
          pv_terms_Db(&P, rr, c, sn, g, hh, &f);
          for (int n = 0; n < 4; n++) {
            aG[n] += g[n] * bd->whi[i];
            aH[n] += hh[n] * bd->whi[i];
          }
          aF += f * bd->whi[i];
        }
        for (int n = 0; n < 4; n++) {
          sG[n] += aG[n] * 0.5 * (rm - DEPS);
          sH[n] += sel * (aH[n] * 0.5 * (rm - DEPS));
        }
        sF += sel * (aF * 0.5 * (rm - DEPS));
      }
      for (int n = 0; n < 4; n++) {
        tG[n] += sG[n] * bd->whi[m];
        tH[n] += sH[n] * bd->whi[m];
      }
      tF += sF * bd->whi[m];
    }
    for (int n = 0; n < 4; n++) {
      G[n] += tG[n] * 0.25 * (th[j + 1] - th[j]);
      H[n] += tH[n] * 0.25 * (th[j + 1] - th[j]);
    }
    F += tF * 0.25 * (th[j + 1] - th[j]);
  }

  /* D_eps: theta over (-pi, pi] in four quarters, r in [0, DEPS], GLN in r */
  {
    double complex tG[4] = {0}, tH[4] = {0};
    static const double off[4] = {-0.75, -0.25, 0.25, 0.75};
    for (int m = 0; m < GHN; m++) {
      double complex sG[4] = {0}, sH[4] = {0};
      for (int q = 0; q < 4; q++) {
        double t = 0.25 * M_PI * bd->xhi[m] + off[q] * M_PI;
        double c = cos(t), sn = sin(t), sel = pv_sel(&P, c, sn);
        double complex aG[4] = {0}, aH0[4] = {0}, aH1[4] = {0};
        for (int i = 0; i < GLN; i++) {
          double rr = 0.5 * DEPS * (bd->xli[i] + 1.0);
          double complex g[4], h0[4], h1[4];
          pv_terms_De(&P, rr, c, sn, g, h0, h1);
          for (int n = 0; n < 4; n++) {
            aG[n] += g[n] * bd->wli[i];
            aH0[n] += h0[n] * bd->wli[i];
            aH1[n] += h1[n] * bd->wli[i];
          }
        }
        for (int n = 0; n < 4; n++) {
          sG[n] += aG[n] * 0.5 * DEPS;
          sH[n] += sel * (aH0[n] * 0.5 * DEPS - I * (aH1[n] * 0.5 * DEPS));
        }
      }
      for (int n = 0; n < 4; n++) {
        tG[n] += sG[n] * bd->whi[m];
        tH[n] += sH[n] * bd->whi[m];
      }
    }
    for (int n = 0; n < 4; n++) {
      G[n] += I * (tG[n] * 0.25 * M_PI);
      H[n] += tH[n] * 0.25 * M_PI;
    }
  }

  /* D_eps principal-value pairs: theta over (0, pi] in two halves */
  {
    double complex tG[4] = {0}, tH[4] = {0};
    double tF = 0.0;
    static const double off[2] = {0.25, 0.75};
    for (int m = 0; m < GHN; m++) {
      double complex sG[4] = {0}, sH[4] = {0};
      double sF = 0.0;
      for (int q = 0; q < 2; q++) {
        double t = 0.25 * M_PI * bd->xhi[m] + off[q] * M_PI;
        double c = cos(t), sn = sin(t), sel = pv_sel(&P, c, sn);
        double complex aG[4] = {0}, aH[4] = {0};
        double aF = 0.0;
        for (int i = 0; i < GLN; i++) {
          double rr = 0.5 * DEPS * (bd->xli[i] + 1.0);
          double complex g[4], hh[4];
          double f;
          pv_terms_DP(&P, rr, c, sn, g, hh, &f);
          for (int n = 0; n < 4; n++) {
            aG[n] += g[n] * bd->wli[i];
            aH[n] += hh[n] * bd->wli[i];
          }
          aF += f * bd->wli[i];
        }
        for (int n = 0; n < 4; n++) {
          sG[n] += aG[n] * 0.5 * DEPS;
          sH[n] += sel * (aH[n] * 0.5 * DEPS);
        }
        sF += sel * (aF * 0.5 * DEPS);
      }
      for (int n = 0; n < 4; n++) {
        tG[n] += sG[n] * bd->whi[m];
        tH[n] += sH[n] * bd->whi[m];
      }
      tF += sF * bd->whi[m];
    }
    for (int n = 0; n < 4; n++) {
      G[n] -= tG[n] * 0.25 * M_PI;
      H[n] -= tH[n] * 0.25 * M_PI;
    }
    F += tF * 0.25 * M_PI;
  }

  for (int n = 0; n < 4; n++) dCC[n] = G[n] * (-I_FP * iT);
  if (plane) {
    for (int l = 4; l < 9; l++) dCC[l] = 0.0;
  } else {
    for (int n = 0; n < 4; n++) dCC[n + 4] = H[n] * (bdC * I_FP * iT);
    dCC[8] = F * (bdC * I_FP * iT);
  }
  if (sig) for (int l = 4; l < 9; l++) dCC[l] = -dCC[l];
}
