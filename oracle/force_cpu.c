/*
 * force_cpu.c -- oracle radiation force and torque, with the boundary field
 * they need.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restated reference routines:
 *   incident plane wave        calc_apw_pv, apw_pv_axz      (d3qd2_src/aw_pw.c:128-150, :176-188)
 *   point-rule coefficients    q0_bil_coef_4p/9p            (q0_bil_coef.c:10-92)
 *   (field types 0 and 1)      bil_coef_4p/9p               (bil_coef.c:25-104)
 *                              q0_lit_coef_4p/7p            (q0_lit_coef.c:10-118)
 *                              lit_coef_4p/7p               (lit_coef.c:26-122)
 *   element dispatch           coef_bd / q0_coef_bd         (d3qd2_elem.c:187-230, :684-726)
 *   total boundary field       pv_t_bd_aqd2                 (bem3_aw_qd2_field.c:432-469)
 *   force and torque           force_FN_aqd2, bil_force_4p/9p, lit_force_4p/7p
 *                                                           (bem3_aw_qd2_force.c:4-42, :73-278)
 * The boundary field of a type-0/1 call takes the element itself on the polar
 * rule about the point (q0_bil_coef_bd / bil_coef_bd and their triangle forms,
 * elem_cpu.c and lit_cpu.c) and the other elements on the 4-, 9- (quads) or
 * 4-, 7-point (triangles) rule.  Each element's force and torque contribution
 * is formed on its own and the contributions are summed in element order, so
 * the result does not depend on the OpenMP team.
 *
 * Pins: tests/test_oracle_force.py -- the type-0 force of the reference's four
 * samples (which reads only the surface solution in ex.dat and the incident
 * wave) against their example1.log; the type-1 boundary field of a closed
 * sphere fed a plane wave's boundary values against that plane wave; the
 * stress-tensor quadrature against the momentum flux of a plane wave through a
 * closed surface (zero) and the point rules' weights against the element area
 * (surface_area_aqd2); the type-1 force of the samples against the log within
 * the quadrature-rule difference the log itself shows between types 0 and 1.
 */
#include <math.h>
#include <complex.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"
#include "oracle_elem.h"

#define I_FP 7.957747154594766788444e-2 /* 1/(4 pi) */
#define I_EP 3.978873577297383394222e-2 /* 1/(8 pi) */

static inline double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline double abs3(const double *a) { return sqrt(dot3(a, a)); }

/* calc_apw_pv (aw_pw.c:128-150) with apw_pv_axz (:176-188): the plane wave
 * p = p0 exp(i k0 z') along the rotated axis z', v = p / (rho0 c0) along z'. */
void orc_apw_pv(const Apw *pw, const double *r, double complex *p, double complex *v)
{
  const double *R = pw->wd.R;
  double rc[3], r0[3];
  for (int i = 0; i < 3; i++) rc[i] = r[i] - pw->tv[i];
  r0[0] = R[0] * rc[0] + R[1] * rc[1] - R[2] * rc[2];
  r0[1] = R[3] * rc[0] + R[4] * rc[1] - R[5] * rc[2];
  r0[2] = -R[6] * rc[0] - R[7] * rc[1] + R[8] * rc[2];
  double arg = pw->wd.k0 * r0[2];
  double complex ce = cos(arg) + I * sin(arg);
  *p = pw->p * ce;
  double complex vt[3] = {0.0, 0.0, *p / (pw->rho0 * pw->c0)};
  for (int i = 0; i < 3; i++) {
    v[i] = 0.0;
    for (int j = 0; j < 3; j++) v[i] += R[i * 3 + j] * vt[j];
  }
}

static double lit_Nn(int n, double z, double e)
{
  const double ia = 0.5 / 0.2, s3 = sqrt(3.0);
  switch (n) {
  case 0: return 1.0 / 3.0 * (1.0 - ia * (z + s3 * e));
  case 1: return 1.0 / 3.0 * (1.0 + 2.0 * ia * z);
  default: return 1.0 / 3.0 * (1.0 - ia * (z - s3 * e));
  }
}

/* Coefficients CC[9] of element s (unsigned id for the periodic domain, signed
 * for the others) seen from rt with the point rules of field types 0 and 1.
 * Returns the lattice-sum term count (> 0), 0 for free space, or a negative
 * Green's function status. */
long orc_coef_pts(const AQD2 *ad, int did, const double *rt, int sd, int type, double complex *CC)
{
  const BOUD *bd = &ad->bd;
  const int periodic = (did == 0 && bd->ps == 1);
  const double k = ad->k0[did];
  int s = abs(sd), sig = sd < 0;
  const int tri = bd->ed[s][3] == 0;
  long terms = 0;
  for (int l = 0; l < 9; l++) CC[l] = 0.0;
  /* quadrature points: (zeta, eta, weight, position, normal) */
  int np;
  double zq[9], eq[9], wq[9], rq[9][3], wv[9][3];
  double cr[3][4], cw[3][3];
  orc_elem_copy(bd, s, cr, cw);
  if (tri) {
    for (int l = 0; l < 3; l++) { cr[l][3] = 0.0; cw[l][1] = cw[l][2] = 0.0; }
  }
  if (type == 0) {
    np = 4;
    for (int i = 0; i < 4; i++) {
      zq[i] = tri ? bd->zt_34[i] : bd->zt_44[i];
      eq[i] = tri ? bd->et_34[i] : bd->et_44[i];
      wq[i] = tri ? bd->wt_34[i] : bd->wt_44[i];
      for (int l = 0; l < 3; l++) { rq[i][l] = bd->ren[s][i][l]; wv[i][l] = bd->wen[s][i][l]; }
    }
  } else {
    np = tri ? 7 : 9;
    for (int i = 0; i < np; i++) {
      zq[i] = tri ? bd->zt_37[i] : bd->zt_49[i];
      eq[i] = tri ? bd->et_37[i] : bd->et_49[i];
      wq[i] = tri ? bd->wt_37[i] : bd->wt_49[i];
      for (int l = 0; l < 3; l++) {
        rq[i][l] = cr[l][0] + cr[l][1] * zq[i] + cr[l][2] * eq[i] + cr[l][3] * zq[i] * eq[i];
        wv[i][l] = cw[l][0] + cw[l][1] * zq[i] + cw[l][2] * eq[i];
      }
    }
  }
  /* kernel at each point: value g (with the reference's scaling) and h = the
   * normal derivative against the point's normal; tD = (R.W)/R^3 */
  double complex g[9], h[9];
  double tD[9];
  double Vt[3]; /* a triangle's constant normal */
  for (int l = 0; l < 3; l++) Vt[l] = cw[l][0];
  for (int i = 0; i < np; i++) {
    double vR[3];
    for (int l = 0; l < 3; l++) vR[l] = rq[i][l] - rt[l];
    const double aR = abs3(vR), iR = 1.0 / aR;
    const double *W = tri ? Vt : wv[i];
    tD[i] = (vR[0] * W[0] + vR[1] * W[1] + vR[2] * W[2]) * iR * iR * iR;
    if (periodic) {
      double complex qG, dqG[3];
      int err = orc_qpgf(2, &bd->qd, vR, MEPS, &qG, dqG, NULL);
      if (err < 0) return err;
      terms += err;
      g[i] = qG;
      h[i] = dqG[0] * W[0] + dqG[1] * W[1] + dqG[2] * W[2];
    } else {
      double complex ca = I * k * aR;
      double complex ce = cos(k * aR) + I * sin(k * aR);
      g[i] = ce * iR;                 /* times I_FP / I_EP below */
      h[i] = (ca - 1.0) * ce * tD[i]; /* likewise */
    }
  }
  if (!tri) {
    if (type == 0) {
      /* q0_bil_coef_4p :28-46, bil_coef_4p :37-54: node values, unit weights */
      for (int i = 0; i < 4; i++) {
        const double aW = abs3(wv[i]);
        if (periodic) {
          CC[i] = g[i] * aW;
          CC[i + 4] = h[i];
        } else {
          CC[i] = I_FP * (g[i] * aW);
          CC[i + 4] = I_FP * h[i];
        }
        CC[8] += bd->wt_44[i] * tD[i];
      }
      CC[8] *= I_FP;
    } else {
      /* q0_bil_coef_9p :66-90, bil_coef_9p :74-99 */
      for (int ep = 0; ep < 4; ep++) {
        for (int i = 0; i < 9; i++) {
          const double aW = abs3(wv[i]);
          const double Np = orc_bil_N(ep, zq[i], eq[i]);
          if (periodic) {
            CC[ep] += wq[i] * g[i] * Np * aW;
            CC[ep + 4] += wq[i] * h[i] * Np;
          } else {
            CC[ep] += wq[i] * g[i] * aW * Np;
            CC[ep + 4] += wq[i] * h[i] * Np;
          }
          if (ep == 0) CC[8] += wq[i] * tD[i];
        }
        if (!periodic) {
          CC[ep] *= I_FP;
          CC[ep + 4] *= I_FP;
        }
        if (ep == 0) CC[8] *= I_FP;
      }
    }
    if (!periodic && sig) {
      for (int l = 4; l < 9; l++) CC[l] = -CC[l];
    }
  } else {
    const double aV = abs3(Vt);
    if (type == 0) {
      /* q0_lit_coef_4p :26-70, lit_coef_4p :38-75: the three nodes, then the
       * centroid shared out by the interpolation functions */
      for (int i = 0; i < 3; i++) {
        CC[i] = bd->wt_34[i] * g[i];
        CC[i + 4] = bd->wt_34[i] * h[i];
        CC[8] += bd->wt_34[i] * tD[i];
      }
      for (int i = 0; i < 3; i++) {
        const double Nn = lit_Nn(i, bd->zt_34[3], bd->et_34[3]);
        CC[i] += bd->wt_34[3] * g[3] * Nn;
        CC[i + 4] += bd->wt_34[3] * h[3] * Nn;
      }
      CC[8] += bd->wt_34[3] * tD[3];
      for (int i = 0; i < 3; i++) {
        if (periodic) {
          CC[i] *= 0.5 * aV;
          CC[i + 4] *= 0.5;
        } else {
          CC[i] *= I_EP * aV;
          CC[i + 4] *= I_EP;
        }
      }
      CC[8] *= I_EP;
    } else {
      /* q0_lit_coef_7p :87-117, lit_coef_7p :96-119 */
      for (int ep = 0; ep < 3; ep++) {
        for (int i = 0; i < 7; i++) {
          const double Np = lit_Nn(ep, zq[i], eq[i]);
          CC[ep] += wq[i] * g[i] * Np;
          CC[ep + 4] += wq[i] * h[i] * Np;
          if (ep == 0) CC[8] += wq[i] * tD[i];
        }
        if (periodic) {
          CC[ep] *= 0.5 * aV;
          CC[ep + 4] *= 0.5;
        } else {
          CC[ep] *= I_EP * aV;
          CC[ep + 4] *= I_EP;
        }
        if (ep == 0) CC[8] *= I_EP;
      }
    }
    if (!periodic && sig) {
      for (int l = 4; l < 7; l++) CC[l] = -CC[l];
      CC[8] = -CC[8];
    }
  }
  return terms;
}

/* pv_t_bd_aqd2 (bem3_aw_qd2_field.c:432-469) for domain 0: scattered boundary
 * field of orc_field_bd plus the incident wave.  Returns 0 or a negative Green's
 * function status. */
int orc_field_bd_total(const AQD2 *ad, int t, double zt, double et, int type, double complex *p,
                       double complex *pv)
{
  int c = orc_field_bd(ad, 0, t, zt, et, type, p, pv);
  if (c < 0) return c;
  const int td = ad->bd.sb[0].sid[t];
  double cr[3][4], cw[3][3], rt[3];
  orc_elem_copy(&ad->bd, abs(td), cr, cw);
  if (ad->bd.ed[abs(td)][3] == 0) for (int l = 0; l < 3; l++) cr[l][3] = 0.0;
  for (int l = 0; l < 3; l++) rt[l] = cr[l][0] + cr[l][1] * zt + cr[l][2] * et + cr[l][3] * zt * et;
  double complex pi, vi[3];
  orc_apw_pv(&ad->pw, rt, &pi, vi);
  *p += pi;
  for (int l = 0; l < 3; l++) pv[l] += vi[l];
  return 0;
}

/* radiation stress tensor of (p, v) contracted with the normal w and weight wt,
 * accumulated into F and into the torque N about rc (bil_force_4p :99-117) */
static void stress_add(const AQD2 *ad, double complex p, const double complex *v, const double *r,
                       const double *w, double wt, const double *rc, double *F, double *N)
{
  const double c1 = 0.25 / (ad->pw.c0 * ad->pw.c0 * ad->pw.rho0);
  const double c2 = 0.25 * ad->pw.rho0, c3 = 0.50 * ad->pw.rho0;
  double T[9];
  const double t1 = c1 * creal(p * conj(p)) - c2 * creal(v[0] * conj(v[0]) + v[1] * conj(v[1]) + v[2] * conj(v[2]));
  for (int l = 0; l < 3; l++)
    for (int m = 0; m < 3; m++) T[l * 3 + m] = c3 * creal(v[l] * conj(v[m]));
  T[0] += t1;
  T[4] += t1;
  T[8] += t1;
  F[0] += (T[0] * w[0] + T[1] * w[1] + T[2] * w[2]) * wt;
  F[1] += (T[3] * w[0] + T[4] * w[1] + T[5] * w[2]) * wt;
  F[2] += (T[6] * w[0] + T[7] * w[1] + T[8] * w[2]) * wt;
  const double x = r[0] - rc[0], y = r[1] - rc[1], z = r[2] - rc[2];
  N[0] += ((y * T[6] - z * T[3]) * w[0] + (y * T[7] - z * T[4]) * w[1] + (y * T[8] - z * T[5]) * w[2]) * wt;
  N[1] += ((z * T[0] - x * T[6]) * w[0] + (z * T[1] - x * T[7]) * w[1] + (z * T[2] - x * T[8]) * w[2]) * wt;
  N[2] += ((x * T[3] - y * T[0]) * w[0] + (x * T[4] - y * T[1]) * w[1] + (x * T[5] - y * T[2]) * w[2]) * wt;
}

/* node values of the total field: incident plus the surface solution
 * (bil_force_4p :95-97) */
static void node_field(const AQD2 *ad, int t, int i, const double *r, double complex *p, double complex *v)
{
  orc_apw_pv(&ad->pw, r, p, v);
  *p += -ad->bd.sb[0].P[t][i] * ad->k2[0];
  for (int l = 0; l < 3; l++) v[l] += ad->bd.sb[0].pv[t][i][l];
}

/* force and torque contribution of element position t of the open domain
 * (bil_force_4p/9p, lit_force_4p/7p).  Returns 0 or a Green's function status. */
int orc_force_elem(const AQD2 *ad, const double *rc, int type, int t, double *F, double *N)
{
  const BOUD *bd = &ad->bd;
  const int s = abs(bd->sb[0].sid[t]);
  const int tri = bd->ed[s][3] == 0;
  double cr[3][4], cw[3][3];
  orc_elem_copy(bd, s, cr, cw);
  if (tri) for (int l = 0; l < 3; l++) { cr[l][3] = 0.0; cw[l][1] = cw[l][2] = 0.0; }
  for (int j = 0; j < 3; j++) F[j] = N[j] = 0.0;
  double complex p, v[3];
  double r[3], w[3];
  if (!tri && type == 0) {
    for (int i = 0; i < 4; i++) {
      node_field(ad, t, i, bd->ren[s][i], &p, v);
      stress_add(ad, p, v, bd->ren[s][i], bd->wen[s][i], bd->wt_44[i], rc, F, N);
    }
    return 0;
  }
  if (!tri) {
    for (int i = 0; i < 9; i++) {
      const double z = bd->zt_49[i], e = bd->et_49[i];
      for (int l = 0; l < 3; l++) {
        r[l] = cr[l][0] + cr[l][1] * z + cr[l][2] * e + cr[l][3] * z * e;
        w[l] = cw[l][0] + cw[l][1] * z + cw[l][2] * e;
      }
      int c = orc_field_bd_total(ad, t, z, e, 1, &p, v);
      if (c < 0) return c;
      stress_add(ad, p, v, r, w, bd->wt_49[i], rc, F, N);
    }
    return 0;
  }
  const int np = type == 0 ? 4 : 7;
  for (int i = 0; i < np; i++) {
    const double z = type == 0 ? bd->zt_34[i] : bd->zt_37[i], e = type == 0 ? bd->et_34[i] : bd->et_37[i];
    const double wt = type == 0 ? bd->wt_34[i] : bd->wt_37[i];
    if (type == 0 && i < 3) {
      for (int l = 0; l < 3; l++) { r[l] = bd->ren[s][i][l]; w[l] = bd->wen[s][i][l]; }
      node_field(ad, t, i, r, &p, v);
    } else {
      for (int l = 0; l < 3; l++) {
        r[l] = cr[l][0] + cr[l][1] * z + cr[l][2] * e;
        w[l] = cw[l][0];
      }
      int c = orc_field_bd_total(ad, t, z, e, type == 0 ? 0 : 1, &p, v);
      if (c < 0) return c;
    }
    stress_add(ad, p, v, r, w, wt, rc, F, N);
  }
  for (int j = 0; j < 3; j++) {
    F[j] *= 0.5;
    N[j] *= 0.5;
  }
  return 0;
}

/* force_FN_aqd2 (bem3_aw_qd2_force.c:4-42).  The elements' contributions are
 * formed in parallel and summed in element order.  Returns 0 or the first
 * Green's function status met. */
int orc_force(const AQD2 *ad, const double *rc, int type, double *F, double *N)
{
  const int ne = ad->bd.sb[0].Ne;
  double *part = calloc((size_t)6 * (ne + 1), sizeof(double));
  int err = 0;
#pragma omp parallel for schedule(dynamic)
  for (int t = 1; t <= ne; t++) {
    int c = orc_force_elem(ad, rc, type, t, part + 6 * t, part + 6 * t + 3);
    if (c < 0) {
#pragma omp critical
      err = c;
    }
  }
  for (int j = 0; j < 3; j++) F[j] = N[j] = 0.0;
  for (int t = 1; t <= ne; t++)
    for (int j = 0; j < 3; j++) {
      F[j] += part[6 * t + j];
      N[j] += part[6 * t + 3 + j];
    }
  free(part);
  return err;
}
