/*
 * field_cpu.c -- oracle scattered field at a point off the boundary.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates pv_s_dbieq_aqd2 and pv_s_bieq_aqd2 (d3qd2_src/bem3_aw_qd2_field.c:
 * 140-201, 227-297) for the element rules of field types 2 (GLN) and 3 (GHN)
 * at a point of the base lattice cell, for bilinear and linear-triangle
 * elements alike, with the element integrals of elem_cpu.c and lit_cpu.c
 * (coef_rt / dcoef_rt_grad, d3qd2_elem.c:148-182, 316-356):
 *
 *   P = sum_s sum_n G_n dP_n - H_n P_n,   F = sum_s F_s,   C = 1 + F (open
 *   domain) or F (closed domain);  p = -k2 P / C;
 *   v = -(grad P - P grad F) / C   (gradient form), or
 *   v = (sum_s sum_n G_n dpv_n - H_n pv_n) / C   (boundary-value form, used
 *   by the reference when |grad F| > CBD_CDF, i.e. near the boundary).
 *
 * In the periodic open domain the oracle has no gradient of G and H over the
 * lattice sum (the reference's gradient form needs the sum's Hessian), so there
 * p is restated exactly (it is the same sum in both forms) and v only in the
 * boundary-value form, which is what the reference uses near the boundary.
 *
 * Pinned by tests/test_oracle_field.py: fed the boundary values of a plane
 * wave inside a closed sphere (triangulated and quadrilateral), the field at
 * interior points is that plane wave to the discretization error, in both
 * forms; the element integrals it sums are pinned by tests/test_oracle_lit.py
 * and tests/test_oracle_sample.py.
 */
#include <complex.h>
#include <math.h>
#include "oracle.h"
#include "oracle_elem.h"

/* returns 0 (gradient form), 1 (boundary-value form, near the boundary),
 * 2 (periodic open domain: p exact, v from the boundary-value form) or a
 * negative Green's function status */
int orc_field_s(const AQD2 *qd, int did, const double *rt, int type, double complex *p, double complex *pv,
                double *gradF)
{
  const BOUD *bd = &qd->bd;
  const SUBD *sb = &bd->sb[did];
  const int periodic = (did == 0 && bd->ps == 1);
  const int n = type == 3 ? GHN : GLN;
  const double *x = type == 3 ? bd->xhi : bd->xli, *w = type == 3 ? bd->whi : bd->wli;
  const double ex[3] = {1, 0, 0}, ey[3] = {0, 1, 0}, ez[3] = {0, 0, 1};
  const double k = qd->k0[did];
  double complex P = 0.0, dP[3] = {0}, V[3] = {0};
  double F = 0.0, dF[3] = {0};
  for (int s = 1; s <= sb->Ne; s++) {
    int sd = sb->sid[s];
    int tri = bd->ed[abs(sd)][3] == 0;
    double complex CC[9], d0[9], d1[9], d2[9], dz[9];
    if (periodic) {
      long c = tri ? orc_q0_lit_pair(bd, rt, ex, ey, sd, n, x, w, CC, d0, d1)
                   : orc_q0_bil_pair(bd, rt, ex, ey, sd, n, x, w, CC, d0, d1);
      if (c < 0) return (int)c;
      /* z derivative of F (the periodic pair carries dF only) */
      double complex C2[9];
      c = tri ? orc_q0_lit_pair(bd, rt, ez, ez, sd, n, x, w, C2, d2, dz)
              : orc_q0_bil_pair(bd, rt, ez, ez, sd, n, x, w, C2, d2, dz);
      if (c < 0) return (int)c;
    } else {
      if (tri) {
        orc_lit_pair(bd, rt, ex, ey, sd, k, n, x, w, CC, d0, d1);
        orc_lit_pair(bd, rt, ez, ez, sd, k, n, x, w, CC, d2, dz);
      } else {
        orc_bil_pair(bd, rt, ex, ey, sd, k, n, x, w, CC, d0, d1);
        orc_bil_pair(bd, rt, ez, ez, sd, k, n, x, w, CC, d2, dz);
      }
    }
    const double complex *dC[3] = {d0, d1, d2};
    for (int m = 0; m < 4; m++) {
      P += CC[m] * sb->dP[s][m] - CC[m + 4] * sb->P[s][m];
      for (int i = 0; i < 3; i++) {
        dP[i] += dC[i][m] * sb->dP[s][m] - dC[i][m + 4] * sb->P[s][m];
        V[i] += CC[m] * sb->dpv[s][m][i] - CC[m + 4] * sb->pv[s][m][i];
      }
    }
    F += creal(CC[8]);
    for (int i = 0; i < 3; i++) dF[i] += creal(dC[i][8]);
  }
  const double iC = 1.0 / (did == 0 ? 1.0 + F : F);
  *p = -qd->k2[did] * P * iC;
  for (int i = 0; i < 3; i++) gradF[i] = dF[i];
  if (periodic || fabs(dF[0]) > CBD_CDF || fabs(dF[1]) > CBD_CDF || fabs(dF[2]) > CBD_CDF) {
    for (int i = 0; i < 3; i++) pv[i] = V[i] * iC;
    return periodic ? 2 : 1;
  }
  for (int i = 0; i < 3; i++) pv[i] = -(dP[i] - P * dF[i]) * iC;
  return 0;
}

/* scattered field at the boundary point (zt, et) of element position t of
 * domain did (pv_s_bd_aqd2, bem3_aw_qd2_field.c:394-430) for field types 0, 1
 * (the 4-/9-point and 4-/7-point rules of orc_coef_pts), 2 (GLN) and 3 (GHN):
 * the other elements on the element rule, the element itself on its polar rule
 * about the point (coef_bd / q0_coef_bd, d3qd2_elem.c:187-230, :684-726).
 * p = -k2 P / C, v = (sum G dpv - H pv) / C.  Returns 0 or a negative Green's
 * function status. */
int orc_field_bd(const AQD2 *qd, int did, int t, double zt, double et, int type, double complex *p,
                 double complex *pv)
{
  const BOUD *bd = &qd->bd;
  const SUBD *sb = &bd->sb[did];
  const int periodic = (did == 0 && bd->ps == 1);
  const int n = type == 3 ? GHN : GLN;
  const double *x = type == 3 ? bd->xhi : bd->xli, *w = type == 3 ? bd->whi : bd->wli;
  const double k = qd->k0[did];
  const int td = sb->sid[t];
  double cr[3][4], cw[3][3], rt[3], v0[3] = {1, 0, 0};
  orc_elem_copy(bd, abs(td), cr, cw);
  for (int l = 0; l < 3; l++) rt[l] = cr[l][0] + cr[l][1] * zt + cr[l][2] * et + cr[l][3] * zt * et;
  double complex P = 0.0, V[3] = {0};
  double F = 0.0;
  for (int s = 1; s <= sb->Ne; s++) {
    int sd = sb->sid[s];
    int tri = bd->ed[abs(sd)][3] == 0;
    double complex CC[9], d0[9], d1[9];
    long c = 0;
    if (sd != td && type <= 1) {
      c = orc_coef_pts(qd, did, rt, sd, type, CC);
    } else if (sd != td) {
      if (periodic)
        c = tri ? orc_q0_lit_pair(bd, rt, v0, v0, sd, n, x, w, CC, d0, d1)
                : orc_q0_bil_pair(bd, rt, v0, v0, sd, n, x, w, CC, d0, d1);
      else if (tri)
        orc_lit_pair(bd, rt, v0, v0, sd, k, n, x, w, CC, d0, d1);
      else
        orc_bil_pair(bd, rt, v0, v0, sd, k, n, x, w, CC, d0, d1);
    } else {
      if (periodic)
        c = tri ? orc_q0_lit_self(bd, sd, zt, et, CC) : orc_q0_bil_self(bd, sd, zt, et, CC);
      else if (tri)
        orc_lit_self(bd, sd, zt, et, k, CC);
      else
        orc_bil_self(bd, sd, zt, et, k, CC);
    }
    if (c < 0) return (int)c;
    for (int m = 0; m < 4; m++) {
      P += CC[m] * sb->dP[s][m] - CC[m + 4] * sb->P[s][m];
      for (int i = 0; i < 3; i++) V[i] += CC[m] * sb->dpv[s][m][i] - CC[m + 4] * sb->pv[s][m][i];
    }
    F += creal(CC[8]);
  }
  const double iC = 1.0 / (did == 0 ? 1.0 + F : F);
  *p = -qd->k2[did] * P * iC;
  for (int i = 0; i < 3; i++) pv[i] = V[i] * iC;
  return 0;
}
