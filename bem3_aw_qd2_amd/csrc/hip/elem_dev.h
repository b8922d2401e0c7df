// elem_dev.h -- device element geometry, interpolation functions and the
// element-rule switch for bilinear (quadrilateral) elements.
//
// Element conventions of the reference (d3qd2_elem.c:903-985): r(z,e) =
// cr0 + cr1 z + cr2 e + cr3 z e, unnormalised normal w(z,e) = cw0 + cw1 z + cw2 e,
// interpolation functions N_n on the collocation nodes at (+-1/sqrt3, +-1/sqrt3).
#pragma once
#include "problem.h"

#pragma clang fp contract(off)

#define BQD2_IFP 7.957747154594766788444e-2  // 1/(4 pi)
#define BQD2_P44 0.5773502691896257645091    // 1/sqrt(3)

struct ElemC {
  double cr[3][4], cw[3][3];
};

__device__ __forceinline__ void load_elem(const DevMesh &M, int s, ElemC &E) {
  const double *c = M.cr + 12 * s;
  const double *w = M.cw + 9 * s;
#pragma unroll
  for (int d = 0; d < 3; d++) {
#pragma unroll
    for (int j = 0; j < 4; j++) E.cr[d][j] = c[4 * d + j];
#pragma unroll
    for (int j = 0; j < 3; j++) E.cw[d][j] = w[3 * d + j];
  }
}

// d3qd2_elem.c:903-915
__device__ __forceinline__ double bil_N(int n, double z, double e) {
  const double ia = 1.0 / BQD2_P44;
  double a = (n == 1 || n == 2) ? 1.0 + z * ia : 1.0 - z * ia;
  double b = (n >= 2) ? 1.0 + e * ia : 1.0 - e * ia;
  return 0.25 * a * b;
}

__device__ __forceinline__ void bil_rw(const ElemC &E, double z, double e, double *r, double *w) {
#pragma unroll
  for (int i = 0; i < 3; i++) {
    r[i] = E.cr[i][0] + E.cr[i][1] * z + E.cr[i][2] * e + E.cr[i][3] * z * e;
    w[i] = E.cw[i][0] + E.cw[i][1] * z + E.cw[i][2] * e;
  }
}

__device__ __forceinline__ void bil_rw_r(const ElemC &E, double z, double e, double *r) {
#pragma unroll
  for (int i = 0; i < 3; i++) r[i] = E.cr[i][0] + E.cr[i][1] * z + E.cr[i][2] * e + E.cr[i][3] * z * e;
}

// free-space Helmholtz kernel e^{ikR}/(4 pi R) and its gradient with respect to
// the offset vR (source minus target), |vR| = aR, 1/|vR| = iR
__device__ __forceinline__ void free_gf(double k, const double *vR, double aR, double iR, cd &g, cd *dg) {
  const cd e = cis(k * aR);
  g = e * (BQD2_IFP * iR);
  const cd f = cd{-1.0, k * aR} * g * (iR * iR);
#pragma unroll
  for (int i = 0; i < 3; i++) dg[i] = f * vR[i];
}

__device__ __forceinline__ double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
__device__ __forceinline__ double abs3(const double *a) { return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }

// unit tangents at a collocation node (d3qd2_elem.c:122-143)
__device__ __forceinline__ void node_tangents(const DevMesh &M, const Tab &T, int s, int nid, double *vtz, double *vte) {
  const double *c = M.cr + 12 * s;
  if (!M.tri[s]) {
#pragma unroll
    for (int i = 0; i < 3; i++) {
      vtz[i] = c[4 * i + 1] + c[4 * i + 3] * T.et_44[nid];
      vte[i] = c[4 * i + 2] + c[4 * i + 3] * T.zt_44[nid];
    }
  } else {
#pragma unroll
    for (int i = 0; i < 3; i++) {
      vtz[i] = c[4 * i + 1];
      vte[i] = c[4 * i + 2];
    }
  }
  double a = abs3(vtz);
  if (a != 0.0) { vtz[0] /= a; vtz[1] /= a; vtz[2] /= a; }
  a = abs3(vte);
  if (a != 0.0) { vte[0] /= a; vte[1] /= a; vte[2] /= a; }
}

// geometry of one quadrature point shared by both kernels
struct PtGeo {
  double vR[3], vW[3], aR, iR, aW, rdW, rdv0, rdv1, Wdv0, Wdv1;
};

__device__ __forceinline__ void point_geo(const ElemC &E, double z, double e, const double *rt, const double *v0,
                                          const double *v1, PtGeo &g) {
  bil_rw(E, z, e, g.vR, g.vW);
  g.vR[0] -= rt[0];
  g.vR[1] -= rt[1];
  g.vR[2] -= rt[2];
  g.aR = abs3(g.vR);
  g.iR = 1.0 / g.aR;
  g.aW = abs3(g.vW);
  g.rdW = dot3(g.vR, g.vW) * g.iR;
  g.rdv0 = dot3(g.vR, v0) * g.iR;
  g.rdv1 = dot3(g.vR, v1) * g.iR;
  g.Wdv0 = dot3(g.vW, v0);
  g.Wdv1 = dot3(g.vW, v1);
}

// An offset whose z component is zero or within rounding of zero relative to its
// length: there the spectral part of the lattice sum ends on rounding residue
// (gf_dev.h), so its value depends on the exact point.  Integrals that would
// move the lattice sum to other points keep the reference's point set instead.
#define BQD2_ZREL 1e-9
__device__ __forceinline__ bool z_degenerate(double x, double y, double z) {
  return fabs(z) <= BQD2_ZREL * sqrt(x * x + y * y + z * z);
}

// every lattice image other than l = 0 of the target at least four element radii
// from the element: then the lattice sum minus the free-space kernel is analytic
// over the element (the split of near and self integrals relies on it)
__device__ __forceinline__ bool images_far(const ElemC &E, const double *rt, const GfParams &P) {
  // element centre and radius
  double c[3], rad = 0.0;
  bil_rw_r(E, 0.0, 0.0, c);
#pragma unroll
  for (int k = 0; k < 4; k++) {
    double q[3];
    bil_rw_r(E, (k == 1 || k == 2) ? 1.0 : -1.0, (k >= 2) ? 1.0 : -1.0, q);
    double d = sqrt((q[0] - c[0]) * (q[0] - c[0]) + (q[1] - c[1]) * (q[1] - c[1]) + (q[2] - c[2]) * (q[2] - c[2]));
    rad = fmax(rad, d);
  }
  for (int l1 = -2; l1 <= 2; l1++) {
    for (int l2 = -2; l2 <= 2; l2++) {
      if (l1 == 0 && l2 == 0) continue;
      double dx = c[0] - rt[0] + l1 * P.d1x + l2 * P.d2x;
      double dy = c[1] - rt[1] + l1 * P.d1y + l2 * P.d2y;
      double dz = c[2] - rt[2];
      if (sqrt(dx * dx + dy * dy + dz * dz) - rad < 4.0 * rad) return false;
    }
  }
  return true;
}

// GL12 dF/dv0, dF/dv1 with the reference's exact operation order and summation
// order (inner j sum, times w_i, then 1/(4 pi)), the quantities that
// d3qd2_elem.c:536-537 / :851-852 compare with CBD_SDF = 0.5.  Products and sums
// are not contracted in this file, so the values match the host reference bit
// for bit and the GL12 / GH64 choice is the same pair by pair.
__device__ __forceinline__ bool needs_gh(const ElemC &E, const Tab &T, const double *rt, const double *v0,
                                         const double *v1) {
  double a0 = 0.0, a1 = 0.0;
  for (int i = 0; i < 12; i++) {
    double t0 = 0.0, t1 = 0.0;
    for (int j = 0; j < 12; j++) {
      PtGeo g;
      point_geo(E, T.xli[i], T.xli[j], rt, v0, v1, g);
      t0 += T.wli[j] * g.iR * g.iR * g.iR * (3.0 * g.rdv0 * g.rdW - g.Wdv0);
      t1 += T.wli[j] * g.iR * g.iR * g.iR * (3.0 * g.rdv1 * g.rdW - g.Wdv1);
    }
    a0 += T.wli[i] * t0;
    a1 += T.wli[i] * t1;
  }
  double dF0 = a0 * BQD2_IFP, dF1 = a1 * BQD2_IFP;
  return fabs(dF0) > 0.5 || fabs(dF1) > 0.5;
}

// wavefront-wide sum (64 lanes) of a double
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ __forceinline__ long long wave_sum_ll(long long v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// corner distances / angles seen from (zt, et) (d3qd2_elem.c:1017-1034)
__device__ __forceinline__ void bil_corners(double zt, double et, double *r, double *th) {
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

__device__ __forceinline__ bool bil_is_plane(const ElemC &E) {
  double bdc = E.cr[0][1] * E.cw[0][2] + E.cr[1][1] * E.cw[1][2] + E.cr[2][1] * E.cw[2][2];
  double ab = sqrt(E.cr[0][1] * E.cr[0][1] + E.cr[1][1] * E.cr[1][1] + E.cr[2][1] * E.cr[2][1]);
  double aC = sqrt(E.cw[0][2] * E.cw[0][2] + E.cw[1][2] * E.cw[1][2] + E.cw[2][2] * E.cw[2][2]);
  return fabs(bdc / (ab * aC)) < 1.0e-15;
}

__device__ __forceinline__ void polar_K(const ElemC &E, double c, double s, double r, double zt, double et, double *K) {
#pragma unroll
  for (int i = 0; i < 3; i++) K[i] = E.cr[i][1] * c + E.cr[i][2] * s + E.cr[i][3] * (r * c * s + zt * s + et * c);
}

// theta node of sub-triangle j, half h, Legendre abscissa x (q0_bil_coef.c:286-300)
__device__ __forceinline__ double theta_node(const double *th, int j, int h, double x) {
  return h == 0 ? 0.25 * (th[j + 1] - th[j]) * x + 0.25 * (th[j + 1] + 3.0 * th[j])
                : 0.25 * (th[j + 1] - th[j]) * x + 0.25 * (3.0 * th[j + 1] + th[j]);
}

__device__ __forceinline__ double polar_rm(const double *r, const double *th, int j, double t) {
  return r[j] * r[j + 1] * sin(th[j + 1] - th[j]) / (r[j] * sin(t - th[j]) - r[j + 1] * sin(t - th[j + 1]));
}

// ------------------------------------------------------------ linear triangles
// Reference triangle: corners (-1/2,-sqrt3/2), (1,0), (-1/2,sqrt3/2); r(z,e) =
// cr0 + cr1 z + cr2 e (cr3 = 0), constant normal V = cw0 (cw1 = cw2 = 0), so the
// bilinear helpers above evaluate triangles unchanged.  Nodes at 2/5 of the way
// to the corners (d3qd2_elem.c:1035-1163).
#define BQD2_SQ3 1.732050807568877293528
#define BQD2_I_SPSQ3 3.062938307898844719507e-2  // 1/(6 pi sqrt3)

__device__ __forceinline__ double lit_N(int n, double z, double e) {
  const double ia = 0.5 / 0.2;
  if (n == 0) return 1.0 / 3.0 * (1.0 - ia * (z + BQD2_SQ3 * e));
  if (n == 1) return 1.0 / 3.0 * (1.0 + 2.0 * ia * z);
  if (n == 2) return 1.0 / 3.0 * (1.0 - ia * (z - BQD2_SQ3 * e));
  return 0.0;
}

// radial derivative term of N_n used by the closed-form disc of the principal
// value (lit_dcoef.c:845-856)
__device__ __forceinline__ double lit_dN(int n, double c, double s) {
  const double ia = 0.5 / 0.2;
  if (n == 0) return -2.0 / 3.0 * ia * (c + BQD2_SQ3 * s);
  if (n == 1) return 4.0 / 3.0 * ia * c;
  return -2.0 / 3.0 * ia * (c - BQD2_SQ3 * s);
}

// the three quadrilateral parts of the triangle as images of (a, b) in [-1,1]^2,
// Jacobian factor 1 + a/4 + b/4 (q0_lit_dcoef_v2_cc.c:196-214)
__device__ __forceinline__ void lit_part(int part, double a, double b, double &z, double &e) {
  if (part == 0) {
    z = 0.125 * (3.0 + 2.0 * a + 2.0 * b + a * b);
    e = 0.125 * BQD2_SQ3 * (-a + b);
  } else if (part == 1) {
    z = 0.0625 * (-3.0 + a - 5.0 * b - a * b);
    e = 0.0625 * BQD2_SQ3 * (3.0 + 3.0 * a + b + a * b);
  } else {
    z = 0.0625 * (-3.0 - 5.0 * a + b - a * b);
    e = 0.0625 * BQD2_SQ3 * (-3.0 - a - 3.0 * b - a * b);
  }
}

// corner distances and angles seen from (zt, et) (d3qd2_elem.c:1149-1163)
__device__ __forceinline__ void lit_corners(double zt, double et, double *r, double *th) {
  double ztp = 1.0 - zt, ztm = -0.5 - zt, etp = 0.5 * BQD2_SQ3 - et, etm = -0.5 * BQD2_SQ3 - et;
  r[0] = sqrt(ztm * ztm + etm * etm);
  r[1] = sqrt(ztp * ztp + et * et);
  r[2] = sqrt(ztm * ztm + etp * etp);
  r[3] = r[0];
  th[0] = atan2(etm, ztm);
  th[1] = atan2(-et, ztp);
  th[2] = atan2(etp, ztm);
  th[3] = th[0] + 2.0 * M_PI;
}

// interpolation function ep of a quadrilateral (0..3) or a triangle (0..2, 3 -> 0)
__device__ __forceinline__ double shape_N(bool tri, int ep, double z, double e) {
  return tri ? lit_N(ep, z, e) : bil_N(ep, z, e);
}

// points of the n x n Gauss rule on an element: (i, j) on the square, or
// (part, i, j) on the triangle's three parts with the part's Jacobian factor
__device__ __forceinline__ int rule_npts(bool tri, int n) { return tri ? 3 * n * n : n * n; }

__device__ __forceinline__ void rule_point(bool tri, int n, const double *x, const double *w, int p, double &z,
                                           double &e, double &wt) {
  if (!tri) {
    const int i = p / n, j = p - i * n;
    z = x[i];
    e = x[j];
    wt = w[i] * w[j];
  } else {
    const int part = p / (n * n), q = p - part * n * n, i = q / n, j = q - i * n;
    const double a = x[i], b = x[j];
    lit_part(part, a, b, z, e);
    wt = (1.0 + 0.25 * a + 0.25 * b) * w[j] * w[i];
  }
}

// points of the n x n Gauss rule for a smooth integrand over the whole element:
// the square's tensor rule, or on the triangle the collapsed (Duffy) rule,
// (z, e) = V0 + s (V1 - V0) + s t (V2 - V1) with s = (1 + u)/2, t = (1 + v)/2
// over the reference corners V0 = (-1/2, -sqrt3/2), V1 = (1, 0), V2 = (-1/2,
// sqrt3/2), weight 3 (1 + u) w_u w_v: the same scale as the three parts' rule
// (its weights also sum to 12)
__device__ __forceinline__ int rem_npts(int n) { return n * n; }
__device__ __forceinline__ void rem_point(bool tri, int n, const double *x, const double *w, int p, double &z,
                                          double &e, double &wt) {
  const int i = p / n, j = p - i * n;
  if (!tri) {
    z = x[i];
    e = x[j];
    wt = w[i] * w[j];
  } else {
    const double s = 0.5 * (1.0 + x[i]), st = s * 0.5 * (1.0 + x[j]);
    z = -0.5 + 1.5 * s - 1.5 * st;
    e = 0.5 * BQD2_SQ3 * (-1.0 + s + st);
    wt = 3.0 * (1.0 + x[i]) * w[i] * w[j];
  }
}

// GL12 dF/dv0, dF/dv1 of a triangle pair with the reference's operation order
// (q0_lit_dcoef_v2_cc.c:196-229, lit_dcoef_v2.c:218-262): three parts summed per
// j, times the part weight and w_j, summed over j, times w_i, then the constant.
__device__ __forceinline__ bool needs_gh_tri(const ElemC &E, const Tab &T, const double *rt, const double *v0,
                                             const double *v1) {
  const double V[3] = {E.cw[0][0], E.cw[1][0], E.cw[2][0]};
  const double Vdv0 = V[0] * v0[0] + V[1] * v0[1] + V[2] * v0[2];
  const double Vdv1 = V[0] * v1[0] + V[1] * v1[1] + V[2] * v1[2];
  double a0 = 0.0, a1 = 0.0;
  for (int i = 0; i < 12; i++) {
    const double a = T.xli[i];
    double t0 = 0.0, t1 = 0.0;
    for (int j = 0; j < 12; j++) {
      const double b = T.xli[j];
      const double wt = 1.0 + 0.25 * a + 0.25 * b;
      double p0 = 0.0, p1 = 0.0;
      for (int part = 0; part < 3; part++) {
        double z, e, vR[3];
        lit_part(part, a, b, z, e);
#pragma unroll
        for (int l = 0; l < 3; l++) vR[l] = E.cr[l][0] + E.cr[l][1] * z + E.cr[l][2] * e - rt[l];
        const double aR = sqrt(vR[0] * vR[0] + vR[1] * vR[1] + vR[2] * vR[2]);
        const double iR = 1.0 / aR;
        const double rdV = (vR[0] * V[0] + vR[1] * V[1] + vR[2] * V[2]) * iR;
        const double rdv0 = (vR[0] * v0[0] + vR[1] * v0[1] + vR[2] * v0[2]) * iR;
        const double rdv1 = (vR[0] * v1[0] + vR[1] * v1[1] + vR[2] * v1[2]) * iR;
        p0 += iR * iR * iR * (3.0 * rdv0 * rdV - Vdv0);
        p1 += iR * iR * iR * (3.0 * rdv1 * rdV - Vdv1);
      }
      t0 += p0 * wt * T.wli[j];
      t1 += p1 * wt * T.wli[j];
    }
    a0 += t0 * T.wli[i];
    a1 += t1 * T.wli[i];
  }
  const double c = 0.0625 * BQD2_SQ3 * BQD2_I_SPSQ3;
  return fabs(a0 * c) > 0.5 || fabs(a1 * c) > 0.5;
}
