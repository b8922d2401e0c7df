// assemble.hip -- coefficient matrices of one domain (bilinear elements).
//
// The reference fills one row (target node t,tn) at a time and, inside it, one
// source element s at a time (bem3_aw_qd2_solve_bieq.c:157-199), calling the
// Green's function once per quadrature point and shape function.  Here:
//   k_decide      one lane per (row, s): the GL12 dF/dv test that picks GL12 or
//                 GH64 (d3qd2_elem.c:536-537, :851-852), bit-exact;
//   k_q0_pairs    eight (row, s) pairs of the periodic domain per wavefront,
//                 16 lanes per pair, pairs grouped by rule on the host: the
//                 chosen Gauss-Legendre rule spread over the pair's lanes, one
//                 Green's function evaluation per point shared by the four
//                 shape functions, cross-lane reduction
//                 (q0_bil_dcoef_v2_cc_GL/GH, q0_bil_dcoef_v2_cc.c:129-281);
//   k_fs_pairs    the same with the free-space kernel and tangential
//                 derivatives (bil_dcoef_v2_GL/GH, bil_dcoef_v2.c:141-325);
//   k_q0_self,    one wavefront per row for the singular self element: polar
//   k_fs_self     rules about the node (q0_bil_coef.c:246-338, bil_coef.c:244-342)
//                 and the principal-value derivatives (bil_dcoef.c:369-740);
//   k_rows        per-row sum of F in source order and the 1+F / F diagonal.
#include <algorithm>
#include <cstring>
#include <vector>
#define BQD2_ERFCX_INLINE 1
#include "elem_dev.h"

#define WAVE 64

// row -> (t, tn); returns false for the unused fourth slot of a triangle
__device__ __forceinline__ bool row_target(const DevMesh &M, const DevDomain &D, int row, int &t, int &tn, int &td) {
  t = row / 4 + 1;
  tn = row % 4;
  td = D.sid[t];
  return !(M.tri[abs(td)] && tn == 3);
}

// A GH64 pair of the periodic domain may be split: the free-space kernel
// e^{ikR}/(4 pi R) (the image l = 0 of the lattice sum) carries the whole
// near-singular behaviour and is integrated with GH64, while the remainder
// G - G0 (the other images and the Ewald corrections) is analytic over the
// element and its GL12 integral equals the GH64 one to rounding.  This needs
// every other image of the target far from the element, and no quadrature point
// with a z offset of exactly zero: there the reference's spectral sum stops on
// rounding residue (see gf_dev.h) and its value depends on the point set, so
// such pairs keep the full GH64 evaluation of the lattice sum.
// offset of the point (z, e) of the element from the target, rounded as point_geo does
__device__ __forceinline__ bool offset_degenerate(const ElemC &E, double z, double e, const double *rt) {
  double r[3];
  bil_rw_r(E, z, e, r);
  return z_degenerate(r[0] - rt[0], r[1] - rt[1], r[2] - rt[2]);
}

__device__ __forceinline__ bool split_ok(const ElemC &E, const Tab &T, const double *rt, const GfParams &P,
                                         bool tri) {
  if (!images_far(E, rt, P)) return false;
  double z, e, wt;
  for (int p = 0; p < rule_npts(tri, 64); p++) {
    rule_point(tri, 64, T.xhi, T.whi, p, z, e, wt);
    if (offset_degenerate(E, z, e, rt)) return false;
  }
  for (int p = 0; p < rule_npts(tri, 12); p++) {
    rule_point(tri, 12, T.xli, T.wli, p, z, e, wt);
    if (offset_degenerate(E, z, e, rt)) return false;
  }
  return true;
}

// Rules of a pair (D.rule):
//   0  the lattice sum on GL12 (the reference's rule)
//   1  the lattice sum on GH64 (the reference's rule for near pairs)
//   2  free-space kernel on GH64, the remainder on GL12
//   3, 4, 5  free-space kernel on GH64, the remainder on 4, 6, 8 points a side
//   6, 7, 8  free-space kernel on GL12, the remainder on 4, 6, 8 points a side
// (2-8 in the periodic domain only).  The remainder G - G0 of the lattice sum is
// analytic over the element up to the other lattice images of the target, so a
// Gauss rule of n points a side integrates it with an error that falls as
// t^-(2n+2), t the distance of the nearest such image in element half-widths.
__host__ __device__ __forceinline__ int rule_free_n(int r) { return r >= 6 ? 12 : (r >= 2 ? 64 : 0); }
__host__ __device__ __forceinline__ int rule_lat_n(int r) {
  if (r == 1) return 64;
  if (r <= 2) return 12;
  return 4 + 2 * ((r - 3) % 3);
}
__device__ __forceinline__ void rule_tab(const Tab &T, int n, const double *&x, const double *&w) {
  if (n == 4) { x = T.xs4; w = T.ws4; }
  else if (n == 6) { x = T.xs6; w = T.ws6; }
  else if (n == 8) { x = T.xs8; w = T.ws8; }
  else if (n == 12) { x = T.xli; w = T.wli; }
  else { x = T.xhi; w = T.whi; }
}

// Gauss points a side for the remainder of the pair (element E, target rt):
// 4, 6 or 8, the smallest whose error estimate 3 (q/6)^(2n) stays below 1e-14
// of the row, q = h max(k, 1/E0, 1/(d - r)) the element's half-width h = r/sqrt 2
// against the remainder's shortest length (wavelength / 2 pi, Ewald parameter,
// distance to the nearest lattice image of the target other than l = 0; r the
// element's radius, d that image's distance from the element's centre); 0 when
// none does.  The estimate is fitted to the measured convergence of the three
// rules to GL12 on the 96- and 384-quad spheres of the 2 mm lattice (q 0.41 and
// 0.21: 5e-10 and 5e-12 with 4 points, 1e-15 with 6 and 8).
__device__ __forceinline__ int smooth_order(const ElemC &E, const double *rt, const GfParams &P, bool tri) {
  double c[3], rad = 0.0;
  bil_rw_r(E, 0.0, 0.0, c);
#pragma unroll
  for (int k = 0; k < 4; k++) {
    double q[3];
    bil_rw_r(E, (k == 1 || k == 2) ? 1.0 : -1.0, (k >= 2) ? 1.0 : -1.0, q);
    rad = fmax(rad, sqrt((q[0] - c[0]) * (q[0] - c[0]) + (q[1] - c[1]) * (q[1] - c[1]) + (q[2] - c[2]) * (q[2] - c[2])));
  }
  const double ox = c[0] - rt[0], oy = c[1] - rt[1], oz = c[2] - rt[2];
  // the lattice point nearest the offset, then the images around it
  const int k1 = (int)floor((ox * P.d2y - oy * P.d2x) * P.idd + 0.5);
  const int k2 = (int)floor(-(ox * P.d1y - oy * P.d1x) * P.idd + 0.5);
  double dmin = 1e300;
  for (int m1 = -2; m1 <= 2; m1++) {
    for (int m2 = -2; m2 <= 2; m2++) {
      const int l1 = m1 - k1, l2 = m2 - k2;
      if (l1 == 0 && l2 == 0) continue;
      const double dx = ox + l1 * P.d1x + l2 * P.d2x, dy = oy + l1 * P.d1y + l2 * P.d2y;
      dmin = fmin(dmin, sqrt(dx * dx + dy * dy + oz * oz));
    }
  }
  if (!(dmin > 2.0 * rad) || rad <= 0.0) return 0;
  const double q = 0.7071067811865476 * rad * fmax(fmax(P.k, 1.0 / P.E0), 1.0 / (dmin - rad));
  // the triangle's collapsed rule is exact to one degree less than the
  // square's: its thresholds are those of the next order down
  if (q <= (tri ? 0.05 : 0.088)) return 4;
  if (q <= (tri ? 0.3 : 0.36)) return 6;
  if (q <= (tri ? 0.6 : 0.73)) return 8;
  return 0;
}

// a point of rule n on the element with a z offset level with the target
__device__ __forceinline__ bool any_degenerate(const ElemC &E, const Tab &T, const double *rt, bool tri, int n,
                                               bool rem = false) {
  const double *x, *w;
  rule_tab(T, n, x, w);
  double z, e, wt;
  const int np = rem ? rem_npts(n) : rule_npts(tri, n);
  for (int p = 0; p < np; p++) {
    if (rem)
      rem_point(tri, n, x, w, p, z, e, wt);
    else
      rule_point(tri, n, x, w, p, z, e, wt);
    if (offset_degenerate(E, z, e, rt)) return true;
  }
  return false;
}

// rule per pair (above); smooth = 0 keeps the remainder on GL12 (rules 0-2 only),
// 1 takes smooth_order's choice, 4, 6 or 8 that order wherever smooth_order
// allows a remainder rule at all
__global__ void __launch_bounds__(256) k_decide(DevMesh M, const Tab *__restrict__ tab, DevDomain D, GfParams P,
                                                int periodic, int smooth) {
  long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (long long)D.N * D.Ne) return;
  int row = (int)(gid / D.Ne), s = (int)(gid % D.Ne) + 1;
  int t, tn, td;
  unsigned char rule = 0;
  float key = 0.0f;
  if (row_target(M, D, row, t, tn, td) && s != t) {
    const Tab &T = *tab;
    int at = abs(td), as = abs(D.sid[s]);
    double rt[3] = {M.ren[12 * at + 3 * tn], M.ren[12 * at + 3 * tn + 1], M.ren[12 * at + 3 * tn + 2]};
    double vtz[3], vte[3];
    node_tangents(M, T, at, tn, vtz, vte);
    ElemC E;
    load_elem(M, as, E);
    const bool tri = M.tri[as] != 0;
    rule = (tri ? needs_gh_tri(E, T, rt, vtz, vte) : needs_gh(E, T, rt, vtz, vte)) ? 1 : 0;
    if (rule && periodic && split_ok(E, T, rt, P, tri)) rule = 2;
    if (periodic && smooth && rule != 1) {
      int n = smooth_order(E, rt, P, tri);
      if (n && smooth > 1) n = smooth;
      if (n && !any_degenerate(E, T, rt, tri, n, true) && (rule == 2 || !any_degenerate(E, T, rt, tri, 12)))
        rule = (unsigned char)((rule == 2 ? 3 : 6) + (n - 4) / 2);
    }
    key = (float)fabs(E.cr[2][0] - rt[2]);
    if (periodic) {
      // the low six mantissa bits of the key carry the in-plane offset of the
      // element centre reduced to the unit cell, in 64 levels up to the cell's
      // half diagonal (a relative change of 2^-17 to the z offset)
      const double dx = E.cr[0][0] - rt[0], dy = E.cr[1][0] - rt[1];
      const double l1 = floor((dx * P.d2y - dy * P.d2x) * P.idd + 0.5);
      const double l2 = floor(-(dx * P.d1y - dy * P.d1x) * P.idd + 0.5);
      const double tx = dx - l1 * P.d1x - l2 * P.d2x, ty = dy - l1 * P.d1y - l2 * P.d2y;
      const double sx = P.d1x + P.d2x, sy = P.d1y + P.d2y, mx = P.d1x - P.d2x, my = P.d1y - P.d2y;
      const double rmax = 0.5 * sqrt(fmax(sx * sx + sy * sy, mx * mx + my * my));
      const unsigned rq = (unsigned)fmin(63.0, sqrt(tx * tx + ty * ty) / rmax * 64.0);
      key = __uint_as_float((__float_as_uint(key) & ~63u) | rq);
    }
  }
  D.rule[gid] = rule;
  if (D.pkey) D.pkey[gid] = key;
}

// ------------------------------------------------------------------ periodic pairs
// Eight (row, s) pairs per wavefront, 8 lanes per pair.  The host sorts the
// pairs by rule and element kind, so the pairs of one wavefront share their
// loop bounds; a pair's rule points are strided over its lanes (two lattice
// evaluations per lane for a 4-point remainder rule, 18 free-space points for
// GL12).  Measured against 4 pairs of 16 lanes and 16 pairs of 4 lanes on
// sphere3456: 1027 ms against 1042 and 1033.
#ifndef PPW
#define PPW 8
#endif
#define LPP (WAVE / PPW)
// the free-space pair kernel's layout: four elements of a row, 16 lanes each
#define FS_PPW 4
#define FS_LPP 16
// single-wave workgroups per launch at most: the dispatch packet holds the grid
// size in work-items as a 32-bit number (2^25 x 64 = 2^31)
#define BQD2_MAX_WG (1LL << 25)
template <int L = LPP>
__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
  for (int off = L / 2; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// the free-space kernel e^{ikR}/(4 pi R) of a split pair on its NF x NF rule
// (x, w), this lane's points of the pair's 16 lanes, added to g, h and the
// solid-angle terms
template <int NF>
__device__ __forceinline__ void free_part(const ElemC &E, const Tab &T, const double *x, const double *w,
                                          const double *rt, const double *vtz, const double *vte, double k, bool tri,
                                          bool on, int sl, cd *g, cd *h, double &F, double &F0, double &F1) {
  const int npf = on ? rule_npts(tri, NF) : 0;
  for (int p = sl; p < npf; p += LPP) {
    double z, e, wt;
    rule_point(tri, NF, x, w, p, z, e, wt);
    PtGeo q;
    point_geo(E, z, e, rt, vtz, vte, q);
    cd g0, dg0[3];
    free_gf(k, q.vR, q.aR, q.iR, g0, dg0);
    cd hn = dg0[0] * q.vW[0] + dg0[1] * q.vW[1] + dg0[2] * q.vW[2];
#pragma unroll
    for (int ep = 0; ep < 4; ep++) {
      double Np = shape_N(tri, ep, z, e);
      g[ep] += g0 * (wt * Np * q.aW);
      h[ep] += hn * (wt * Np);
    }
    F += wt * q.rdW * q.iR * q.iR;
    F0 += wt * q.iR * q.iR * q.iR * (3.0 * q.rdv0 * q.rdW - q.Wdv0);
    F1 += wt * q.iR * q.iR * q.iR * (3.0 * q.rdv1 * q.rdW - q.Wdv1);
  }
}

// The pairs PPW wg .. PPW wg + PPW - 1 of the list (one wavefront); wg is the
// workgroup of a launch over one domain, or a task of the batched launch.
__device__ __forceinline__ void q0_pairs_wave(const DevMesh &M, const Tab *__restrict__ tab, const DevDomain &D,
                                              const GfParams &P, const int *__restrict__ list, long long nlist,
                                              long long *__restrict__ counters, long long wg) {
  const int lane = threadIdx.x, sl = lane % LPP;
  const long long li = wg * PPW + lane / LPP;
  // padding entries (the end of a class) take the wavefront's first pair and
  // add nothing: their loops are empty and they write nothing
  const int pair = (li < nlist) ? list[li] : -1;
  const bool on = pair >= 0;
  const int gid = on ? pair : list[wg * PPW];
  const int row = gid / D.Ne, s = gid % D.Ne + 1;
  int t, tn, td;
  row_target(M, D, row, t, tn, td);
  const Tab &T = *tab;
  const int at = abs(td), as = D.sid[s];  // periodic domain: source ids are positive
  const int rule = D.rule[gid];
  double rt[3] = {M.ren[12 * at + 3 * tn], M.ren[12 * at + 3 * tn + 1], M.ren[12 * at + 3 * tn + 2]};
  double vtz[3], vte[3];
  node_tangents(M, T, at, tn, vtz, vte);
  ElemC E;
  load_elem(M, as, E);
  const bool tri = M.tri[as] != 0;

  cd g[4] = {}, h[4] = {};
  double F = 0.0, F0 = 0.0, F1 = 0.0;
  GfTally tl{0, 0, 0, 0};
  int err = 0;
  long long nev = 0;
  const int nfree = rule_free_n(rule);
  // the image l = 0 with the free-space kernel, on GH64 (near pairs) or GL12
  if (nfree == 64)
    free_part<64>(E, T, T.xhi, T.whi, rt, vtz, vte, P.k, tri, on, sl, g, h, F, F0, F1);
  else if (nfree == 12)
    free_part<12>(E, T, T.xli, T.wli, rt, vtz, vte, P.k, tri, on, sl, g, h, F, F0, F1);
  {
    // lattice sum on the pair's rule (the remainder's rule for a split pair).
    // Only the offset and the accumulators stay in registers across the lattice
    // sum: the pair's element, target node and tangents sit in LDS per pair, and
    // the point's weights for the accumulation are formed before the sum and
    // parked in LDS per lane (the empty asm statements keep the compiler from
    // holding LDS values in registers across the sum instead).
    __shared__ double sgeo[PPW][30];  // cr 12, cw 9, rt 3, vtz 3, vte 3
    __shared__ double spt[11][WAVE];  // per lane: g weights 4, h weights 4, W 3
    __shared__ double sacc[16][WAVE]; // per lane: g (re, im) 4, h (re, im) 4
    const int grp = lane / LPP;
#pragma unroll
    for (int ep = 0; ep < 4; ep++) {
      sacc[4 * ep][lane] = g[ep].x; sacc[4 * ep + 1][lane] = g[ep].y;
      sacc[4 * ep + 2][lane] = h[ep].x; sacc[4 * ep + 3][lane] = h[ep].y;
    }
    if (sl == 0) {
      double *G = sgeo[grp];
#pragma unroll
      for (int d = 0; d < 3; d++) {
#pragma unroll
        for (int j = 0; j < 4; j++) G[4 * d + j] = E.cr[d][j];
#pragma unroll
        for (int j = 0; j < 3; j++) G[12 + 3 * d + j] = E.cw[d][j];
        G[21 + d] = rt[d];
        G[24 + d] = vtz[d];
        G[27 + d] = vte[d];
      }
    }
    __syncthreads();
    const int n = rule_lat_n(rule);
    const double *x, *w;
    rule_tab(T, n, x, w);
    // rules 3-8: the remainder on the whole-element rule; 0-2 the reference's
    const bool rem = rule >= 3;
    const int npts = on ? (rem ? rem_npts(n) : rule_npts(tri, n)) : 0;
    nev = (sl == 0) ? npts : 0;
    for (int p = sl; p < npts; p += LPP) {
      asm volatile("" ::: "memory");
      const double *G = sgeo[grp];
      ElemC Ep;
#pragma unroll
      for (int d = 0; d < 3; d++) {
#pragma unroll
        for (int j = 0; j < 4; j++) Ep.cr[d][j] = G[4 * d + j];
#pragma unroll
        for (int j = 0; j < 3; j++) Ep.cw[d][j] = G[12 + 3 * d + j];
      }
      double z, e, wt;
      if (rem)
        rem_point(tri, n, x, w, p, z, e, wt);
      else
        rule_point(tri, n, x, w, p, z, e, wt);
      PtGeo q;
      point_geo(Ep, z, e, G + 21, G + 24, G + 27, q);
#pragma unroll
      for (int ep = 0; ep < 4; ep++) {
        const double Np = shape_N(tri, ep, z, e);
        spt[ep][lane] = wt * Np * q.aW;
        spt[4 + ep][lane] = wt * Np;
      }
#pragma unroll
      for (int d = 0; d < 3; d++) spt[8 + d][lane] = q.vW[d];
      if (!nfree) {
        F += wt * q.rdW * q.iR * q.iR;
        F0 += wt * q.iR * q.iR * q.iR * (3.0 * q.rdv0 * q.rdW - q.Wdv0);
        F1 += wt * q.iR * q.iR * q.iR * (3.0 * q.rdv1 * q.rdW - q.Wdv1);
      }
      const double vR[3] = {q.vR[0], q.vR[1], q.vR[2]};
      asm volatile("" ::: "memory");
      Sum10 gf;
      int c = qpgf_eval<GF_V1>(P, vR[0], vR[1], vR[2], 1.0e-15, gf, tl);
      if (c < 0) err = c;
      asm volatile("" ::: "memory");
      if (nfree) {
        cd g0, dg0[3];
        const double aR = abs3(vR);
        free_gf(P.k, vR, aR, 1.0 / aR, g0, dg0);
        gf.g -= g0;
#pragma unroll
        for (int d = 0; d < 3; d++) gf.dg[d] -= dg0[d];
      }
      cd hn = gf.dg[0] * spt[8][lane] + gf.dg[1] * spt[9][lane] + gf.dg[2] * spt[10][lane];
#pragma unroll
      for (int ep = 0; ep < 4; ep++) {
        const double cg = spt[ep][lane], ch = spt[4 + ep][lane];
        sacc[4 * ep][lane] += gf.g.x * cg; sacc[4 * ep + 1][lane] += gf.g.y * cg;
        sacc[4 * ep + 2][lane] += hn.x * ch; sacc[4 * ep + 3][lane] += hn.y * ch;
      }
    }
    asm volatile("" ::: "memory");
#pragma unroll
    for (int ep = 0; ep < 4; ep++) {
      g[ep] = cd{sacc[4 * ep][lane], sacc[4 * ep + 1][lane]};
      h[ep] = cd{sacc[4 * ep + 2][lane], sacc[4 * ep + 3][lane]};
    }
  }
#pragma unroll
  for (int ep = 0; ep < 4; ep++) {
    g[ep].x = group_sum(g[ep].x); g[ep].y = group_sum(g[ep].y);
    h[ep].x = group_sum(h[ep].x); h[ep].y = group_sum(h[ep].y);
  }
  F = group_sum(F); F0 = group_sum(F0); F1 = group_sum(F1);
  long long tsp = wave_sum_ll(tl.sp), tpr = wave_sum_ll(tl.pr), tev = wave_sum_ll(tl.ev), tfd = wave_sum_ll(tl.fd);
  nev = wave_sum_ll(nev);
  int anyerr = __any(err != 0);
  if (on && sl == 0) {
    // the triangle rules carry 1/24 of the parts' area (q0_lit_dcoef_v2_cc.c:231-237)
    const double sc = tri ? 1.0 / 24.0 : 1.0;
    const size_t base = dom_gh(D, row, (size_t)(s - 1) * 4) * 2;
#pragma unroll
    for (int ep = 0; ep < 4; ep++) {
      D.G[base + 2 * ep] = g[ep].x * sc; D.G[base + 2 * ep + 1] = g[ep].y * sc;
      D.H[base + 2 * ep] = h[ep].x * sc; D.H[base + 2 * ep + 1] = h[ep].y * sc;
    }
    double *fp = D.Fpair + dom_fp(D, row, s);
    fp[0] = F * (BQD2_IFP * sc); fp[1] = F0 * (BQD2_IFP * sc); fp[2] = F1 * (BQD2_IFP * sc);
  }
  if (lane == 0) {
    tally_add(counters + CNT_PAIRS, tsp, tpr, tev, tfd, nev);
    if (anyerr) atomicAdd((unsigned long long *)&counters[CNT_ERR], 1ull);
  }
}

__global__ void __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(2, 2))) k_q0_pairs(DevMesh M, const Tab *__restrict__ tab, DevDomain D, GfParams P,
                                                   const int *__restrict__ list, long long nlist,
                                                   long long *__restrict__ counters, long long wg0) {
  q0_pairs_wave(M, tab, D, P, list, nlist, counters, wg0 + blockIdx.x);
}

// ------------------------------------------------------------------ free-space pairs
// Four consecutive source elements s0 .. s0 + 3 of one row per wavefront, 16
// lanes per pair (the pair's rule points strided over its lanes, sums over the
// 16 lanes); a GH64 pair keeps its group's lanes longer than a GL12 one.
// One pair (row, s) per group of FS_LPP lanes (the pair's rule points strided
// over its lanes, sums over them); a group with on = false adds nothing.  The
// arithmetic of a pair does not depend on the other groups of its wavefront.
__device__ __forceinline__ void fs_pair_group(const DevMesh &M, const Tab *__restrict__ tab, const DevDomain &D,
                                              int row, int s, bool on) {
  const int lane = threadIdx.x, sl = lane % FS_LPP;
  int t, tn, td;
  on = row_target(M, D, row, t, tn, td) && on && s != t;
  const int sv = (s >= 1 && s <= D.Ne) ? s : D.Ne;  // a valid position for an idle group's loads
  const Tab &T = *tab;
  const int at = abs(td), sd = D.sid[sv], as = abs(sd);
  const bool gh = D.rule[(long long)row * D.Ne + (sv - 1)] != 0;
  const int n = gh ? 64 : 12;
  const double *x = gh ? T.xhi : T.xli;
  const double *w = gh ? T.whi : T.wli;
  const double k = D.k;
  double rt[3] = {M.ren[12 * at + 3 * tn], M.ren[12 * at + 3 * tn + 1], M.ren[12 * at + 3 * tn + 2]};
  double vtz[3], vte[3];
  node_tangents(M, T, at, tn, vtz, vte);
  ElemC E;
  load_elem(M, as, E);
  const bool tri = M.tri[as] != 0;

  cd g[4] = {}, h[4] = {}, dg0[4] = {}, dg1[4] = {}, dh0[4] = {}, dh1[4] = {};
  double F = 0.0, F0 = 0.0, F1 = 0.0;
  const int npts = on ? rule_npts(tri, n) : 0;
  for (int p = sl; p < npts; p += FS_LPP) {
    double z, e, wt;
    rule_point(tri, n, x, w, p, z, e, wt);
    PtGeo q;
    point_geo(E, z, e, rt, vtz, vte, q);
    cd ce = cis(k * q.aR);
    cd cb = cd{-1.0, k * q.aR};            // i k R - 1
    double tD = q.rdW * q.iR * q.iR;
    double kk = k * k * q.aR * q.aR;
    cd cg = ce * (q.iR * q.aW);                          // e^{ikR}/R |W|
    cd chh = cb * ce * tD;                                // (ikR-1) e^{ikR} R.W/R^3
    cd cdg = cb * ce * (q.iR * q.iR * q.aW);              // (ikR-1) e^{ikR}/R^2 |W|
    cd c3 = cd{kk + 3.0 * cb.x, 3.0 * cb.y};               // k^2R^2 + 3(ikR-1)
    cd cdh0 = ce * (q.iR * q.iR * q.iR) * (c3 * (q.rdv0 * q.rdW) - cb * q.Wdv0);
    cd cdh1 = ce * (q.iR * q.iR * q.iR) * (c3 * (q.rdv1 * q.rdW) - cb * q.Wdv1);
#pragma unroll
    for (int ep = 0; ep < 4; ep++) {
      double Nw = wt * shape_N(tri, ep, z, e);
      g[ep] += cg * Nw;
      h[ep] += chh * Nw;
      dg0[ep] += cdg * (Nw * q.rdv0);
      dg1[ep] += cdg * (Nw * q.rdv1);
      dh0[ep] += cdh0 * Nw;
      dh1[ep] += cdh1 * Nw;
    }
    F += wt * tD;
    F0 += wt * q.iR * q.iR * q.iR * (3.0 * q.rdv0 * q.rdW - q.Wdv0);
    F1 += wt * q.iR * q.iR * q.iR * (3.0 * q.rdv1 * q.rdW - q.Wdv1);
  }
  // scaling (bil_dcoef_v2.c:207-218; triangles carry 1/24 of the parts' area,
  // lit_dcoef_v2.c:248-262) and the sign of the sub side (:220-224)
  const double sg = sd < 0 ? -1.0 : 1.0;
  const double ifp = BQD2_IFP * (tri ? 1.0 / 24.0 : 1.0);
#pragma unroll
  for (int ep = 0; ep < 4; ep++) {
    g[ep] = cd{group_sum<FS_LPP>(g[ep].x), group_sum<FS_LPP>(g[ep].y)} * ifp;
    h[ep] = cd{group_sum<FS_LPP>(h[ep].x), group_sum<FS_LPP>(h[ep].y)} * (ifp * sg);
    dg0[ep] = cd{group_sum<FS_LPP>(dg0[ep].x), group_sum<FS_LPP>(dg0[ep].y)} * (-ifp);
    dg1[ep] = cd{group_sum<FS_LPP>(dg1[ep].x), group_sum<FS_LPP>(dg1[ep].y)} * (-ifp);
    dh0[ep] = cd{group_sum<FS_LPP>(dh0[ep].x), group_sum<FS_LPP>(dh0[ep].y)} * (ifp * sg);
    dh1[ep] = cd{group_sum<FS_LPP>(dh1[ep].x), group_sum<FS_LPP>(dh1[ep].y)} * (ifp * sg);
  }
  F = group_sum<FS_LPP>(F) * ifp * sg;
  F0 = group_sum<FS_LPP>(F0) * ifp * sg;
  F1 = group_sum<FS_LPP>(F1) * ifp * sg;
  if (on && sl == 0) {
    const size_t col = (size_t)(s - 1) * 4;
    const size_t base = dom_gh(D, row, col) * 2;
    const size_t b0 = dom_dgh(D, row, 0, col) * 2, b1 = dom_dgh(D, row, 1, col) * 2;
#pragma unroll
    for (int ep = 0; ep < 4; ep++) {
      D.G[base + 2 * ep] = g[ep].x; D.G[base + 2 * ep + 1] = g[ep].y;
      D.H[base + 2 * ep] = h[ep].x; D.H[base + 2 * ep + 1] = h[ep].y;
    }
    if (D.dG) {
#pragma unroll
      for (int ep = 0; ep < 4; ep++) {
        D.dG[b0 + 2 * ep] = dg0[ep].x; D.dG[b0 + 2 * ep + 1] = dg0[ep].y;
        D.dG[b1 + 2 * ep] = dg1[ep].x; D.dG[b1 + 2 * ep + 1] = dg1[ep].y;
        D.dH[b0 + 2 * ep] = dh0[ep].x; D.dH[b0 + 2 * ep + 1] = dh0[ep].y;
        D.dH[b1 + 2 * ep] = dh1[ep].x; D.dH[b1 + 2 * ep + 1] = dh1[ep].y;
      }
    }
    double *fp = D.Fpair + dom_fp(D, row, s);
    fp[0] = F; fp[1] = F0; fp[2] = F1;
  }
}

// element groups of four per row
__host__ __device__ __forceinline__ int fs_groups(int ne) { return (ne + FS_PPW - 1) / FS_PPW; }

// the elements s0 .. s0 + 3 of one row
__device__ __forceinline__ void fs_pairs_body(const DevMesh &M, const Tab *__restrict__ tab, const DevDomain &D,
                                              int s0, int row) {
  int t, tn, td;
  if (!row_target(M, D, row, t, tn, td)) return;
  const int s = s0 + (int)threadIdx.x / FS_LPP;
  fs_pair_group(M, tab, D, row, s, s <= D.Ne);
}

__global__ void __launch_bounds__(WAVE) k_fs_pairs(DevMesh M, const Tab *__restrict__ tab, DevDomain D) {
  fs_pairs_body(M, tab, D, FS_PPW * (int)blockIdx.x + 1, D.r0 + (int)blockIdx.y);
}

// the pairs FS_PPW wg .. FS_PPW wg + FS_PPW - 1 of the domain's rule-sorted
// list (padding entries -1 add nothing), so that the GH64 pairs share their
// wavefronts with each other
__global__ void __launch_bounds__(WAVE) k_fs_pairs_list(DevMesh M, const Tab *__restrict__ tab, DevDomain D,
                                                        long long nlist, long long wg0) {
  const long long wg = wg0 + blockIdx.x;
  // classes are padded to multiples of PPW: a group of FS_PPW entries that
  // starts with padding holds padding only
  const int first = D.plist[wg * FS_PPW];
  if (first < 0) return;
  const long long li = wg * FS_PPW + threadIdx.x / FS_LPP;
  const int pair = li < nlist ? D.plist[li] : -1;
  const int gid = pair >= 0 ? pair : first;
  fs_pair_group(M, tab, D, gid / D.Ne, gid % D.Ne + 1, pair >= 0);
}

// ------------------------------------------------------------------ self elements
// static F of the self element at (zt, et): polar rule, wave-parallel over the
// 4 x 128 x 12 points (bil_coef.c:525-536 integrand, q0_bil_coef.c:315-337 rule)
__device__ double self_F_wave(const ElemC &E, const Tab &T, double zt, double et, int lane) {
  if (bil_is_plane(E)) return 0.0;
  double r[5], th[5];
  bil_corners(zt, et, r, th);
  const double bdC = E.cr[0][1] * E.cw[0][2] + E.cr[1][1] * E.cw[1][2] + E.cr[2][1] * E.cw[2][2];
  double acc = 0.0;
  // point index p = ((j*64 + m)*2 + h)*12 + i
  for (int p = lane; p < 4 * 64 * 2 * 12; p += WAVE) {
    int i = p % 12, q = p / 12, h = q % 2, m = (q / 2) % 64, j = q / 128;
    double t = theta_node(th, j, h, T.xhi[m]);
    double rm = polar_rm(r, th, j, t);
    double c = cos(t), sn = sin(t);
    double rr = 0.5 * rm * (T.xli[i] + 1.0), K[3];
    polar_K(E, c, sn, rr, zt, et, K);
    double aK = abs3(K);
    double wgt = 0.25 * (th[j + 1] - th[j]) * T.whi[m] * 0.5 * rm * T.wli[i];
    acc += wgt * (1.0 / (aK * aK * aK) * c * sn);
  }
  return wave_sum(acc) * BQD2_IFP * bdC;
}

__device__ __forceinline__ void q0_self_body(const DevMesh &M, const Tab *__restrict__ tab, const DevDomain &D, const GfParams &P, long long *__restrict__ counters, int row) {
  const int lane = threadIdx.x;
  int t, tn, td;
  if (!row_target(M, D, row, t, tn, td)) return;
  const Tab &T = *tab;
  const int s = abs(td);
  if (M.tri[s]) return;  // k_q0_self_tri
  ElemC E;
  load_elem(M, s, E);
  const double zt = T.zt_44[tn], et = T.et_44[tn];
  double r[5], th[5];
  bil_corners(zt, et, r, th);
  cd g[4] = {}, h[4] = {};
  GfTally tl{0, 0, 0, 0};
  int err = 0;
  const int npts = 4 * 64 * 2 * 12;
  // Same split as for near pairs: the free-space image on the polar rule, the
  // smooth remainder of the lattice sum on GL12 over the element, unless another
  // image is close or a point has a z offset of exactly zero.
  const double rt[3] = {M.ren[12 * s + 3 * tn], M.ren[12 * s + 3 * tn + 1], M.ren[12 * s + 3 * tn + 2]};
  bool zero = false;
  for (int p = lane; p < npts; p += WAVE) {
    int i = p % 12, q = p / 12, hh = q % 2, m = (q / 2) % 64, j = q / 128;
    double tt = theta_node(th, j, hh, T.xhi[m]);
    double rm = polar_rm(r, th, j, tt);
    double c = cos(tt), sn = sin(tt);
    double rr = 0.5 * rm * (T.xli[i] + 1.0);
    double v[3];
#pragma unroll
    for (int l = 0; l < 3; l++)
      v[l] = rr * (E.cr[l][1] * c + E.cr[l][2] * sn + E.cr[l][3] * (rr * c * sn + zt * sn + et * c));
    zero |= z_degenerate(v[0], v[1], v[2]);
  }
  for (int p = lane; p < 144; p += WAVE) zero |= offset_degenerate(E, T.xli[p / 12], T.xli[p % 12], rt);
  const bool split = images_far(E, rt, P) && !__any(zero);
  long long evals = split ? 144 : npts;
  for (int p = lane; p < npts; p += WAVE) {
    int i = p % 12, q = p / 12, hh = q % 2, m = (q / 2) % 64, j = q / 128;
    double tt = theta_node(th, j, hh, T.xhi[m]);
    double rm = polar_rm(r, th, j, tt);
    double c = cos(tt), sn = sin(tt);
    double rr = 0.5 * rm * (T.xli[i] + 1.0);
    double z = rr * c + zt, e = rr * sn + et, vR[3], vW[3];
#pragma unroll
    for (int l = 0; l < 3; l++) {
      vR[l] = rr * (E.cr[l][1] * c + E.cr[l][2] * sn + E.cr[l][3] * (rr * c * sn + zt * sn + et * c));
      vW[l] = E.cw[l][0] + E.cw[l][1] * z + E.cw[l][2] * e;
    }
    double aW = abs3(vW);
    Sum10 gf;
    if (split) {
      const double aR = abs3(vR);
      free_gf(P.k, vR, aR, 1.0 / aR, gf.g, gf.dg);
    } else {
      int cnt = qpgf_eval_g_dg(P, vR[0], vR[1], vR[2], 1.0e-15, gf, tl);
      if (cnt < 0) err = cnt;
    }
    cd hn = gf.dg[0] * vW[0] + gf.dg[1] * vW[1] + gf.dg[2] * vW[2];
    double wgt = 0.25 * (th[j + 1] - th[j]) * T.whi[m] * 0.5 * rm * T.wli[i] * rr;
#pragma unroll
    for (int n = 0; n < 4; n++) {
      double Nn = bil_N(n, z, e) * wgt;
      g[n] += gf.g * (Nn * aW);
      h[n] += hn * Nn;
    }
  }
  if (split) {
    for (int p = lane; p < 144; p += WAVE) {
      const int i = p / 12, j = p - i * 12;
      const double z = T.xli[i], e = T.xli[j], wt = T.wli[i] * T.wli[j];
      double vR[3], vW[3];
      bil_rw(E, z, e, vR, vW);
#pragma unroll
      for (int l = 0; l < 3; l++) vR[l] -= rt[l];
      const double aR = abs3(vR), aW = abs3(vW);
      Sum10 gf;
      int cnt = qpgf_eval<GF_V1>(P, vR[0], vR[1], vR[2], 1.0e-15, gf, tl);
      if (cnt < 0) err = cnt;
      cd g0, dg0[3];
      free_gf(P.k, vR, aR, 1.0 / aR, g0, dg0);
      const cd gs = gf.g - g0;
      const cd hn = (gf.dg[0] - dg0[0]) * vW[0] + (gf.dg[1] - dg0[1]) * vW[1] + (gf.dg[2] - dg0[2]) * vW[2];
#pragma unroll
      for (int n = 0; n < 4; n++) {
        const double Nn = bil_N(n, z, e) * wt;
        g[n] += gs * (Nn * aW);
        h[n] += hn * Nn;
      }
    }
  }
#pragma unroll
  for (int n = 0; n < 4; n++) {
    g[n] = cd{wave_sum(g[n].x), wave_sum(g[n].y)};
    h[n] = cd{wave_sum(h[n].x), wave_sum(h[n].y)};
  }
  long long tsp = wave_sum_ll(tl.sp), tpr = wave_sum_ll(tl.pr), tev = wave_sum_ll(tl.ev), tfd = wave_sum_ll(tl.fd);
  int anyerr = __any(err != 0);
  const double F = self_F_wave(E, T, zt, et, lane);
  // central difference of F along the element parameters (q0_bil_dcoef_v2_cc.c:336-390)
  double Tz[3], Te[3];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    Tz[i] = E.cr[i][1] + E.cr[i][3] * et;
    Te[i] = E.cr[i][2] + E.cr[i][3] * zt;
  }
  const double h_ = 1.0e-3;
  double Fzp = self_F_wave(E, T, zt + h_, et, lane), Fzm = self_F_wave(E, T, zt - h_, et, lane);
  double Fep = self_F_wave(E, T, zt, et + h_, lane), Fem = self_F_wave(E, T, zt, et - h_, lane);
  if (lane == 0) {
    const size_t base = dom_gh(D, row, (size_t)(t - 1) * 4) * 2;
#pragma unroll
    for (int n = 0; n < 4; n++) {
      D.G[base + 2 * n] = g[n].x; D.G[base + 2 * n + 1] = g[n].y;
      D.H[base + 2 * n] = h[n].x; D.H[base + 2 * n + 1] = h[n].y;
    }
    double *fp = D.Fpair + dom_fp(D, row, t);
    fp[0] = F;
    fp[1] = 1.0 / abs3(Tz) * (Fzp - Fzm) / (2.0 * h_);
    fp[2] = 1.0 / abs3(Te) * (Fep - Fem) / (2.0 * h_);
    tally_add(counters + CNT_SELF, tsp, tpr, tev, tfd, evals);
    if (anyerr) atomicAdd((unsigned long long *)&counters[CNT_ERR], 1ull);
  }
}

__global__ void __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(2, 2))) k_q0_self(DevMesh M, const Tab *__restrict__ tab, DevDomain D, GfParams P,
                                                  long long *__restrict__ counters) {
  q0_self_body(M, tab, D, P, counters, D.r0 + (int)blockIdx.x);
}

// free-space self element: G, H, F (bil_coef_bd) and the principal-value
// tangential derivatives dG, dH, dF along t_zeta and t_eta (bil_dcoef.c:369-740)
__device__ __forceinline__ void fs_self_body(const DevMesh &M, const Tab *__restrict__ tab, const DevDomain &D, int row) {
  const int lane = threadIdx.x;
  int t, tn, td;
  if (!row_target(M, D, row, t, tn, td)) return;
  const Tab &T = *tab;
  const int s = abs(td);
  if (M.tri[s]) return;  // k_fs_self_tri
  const double sg = td < 0 ? -1.0 : 1.0;
  const double k = D.k;
  ElemC E;
  load_elem(M, s, E);
  const double zt = T.zt_44[tn], et = T.et_44[tn];
  double r[5], th[5];
  bil_corners(zt, et, r, th);
  const bool plane = bil_is_plane(E);
  const double bdC = E.cr[0][1] * E.cw[0][2] + E.cr[1][1] * E.cw[1][2] + E.cr[2][1] * E.cw[2][2];
  // unnormalised tangents T_zeta, T_eta at the node
  double Tv[2][3];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    Tv[0][i] = E.cr[i][1] + E.cr[i][3] * et;
    Tv[1][i] = E.cr[i][2] + E.cr[i][3] * zt;
  }

  // ---- G, H on the polar rule (12 points in r)
  cd g[4] = {}, h[4] = {};
  for (int p = lane; p < 4 * 64 * 2 * 12; p += WAVE) {
    int i = p % 12, q = p / 12, hh = q % 2, m = (q / 2) % 64, j = q / 128;
    double tt = theta_node(th, j, hh, T.xhi[m]);
    double rm = polar_rm(r, th, j, tt);
    double c = cos(tt), sn = sin(tt);
    double rr = 0.5 * rm * (T.xli[i] + 1.0);
    double z = rr * c + zt, e = rr * sn + et, K[3], W[3];
    polar_K(E, c, sn, rr, zt, et, K);
#pragma unroll
    for (int l = 0; l < 3; l++) W[l] = E.cw[l][0] + E.cw[l][1] * z + E.cw[l][2] * e;
    double aK = abs3(K), aW = abs3(W);
    cd ce = cis(k * aK * rr);
    cd cb = cd{-1.0, k * rr * aK};
    double wgt = 0.25 * (th[j + 1] - th[j]) * T.whi[m] * 0.5 * rm * T.wli[i];
    cd cg = ce * (aW / aK * wgt);
    cd chh = cb * ce * (1.0 / (aK * aK * aK) * c * sn * wgt);
#pragma unroll
    for (int n = 0; n < 4; n++) {
      double Nn = bil_N(n, z, e);
      g[n] += cg * Nn;
      h[n] += chh * Nn;
    }
  }

  // ---- principal-value derivatives, both tangent directions at once
  cd dG[2][4] = {}, dH[2][4] = {};
  double dFv[2] = {0.0, 0.0};
  // D_bar(eps): 4 sub-triangles x 128 theta x 64 r on [DEPS, rm]
  for (int p = lane; p < 4 * 64 * 2 * 64; p += WAVE) {
    int i = p % 64, q = p / 64, hh = q % 2, m = (q / 2) % 64, j = q / 128;
    double tt = theta_node(th, j, hh, T.xhi[m]);
    double rm = polar_rm(r, th, j, tt);
    double c = cos(tt), sn = sin(tt);
    double rr = 0.5 * (rm - 0.1) * T.xhi[i] + 0.5 * (rm + 0.1);
    double wgt = 0.25 * (th[j + 1] - th[j]) * T.whi[m] * 0.5 * (rm - 0.1) * T.whi[i];
    double z = rr * c + zt, e = rr * sn + et, K[3], W[3];
    polar_K(E, c, sn, rr, zt, et, K);
#pragma unroll
    for (int l = 0; l < 3; l++) W[l] = E.cw[l][0] + E.cw[l][1] * z + E.cw[l][2] * e;
    double aK = abs3(K), iK = 1.0 / aK, aW = abs3(W);
    double kR = k * rr * aK;
    cd ce = cis(kR), cb = cd{-1.0, kR};
    cd c3 = cd{kR * kR + 3.0 * cb.x, 3.0 * cb.y};
    double aK2 = aK * aK, iK5 = 1.0 / (aK2 * aK2 * aK);
#pragma unroll
    for (int dir = 0; dir < 2; dir++) {
      double KdT = dot3(K, Tv[dir]);
      double sel = dir == 0 ? c : sn;    // inner factor: cos for zeta, sin for eta
      double outer = dir == 0 ? sn : c;  // outer factor of the H and F kernels
      cd cg = cb * ce * (iK * iK / rr * (KdT * iK) * aW * wgt);
      cd chh = ce * (iK5 / rr) * (c3 * (KdT * sel) - cb * aK2) * (outer * wgt);
#pragma unroll
      for (int n = 0; n < 4; n++) {
        double Nn = bil_N(n, z, e);
        dG[dir][n] += cg * Nn;
        dH[dir][n] += chh * Nn;
      }
      dFv[dir] += outer * wgt * (iK5 / rr * (3.0 * KdT * sel - aK2));
    }
  }
  // D_eps: theta in four quarters of (-pi, pi], r in [0, DEPS] (GL12)
  for (int p = lane; p < 4 * 64 * 12; p += WAVE) {
    int i = p % 12, q = p / 12, qq = q % 4, m = q / 4;
    const double off = -0.75 + 0.5 * qq;
    double tt = 0.25 * M_PI * T.xhi[m] + off * M_PI;
    double c = cos(tt), sn = sin(tt);
    double rr = 0.5 * 0.1 * (T.xli[i] + 1.0);
    double wgt = 0.25 * M_PI * T.whi[m] * 0.5 * 0.1 * T.wli[i];
    double z = rr * c + zt, e = rr * sn + et, K[3], W[3];
    polar_K(E, c, sn, rr, zt, et, K);
#pragma unroll
    for (int l = 0; l < 3; l++) W[l] = E.cw[l][0] + E.cw[l][1] * z + E.cw[l][2] * e;
    double aK = abs3(K), iK = 1.0 / aK, aW = abs3(W);
    double kR = k * aK * rr;
    double cc, cs;
    sincos(kR, &cs, &cc);
    cd ce = cd{cc, cs};
    cd cb = cd{-1.0, kR};
    double aK2 = aK * aK;
#pragma unroll
    for (int dir = 0; dir < 2; dir++) {
      double KdT = dot3(K, Tv[dir]);
      double sel = dir == 0 ? c : sn, outer = dir == 0 ? sn : c;
      // i * [ (cb sin kR + kR cos kR) / (K^2 r) (K.T/|K|) |W| ]
      cd cg0 = cd{cb.x * cs + kR * cc, cb.y * cs} * (iK * iK / rr * (KdT * iK) * aW * wgt);
      cd cg = mul_i(cg0);
      // De0 - i De1
      cd cbk = cd{0.0, k};
      cd h0 = ce * (1.0 / (aK2 * aK2)) * (cd{k * kR, 0.0} * (KdT * sel) + cbk * (3.0 * KdT * sel) - cbk * aK2);
      double h1 = cs / (aK2 * aK2 * aK * rr) * (3.0 * KdT * sel - aK2);
      cd chh = (h0 - mul_i(cd{h1, 0.0})) * (outer * wgt);
#pragma unroll
      for (int n = 0; n < 4; n++) {
        double Nn = bil_N(n, z, e);
        dG[dir][n] += cg * Nn;
        dH[dir][n] += chh * Nn;
      }
    }
  }
  // D_eps principal-value pairs (+r, -r): theta in (0, pi], r in [0, DEPS]
  for (int p = lane; p < 2 * 64 * 12; p += WAVE) {
    int i = p % 12, q = p / 12, qq = q % 2, m = q / 2;
    const double off = qq == 0 ? 0.25 : 0.75;
    double tt = 0.25 * M_PI * T.xhi[m] + off * M_PI;
    double c = cos(tt), sn = sin(tt);
    double rr = 0.5 * 0.1 * (T.xli[i] + 1.0);
    double wgt = 0.25 * M_PI * T.whi[m] * 0.5 * 0.1 * T.wli[i];
    double zp = rr * c + zt, ep = rr * sn + et, zm = -rr * c + zt, em = -rr * sn + et;
    double Kp[3], Km[3], Wp[3], Wm[3];
#pragma unroll
    for (int l = 0; l < 3; l++) {
      Kp[l] = E.cr[l][1] * c + E.cr[l][2] * sn + E.cr[l][3] * (rr * c * sn + zt * sn + et * c);
      Km[l] = E.cr[l][1] * c + E.cr[l][2] * sn + E.cr[l][3] * (-rr * c * sn + zt * sn + et * c);
      Wp[l] = E.cw[l][0] + E.cw[l][1] * zp + E.cw[l][2] * ep;
      Wm[l] = E.cw[l][0] + E.cw[l][1] * zm + E.cw[l][2] * em;
    }
    double aKp = abs3(Kp), aKm = abs3(Km), iKp = 1.0 / aKp, iKm = 1.0 / aKm, aWp = abs3(Wp), aWm = abs3(Wm);
    double ccp = cos(k * aKp * rr), ccm = cos(-k * aKm * rr);
    double iKp5 = 1.0 / (aKp * aKp * aKp * aKp * aKp), iKm5 = 1.0 / (aKm * aKm * aKm * aKm * aKm);
#pragma unroll
    for (int dir = 0; dir < 2; dir++) {
      double KpT = dot3(Kp, Tv[dir]), KmT = dot3(Km, Tv[dir]);
      double sel = dir == 0 ? c : sn, outer = dir == 0 ? sn : c;
#pragma unroll
      for (int n = 0; n < 4; n++) {
        double Np = bil_N(n, zp, ep), Nm = bil_N(n, zm, em);
        double vg = 1.0 / rr * (ccp * iKp * iKp * (KpT * iKp) * Np * aWp - ccm * iKm * iKm * (KmT * iKm) * Nm * aWm);
        double vh = 1.0 / rr * (ccp * iKp5 * (3.0 * KpT * sel - aKp * aKp) * Np - ccm * iKm5 * (3.0 * KmT * sel - aKm * aKm) * Nm);
        dG[dir][n] -= cd{vg * wgt, 0.0};
        dH[dir][n] -= cd{vh * outer * wgt, 0.0};
      }
      dFv[dir] += outer * wgt * (1.0 / rr * (iKp5 * (3.0 * KpT * sel - aKp * aKp) - iKm5 * (3.0 * KmT * sel - aKm * aKm)));
    }
  }

  const double F = self_F_wave(E, T, zt, et, lane);
#pragma unroll
  for (int n = 0; n < 4; n++) {
    g[n] = cd{wave_sum(g[n].x), wave_sum(g[n].y)} * BQD2_IFP;
    h[n] = plane ? cd{0.0, 0.0} : cd{wave_sum(h[n].x), wave_sum(h[n].y)} * (BQD2_IFP * bdC * sg);
  }
  double iT[2] = {1.0 / abs3(Tv[0]), 1.0 / abs3(Tv[1])};
#pragma unroll
  for (int dir = 0; dir < 2; dir++) {
#pragma unroll
    for (int n = 0; n < 4; n++) {
      dG[dir][n] = cd{wave_sum(dG[dir][n].x), wave_sum(dG[dir][n].y)} * (-BQD2_IFP * iT[dir]);
      dH[dir][n] = plane ? cd{0.0, 0.0}
                         : cd{wave_sum(dH[dir][n].x), wave_sum(dH[dir][n].y)} * (bdC * BQD2_IFP * iT[dir] * sg);
    }
    dFv[dir] = plane ? 0.0 : wave_sum(dFv[dir]) * (bdC * BQD2_IFP * iT[dir] * sg);
  }
  if (lane == 0) {
    const size_t col = (size_t)(t - 1) * 4;
    const size_t base = dom_gh(D, row, col) * 2;
    const size_t b0 = dom_dgh(D, row, 0, col) * 2, b1 = dom_dgh(D, row, 1, col) * 2;
#pragma unroll
    for (int n = 0; n < 4; n++) {
      D.G[base + 2 * n] = g[n].x; D.G[base + 2 * n + 1] = g[n].y;
      D.H[base + 2 * n] = h[n].x; D.H[base + 2 * n + 1] = h[n].y;
    }
    if (D.dG) {
#pragma unroll
      for (int n = 0; n < 4; n++) {
        D.dG[b0 + 2 * n] = dG[0][n].x; D.dG[b0 + 2 * n + 1] = dG[0][n].y;
        D.dG[b1 + 2 * n] = dG[1][n].x; D.dG[b1 + 2 * n + 1] = dG[1][n].y;
        D.dH[b0 + 2 * n] = dH[0][n].x; D.dH[b0 + 2 * n + 1] = dH[0][n].y;
        D.dH[b1 + 2 * n] = dH[1][n].x; D.dH[b1 + 2 * n + 1] = dH[1][n].y;
      }
    }
    double *fp = D.Fpair + dom_fp(D, row, t);
    fp[0] = F * sg; fp[1] = dFv[0]; fp[2] = dFv[1];
  }
}

__global__ void __launch_bounds__(WAVE) k_fs_self(DevMesh M, const Tab *__restrict__ tab, DevDomain D) {
  fs_self_body(M, tab, D, D.r0 + (int)blockIdx.x);
}

// ------------------------------------------------------------------ triangle self elements
// Polar rules about the node over the three sub-triangles to the edges, two
// GH64 halves in theta and GL12 in r.  The element is flat: F, dF and the
// free-space H vanish, and the area Jacobian is the constant
// J = |cr1 x cr2| = 2|V|/(3 sqrt3), which takes the place of |W(z,e)|.

// periodic self element (q0_lit_coef_bd, q0_lit_coef.c:297-372; dC = 0,
// q0_lit_dcoef_v2_cc.c:331-343), with the same near/smooth split as the
// quadrilateral self element
__device__ __forceinline__ void q0_self_tri_body(const DevMesh &M, const Tab *__restrict__ tab, const DevDomain &D, const GfParams &P, long long *__restrict__ counters, int row) {
  const int lane = threadIdx.x;
  int t, tn, td;
  if (!row_target(M, D, row, t, tn, td)) return;
  const int s = abs(td);
  if (!M.tri[s]) return;
  const Tab &T = *tab;
  ElemC E;
  load_elem(M, s, E);
  const double zt = T.zt_34[tn], et = T.et_34[tn];
  double r[4], th[4];
  lit_corners(zt, et, r, th);
  const double V[3] = {E.cw[0][0], E.cw[1][0], E.cw[2][0]};
  const double aV = abs3(V), J = 2.0 / (3.0 * BQD2_SQ3) * aV;
  const double Wn[3] = {V[0] * (J / aV), V[1] * (J / aV), V[2] * (J / aV)};  // unit normal times J
  const double rt[3] = {M.ren[12 * s + 3 * tn], M.ren[12 * s + 3 * tn + 1], M.ren[12 * s + 3 * tn + 2]};
  const int npts = 3 * 64 * 2 * 12;
  bool zero = false;
  for (int p = lane; p < npts; p += WAVE) {
    int i = p % 12, q = p / 12, hh = q % 2, m = (q / 2) % 64, j = q / 128;
    double tt = theta_node(th, j, hh, T.xhi[m]);
    double rm = polar_rm(r, th, j, tt);
    double c = cos(tt), sn = sin(tt);
    double rr = 0.5 * rm * (T.xli[i] + 1.0), v[3];
#pragma unroll
    for (int l = 0; l < 3; l++) v[l] = rr * (E.cr[l][1] * c + E.cr[l][2] * sn);
    zero |= z_degenerate(v[0], v[1], v[2]);
  }
  for (int p = lane; p < 3 * 144; p += WAVE) {
    double z, e, wt;
    rule_point(true, 12, T.xli, T.wli, p, z, e, wt);
    zero |= offset_degenerate(E, z, e, rt);
  }
  const bool split = images_far(E, rt, P) && !__any(zero);
  cd g[3] = {}, h[3] = {};
  GfTally tl{0, 0, 0, 0};
  int err = 0;
  for (int p = lane; p < npts; p += WAVE) {
    int i = p % 12, q = p / 12, hh = q % 2, m = (q / 2) % 64, j = q / 128;
    double tt = theta_node(th, j, hh, T.xhi[m]);
    double rm = polar_rm(r, th, j, tt);
    double c = cos(tt), sn = sin(tt);
    double rr = 0.5 * rm * (T.xli[i] + 1.0);
    double z = rr * c + zt, e = rr * sn + et, vR[3];
#pragma unroll
    for (int l = 0; l < 3; l++) vR[l] = rr * (E.cr[l][1] * c + E.cr[l][2] * sn);
    Sum10 gf;
    if (split) {
      const double aR = abs3(vR);
      free_gf(P.k, vR, aR, 1.0 / aR, gf.g, gf.dg);
    } else {
      int cnt = qpgf_eval_g_dg(P, vR[0], vR[1], vR[2], 1.0e-15, gf, tl);
      if (cnt < 0) err = cnt;
    }
    cd hn = gf.dg[0] * Wn[0] + gf.dg[1] * Wn[1] + gf.dg[2] * Wn[2];
    double wgt = 0.25 * (th[j + 1] - th[j]) * T.whi[m] * 0.5 * rm * T.wli[i] * rr;
#pragma unroll
    for (int n = 0; n < 3; n++) {
      double Nn = lit_N(n, z, e) * wgt;
      g[n] += gf.g * (Nn * J);
      h[n] += hn * Nn;
    }
  }
  if (split) {
    // smooth remainder of the lattice sum on the triangle's GL12 parts
    for (int p = lane; p < 3 * 144; p += WAVE) {
      double z, e, wt;
      rule_point(true, 12, T.xli, T.wli, p, z, e, wt);
      double vR[3], vW[3];
      bil_rw(E, z, e, vR, vW);
#pragma unroll
      for (int l = 0; l < 3; l++) vR[l] -= rt[l];
      const double aR = abs3(vR);
      Sum10 gf;
      int cnt = qpgf_eval<GF_V1>(P, vR[0], vR[1], vR[2], 1.0e-15, gf, tl);
      if (cnt < 0) err = cnt;
      cd g0, dg0[3];
      free_gf(P.k, vR, aR, 1.0 / aR, g0, dg0);
      const cd gs = gf.g - g0;
      const cd hs = (gf.dg[0] - dg0[0]) * V[0] + (gf.dg[1] - dg0[1]) * V[1] + (gf.dg[2] - dg0[2]) * V[2];
#pragma unroll
      for (int n = 0; n < 3; n++) {
        const double Nn = lit_N(n, z, e) * wt * (1.0 / 24.0);
        g[n] += gs * (Nn * aV);
        h[n] += hs * Nn;
      }
    }
  }
#pragma unroll
  for (int n = 0; n < 3; n++) {
    g[n] = cd{wave_sum(g[n].x), wave_sum(g[n].y)};
    h[n] = cd{wave_sum(h[n].x), wave_sum(h[n].y)};
  }
  long long tsp = wave_sum_ll(tl.sp), tpr = wave_sum_ll(tl.pr), tev = wave_sum_ll(tl.ev), tfd = wave_sum_ll(tl.fd);
  int anyerr = __any(err != 0);
  if (lane == 0) {
    const size_t base = dom_gh(D, row, (size_t)(t - 1) * 4) * 2;
#pragma unroll
    for (int n = 0; n < 3; n++) {
      D.G[base + 2 * n] = g[n].x; D.G[base + 2 * n + 1] = g[n].y;
      D.H[base + 2 * n] = h[n].x; D.H[base + 2 * n + 1] = h[n].y;
    }
    double *fp = D.Fpair + dom_fp(D, row, t);
    fp[0] = 0.0; fp[1] = 0.0; fp[2] = 0.0;
    tally_add(counters + CNT_SELF, tsp, tpr, tev, tfd, split ? 3 * 144 : npts);
    if (anyerr) atomicAdd((unsigned long long *)&counters[CNT_ERR], 1ull);
  }
}

__global__ void __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(2, 2))) k_q0_self_tri(DevMesh M, const Tab *__restrict__ tab, DevDomain D, GfParams P,
                                                      long long *__restrict__ counters) {
  q0_self_tri_body(M, tab, D, P, counters, D.r0 + (int)blockIdx.x);
}

// free-space self element: G on the polar rule (lit_coef_bd, lit_coef.c:290-350)
// and the principal-value dG/dv along both element parameters
// (lit_dcoef_bd_tz0_PV / _te0_PV, lit_dcoef.c:461-637, :845-952)
__device__ __forceinline__ void fs_self_tri_body(const DevMesh &M, const Tab *__restrict__ tab, const DevDomain &D, int row) {
  const int lane = threadIdx.x;
  int t, tn, td;
  if (!row_target(M, D, row, t, tn, td)) return;
  const int s = abs(td);
  if (!M.tri[s]) return;
  const Tab &T = *tab;
  const double k = D.k;
  ElemC E;
  load_elem(M, s, E);
  const double zt = T.zt_34[tn], et = T.et_34[tn];
  double r[4], th[4];
  lit_corners(zt, et, r, th);
  const double V[3] = {E.cw[0][0], E.cw[1][0], E.cw[2][0]};
  const double J = 2.0 / (3.0 * BQD2_SQ3) * abs3(V);
  double Tv[2][3];
#pragma unroll
  for (int l = 0; l < 3; l++) {
    Tv[0][l] = E.cr[l][1];
    Tv[1][l] = E.cr[l][2];
  }
  cd g[3] = {};
  for (int p = lane; p < 3 * 64 * 2 * 12; p += WAVE) {
    int i = p % 12, q = p / 12, hh = q % 2, m = (q / 2) % 64, j = q / 128;
    double tt = theta_node(th, j, hh, T.xhi[m]);
    double rm = polar_rm(r, th, j, tt);
    double c = cos(tt), sn = sin(tt);
    double rr = 0.5 * rm * (T.xli[i] + 1.0), K[3];
#pragma unroll
    for (int l = 0; l < 3; l++) K[l] = E.cr[l][1] * c + E.cr[l][2] * sn;
    double aK = abs3(K);
    cd ce = cis(k * aK * rr);
    double wgt = 0.25 * (th[j + 1] - th[j]) * T.whi[m] * 0.5 * rm * T.wli[i];
    cd cg = ce * (J / aK * wgt);
#pragma unroll
    for (int n = 0; n < 3; n++) g[n] += cg * lit_N(n, rr * c + zt, rr * sn + et);
  }
  cd dG[2][3] = {};
  // D_bar(eps): 3 sub-triangles x 128 theta x 64 r on [DEPS, rm]
  for (int p = lane; p < 3 * 64 * 2 * 64; p += WAVE) {
    int i = p % 64, q = p / 64, hh = q % 2, m = (q / 2) % 64, j = q / 128;
    double tt = theta_node(th, j, hh, T.xhi[m]);
    double rm = polar_rm(r, th, j, tt);
    double c = cos(tt), sn = sin(tt);
    double rr = 0.5 * (rm - 0.1) * T.xhi[i] + 0.5 * (rm + 0.1);
    double wgt = 0.25 * (th[j + 1] - th[j]) * T.whi[m] * 0.5 * (rm - 0.1) * T.whi[i];
    double K[3];
#pragma unroll
    for (int l = 0; l < 3; l++) K[l] = E.cr[l][1] * c + E.cr[l][2] * sn;
    double aK = abs3(K), iK = 1.0 / aK;
    double kR = k * rr * aK;
    cd ce = cis(kR), cb = cd{-1.0, kR};
    const double z = rr * c + zt, e = rr * sn + et;
#pragma unroll
    for (int dir = 0; dir < 2; dir++) {
      cd cg = cb * ce * (iK * iK / rr * (dot3(K, Tv[dir]) * iK) * J * wgt);
#pragma unroll
      for (int n = 0; n < 3; n++) dG[dir][n] += cg * lit_N(n, z, e);
    }
  }
  // the disc r < DEPS in closed form (four quarters) and its principal value (two)
  for (int p = lane; p < 6 * 64; p += WAVE) {
    const int m = p % 64, qq = p / 64;
    const bool pv = qq >= 4;
    const double off = pv ? (qq == 4 ? 0.25 : 0.75) : (-0.75 + 0.5 * qq);
    double tt = 0.25 * M_PI * T.xhi[m] + off * M_PI;
    double c = cos(tt), sn = sin(tt), K[3];
#pragma unroll
    for (int l = 0; l < 3; l++) K[l] = E.cr[l][1] * c + E.cr[l][2] * sn;
    double aK = abs3(K), iK = 1.0 / aK;
    double keK = k * 0.1 * aK, cs, cc;
    sincos(keK, &cs, &cc);
    const double wq = 0.25 * M_PI * T.whi[m] * J;
#pragma unroll
    for (int dir = 0; dir < 2; dir++) {
      const double Kdv = dot3(K, Tv[dir]) * iK;
#pragma unroll
      for (int n = 0; n < 3; n++) {
        const double base = lit_dN(n, c, sn) * iK * iK * iK * Kdv * wq;
        if (!pv) {
          // i/k [ (keK + i) sin + (2 - i keK) cos - 2 ]
          const cd f = cd{keK * cs + 2.0 * cc - 2.0, cs - keK * cc};
          dG[dir][n] += mul_i(f) * (base / k);
        } else {
          dG[dir][n] -= cd{base * cs / k, 0.0};
        }
      }
    }
  }
#pragma unroll
  for (int n = 0; n < 3; n++) g[n] = cd{wave_sum(g[n].x), wave_sum(g[n].y)} * BQD2_IFP;
  const double iT[2] = {1.0 / abs3(Tv[0]), 1.0 / abs3(Tv[1])};
#pragma unroll
  for (int dir = 0; dir < 2; dir++)
#pragma unroll
    for (int n = 0; n < 3; n++)
      dG[dir][n] = cd{wave_sum(dG[dir][n].x), wave_sum(dG[dir][n].y)} * (-BQD2_IFP * iT[dir]);
  if (lane == 0) {
    const size_t col = (size_t)(t - 1) * 4;
    const size_t base = dom_gh(D, row, col) * 2;
    const size_t b0 = dom_dgh(D, row, 0, col) * 2, b1 = dom_dgh(D, row, 1, col) * 2;
#pragma unroll
    for (int n = 0; n < 3; n++) {
      D.G[base + 2 * n] = g[n].x; D.G[base + 2 * n + 1] = g[n].y;
      D.H[base + 2 * n] = 0.0; D.H[base + 2 * n + 1] = 0.0;
    }
    if (D.dG) {
#pragma unroll
      for (int n = 0; n < 3; n++) {
        D.dG[b0 + 2 * n] = dG[0][n].x; D.dG[b0 + 2 * n + 1] = dG[0][n].y;
        D.dG[b1 + 2 * n] = dG[1][n].x; D.dG[b1 + 2 * n + 1] = dG[1][n].y;
        D.dH[b0 + 2 * n] = 0.0; D.dH[b0 + 2 * n + 1] = 0.0;
        D.dH[b1 + 2 * n] = 0.0; D.dH[b1 + 2 * n + 1] = 0.0;
      }
    }
    double *fp = D.Fpair + dom_fp(D, row, t);
    fp[0] = 0.0; fp[1] = 0.0; fp[2] = 0.0;
  }
}

__global__ void __launch_bounds__(WAVE) k_fs_self_tri(DevMesh M, const Tab *__restrict__ tab, DevDomain D) {
  fs_self_tri_body(M, tab, D, D.r0 + (int)blockIdx.x);
}

// ------------------------------------------------------------------ batched assembly
// One domain of one problem in a batched assembly.  Its tasks fall in six
// kinds, each run for all instances by one persistent launch: periodic pairs
// (the wavefronts of its pair list), periodic self rows (quadrilateral rule,
// triangle rule), free-space pairs (one (element, row) pair each, row-major) and
// free-space self rows (quadrilateral rule, triangle rule).
enum { BK_Q0_PAIRS, BK_Q0_SELF, BK_Q0_SELF_TRI, BK_FS_PAIRS, BK_FS_SELF, BK_FS_SELF_TRI, BK_N };
struct AsmBatch {
  DevMesh M;
  DevDomain D;
  GfParams P;
  const Tab *tab;
  long long *counters;
  long long t_end[BK_N];  // tasks of each kind in this and the earlier instances
};

// Task t of one kind among the concatenated task lists of the instances; a task
// is computed exactly as the per-domain kernels compute it.
template <int KIND>
__device__ __forceinline__ void batch_task(const AsmBatch *__restrict__ inst, long long t) {
  int i = 0;
  while (t >= inst[i].t_end[KIND]) i++;
  // one instance per wavefront: a uniform index keeps the instance's fields in
  // scalar registers, as kernel arguments would be
  i = __builtin_amdgcn_readfirstlane(i);
  const long long u = t - (i ? inst[i - 1].t_end[KIND] : 0);
  const DevMesh &M = inst[i].M;
  const DevDomain &D = inst[i].D;
  const GfParams &P = inst[i].P;
  const Tab *tab = inst[i].tab;
  long long *counters = inst[i].counters;
  // row and element indices uniform too, as blockIdx is in the per-domain kernels
  const int row = __builtin_amdgcn_readfirstlane((int)(KIND == BK_FS_PAIRS ? u / fs_groups(D.Ne) : u));
  if (KIND == BK_Q0_PAIRS) q0_pairs_wave(M, tab, D, P, D.plist, (long long)D.plist_n, counters, u);
  if (KIND == BK_Q0_SELF) q0_self_body(M, tab, D, P, counters, row);
  if (KIND == BK_Q0_SELF_TRI) q0_self_tri_body(M, tab, D, P, counters, row);
  if (KIND == BK_FS_PAIRS)
    fs_pairs_body(M, tab, D, FS_PPW * __builtin_amdgcn_readfirstlane((int)(u % fs_groups(D.Ne))) + 1, row);
  if (KIND == BK_FS_SELF) fs_self_body(M, tab, D, row);
  if (KIND == BK_FS_SELF_TRI) fs_self_tri_body(M, tab, D, row);
}

// The tasks of one kind for several problems' domains in one launch, one
// workgroup per task over the concatenated task lists.  (A persistent form, a
// grid of 8 workgroups per CU striding over the tasks, measured 11-15 % slower
// on the pair kernels than this grid, which the hardware dispatches as
// workgroups finish.)
template <int KIND>
__global__ void __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(2, 2))) k_asm_batch(const AsmBatch *__restrict__ inst, int ninst, long long t0) {
  const long long t = t0 + blockIdx.x;
  if (t < inst[ninst - 1].t_end[KIND]) batch_task<KIND>(inst, t);
}
// the free-space kinds with the per-domain kernels' register bound (no wave
// count fixed: the self rows then keep their registers, 0 B spilled)
template <int KIND>
__global__ void __launch_bounds__(WAVE) k_asm_batch_fs(const AsmBatch *__restrict__ inst, int ninst, long long t0) {
  const long long t = t0 + blockIdx.x;
  if (t < inst[ninst - 1].t_end[KIND]) batch_task<KIND>(inst, t);
}

// per-row F sums in source order and the diagonal term (bem3_aw_qd2_solve_bieq.c:185-198)
__global__ void __launch_bounds__(256) k_rows(DevMesh M, DevDomain D) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= D.nr) return;
  const int row = D.r0 + i;
  int t, tn, td;
  if (!row_target(M, D, row, t, tn, td)) {
    D.dF[3 * row] = D.dF[3 * row + 1] = D.dF[3 * row + 2] = 0.0;
    return;
  }
  double F = 0.0, Fz = 0.0, Fe = 0.0;
  const double *fp = D.Fpair + dom_fp(D, row, 1);
  for (int s = 0; s < D.Ne; s++) {
    F += fp[3 * s];
    Fz += fp[3 * s + 1];
    Fe += fp[3 * s + 2];
  }
  double diag = D.did == 0 ? 1.0 + F : F;
  D.H[dom_gh(D, row, row) * 2] += diag;
  D.dF[3 * row] = diag;
  D.dF[3 * row + 1] = Fz;
  D.dF[3 * row + 2] = Fe;
}

// ------------------------------------------------------------------ pair list
// (row, s) pairs of the periodic domain by class (rule, element kind), each
// class padded with -1 to a multiple of PPW so that a wavefront never mixes
// classes.  Within a class the pairs are ordered by the z offset of the element
// centre from the target (256 levels), then by its in-plane offset reduced to
// the unit cell (64 levels, in the key's low mantissa bits), ties in row-major
// order: the pairs of one wavefront then take the same branch of the lattice
// sum (pure spectral beyond a height, Ewald below) and about as many terms.
static void pair_list(const DevDomain &D, const std::vector<unsigned char> &rule, const std::vector<float> &key,
                      const std::vector<char> &tripos, std::vector<int> &out) {
  std::vector<int> cls[18];
  for (int row = 0; row < D.N; row++) {
    const int t = row / 4 + 1, tn = row % 4;
    if (tripos[t] && tn == 3) continue;
    for (int s = 1; s <= D.Ne; s++) {
      if (s == t) continue;
      const int gid = row * D.Ne + (s - 1);
      cls[2 * rule[gid] + (tripos[s] ? 1 : 0)].push_back(gid);
    }
  }
  out.clear();
  for (auto &c : cls) {
    // stable counting sort on (z offset in 256 levels, in-plane offset in 64)
    float kmax = 0.0f;
    for (int g : c) kmax = key[g] > kmax ? key[g] : kmax;
    if (kmax > 0.0f && c.size() > 1) {
      auto bucket = [&](int g) {
        unsigned bits;
        std::memcpy(&bits, &key[g], sizeof(bits));
        int b = (int)(key[g] / kmax * 256.0f);
        return 64 * (b > 255 ? 255 : b) + (int)(bits & 63u);
      };
      std::vector<size_t> start(256 * 64 + 1, 0);
      for (int g : c) start[bucket(g) + 1]++;
      for (int b = 0; b < 256 * 64; b++) start[b + 1] += start[b];
      std::vector<int> sorted(c.size());
      for (int g : c) sorted[start[bucket(g)]++] = g;
      c.swap(sorted);
    }
    out.insert(out.end(), c.begin(), c.end());
    while (out.size() % PPW) out.push_back(-1);
  }
}

// ------------------------------------------------------------------ entry point
// Assembly in three phases: clear the outputs and (first time) decide the rules
// and build the pair list; the off-diagonal pairs; the self blocks, row sums and
// counters.  bqd2_assemble runs them for one domain, bqd2_assemble_batch runs
// the pairs of several problems' periodic domains as one launch.
struct AsmPhase {
  hipEvent_t e0, e1, e1b, e2, e3;
};

// zero the rows [r0, r0 + nr) of G, H, dG, dH and Fpair
int bqd2_dom_clear(const DevDomain &D, hipStream_t st) {
  const size_t nr = (size_t)D.nr, rowb = (size_t)D.N * 16;
  if (D.ldm == D.N) {
    BQD2_HIP(hipMemsetAsync(D.G, 0, nr * rowb, st));
    BQD2_HIP(hipMemsetAsync(D.H, 0, nr * rowb, st));
  } else if (D.ldm == 2 * (long long)D.N && D.G == D.H + 2 * (size_t)D.N) {
    // system layout: the rows of H and G are whole rows of one array
    BQD2_HIP(hipMemsetAsync(D.H, 0, 2 * nr * rowb, st));
  } else {
    BQD2_HIP(hipMemset2DAsync(D.G, (size_t)D.ldm * 16, 0, rowb, nr, st));
    BQD2_HIP(hipMemset2DAsync(D.H, (size_t)D.ldm * 16, 0, rowb, nr, st));
  }
  if (D.dG) {
    BQD2_HIP(hipMemsetAsync(D.dG, 0, 2 * nr * rowb, st));
    BQD2_HIP(hipMemsetAsync(D.dH, 0, 2 * nr * rowb, st));
  }
  BQD2_HIP(hipMemsetAsync(D.Fpair, 0, nr * (size_t)D.Ne * 3 * 8, st));
  return BQD2_OK;
}

// the free-space rows [D.r0, D.r0 + D.nr) of domain D (rules decided): pairs,
// self blocks and row sums
int bqd2_fs_rows(bqd2_problem *p, const DevDomain &D, hipStream_t st) {
  if (D.nr <= 0) return BQD2_OK;
  int rc = bqd2_dom_clear(D, st);
  if (rc) return rc;
  hipLaunchKernelGGL(k_fs_pairs, dim3(fs_groups(D.Ne), D.nr), dim3(WAVE), 0, st, p->m, p->tab, D);
  hipLaunchKernelGGL(k_fs_self, dim3(D.nr), dim3(WAVE), 0, st, p->m, p->tab, D);
  if (D.has_tri) hipLaunchKernelGGL(k_fs_self_tri, dim3(D.nr), dim3(WAVE), 0, st, p->m, p->tab, D);
  hipLaunchKernelGGL(k_rows, dim3(bqd2_blocks(D.nr, 256)), dim3(256), 0, st, p->m, D);
  BQD2_HIP(hipGetLastError());
  return BQD2_OK;
}

static int asm_begin(bqd2_problem *p, int did, AsmPhase &ev) {
  DevDomain &D = p->dom[did];
  const bool periodic = (did == 0 && p->m.ps == 1);
  hipStream_t st_ = p->stream;
  BQD2_CHECK(D.r0 == 0 && D.nr == D.N, BQD2_E_STATE, "bqd2_assemble: domain %d holds a window of its rows", did);
  BQD2_HIP(hipEventCreate(&ev.e0));
  BQD2_HIP(hipEventCreate(&ev.e1));
  BQD2_HIP(hipEventCreate(&ev.e1b));
  BQD2_HIP(hipEventCreate(&ev.e2));
  BQD2_HIP(hipEventCreate(&ev.e3));
  BQD2_HIP(hipMemsetAsync(p->d_counters, 0, CNT_N * sizeof(long long), st_));
  int rc = bqd2_dom_clear(D, st_);
  if (rc) return rc;
  BQD2_HIP(hipEventRecord(ev.e0, st_));
  long long npairs = (long long)D.N * D.Ne;
  // the rules and the pair list depend on the geometry and the lattice only:
  // decided at the first assembly of the domain, reused afterwards
  const bool first = (D.plist_n == 0);
  if (first) {
    hipLaunchKernelGGL(k_decide, dim3(bqd2_blocks(npairs, 256)), dim3(256), 0, st_, p->m, p->tab, D, p->gf,
                       periodic ? 1 : 0, p->smooth_mode);
    BQD2_HIP(hipGetLastError());
  }
  BQD2_HIP(hipEventRecord(ev.e1, st_));
  if (first) {
    std::vector<unsigned char> rule((size_t)npairs);
    BQD2_HIP(hipMemcpyAsync(rule.data(), D.rule, (size_t)npairs, hipMemcpyDeviceToHost, st_));
    // the pair list: both domains in the default layout (the periodic domain
    // in the system layout, whose interior domain goes by row windows)
    const bool listed = D.plist != nullptr && D.pkey != nullptr;
    std::vector<float> key;
    if (listed) {
      key.resize((size_t)npairs);
      BQD2_HIP(hipMemcpyAsync(key.data(), D.pkey, (size_t)npairs * sizeof(float), hipMemcpyDeviceToHost, st_));
    }
    BQD2_HIP(hipStreamSynchronize(st_));
    D.n_gh = D.n_split = D.n_smooth = D.n_free_pts = 0;
    for (long long gid = 0; gid < npairs; gid++) {
      const int r = rule[gid];
      D.n_gh += (r >= 1 && r <= 5);
      D.n_split += (r >= 2 && r <= 5);
      D.n_smooth += (r >= 3);
      if (r >= 2) {
        const int nf = rule_free_n(r), s = (int)(gid % D.Ne) + 1;
        D.n_free_pts += (p->tripos[did][s] ? 3 : 1) * (long long)nf * nf;
      }
    }
    std::vector<int> plist;
    if (listed) {
      pair_list(D, rule, key, p->tripos[did], plist);
      BQD2_CHECK(plist.size() <= (size_t)D.N * D.Ne + PLIST_PAD, BQD2_E_STATE,
                 "bqd2_assemble: pair list of %zu entries exceeds its buffer", plist.size());
      if (!plist.empty()) {
        BQD2_HIP(hipMemcpyAsync(D.plist, plist.data(), plist.size() * sizeof(int), hipMemcpyHostToDevice, st_));
        BQD2_HIP(hipStreamSynchronize(st_));  // the host list is released on return
      }
    }
    D.plist_n = listed ? (int)plist.size() : 1;
  }
  BQD2_HIP(hipEventRecord(ev.e1b, st_));
  return BQD2_OK;
}

static int asm_end(bqd2_problem *p, int did, AsmPhase &ev, bqd2_assembly_stats *st, bool self_done = false,
                   bool own_counters = true) {
  DevDomain &D = p->dom[did];
  const bool periodic = (did == 0 && p->m.ps == 1);
  hipStream_t st_ = p->stream;
  if (self_done) {
  } else if (periodic) {
    hipLaunchKernelGGL(k_q0_self, dim3(D.N), dim3(WAVE), 0, st_, p->m, p->tab, D, p->gf, p->d_counters);
    if (D.has_tri)
      hipLaunchKernelGGL(k_q0_self_tri, dim3(D.N), dim3(WAVE), 0, st_, p->m, p->tab, D, p->gf, p->d_counters);
  } else {
    hipLaunchKernelGGL(k_fs_self, dim3(D.N), dim3(WAVE), 0, st_, p->m, p->tab, D);
    if (D.has_tri) hipLaunchKernelGGL(k_fs_self_tri, dim3(D.N), dim3(WAVE), 0, st_, p->m, p->tab, D);
  }
  BQD2_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_rows, dim3(bqd2_blocks(D.N, 256)), dim3(256), 0, st_, p->m, D);
  BQD2_HIP(hipGetLastError());
  BQD2_HIP(hipEventRecord(ev.e3, st_));
  BQD2_HIP(hipEventSynchronize(ev.e3));
  long long cnt[CNT_N];
  BQD2_HIP(hipMemcpyAsync(cnt, p->d_counters, sizeof(cnt), hipMemcpyDeviceToHost, st_));
  BQD2_HIP(hipStreamSynchronize(st_));
  const long long ngh = D.n_gh, nsplit = D.n_split;
  if (st) {
    memset(st, 0, sizeof(*st));
    float a, b, c;
    BQD2_HIP(hipEventElapsedTime(&a, ev.e0, ev.e1));
    BQD2_HIP(hipEventElapsedTime(&b, ev.e1b, ev.e2));
    BQD2_HIP(hipEventElapsedTime(&c, ev.e2, ev.e3));
    st->ms_decide = a; st->ms_pairs = b; st->ms_self = c;
    if (!own_counters && !periodic) memset(cnt + CNT_PAIRS, 0, (CNT_N - CNT_PAIRS) * sizeof(long long));
    for (int i = 0; i < 4; i++) {
      st->pair_terms[i] = cnt[CNT_PAIRS + i];
      st->self_terms[i] = cnt[CNT_SELF + i];
    }
    st->pair_evals = cnt[CNT_PAIRS + 4];
    st->self_evals = cnt[CNT_SELF + 4];
    st->gf_evals = st->pair_evals + st->self_evals;
    st->gf_terms = 0;
    for (int i = 0; i < 4; i++) st->gf_terms += st->pair_terms[i] + st->self_terms[i];
    st->pairs_gh = ngh;
    st->pairs_split = nsplit;
    st->pairs_smooth = D.n_smooth;
    st->free_points = D.n_free_pts;
    (void)periodic;
    const long long valid_rows = D.nn;  // collocation nodes: 4 per quadrilateral, 3 per triangle
    st->self_rows = valid_rows;
    st->pairs_gl = valid_rows * (D.Ne - 1) - ngh;
  }
  (void)hipEventDestroy(ev.e0); (void)hipEventDestroy(ev.e1); (void)hipEventDestroy(ev.e1b);
  (void)hipEventDestroy(ev.e2); (void)hipEventDestroy(ev.e3);
  BQD2_CHECK(cnt[CNT_ERR] == 0, BQD2_E_GF, "bqd2_assemble: %lld Green's function evaluations failed (domain %d)",
             cnt[CNT_ERR], did);
  D.assembled = true;
  D.g_lu_valid = false;
  p->lu32_valid = false;
  p->lu64_valid = false;
  return BQD2_OK;
}

// the rules of domain did's pairs (first time only) and, for the periodic
// domain, its pair list, without the pairs themselves (the streamed interior
// domain of the system layout assembles its rows later, window by window)
int bqd2_decide_only(bqd2_problem *p, int did, bqd2_assembly_stats *st) {
  DevDomain &D = p->dom[did];
  AsmPhase ev;
  BQD2_HIP(hipEventCreate(&ev.e0));
  BQD2_HIP(hipEventCreate(&ev.e1));
  BQD2_HIP(hipEventRecord(ev.e0, p->stream));
  if (D.plist_n == 0) {
    const long long npairs = (long long)D.N * D.Ne;
    hipLaunchKernelGGL(k_decide, dim3(bqd2_blocks(npairs, 256)), dim3(256), 0, p->stream, p->m, p->tab, D, p->gf, 0,
                       0);
    BQD2_HIP(hipGetLastError());
    D.plist_n = 1;
  }
  BQD2_HIP(hipEventRecord(ev.e1, p->stream));
  BQD2_HIP(hipEventSynchronize(ev.e1));
  if (st) {
    memset(st, 0, sizeof(*st));
    float a = 0.0f;
    BQD2_HIP(hipEventElapsedTime(&a, ev.e0, ev.e1));
    st->ms_decide = a;
  }
  (void)hipEventDestroy(ev.e0);
  (void)hipEventDestroy(ev.e1);
  return BQD2_OK;
}

int bqd2_lg_assemble(bqd2_problem *p, int did, bqd2_assembly_stats *st);
int bqd2_assemble_full(bqd2_problem *p, int did, bqd2_assembly_stats *st);

// The remainder rules 3-8 on (1), off (0: rules 0-2 only, the lattice sum on
// every GL12 point as the reference sums it) or on with one order (4, 6, 8)
// wherever a remainder rule is allowed; the rules are decided again at the
// next assembly of each domain
extern "C" int bqd2_set_remainder_rules(bqd2_problem *p, int on) {
  BQD2_CHECK(p && (on == 0 || on == 1 || on == 4 || on == 6 || on == 8), BQD2_E_ARG,
             "bqd2_set_remainder_rules: bad arguments");
  p->smooth_mode = on;
  for (auto &D : p->dom) D.plist_n = 0;
  return BQD2_OK;
}

extern "C" int bqd2_assemble(bqd2_problem *p, int did, bqd2_assembly_stats *st) {
  BQD2_CHECK(p && did >= 0 && did <= p->m.MN, BQD2_E_ARG, "bqd2_assemble: bad domain %d", did);
  if (p->large) return bqd2_lg_assemble(p, did, st);
  return bqd2_assemble_full(p, did, st);
}

int bqd2_assemble_full(bqd2_problem *p, int did, bqd2_assembly_stats *st) {
  DevDomain &D = p->dom[did];
  if (D.Ne == 0) return BQD2_OK;
  const bool periodic = (did == 0 && p->m.ps == 1);
  hipStream_t st_ = p->stream;
  AsmPhase ev;
  int rc = asm_begin(p, did, ev);
  if (rc) return rc;
  // with a progress callback the pairs go in 16 launches, reported as each ends;
  // in any case in parts of at most BQD2_MAX_WG workgroups (the dispatch holds
  // the grid's size in work-items as 32 bits: 13 824 quads need 1.9e8)
  std::vector<hipEvent_t> parts;
  if (periodic) {
    const long long nwg = D.plist_n / PPW;
    const int np = std::max<int>((bqd2_progress_on() && nwg >= 16) ? 16 : 1,
                                 (int)((nwg + BQD2_MAX_WG - 1) / BQD2_MAX_WG));
    for (int k = 0; k < np; k++) {
      const long long a = nwg * k / np, b = nwg * (k + 1) / np;
      if (b > a)
        hipLaunchKernelGGL(k_q0_pairs, dim3((unsigned)(b - a)), dim3(WAVE), 0, st_, p->m, p->tab, D, p->gf, D.plist,
                           (long long)D.plist_n, p->d_counters, a);
      if (np > 1 && bqd2_progress_on()) {
        hipEvent_t e;
        BQD2_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        BQD2_HIP(hipEventRecord(e, st_));
        parts.push_back(e);
      }
    }
  } else if (D.plist_n > 1) {
    // the rule-sorted list, in launches of at most BQD2_MAX_WG wavefronts
    const long long nwg = (D.plist_n + FS_PPW - 1) / FS_PPW;
    for (long long a = 0; a < nwg; a += BQD2_MAX_WG) {
      const long long b = std::min(nwg, a + BQD2_MAX_WG);
      hipLaunchKernelGGL(k_fs_pairs_list, dim3((unsigned)(b - a)), dim3(WAVE), 0, st_, p->m, p->tab, D,
                         (long long)D.plist_n, a);
    }
  } else {
    hipLaunchKernelGGL(k_fs_pairs, dim3(fs_groups(D.Ne), D.N), dim3(WAVE), 0, st_, p->m, p->tab, D);
  }
  BQD2_HIP(hipGetLastError());
  BQD2_HIP(hipEventRecord(ev.e2, st_));
  for (size_t k = 0; k < parts.size(); k++) {
    BQD2_HIP(hipEventSynchronize(parts[k]));
    (void)hipEventDestroy(parts[k]);
    bqd2_progress("periodic pairs", (long long)k + 1, (long long)parts.size());
  }
  return asm_end(p, did, ev, st);
}

// No counterpart in the reference.  Domain did of every problem (did = -1:
// every domain of every problem): the set-up of each (rules and pair list at
// the first assembly) on its problem's stream, then the pairs and self rows of
// all of them, one launch per task kind on the first problem's stream, then the
// row sums of each.
extern "C" int bqd2_assemble_batch(bqd2_problem **ps, int count, int did, bqd2_assembly_stats *st) {
  BQD2_CHECK(ps && count > 0, BQD2_E_ARG, "bqd2_assemble_batch: no problems");
  for (int i = 0; i < count; i++) {
    BQD2_CHECK(ps[i] && did >= -1 && did <= ps[i]->m.MN, BQD2_E_ARG, "bqd2_assemble_batch: bad problem %d or domain",
               i);
    BQD2_CHECK(!ps[i]->large, BQD2_E_ARG, "bqd2_assemble_batch: problem %d is in the system layout", i);
  }
  struct Job {
    bqd2_problem *p;
    int did;
    bqd2_assembly_stats *st;
    AsmPhase ev;
  };
  std::vector<Job> jobs;
  size_t si = 0;
  for (int i = 0; i < count; i++) {
    const int d0 = did < 0 ? 0 : did, d1 = did < 0 ? ps[i]->m.MN : did;
    for (int d = d0; d <= d1; d++, si++) {
      bqd2_assembly_stats *sp = st ? st + si : nullptr;
      if (ps[i]->dom[d].Ne == 0) {
        if (sp) memset(sp, 0, sizeof(*sp));
        continue;
      }
      jobs.push_back(Job{ps[i], d, sp, AsmPhase{}});
    }
  }
  if (jobs.empty()) return BQD2_OK;
  // every set-up clears its problem's counters; all of them precede the launch
  for (auto &j : jobs) {
    int rc = asm_begin(j.p, j.did, j.ev);
    if (rc) return rc;
  }
  std::vector<AsmBatch> hb(jobs.size());
  long long tasks[BK_N] = {0, 0, 0, 0, 0, 0};
  for (size_t j = 0; j < jobs.size(); j++) {
    bqd2_problem *p = jobs[j].p;
    const DevDomain &D = p->dom[jobs[j].did];
    const bool periodic = (jobs[j].did == 0 && p->m.ps == 1);
    hb[j].M = p->m;
    hb[j].D = D;
    hb[j].P = p->gf;
    hb[j].tab = p->tab;
    hb[j].counters = p->d_counters;
    const long long rows = D.N, rows_tri = D.has_tri ? D.N : 0;
    if (periodic) {
      tasks[BK_Q0_PAIRS] += D.plist_n / PPW;
      tasks[BK_Q0_SELF] += rows;
      tasks[BK_Q0_SELF_TRI] += rows_tri;
    } else {
      tasks[BK_FS_PAIRS] += (long long)fs_groups(D.Ne) * D.N;
      tasks[BK_FS_SELF] += rows;
      tasks[BK_FS_SELF_TRI] += rows_tri;
    }
    for (int k = 0; k < BK_N; k++) hb[j].t_end[k] = tasks[k];
  }
  hipStream_t bs = jobs[0].p->stream;
  // the instance table in the first problem's workspace (its stream runs the launch)
  char *wsb = nullptr;
  int rc = bqd2_workspace(&jobs[0].p->main, hb.size() * sizeof(AsmBatch), &wsb);
  if (rc) return rc;
  AsmBatch *d_inst = (AsmBatch *)wsb;
  BQD2_HIP(hipMemcpyAsync(d_inst, hb.data(), hb.size() * sizeof(AsmBatch), hipMemcpyHostToDevice, bs));
  // the launch waits for every set-up; every problem waits for the launch
  for (auto &j : jobs) {
    BQD2_HIP(hipStreamWaitEvent(bs, j.ev.e1b, 0));
    BQD2_HIP(hipEventRecord(j.ev.e1b, bs));  // the pair time is the batch's
  }
  // one launch per task kind over every instance's tasks (in parts of at most
  // BQD2_MAX_WG single-wave workgroups: the dispatch holds the grid's size in
  // work-items as 32 bits)
  auto launch = [&](int kind, void (*kern)(const AsmBatch *, int, long long)) {
    for (long long t0 = 0; t0 < tasks[kind]; t0 += BQD2_MAX_WG)
      hipLaunchKernelGGL(kern, dim3((unsigned)std::min<long long>(BQD2_MAX_WG, tasks[kind] - t0)), dim3(WAVE), 0, bs,
                         d_inst, (int)hb.size(), t0);
  };
  launch(BK_Q0_PAIRS, k_asm_batch<BK_Q0_PAIRS>);
  launch(BK_FS_PAIRS, k_asm_batch_fs<BK_FS_PAIRS>);
  launch(BK_Q0_SELF, k_asm_batch<BK_Q0_SELF>);
  launch(BK_Q0_SELF_TRI, k_asm_batch<BK_Q0_SELF_TRI>);
  launch(BK_FS_SELF, k_asm_batch_fs<BK_FS_SELF>);
  launch(BK_FS_SELF_TRI, k_asm_batch_fs<BK_FS_SELF_TRI>);
  BQD2_HIP(hipGetLastError());
  for (auto &j : jobs) {
    BQD2_HIP(hipEventRecord(j.ev.e2, bs));
    BQD2_HIP(hipStreamWaitEvent(j.p->stream, j.ev.e2, 0));
  }
  BQD2_HIP(hipStreamSynchronize(bs));
  for (auto &j : jobs) {
    // several domains of one problem share its counters: each domain's
    // statistics keep the periodic domain's terms only
    const bool shared = did < 0 && j.p->m.MN > 0;
    if ((rc = asm_end(j.p, j.did, j.ev, j.st, true, !shared))) return rc;
  }
  return BQD2_OK;
}
