// field.hip -- pressure and particle velocity off and on the boundary, and the
// radiation force and torque (bilinear and linear-triangle elements).
//
//   k_domain_id   point classification by signed solid angles (domain_id,
//                 bem3_aw_qd2.c:1148-1257), one wavefront per point, the
//                 reference's domain order and early exit
//   k_field_F     F and grad F of every chunk of a field point's items (the
//                 chunks of k_field_part that lie wholly in element bands taken
//                 from the far field's amplitudes end at once); k_field_bd_F
//                 likewise for boundary points
//   k_field_part  partial sums of the representation formula and its gradient
//                 (pv_s_dbieq_aqd2, bem3_aw_qd2_field.c:227-297 with
//                 q0_bil_dcoef_grad_* / bil_dcoef_grad_*) over chunks of a
//                 point's items, one wavefront per (point, chunk); k_field_fin
//                 sums them, falls back to the boundary-value form
//                 (pv_s_bieq_aqd2, :140-201) when |grad F| > CBD_CDF, applies the
//                 Bloch phase and the incident field
//   k_field_bd    chunks of a boundary point's (element, point) items, one
//                 wavefront each, plus one wavefront for the point's own element
//                 in polar form (q0_bil_coef_bd / bil_coef_bd); k_field_bd_fin
//                 sums them as pv_s_bd_aqd2 / pv_t_bd_aqd2 (:394-469) do
//   k_force       radiation stress tensor on the boundary of the open domain
//                 (force_FN_aqd2, bem3_aw_qd2_force.c:4-165), per element,
//                 then an ordered sum
#include <cstring>
#include <algorithm>
#include <mutex>
#include "elem_dev.h"

#define CBD_CDF_D 10.0

struct DomView {
  int Ne;
  const int *sid;
  const int *fperm;  // element positions in visiting order
  const double *P, *dP, *pv, *dpv;
  double k, k2x, k2y;
};

struct FieldCtx {
  DomView dom[16];
  int MN, ps;
  int any_tri;  // the mesh has linear triangles: rules are sized for both kinds
  int far_on;   // points of the periodic domain take the element bands at least
                // far_h below or above them from plane-wave amplitudes
  double far_ztop, far_zbot, far_h;
  const FarBand *far_bands;  // FAR_NB bands of domain 0's field order
  double pw_p[2], pw_rho0, pw_c0, pw_k0, pw_tv[3], pw_R[9];
  double d1x, d1y, d2x, d2y, idd, vkx, vky;
};

__device__ __forceinline__ void incident_pv(const FieldCtx &X, const double *r, cd &p, cd *v) {
  const double *R = X.pw_R;
  double rc0 = r[0] - X.pw_tv[0], rc1 = r[1] - X.pw_tv[1], rc2 = r[2] - X.pw_tv[2];
  double z = -R[6] * rc0 - R[7] * rc1 + R[8] * rc2;
  cd ce = cis(X.pw_k0 * z);
  p = cd{X.pw_p[0], X.pw_p[1]} * ce;
  cd vz = p * (1.0 / (X.pw_rho0 * X.pw_c0));
  for (int i = 0; i < 3; i++) v[i] = vz * R[3 * i + 2];
}

// ------------------------------------------------------------ domain id
__device__ __forceinline__ int unit3(double *v) {
  double a = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  if (a == 0.0) return -1;
  v[0] /= a; v[1] /= a; v[2] /= a;
  return 0;
}
__device__ __forceinline__ void cross3(double *r, const double *a, const double *b) {
  r[0] = a[1] * b[2] - a[2] * b[1];
  r[1] = a[2] * b[0] - a[0] * b[2];
  r[2] = a[0] * b[1] - a[1] * b[0];
}

// signed solid angle of a triangle / quadrangle seen from the origin of the
// corner vectors r (fid_calc_solid_angle, bem3_aw_qd2.c:1196-1257)
__device__ double solid_angle(bool quad, double r[4][3], int &flag) {
  flag = 0;
  int st = 0;
  for (int i = 0; i < (quad ? 4 : 3); i++) st += unit3(r[i]);
  if (st < 0) { flag = -1; return 0.0; }
  double n01[3], n12[3], n20[3], nt[3], om = 0.0;
  cross3(n01, r[0], r[1]);
  cross3(n12, r[1], r[2]);
  cross3(n20, r[2], r[0]);
  cross3(nt, n01, n20);
  double a0 = atan2(abs3(nt), -dot3(n01, n20));
  cross3(nt, n12, n01);
  double a1 = atan2(abs3(nt), -dot3(n12, n01));
  cross3(nt, n20, n12);
  double a2 = atan2(abs3(nt), -dot3(n20, n12));
  double D = dot3(r[0], n12);
  if (D > 0.0) om += a0 + a1 + a2 - M_PI;
  else if (D < 0.0) om -= a0 + a1 + a2 - M_PI;
  else { flag = -1; return 0.0; }
  if (quad) {
    n01[0] = -n20[0]; n01[1] = -n20[1]; n01[2] = -n20[2];
    cross3(n12, r[2], r[3]);
    cross3(n20, r[3], r[0]);
    cross3(nt, n01, n20);
    a0 = atan2(abs3(nt), -dot3(n01, n20));
    cross3(nt, n12, n01);
    a1 = atan2(abs3(nt), -dot3(n12, n01));
    cross3(nt, n20, n12);
    a2 = atan2(abs3(nt), -dot3(n20, n12));
    D = dot3(r[0], n12);
    if (D > 0.0) om += a0 + a1 + a2 - M_PI;
    else if (D < 0.0) om -= a0 + a1 + a2 - M_PI;
    else { flag = -1; return 0.0; }
  }
  return om;
}

// corner positions come from the mesh nodes: rn (Nn+1) x 3 and ed (Ne+1) x 4.
// One wavefront per point (four per workgroup): the lanes take a domain's
// elements in turn and keep their shares of the signed solid angles in LDS
// (the element's domain and the domain on its other side, any of MN + 1);
// the domain's total is the wavefront's sum.  A point on an element (solid
// angle undefined) belongs to the domain being summed, as in the reference's
// loop, which stops at the first such element: which one does not matter.
__global__ void __launch_bounds__(256) k_domain_id(DevMesh M, FieldCtx X, const double *__restrict__ rn,
                                                   const int *__restrict__ ed, int n, const double *__restrict__ pts,
                                                   int *__restrict__ did_out, double *__restrict__ r0_out,
                                                   int *__restrict__ l_out) {
  __shared__ double sOg[4][16][64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int ip = blockIdx.x * 4 + wv;
  if (ip >= n) return;
  double rt[3] = {pts[3 * ip], pts[3 * ip + 1], pts[3 * ip + 2]};
  int l1 = 0, l2 = 0;
  if (X.ps == 1) {
    l1 = (int)floor((rt[0] * X.d2y - rt[1] * X.d2x) * X.idd + 0.5);
    l2 = (int)floor(-(rt[0] * X.d1y - rt[1] * X.d1x) * X.idd + 0.5);
    rt[0] = rt[0] - (double)l1 * X.d1x - (double)l2 * X.d2x;
    rt[1] = rt[1] - (double)l1 * X.d1y - (double)l2 * X.d2y;
  }
  if (lane == 0) {
    r0_out[3 * ip] = rt[0]; r0_out[3 * ip + 1] = rt[1]; r0_out[3 * ip + 2] = rt[2];
    l_out[2 * ip] = l1; l_out[2 * ip + 1] = l2;
  }
  double *Og = &sOg[wv][0][lane];  // Og[d] at Og[64 d]
  for (int d = 0; d <= X.MN; d++) Og[64 * d] = 0.0;
  int result = -1;
  for (int d = 0; d <= X.MN && result < 0; d++) {
    const DomView &V = X.dom[d];
    bool hit = false;
    for (int i = 1 + lane; i <= V.Ne && !hit; i += 64) {
      int s = V.sid[i];
      if (s <= 0) continue;
      double rv[4][3];
      for (int j = 0; j < 4; j++)
        for (int k = 0; k < 3; k++) rv[j][k] = rn[3 * ed[4 * s + j] + k] - rt[k];
      int flag;
      double om = solid_angle(!M.tri[s], rv, flag);
      if (flag < 0) { hit = true; break; }
      Og[64 * d] += om;
      Og[64 * M.sd[s]] -= om;
    }
    if (__any(hit)) { result = d; break; }
    const double tot = wave_sum(Og[64 * d]);
    if (d == 0 && fabs(tot) < 2.0 * M_PI) result = d;
    else if (tot > 2.0 * M_PI) result = d;
  }
  if (lane == 0) did_out[ip] = result;
}

// ------------------------------------------------------------ source points
// position, normal, weight and interpolation weights of quadrature point q of
// element s for rule type (0: nodes, 1: 9-point / 7-point, 2: GLN^2, 3: GHN^2;
// triangles use their three GL parts for 2 and 3).  Triangle weights carry the
// reference's scale (q0_lit_coef.c:10-185: 1/2 for the node and 7-point rules,
// 1/24 for the parts) so that every rule is summed the same way.  Returns false
// for the unused tail of a rule sized for the other element kind.
__device__ __forceinline__ bool source_point(const DevMesh &M, const Tab &T, const ElemC &E, int s, int type, int q,
                                             double *r, double *w, double &wt, double *Nv) {
  const bool tri = M.tri[s] != 0;
  if (type == 0) {
    if (q >= 4) return false;
    for (int d = 0; d < 3; d++) {
      r[d] = M.ren[12 * s + 3 * q + d];
      w[d] = M.wen[12 * s + 3 * q + d];
    }
    if (!tri) {
      wt = T.wt_44[q];
      for (int n = 0; n < 4; n++) Nv[n] = (n == q) ? 1.0 : 0.0;
    } else {
      wt = 0.5 * T.wt_34[q];
      for (int n = 0; n < 4; n++) Nv[n] = (q < 3) ? ((n == q) ? 1.0 : 0.0) : lit_N(n, T.zt_34[3], T.et_34[3]);
    }
    return true;
  }
  double z, e;
  if (type == 1) {
    if (q >= (tri ? 7 : 9)) return false;
    if (tri) {
      z = T.zt_37[q]; e = T.et_37[q]; wt = 0.5 * T.wt_37[q];
    } else {
      z = T.zt_49[q]; e = T.et_49[q]; wt = T.wt_49[q];
    }
  } else {
    const int n = type == 2 ? 12 : 64;
    if (q >= rule_npts(tri, n)) return false;
    rule_point(tri, n, type == 2 ? T.xli : T.xhi, type == 2 ? T.wli : T.whi, q, z, e, wt);
    if (tri) wt *= 1.0 / 24.0;
  }
  bil_rw(E, z, e, r, w);
  for (int n = 0; n < 4; n++) Nv[n] = shape_N(tri, n, z, e);
  return true;
}

// points per element of rule type, sized for the mesh's element kinds
__host__ __device__ __forceinline__ int rule_points(int type, int any_tri) {
  if (type == 0) return 4;
  if (type == 1) return 9;
  const int n2 = type == 2 ? 144 : 4096;
  return any_tri ? 3 * n2 : n2;
}

__device__ __forceinline__ cd ld(const double *a, size_t i) { return cd{a[2 * i], a[2 * i + 1]}; }

// The element bands of domain 0 that a point at height z takes from the
// plane-wave amplitudes: bands 0..b0 lie at least far_h below it (b0 = -1: none),
// bands b1..FAR_NB-1 at least far_h above it (b1 = FAR_NB: none); the lattice
// sums cover the elements at field-order positions [es, ee)
__device__ __forceinline__ void far_split(const FieldCtx &X, int did, double z, int &b0, int &b1, int &es, int &ee) {
  b0 = -1;
  b1 = FAR_NB;
  es = 0;
  ee = X.dom[did].Ne;
  if (!X.far_on || did != 0) return;
  const FarBand *B = X.far_bands;
  while (b0 + 1 < FAR_NB && B[b0 + 1].zpre <= z - X.far_h) b0++;
  while (b1 - 1 > b0 && B[b1 - 1].zsuf >= z + X.far_h) b1--;
  es = b0 >= 0 ? B[b0].ee : 0;
  ee = b1 < FAR_NB ? B[b1].es : X.dom[did].Ne;
}

// ------------------------------------------------------------ off-boundary field
// The (element, quadrature point) items of one field point are cut into chunks
// of a fixed size that depends only on the mesh and the rule (not on how many
// points a call asks for), one wavefront per (point, chunk); k_field_fin sums a
// point's chunks in order.  A single point is then spread over up to 64
// wavefronts and gives the same bits alone as inside a batch.
enum { FP_N = 18 };  // per chunk: Pa (2), dPa (6), PVa (6), F, dF (3)
enum { BD_CHUNK = 1024, BP_N = 9 };  // k_field_bd's chunks: Pa (2), PVa (6), F

__host__ __device__ __forceinline__ long long field_chunk_items(long long items) {
  const long long per = (items + 64 * 64 - 1) / (64 * 64);
  return 64 * (per > 0 ? per : 1);
}

#ifndef BQD2_FIELD_WPE
#define BQD2_FIELD_WPE 2
#endif
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(BQD2_FIELD_WPE, BQD2_FIELD_WPE))) k_field_part(DevMesh M, const Tab *__restrict__ tab, FieldCtx X, GfParams P,
                                                   int type, int n, int nchunk, const double *__restrict__ r0s,
                                                   const int *__restrict__ did_io,
                                                   double *__restrict__ part, long long *__restrict__ counters) {
  const int ip = blockIdx.x, ic = blockIdx.y, lane = threadIdx.x;
  if (ip >= n) return;
  const int did = did_io[ip];
  double *dst = part + ((size_t)ip * nchunk + ic) * FP_N;
  if (did < 0) {
    if (lane < 14) dst[lane] = 0.0;
    return;
  }
  const Tab &T = *tab;
  const DomView &V = X.dom[did];
  const bool periodic = (X.ps == 1 && did == 0);
  const double r0[3] = {r0s[3 * ip], r0s[3 * ip + 1], r0s[3 * ip + 2]};
  // elements of bands far below or above the point: their lattice sums from
  // k_far_eval (F and grad F of every item from k_field_F)
  int fb0, fb1, fes, fee;
  far_split(X, did, r0[2], fb0, fb1, fes, fee);
  const int npt = rule_points(type, X.any_tri);
  const long long items = (long long)V.Ne * npt;
  const long long chunk = field_chunk_items(items);
  const long long i0 = (long long)ic * chunk, i1 = min(items, i0 + chunk);
  if (i1 <= (long long)fes * npt || i0 >= (long long)fee * npt) {
    // nothing of this chunk is summed here
    if (lane < 14) dst[lane] = 0.0;
    return;
  }
  const double k = V.k;
  // Accumulators and the per-item coefficients live in LDS (per lane), so that
  // only the offset stays in registers across the Hessian lattice sum; the empty
  // asm statements keep the compiler from caching LDS values across it.
  //   sacc: Pa, dPa[3], PVa[3] (re, im)
  //   scf:  alpha = wt aW dPi, beta = -wt sg Pi, gamma[c] = wt aW dpv_c,
  //         delta[c] = -wt sg pv_c (re, im), W[3], the offset r[3]
  __shared__ double sacc[14][64];
  __shared__ double scf[22][64];
#pragma unroll
  for (int i = 0; i < 14; i++) sacc[i][lane] = 0.0;
  GfTally tl{0, 0, 0, 0};
  long long evals = 0;
  int err = 0;
  for (long long it = i0 + lane; it < i1; it += 64) {
    const long long fpos = it / npt;
    if (fpos < fes || fpos >= fee) continue;
    const int s = V.fperm[fpos], q = (int)(it % npt);
    const int sd = V.sid[s], as = abs(sd);
    ElemC E;
    if (type != 0) load_elem(M, as, E);
    double r[3], w[3], wt, Nv[4];
    if (!source_point(M, T, E, as, type, q, r, w, wt, Nv)) continue;
    double vR[3] = {r[0] - r0[0], r[1] - r0[1], r[2] - r0[2]};
    double aR = abs3(vR), iR = 1.0 / aR, aW = abs3(w);
    double rdW = dot3(vR, w) * iR;
    const double sg = (!periodic && sd < 0) ? -1.0 : 1.0;
    cd Pi = {}, dPi = {}, pvi[3] = {}, dpvi[3] = {};
    for (int m = 0; m < 4; m++) {
      if (Nv[m] == 0.0) continue;
      Pi += ld(V.P, 4 * s + m) * Nv[m];
      dPi += ld(V.dP, 4 * s + m) * Nv[m];
      for (int c = 0; c < 3; c++) {
        pvi[c] += ld(V.pv, 12 * s + 3 * m + c) * Nv[m];
        dpvi[c] += ld(V.dpv, 12 * s + 3 * m + c) * Nv[m];
      }
    }
    // per item: Pa += g alpha + (dg.W) beta, dPa_c += -dg_c alpha - (H W)_c beta,
    // PVa_c += g gamma_c + (dg.W) delta_c, with G = g aW and H = dg.W the kernel
    // values of the representation formula (free space or lattice sum)
    const cd alpha = dPi * (wt * aW), beta = Pi * (-wt * sg);
    // one item's share: g, dg.W, dg_c and (H W)_c against the coefficients
    // al, be (scalar part) and gamma_c, delta_c (LDS)
    auto accumulate = [&](const cd &gv, const cd &dgw, const cd *dgp, const cd *hwp, const cd &al, const cd &be) {
      {
        const cd v = gv * al + dgw * be;
        sacc[0][lane] += v.x; sacc[1][lane] += v.y;
      }
      for (int c2 = 0; c2 < 3; c2++) {
        const cd d = -(dgp[c2] * al) - hwp[c2] * be;
        sacc[2 + 2 * c2][lane] += d.x; sacc[3 + 2 * c2][lane] += d.y;
        const cd gm = cd{scf[4 + 2 * c2][lane], scf[5 + 2 * c2][lane]};
        const cd dl = cd{scf[10 + 2 * c2][lane], scf[11 + 2 * c2][lane]};
        const cd pvv = gv * gm + dgw * dl;
        sacc[8 + 2 * c2][lane] += pvv.x; sacc[9 + 2 * c2][lane] += pvv.y;
      }
    };
    if (periodic) {
      scf[0][lane] = alpha.x; scf[1][lane] = alpha.y; scf[2][lane] = beta.x; scf[3][lane] = beta.y;
      for (int c2 = 0; c2 < 3; c2++) {
        const cd gm = dpvi[c2] * (wt * aW), dl = pvi[c2] * (-wt * sg);
        scf[4 + 2 * c2][lane] = gm.x; scf[5 + 2 * c2][lane] = gm.y;
        scf[10 + 2 * c2][lane] = dl.x; scf[11 + 2 * c2][lane] = dl.y;
        scf[16 + c2][lane] = w[c2];
      }
      scf[19][lane] = vR[0]; scf[20][lane] = vR[1]; scf[21][lane] = vR[2];
      // the image share and the spectral share of the Ewald split in two
      // evaluations (the result is linear in them), so that neither share is
      // held in registers across the other's sum
      for (int part = 1; part <= 2; part++) {
        asm volatile("" ::: "memory");
        const double x0 = scf[19][lane], x1 = scf[20][lane], x2 = scf[21][lane];
        Sum10 gf;
        // part 1: the image share or the whole pure spectral sum; part 2: the spectral share
        int c = part == 1 ? qpgf_eval<GF_V2, 5>(P, x0, x1, x2, 1.0e-15, gf, tl)
                          : qpgf_eval<GF_V2, 2>(P, x0, x1, x2, 1.0e-15, gf, tl);
        if (c < 0) err = c;
        asm volatile("" ::: "memory");
        const double W0 = scf[16][lane], W1 = scf[17][lane], W2 = scf[18][lane];
        // Hessian as a symmetric matrix: xx, yy, zz, xy, yz, zx
        const cd *h = gf.d2g;
        const cd dgw = gf.dg[0] * W0 + gf.dg[1] * W1 + gf.dg[2] * W2;
        cd hwp[3];
        hwp[0] = h[0] * W0 + h[3] * W1 + h[5] * W2;
        hwp[1] = h[3] * W0 + h[1] * W1 + h[4] * W2;
        hwp[2] = h[5] * W0 + h[4] * W1 + h[2] * W2;
        accumulate(gf.g, dgw, gf.dg, hwp, cd{scf[0][lane], scf[1][lane]}, cd{scf[2][lane], scf[3][lane]});
      }
      evals++;
    } else {
      // free space: G = e^{ikR}/(4 pi R) aW, dG/dr_c and (H W)_c in closed form
      for (int c2 = 0; c2 < 3; c2++) {
        const cd gm = dpvi[c2] * (wt * aW), dl = pvi[c2] * (-wt * sg);
        scf[4 + 2 * c2][lane] = gm.x; scf[5 + 2 * c2][lane] = gm.y;
        scf[10 + 2 * c2][lane] = dl.x; scf[11 + 2 * c2][lane] = dl.y;
      }
      cd ce = cis(k * aR), cb = cd{-1.0, k * aR};
      const cd gv = ce * (BQD2_IFP * iR);
      const cd dgw = cb * ce * (BQD2_IFP * rdW * iR * iR);
      cd c3 = cd{k * k * aR * aR + 3.0 * cb.x, 3.0 * cb.y};
      cd dgp[3], hwp[3];
      for (int c2 = 0; c2 < 3; c2++) {
        dgp[c2] = cb * ce * (BQD2_IFP * iR * iR * vR[c2] * iR);
        hwp[c2] = -(ce * (BQD2_IFP * iR * iR * iR) * (c3 * (vR[c2] * iR * rdW) - cb * w[c2]));
      }
      accumulate(gv, dgw, dgp, hwp, alpha, beta);
    }
  }
  asm volatile("" ::: "memory");
  cd Pa = cd{sacc[0][lane], sacc[1][lane]}, dPa[3], PVa[3];
  for (int c2 = 0; c2 < 3; c2++) {
    dPa[c2] = cd{sacc[2 + 2 * c2][lane], sacc[3 + 2 * c2][lane]};
    PVa[c2] = cd{sacc[8 + 2 * c2][lane], sacc[9 + 2 * c2][lane]};
  }
  Pa = cd{wave_sum(Pa.x), wave_sum(Pa.y)};
  for (int c2 = 0; c2 < 3; c2++) {
    dPa[c2] = cd{wave_sum(dPa[c2].x), wave_sum(dPa[c2].y)};
    PVa[c2] = cd{wave_sum(PVa[c2].x), wave_sum(PVa[c2].y)};
  }
  long long tsp = wave_sum_ll(tl.sp), tpr = wave_sum_ll(tl.pr), tev = wave_sum_ll(tl.ev), tfd = wave_sum_ll(tl.fd);
  evals = wave_sum_ll(evals);
  int anyerr = __any(err != 0);
  if (lane != 0) return;
  if (evals) tally_add(counters + CNT_FIELD, tsp, tpr, tev, tfd, evals);
  if (anyerr) atomicAdd((unsigned long long *)&counters[CNT_ERR], 1ull);
  dst[0] = Pa.x;
  dst[1] = Pa.y;
  for (int c2 = 0; c2 < 3; c2++) {
    dst[2 + 2 * c2] = dPa[c2].x;
    dst[3 + 2 * c2] = dPa[c2].y;
    dst[8 + 2 * c2] = PVa[c2].x;
    dst[9 + 2 * c2] = PVa[c2].y;
  }
}

// F and grad F (the boundary's solid angle seen from the point, which picks
// the form of the representation formula, and its gradient) of every chunk of
// a field point's items, one wavefront per point, into slots 14..17 of
// k_field_part's chunk records.  Apart from the lattice sums this is the only
// work on the elements far below or above a point, so it runs apart from
// k_field_part's register-heavy wavefronts.
__global__ void __launch_bounds__(64) k_field_F(DevMesh M, const Tab *__restrict__ tab, FieldCtx X, int type, int n,
                                                int nchunk, const double *__restrict__ r0s,
                                                const int *__restrict__ did_io, double *__restrict__ part) {
  const int ip = blockIdx.x, lane = threadIdx.x;
  if (ip >= n) return;
  const int did = did_io[ip];
  double *rec = part + (size_t)ip * nchunk * FP_N;
  if (did < 0) {
    for (int c = lane; c < nchunk; c += 64)
      for (int i = 14; i < FP_N; i++) rec[(size_t)c * FP_N + i] = 0.0;
    return;
  }
  const Tab &T = *tab;
  const DomView &V = X.dom[did];
  const bool periodic = (X.ps == 1 && did == 0);
  const double r0[3] = {r0s[3 * ip], r0s[3 * ip + 1], r0s[3 * ip + 2]};
  const int npt = rule_points(type, X.any_tri);
  const long long items = (long long)V.Ne * npt;
  const long long chunk = field_chunk_items(items);
  for (int ic = 0; ic < nchunk; ic++) {
    const long long i0 = (long long)ic * chunk, i1 = min(items, i0 + chunk);
    double F = 0.0, dF[3] = {0, 0, 0};
    for (long long it = i0 + lane; it < i1; it += 64) {
      const int s = V.fperm[it / npt], q = (int)(it % npt);
      const int sd = V.sid[s], as = abs(sd);
      ElemC E;
      if (type != 0) load_elem(M, as, E);
      double r[3], w[3], wt, Nv[4];
      if (!source_point(M, T, E, as, type, q, r, w, wt, Nv)) continue;
      double vR[3] = {r[0] - r0[0], r[1] - r0[1], r[2] - r0[2]};
      double aR = abs3(vR), iR = 1.0 / aR;
      double rdW = dot3(vR, w) * iR;
      const double sg = (!periodic && sd < 0) ? -1.0 : 1.0;
      F += sg * wt * rdW * iR * iR;
      for (int c2 = 0; c2 < 3; c2++) dF[c2] += sg * wt * iR * iR * iR * (3.0 * vR[c2] * iR * rdW - w[c2]);
    }
    F = wave_sum(F);
    for (int c2 = 0; c2 < 3; c2++) dF[c2] = wave_sum(dF[c2]);
    if (lane == 0) {
      double *dst = rec + (size_t)ic * FP_N;
      dst[14] = F;
      for (int c2 = 0; c2 < 3; c2++) dst[15 + c2] = dF[c2];
    }
  }
}

// one thread per field point: chunk sums in order, then the representation
// formula's normalisation, Bloch phase and incident field (out: p, pv per point)
__global__ void __launch_bounds__(256) k_field_fin(FieldCtx X, int total, int n, int nchunk,
                                                   const double *__restrict__ part, const int *__restrict__ ls,
                                                   const double *__restrict__ pts, int *__restrict__ did_io,
                                                   double *__restrict__ out) {
  const int ip = blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= n) return;
  const int did = did_io[ip];
  if (did < 0) {
    for (int i = 0; i < 8; i++) out[8 * ip + i] = 0.0;
    return;
  }
  const DomView &V = X.dom[did];
  double acc[FP_N];
  for (int i = 0; i < FP_N; i++) acc[i] = 0.0;
  for (int c = 0; c < nchunk; c++) {
    const double *q = part + ((size_t)ip * nchunk + c) * FP_N;
    for (int i = 0; i < FP_N; i++) acc[i] += q[i];
  }
  const cd Pa = cd{acc[0], acc[1]};
  cd dPa[3], PVa[3];
  double dF[3];
  for (int c2 = 0; c2 < 3; c2++) {
    dPa[c2] = cd{acc[2 + 2 * c2], acc[3 + 2 * c2]};
    PVa[c2] = cd{acc[8 + 2 * c2], acc[9 + 2 * c2]};
    dF[c2] = acc[15 + c2] * BQD2_IFP;
  }
  const double F = acc[14] * BQD2_IFP;
  const double iC = did == 0 ? 1.0 / (1.0 + F) : 1.0 / F;
  const cd mk2 = cd{-V.k2x, -V.k2y};
  cd p, pv[3];
  if (fabs(dF[0]) > CBD_CDF_D || fabs(dF[1]) > CBD_CDF_D || fabs(dF[2]) > CBD_CDF_D) {
    // near the boundary: boundary-value form with the stored pv, dpv
    p = mk2 * Pa * iC;
    for (int c2 = 0; c2 < 3; c2++) pv[c2] = PVa[c2] * iC;
  } else {
    p = mk2 * Pa * iC;
    for (int c2 = 0; c2 < 3; c2++) pv[c2] = -(dPa[c2] - Pa * dF[c2]) * iC;
  }
  const int l1 = ls[2 * ip], l2 = ls[2 * ip + 1];
  if (l1 != 0 || l2 != 0) {
    double arg = X.vkx * ((double)l1 * X.d1x + (double)l2 * X.d2x) + X.vky * ((double)l1 * X.d1y + (double)l2 * X.d2y);
    cd ce = cis(arg);
    p *= ce;
    for (int c2 = 0; c2 < 3; c2++) pv[c2] *= ce;
  }
  if (total && did == 0) {
    const double rr[3] = {pts[3 * ip], pts[3 * ip + 1], pts[3 * ip + 2]};
    cd pi, vi[3];
    incident_pv(X, rr, pi, vi);
    p += pi;
    for (int c2 = 0; c2 < 3; c2++) pv[c2] += vi[c2];
  }
  out[8 * ip] = p.x;
  out[8 * ip + 1] = p.y;
  for (int c2 = 0; c2 < 3; c2++) {
    out[8 * ip + 2 + 2 * c2] = pv[c2].x;
    out[8 * ip + 3 + 2 * c2] = pv[c2].y;
  }
}

// ------------------------------------------------------------ far field
// The source points y of the periodic boundary that lie at least h below a
// field point x (or at least h above it) have r_z = y_z - x_z of one sign.  For
// them the lattice sum is the pure spectral series (d3_qpgf_d2.c:484-533,
// :887-950, what fd_cell sums per pair)
//   G(r) = (i/2A) sum_n exp(i a_n . r_t + i b_n |r_z|) / b_n,   r = y - x,
//   a_n = -k_t + n1 g1 + n2 g2,  b_n = sqrt(k^2 - |a_n|^2), Im b_n >= 0,
// whose gradient takes i kappa_n and Hessian -kappa_n kappa_n^T per term,
// kappa_n = (a_n, b_n sgn r_z), and every term is a target factor times a source
// factor.  So the representation formula's sums over those items (k_field_part's
// alpha, beta, gamma_c, delta_c, weight vector w) become per order four
// amplitudes, summed once for all points,
//   A_n   = (i/2A)/b_n sum_items Y_n (alpha   + i (kappa_n . w) beta),
//   B_n,c = (i/2A)/b_n sum_items Y_n (gamma_c + i (kappa_n . w) delta_c),
// and per point a sum over the orders
//   Pa = sum_n X_n A_n,  dPa_c = sum_n X_n (-i kappa_n,c) A_n,
//   PVa_c = sum_n X_n B_n,c.
// The items are the element bands of the field order (FarBand): the amplitudes
// of bands 0..b (side 0, sources below, reference height zpre_b) and of bands
// b..FAR_NB-1 (side 1, sources above, reference height zsuf_b) are kept for
// every b, and a point takes the largest set below and above it that clears h.
// Side 0: Y_n = exp(i a_n . y_t + i b_n (zpre_b - y_z)),
//         X_n = exp(-i a_n . x_t + i b_n (x_z - zpre_b)), kappa_z = -b_n;
// side 1: Y_n = exp(i a_n . y_t + i b_n (y_z - zsuf_b)),
//         X_n = exp(-i a_n . x_t + i b_n (zsuf_b - x_z)), kappa_z = b_n:
// no factor exceeds 1 in size.  The orders are the grid of FarGeom, which keeps
// every order that has not decayed by 1e-20 over the height h.

// order j of the grid: a_n and b_n; false at a Wood anomaly (b_n = 0)
__device__ __forceinline__ bool far_order(const GfParams &P, const FarGeom &F, int j, double &ax, double &ay, cd &bn) {
  const int n1 = F.n1lo + j / F.n2n, n2 = F.n2lo + j % F.n2n;
  ax = -P.vkx + n1 * P.g1x + n2 * P.g2x;
  ay = -P.vky + n1 * P.g1y + n2 * P.g2y;
  const double b2 = P.k * P.k - ax * ax - ay * ay;
  bn = b2 > 0.0 ? cd{sqrt(b2), 0.0} : cd{0.0, sqrt(-b2)};
  return b2 != 0.0;
}

// exp(i ph) exp(i bn d), d >= 0
__device__ __forceinline__ cd far_phase(double ph, const cd &bn, double d) {
  if (bn.y == 0.0) return cis(ph + bn.x * d);
  return cis(ph) * exp(-bn.y * d);
}

// band amplitudes: blockIdx.x blocks of 64 orders, blockIdx.y the side,
// blockIdx.z the band; the band's items pass through LDS 64 at a time.  The
// source factors of side 0 refer to the band's zpre, those of side 1 to its
// zsuf, so none exceeds 1 in size.  amp: (side, band, order) x 8 doubles.
__global__ void __launch_bounds__(64) k_far_amp(DevMesh M, const Tab *__restrict__ tab, FieldCtx X, GfParams P,
                                                FarGeom F, int type, double *__restrict__ amp) {
  const int lane = threadIdx.x, j = blockIdx.x * 64 + lane, side = blockIdx.y, band = blockIdx.z;
  const Tab &T = *tab;
  const DomView &V = X.dom[0];
  const FarBand Bd = X.far_bands[band];
  const int npt = rule_points(type, X.any_tri);
  const long long i0 = (long long)Bd.es * npt, i1 = (long long)Bd.ee * npt;
  double ax = 0.0, ay = 0.0;
  cd bn = {0.0, 0.0};
  const bool ok = j < F.G;
  if (ok) far_order(P, F, j, ax, ay, bn);
  const cd kz = side == 0 ? -bn : bn;
  const double zref = side == 0 ? Bd.zpre : Bd.zsuf;
  // per item: y (3), w (3), alpha, beta, gamma[3], delta[3] (re, im)
  __shared__ double sit[22][64];
  cd A = {0.0, 0.0}, B[3] = {};
  for (long long b0 = i0; b0 < i1; b0 += 64) {
    const long long it = b0 + lane;
    double y[3] = {0.0, 0.0, zref}, w[3] = {0.0, 0.0, 0.0};
    cd al = {}, be = {}, gm[3] = {}, dl[3] = {};
    if (it < i1) {
      const int s = V.fperm[it / npt], q = (int)(it % npt);
      const int as = abs(V.sid[s]);
      ElemC E;
      if (type != 0) load_elem(M, as, E);
      double r[3], wv[3], wt, Nv[4];
      if (source_point(M, T, E, as, type, q, r, wv, wt, Nv)) {
        cd Pi = {}, dPi = {}, pvi[3] = {}, dpvi[3] = {};
        for (int m = 0; m < 4; m++) {
          if (Nv[m] == 0.0) continue;
          Pi += ld(V.P, 4 * s + m) * Nv[m];
          dPi += ld(V.dP, 4 * s + m) * Nv[m];
          for (int c = 0; c < 3; c++) {
            pvi[c] += ld(V.pv, 12 * s + 3 * m + c) * Nv[m];
            dpvi[c] += ld(V.dpv, 12 * s + 3 * m + c) * Nv[m];
          }
        }
        const double aW = abs3(wv);
        al = dPi * (wt * aW);
        be = Pi * (-wt);
        for (int c = 0; c < 3; c++) {
          gm[c] = dpvi[c] * (wt * aW);
          dl[c] = pvi[c] * (-wt);
          y[c] = r[c];
          w[c] = wv[c];
        }
      }
    }
    for (int c = 0; c < 3; c++) {
      sit[c][lane] = y[c];
      sit[3 + c][lane] = w[c];
      sit[10 + 2 * c][lane] = gm[c].x;
      sit[11 + 2 * c][lane] = gm[c].y;
      sit[16 + 2 * c][lane] = dl[c].x;
      sit[17 + 2 * c][lane] = dl[c].y;
    }
    sit[6][lane] = al.x;
    sit[7][lane] = al.y;
    sit[8][lane] = be.x;
    sit[9][lane] = be.y;
    __syncthreads();
    const int nb = (int)min((long long)64, i1 - b0);
    if (ok) {
      for (int i = 0; i < nb; i++) {
        const double d = side == 0 ? zref - sit[2][i] : sit[2][i] - zref;
        const cd Y = far_phase(ax * sit[0][i] + ay * sit[1][i], bn, d);
        const cd ikw = mul_i(cd{ax * sit[3][i] + ay * sit[4][i], 0.0} + kz * sit[5][i]);
        A += Y * (cd{sit[6][i], sit[7][i]} + ikw * cd{sit[8][i], sit[9][i]});
        for (int c = 0; c < 3; c++)
          B[c] += Y * (cd{sit[10 + 2 * c][i], sit[11 + 2 * c][i]} + ikw * cd{sit[16 + 2 * c][i], sit[17 + 2 * c][i]});
      }
    }
    __syncthreads();
  }
  if (!ok) return;
  double *dst = amp + (((size_t)side * FAR_NB + band) * F.G + j) * 8;
  dst[0] = A.x;
  dst[1] = A.y;
  for (int c = 0; c < 3; c++) {
    dst[2 + 2 * c] = B[c].x;
    dst[3 + 2 * c] = B[c].y;
  }
}

// the bands' amplitudes summed over the bands below (side 0, in rising band
// order) or above (side 1, falling), each sum moved to the next band's reference
// height on the way, in place, then times (i/2A)/b_n.  A Wood anomaly (b_n = 0)
// counts as an error, as in fd_cell.
__global__ void __launch_bounds__(256) k_far_fin(GfParams P, FarGeom F, const FarBand *__restrict__ bands,
                                                 double *__restrict__ amp, long long amp_terms,
                                                 long long *__restrict__ counters) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx == 0) atomicAdd((unsigned long long *)&counters[CNT_FAR + 2], (unsigned long long)amp_terms);
  if (idx >= 2 * F.G) return;
  const int side = idx / F.G, j = idx % F.G;
  double ax, ay;
  cd bn;
  const bool good = far_order(P, F, j, ax, ay, bn);
  if (!good) atomicAdd((unsigned long long *)&counters[CNT_ERR], 1ull);
  const cd cf = good ? cd{0.0, 0.5 * P.aidd} * crecip(bn) : cd{0.0, 0.0};
  cd C[4] = {};
  for (int k = 0; k < FAR_NB; k++) {
    const int b = side == 0 ? k : FAR_NB - 1 - k;
    double *a = amp + (((size_t)side * FAR_NB + b) * F.G + j) * 8;
    if (k) {
      const int bp = side == 0 ? b - 1 : b + 1;
      const double shift = side == 0 ? bands[b].zpre - bands[bp].zpre : bands[bp].zsuf - bands[b].zsuf;
      const cd f = far_phase(0.0, bn, shift);
      for (int i = 0; i < 4; i++) C[i] = C[i] * f;
    }
    for (int i = 0; i < 4; i++) {
      C[i] += cd{a[2 * i], a[2 * i + 1]};
      const cd v = C[i] * cf;
      a[2 * i] = v.x;
      a[2 * i + 1] = v.y;
    }
  }
}

// the amplitude parts of the points' sums, one wavefront per point with the
// orders strided over its lanes and summed across them in a fixed order (bands
// below, then above), added to the point's first chunk of partial sums (which
// hold the lattice sums of the other bands and F, grad F).
//   BD = 0: field points (r0s, did_io), k_field_part's chunks of FP_N doubles;
//   BD = 1: boundary points of domain bd_did (element position tpos, local
//           coordinates ze), k_field_bd's chunks of BP_N doubles (no gradient)
template <int BD>
__global__ void __launch_bounds__(64) k_far_eval(DevMesh M, FieldCtx X, GfParams P, FarGeom F, int n, int nchunk,
                                                 const double *__restrict__ r0s, const int *__restrict__ did_io,
                                                 const int *__restrict__ tpos, const double *__restrict__ ze,
                                                 int bd_did, const double *__restrict__ amp,
                                                 double *__restrict__ part, long long *__restrict__ counters) {
  const int ip = blockIdx.x, lane = threadIdx.x;
  if (ip >= n) return;
  const int did = BD ? bd_did : did_io[ip];
  if (did < 0) return;
  double x0, x1, x2;
  if (BD) {
    // the boundary point as k_field_bd places it
    ElemC Et;
    load_elem(M, abs(X.dom[did].sid[tpos[ip]]), Et);
    double rt[3], wt_[3];
    bil_rw(Et, ze[2 * ip], ze[2 * ip + 1], rt, wt_);
    x0 = rt[0];
    x1 = rt[1];
    x2 = rt[2];
  } else {
    x0 = r0s[3 * ip];
    x1 = r0s[3 * ip + 1];
    x2 = r0s[3 * ip + 2];
  }
  int b0, b1, es, ee;
  far_split(X, did, x2, b0, b1, es, ee);
  const int nsides = (b0 >= 0) + (b1 < FAR_NB);
  if (!nsides) return;
  if (lane == 0) {
    atomicAdd((unsigned long long *)&counters[CNT_FAR], 1ull);
    atomicAdd((unsigned long long *)&counters[CNT_FAR + 1], (unsigned long long)F.G * nsides);
  }
  cd Pa = {}, dPa[3] = {}, PVa[3] = {};
  for (int side = 0; side < 2; side++) {
    const int b = side == 0 ? b0 : b1;
    if (b < 0 || b >= FAR_NB) continue;
    const double d = side == 0 ? x2 - X.far_bands[b].zpre : X.far_bands[b].zsuf - x2;
    const double *am = amp + ((size_t)side * FAR_NB + b) * F.G * 8;
    for (int j = lane; j < F.G; j += 64) {
      double ax, ay;
      cd bn;
      far_order(P, F, j, ax, ay, bn);
      const cd Xn = far_phase(-(ax * x0 + ay * x1), bn, d);
      const double *a = am + (size_t)j * 8;
      const cd XA = Xn * cd{a[0], a[1]};
      Pa += XA;
      if (!BD) {
        // -i kappa_c X_n A_n
        dPa[0] += mul_i(XA) * (-ax);
        dPa[1] += mul_i(XA) * (-ay);
        dPa[2] += mul_i(XA * (side == 0 ? bn : -bn));
      }
      for (int c = 0; c < 3; c++) PVa[c] += Xn * cd{a[2 + 2 * c], a[3 + 2 * c]};
    }
  }
  Pa = cd{wave_sum(Pa.x), wave_sum(Pa.y)};
  for (int c = 0; c < 3; c++) {
    if (!BD) dPa[c] = cd{wave_sum(dPa[c].x), wave_sum(dPa[c].y)};
    PVa[c] = cd{wave_sum(PVa[c].x), wave_sum(PVa[c].y)};
  }
  if (lane != 0) return;
  if (BD) {
    double *dst = part + (size_t)ip * nchunk * BP_N;
    dst[0] += Pa.x;
    dst[1] += Pa.y;
    for (int c = 0; c < 3; c++) {
      dst[2 + 2 * c] += PVa[c].x;
      dst[3 + 2 * c] += PVa[c].y;
    }
    return;
  }
  double *dst = part + (size_t)ip * nchunk * FP_N;
  dst[0] += Pa.x;
  dst[1] += Pa.y;
  for (int c = 0; c < 3; c++) {
    dst[2 + 2 * c] += dPa[c].x;
    dst[3 + 2 * c] += dPa[c].y;
    dst[8 + 2 * c] += PVa[c].x;
    dst[9 + 2 * c] += PVa[c].y;
  }
}

// ------------------------------------------------------------ boundary points
// Boundary point (domain did, element position t, zeta, eta): the regular
// (element, quadrature point) items are cut into chunks of BD_CHUNK, one
// wavefront per (point, chunk), and one more wavefront per point takes the
// point's own element on its polar rule; k_field_bd_fin sums the chunks in
// order.  The chunk count depends only on the mesh and the rule.
__host__ __device__ __forceinline__ int bd_item_chunks(long long items) {
  return (int)((items + BD_CHUNK - 1) / BD_CHUNK);
}

// SELF = 0: the item chunks (blockIdx.y = chunk); SELF = 1: the point's own
// element (one wavefront per point, chunk slot nci).  Two instantiations keep
// the polar rule's register needs out of the chunks' lattice-sum loop.
template <int SELF>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 2))) k_field_bd(DevMesh M, const Tab *__restrict__ tab, FieldCtx X, GfParams P,
                                                 int type, int did, int n, const int *__restrict__ tpos,
                                                 const double *__restrict__ ze, double *__restrict__ part,
                                                 long long *__restrict__ counters) {
  const int ip = blockIdx.x, lane = threadIdx.x;
  if (ip >= n) return;
  const Tab &T = *tab;
  const DomView &V = X.dom[did];
  const bool periodic = (X.ps == 1 && did == 0);
  const int t = tpos[ip];
  const double zt = ze[2 * ip], et = ze[2 * ip + 1];
  const int td = V.sid[t], at = abs(td);
  double rt[3];
  {
    ElemC Et;
    load_elem(M, at, Et);
    double wt_[3];
    bil_rw(Et, zt, et, rt, wt_);
  }
  const int npt = rule_points(type, X.any_tri);
  const double k = V.k;
  // Regular elements: accumulators and per-item coefficients in LDS (per lane),
  // so that only the offset stays in registers across the lattice sum; the empty
  // asm statements keep the compiler from caching LDS values across it.
  //   bacc: Pa, PVa[3] (re, im)
  //   bcf:  alpha = wt aW dPi, beta = -wt sg Pi, gamma[c] = wt aW dpv_c,
  //         delta[c] = -wt sg pv_c (re, im), W[3]
  __shared__ double bacc[8][64];
  __shared__ double bcf[19][64];
#pragma unroll
  for (int i = 0; i < 8; i++) bacc[i][lane] = 0.0;
  double F = 0.0;
  GfTally tl{0, 0, 0, 0};
  long long evals = 0;
  int err = 0;
  // regular elements
  const long long items = (long long)V.Ne * npt;
  const int nci = bd_item_chunks(items);
  const int ic = SELF ? nci : (int)blockIdx.y;
  const long long i0 = (long long)ic * BD_CHUNK, i1 = SELF ? 0 : min(items, i0 + BD_CHUNK);
  // elements of bands far below or above the point: their lattice sums from
  // k_far_eval; F of every item of the chunks from k_field_bd_F
  int fb0, fb1, fes, fee;
  far_split(X, did, rt[2], fb0, fb1, fes, fee);
  if (!SELF && (i1 <= (long long)fes * npt || i0 >= (long long)fee * npt)) {
    // nothing of this chunk is summed here
    if (lane < 8) part[((size_t)ip * (nci + 1) + ic) * BP_N + lane] = 0.0;
    return;
  }
  for (long long it = i0 + lane; it < i1; it += 64) {
    const long long fpos = it / npt;
    if (fpos < fes || fpos >= fee) continue;
    const int s = V.fperm[fpos], q = (int)(it % npt);
    if (s == t) continue;
    const int sd = V.sid[s], as = abs(sd);
    ElemC E;
    if (type != 0) load_elem(M, as, E);
    double r[3], w[3], wt, Nv[4];
    if (!source_point(M, T, E, as, type, q, r, w, wt, Nv)) continue;
    double vR[3] = {r[0] - rt[0], r[1] - rt[1], r[2] - rt[2]};
    double aR = abs3(vR), iR = 1.0 / aR, aW = abs3(w);
    double tD = dot3(vR, w) * iR * iR * iR;
    const double sg = (!periodic && sd < 0) ? -1.0 : 1.0;
    cd Pi = {}, dPi = {}, pvi[3] = {}, dpvi[3] = {};
    for (int m = 0; m < 4; m++) {
      if (Nv[m] == 0.0) continue;
      Pi += ld(V.P, 4 * s + m) * Nv[m];
      dPi += ld(V.dP, 4 * s + m) * Nv[m];
      for (int c = 0; c < 3; c++) {
        pvi[c] += ld(V.pv, 12 * s + 3 * m + c) * Nv[m];
        dpvi[c] += ld(V.dpv, 12 * s + 3 * m + c) * Nv[m];
      }
    }
    // Pa += g alpha + (dg.W) beta, PVa_c += g gamma_c + (dg.W) delta_c with the
    // kernel value g and its normal derivative dg.W (free space or lattice sum)
    {
      const cd al = dPi * (wt * aW), be = Pi * (-wt * sg);
      bcf[0][lane] = al.x; bcf[1][lane] = al.y; bcf[2][lane] = be.x; bcf[3][lane] = be.y;
      for (int c2 = 0; c2 < 3; c2++) {
        const cd gm = dpvi[c2] * (wt * aW), dl = pvi[c2] * (-wt * sg);
        bcf[4 + 2 * c2][lane] = gm.x; bcf[5 + 2 * c2][lane] = gm.y;
        bcf[10 + 2 * c2][lane] = dl.x; bcf[11 + 2 * c2][lane] = dl.y;
        bcf[16 + c2][lane] = w[c2];
      }
    }
    cd gv, dgw;
    if (periodic) {
      const double x0 = vR[0], x1 = vR[1], x2 = vR[2];
      Sum10 gf;
      int c = qpgf_eval<GF_V1>(P, x0, x1, x2, 1.0e-15, gf, tl);
      if (c < 0) err = c;
      evals++;
      gv = gf.g;
      dgw = gf.dg[0] * bcf[16][lane] + gf.dg[1] * bcf[17][lane] + gf.dg[2] * bcf[18][lane];
    } else {
      cd ce = cis(k * aR), cb = cd{-1.0, k * aR};
      gv = ce * (BQD2_IFP * iR);
      dgw = cb * ce * (BQD2_IFP * tD);
    }
    {
      const cd v = gv * cd{bcf[0][lane], bcf[1][lane]} + dgw * cd{bcf[2][lane], bcf[3][lane]};
      bacc[0][lane] += v.x; bacc[1][lane] += v.y;
    }
    for (int c2 = 0; c2 < 3; c2++) {
      const cd v = gv * cd{bcf[4 + 2 * c2][lane], bcf[5 + 2 * c2][lane]} +
                   dgw * cd{bcf[10 + 2 * c2][lane], bcf[11 + 2 * c2][lane]};
      bacc[2 + 2 * c2][lane] += v.x; bacc[3 + 2 * c2][lane] += v.y;
    }
  }
  F *= BQD2_IFP;
  ElemC Et;
  load_elem(M, at, Et);
  if (!SELF) {
    // item chunk: no self element
  } else if (M.tri[at]) {
    // self triangle: polar rule over the three sub-triangles about (zt, et)
    // (lit_coef_bd, lit_coef.c:290-350; q0_lit_coef_bd, q0_lit_coef.c:187-262).
    // The element is flat, so the free-space H and F vanish; the periodic H
    // comes from the other lattice images.
    const int s = t;
    double r[4], th[4];
    lit_corners(zt, et, r, th);
    const double Vn[3] = {Et.cw[0][0], Et.cw[1][0], Et.cw[2][0]};
    const double aV = abs3(Vn), J = 2.0 / (3.0 * BQD2_SQ3) * aV;
    const double Wn[3] = {Vn[0] * (J / aV), Vn[1] * (J / aV), Vn[2] * (J / aV)};
    const int npts = 3 * 64 * 2 * 12;
    bool split = false;
    if (periodic) {
      bool zero = false;
      for (int p = lane; p < npts; p += 64) {
        int i = p % 12, q = p / 12, hh = q % 2, m = (q / 2) % 64, j = q / 128;
        double tt = theta_node(th, j, hh, T.xhi[m]);
        double rm = polar_rm(r, th, j, tt);
        double c = cos(tt), sn = sin(tt), v[3];
        double rr = 0.5 * rm * (T.xli[i] + 1.0);
        for (int l = 0; l < 3; l++) v[l] = rr * (Et.cr[l][1] * c + Et.cr[l][2] * sn);
        zero |= z_degenerate(v[0], v[1], v[2]);
      }
      for (int p = lane; p < 3 * 144; p += 64) {
        double z, e, wq, q[3];
        rule_point(true, 12, T.xli, T.wli, p, z, e, wq);
        bil_rw_r(Et, z, e, q);
        zero |= z_degenerate(q[0] - rt[0], q[1] - rt[1], q[2] - rt[2]);
      }
      split = images_far(Et, rt, P) && !__any(zero);
    }
    cd Ps = {}, PVs[3] = {};
    for (int p = lane; p < npts; p += 64) {
      int i = p % 12, q = p / 12, hh = q % 2, m = (q / 2) % 64, j = q / 128;
      double tt = theta_node(th, j, hh, T.xhi[m]);
      double rm = polar_rm(r, th, j, tt);
      double c = cos(tt), sn = sin(tt);
      double rr = 0.5 * rm * (T.xli[i] + 1.0);
      double z = rr * c + zt, e = rr * sn + et, K[3];
      for (int l = 0; l < 3; l++) K[l] = Et.cr[l][1] * c + Et.cr[l][2] * sn;
      double wgt = 0.25 * (th[j + 1] - th[j]) * T.whi[m] * 0.5 * rm * T.wli[i];
      cd Pi = {}, dPi = {}, pvi[3] = {}, dpvi[3] = {};
      for (int mm = 0; mm < 3; mm++) {
        double Nn = lit_N(mm, z, e);
        Pi += ld(V.P, 4 * s + mm) * Nn;
        dPi += ld(V.dP, 4 * s + mm) * Nn;
        for (int cc = 0; cc < 3; cc++) {
          pvi[cc] += ld(V.pv, 12 * s + 3 * mm + cc) * Nn;
          dpvi[cc] += ld(V.dpv, 12 * s + 3 * mm + cc) * Nn;
        }
      }
      cd a, b = {};
      if (periodic) {
        double vR[3] = {rr * K[0], rr * K[1], rr * K[2]};
        Sum10 gf;
        if (split) {
          const double aR = abs3(vR);
          free_gf(P.k, vR, aR, 1.0 / aR, gf.g, gf.dg);
        } else {
          int cnt = qpgf_eval_g_dg(P, vR[0], vR[1], vR[2], 1.0e-15, gf, tl);
          if (cnt < 0) err = cnt;
          evals++;
        }
        a = gf.g * (J * rr);
        b = (gf.dg[0] * Wn[0] + gf.dg[1] * Wn[1] + gf.dg[2] * Wn[2]) * rr;
      } else {
        const double aK = abs3(K);
        a = cis(k * aK * rr) * (BQD2_IFP * J / aK);
      }
      Ps += (a * dPi - b * Pi) * wgt;
      for (int c2 = 0; c2 < 3; c2++) PVs[c2] += (a * dpvi[c2] - b * pvi[c2]) * wgt;
    }
    if (split) {
      // smooth remainder of the lattice sum on the triangle's GL12 parts
      for (int p = lane; p < 3 * 144; p += 64) {
        double z, e, wgt, vR[3], W[3];
        rule_point(true, 12, T.xli, T.wli, p, z, e, wgt);
        wgt *= 1.0 / 24.0;
        bil_rw(Et, z, e, vR, W);
        for (int l = 0; l < 3; l++) vR[l] -= rt[l];
        const double aR = abs3(vR);
        Sum10 gf;
        int cnt = qpgf_eval<GF_V1>(P, vR[0], vR[1], vR[2], 1.0e-15, gf, tl);
        if (cnt < 0) err = cnt;
        evals++;
        cd g0, dg0[3];
        free_gf(P.k, vR, aR, 1.0 / aR, g0, dg0);
        const cd a = (gf.g - g0) * aV;
        const cd b = (gf.dg[0] - dg0[0]) * Vn[0] + (gf.dg[1] - dg0[1]) * Vn[1] + (gf.dg[2] - dg0[2]) * Vn[2];
        cd Pi = {}, dPi = {}, pvi[3] = {}, dpvi[3] = {};
        for (int mm = 0; mm < 3; mm++) {
          double Nn = lit_N(mm, z, e);
          Pi += ld(V.P, 4 * s + mm) * Nn;
          dPi += ld(V.dP, 4 * s + mm) * Nn;
          for (int cc = 0; cc < 3; cc++) {
            pvi[cc] += ld(V.pv, 12 * s + 3 * mm + cc) * Nn;
            dpvi[cc] += ld(V.dpv, 12 * s + 3 * mm + cc) * Nn;
          }
        }
        Ps += (a * dPi - b * Pi) * wgt;
        for (int c2 = 0; c2 < 3; c2++) PVs[c2] += (a * dpvi[c2] - b * pvi[c2]) * wgt;
      }
    }
    bacc[0][lane] += Ps.x; bacc[1][lane] += Ps.y;
    for (int c2 = 0; c2 < 3; c2++) { bacc[2 + 2 * c2][lane] += PVs[c2].x; bacc[3 + 2 * c2][lane] += PVs[c2].y; }
  } else {
  // self element: polar rule about (zt, et)
    const int sd = td, s = t;
    const double sg = (!periodic && sd < 0) ? -1.0 : 1.0;
    double r[5], th[5];
    bil_corners(zt, et, r, th);
    const bool plane = bil_is_plane(Et);
    const double bdC = Et.cr[0][1] * Et.cw[0][2] + Et.cr[1][1] * Et.cw[1][2] + Et.cr[2][1] * Et.cw[2][2];
    cd Ps = {}, PVs[3] = {};
    double Fs = 0.0;
    // periodic domain: free-space image on the polar rule plus the smooth
    // remainder of the lattice sum on GL12 (as in the assembly), unless another
    // image is close or a point has a z offset of exactly zero
    bool split = false;
    if (periodic) {
      bool zero = false;
      for (int p = lane; p < 4 * 64 * 2 * 12; p += 64) {
        int i = p % 12, q = p / 12, hh = q % 2, m = (q / 2) % 64, j = q / 128;
        double tt = theta_node(th, j, hh, T.xhi[m]);
        double rm = polar_rm(r, th, j, tt);
        double c = cos(tt), sn = sin(tt), K[3];
        double rr = 0.5 * rm * (T.xli[i] + 1.0);
        polar_K(Et, c, sn, rr, zt, et, K);
        zero |= z_degenerate(rr * K[0], rr * K[1], rr * K[2]);
      }
      for (int p = lane; p < 144; p += 64) {
        double q[3];
        bil_rw_r(Et, T.xli[p / 12], T.xli[p % 12], q);
        zero |= z_degenerate(q[0] - rt[0], q[1] - rt[1], q[2] - rt[2]);
      }
      split = images_far(Et, rt, P) && !__any(zero);
    }
    for (int p = lane; p < 4 * 64 * 2 * 12; p += 64) {
      int i = p % 12, q = p / 12, hh = q % 2, m = (q / 2) % 64, j = q / 128;
      double tt = theta_node(th, j, hh, T.xhi[m]);
      double rm = polar_rm(r, th, j, tt);
      double c = cos(tt), sn = sin(tt);
      double rr = 0.5 * rm * (T.xli[i] + 1.0);
      double z = rr * c + zt, e = rr * sn + et, K[3], W[3];
      polar_K(Et, c, sn, rr, zt, et, K);
      for (int l = 0; l < 3; l++) W[l] = Et.cw[l][0] + Et.cw[l][1] * z + Et.cw[l][2] * e;
      double aK = abs3(K), aW = abs3(W);
      double wgt = 0.25 * (th[j + 1] - th[j]) * T.whi[m] * 0.5 * rm * T.wli[i];
      cd Pi = {}, dPi = {}, pvi[3] = {}, dpvi[3] = {};
      for (int mm = 0; mm < 4; mm++) {
        double Nn = bil_N(mm, z, e);
        Pi += ld(V.P, 4 * s + mm) * Nn;
        dPi += ld(V.dP, 4 * s + mm) * Nn;
        for (int cc = 0; cc < 3; cc++) {
          pvi[cc] += ld(V.pv, 12 * s + 3 * mm + cc) * Nn;
          dpvi[cc] += ld(V.dpv, 12 * s + 3 * mm + cc) * Nn;
        }
      }
      cd a, b;
      if (periodic) {
        double vR[3] = {rr * K[0], rr * K[1], rr * K[2]};
        Sum10 gf;
        if (split) {
          const double aR = abs3(vR);
          free_gf(P.k, vR, aR, 1.0 / aR, gf.g, gf.dg);
        } else {
          int cnt = qpgf_eval_g_dg(P, vR[0], vR[1], vR[2], 1.0e-15, gf, tl);
          if (cnt < 0) err = cnt;
          evals++;
        }
        a = gf.g * (aW * rr);
        b = (gf.dg[0] * W[0] + gf.dg[1] * W[1] + gf.dg[2] * W[2]) * rr;
      } else {
        cd ce = cis(k * aK * rr), cb = cd{-1.0, k * rr * aK};
        a = ce * (BQD2_IFP * aW / aK);
        b = plane ? cd{0.0, 0.0} : cb * ce * (BQD2_IFP * bdC / (aK * aK * aK) * c * sn);
      }
      Ps += (a * dPi - b * Pi * sg) * wgt;
      for (int c2 = 0; c2 < 3; c2++) PVs[c2] += (a * dpvi[c2] - b * pvi[c2] * sg) * wgt;
      if (!plane) Fs += wgt * (1.0 / (aK * aK * aK) * c * sn);
    }
    if (split) {
      for (int p = lane; p < 144; p += 64) {
        const int i = p / 12, j = p - i * 12;
        const double z = T.xli[i], e = T.xli[j], wgt = T.wli[i] * T.wli[j];
        double vR[3], W[3];
        bil_rw(Et, z, e, vR, W);
        for (int l = 0; l < 3; l++) vR[l] -= rt[l];
        const double aR = abs3(vR), aW = abs3(W);
        Sum10 gf;
        int cnt = qpgf_eval<GF_V1>(P, vR[0], vR[1], vR[2], 1.0e-15, gf, tl);
        if (cnt < 0) err = cnt;
        evals++;
        cd g0, dg0[3];
        free_gf(P.k, vR, aR, 1.0 / aR, g0, dg0);
        const cd a = (gf.g - g0) * aW;
        const cd b = (gf.dg[0] - dg0[0]) * W[0] + (gf.dg[1] - dg0[1]) * W[1] + (gf.dg[2] - dg0[2]) * W[2];
        cd Pi = {}, dPi = {}, pvi[3] = {}, dpvi[3] = {};
        for (int mm = 0; mm < 4; mm++) {
          double Nn = bil_N(mm, z, e);
          Pi += ld(V.P, 4 * s + mm) * Nn;
          dPi += ld(V.dP, 4 * s + mm) * Nn;
          for (int cc = 0; cc < 3; cc++) {
            pvi[cc] += ld(V.pv, 12 * s + 3 * mm + cc) * Nn;
            dpvi[cc] += ld(V.dpv, 12 * s + 3 * mm + cc) * Nn;
          }
        }
        Ps += (a * dPi - b * Pi * sg) * wgt;
        for (int c2 = 0; c2 < 3; c2++) PVs[c2] += (a * dpvi[c2] - b * pvi[c2] * sg) * wgt;
      }
    }
    bacc[0][lane] += Ps.x; bacc[1][lane] += Ps.y;
    for (int c2 = 0; c2 < 3; c2++) { bacc[2 + 2 * c2][lane] += PVs[c2].x; bacc[3 + 2 * c2][lane] += PVs[c2].y; }
    F += sg * Fs * BQD2_IFP * bdC;
  }
  asm volatile("" ::: "memory");
  cd Pa = cd{bacc[0][lane], bacc[1][lane]}, PVa[3];
  for (int c2 = 0; c2 < 3; c2++) PVa[c2] = cd{bacc[2 + 2 * c2][lane], bacc[3 + 2 * c2][lane]};
  Pa = cd{wave_sum(Pa.x), wave_sum(Pa.y)};
  for (int c2 = 0; c2 < 3; c2++) PVa[c2] = cd{wave_sum(PVa[c2].x), wave_sum(PVa[c2].y)};
  F = wave_sum(F);
  long long tsp = wave_sum_ll(tl.sp), tpr = wave_sum_ll(tl.pr), tev = wave_sum_ll(tl.ev), tfd = wave_sum_ll(tl.fd);
  evals = wave_sum_ll(evals);
  int anyerr = __any(err != 0);
  if (lane != 0) return;
  if (evals) tally_add(counters + CNT_FIELD, tsp, tpr, tev, tfd, evals);
  if (anyerr) atomicAdd((unsigned long long *)&counters[CNT_ERR], 1ull);
  double *dst = part + ((size_t)ip * (nci + 1) + ic) * BP_N;
  dst[0] = Pa.x;
  dst[1] = Pa.y;
  for (int c2 = 0; c2 < 3; c2++) {
    dst[2 + 2 * c2] = PVa[c2].x;
    dst[3 + 2 * c2] = PVa[c2].y;
  }
  if (SELF) dst[8] = F;  // the item chunks' F comes from k_field_bd_F
}

// F (the boundary's solid angle seen from the boundary point, without its own
// element) of every item chunk of k_field_bd, one wavefront per point, into
// slot 8 of the chunk records: the work on the elements far below or above the
// point apart from the lattice sums, kept out of k_field_bd's register-heavy
// wavefronts
__global__ void __launch_bounds__(64) k_field_bd_F(DevMesh M, const Tab *__restrict__ tab, FieldCtx X, int type,
                                                   int did, int n, const int *__restrict__ tpos,
                                                   const double *__restrict__ ze, double *__restrict__ part) {
  const int ip = blockIdx.x, lane = threadIdx.x;
  if (ip >= n) return;
  const Tab &T = *tab;
  const DomView &V = X.dom[did];
  const bool periodic = (X.ps == 1 && did == 0);
  const int t = tpos[ip];
  double rt[3];
  {
    ElemC Et;
    load_elem(M, abs(V.sid[t]), Et);
    double wt_[3];
    bil_rw(Et, ze[2 * ip], ze[2 * ip + 1], rt, wt_);
  }
  const int npt = rule_points(type, X.any_tri);
  const long long items = (long long)V.Ne * npt;
  const int nci = bd_item_chunks(items);
  for (int ic = 0; ic < nci; ic++) {
    const long long i0 = (long long)ic * BD_CHUNK, i1 = min(items, i0 + BD_CHUNK);
    double F = 0.0;
    for (long long it = i0 + lane; it < i1; it += 64) {
      const int s = V.fperm[it / npt], q = (int)(it % npt);
      if (s == t) continue;
      const int sd = V.sid[s], as = abs(sd);
      ElemC E;
      if (type != 0) load_elem(M, as, E);
      double r[3], w[3], wt, Nv[4];
      if (!source_point(M, T, E, as, type, q, r, w, wt, Nv)) continue;
      double vR[3] = {r[0] - rt[0], r[1] - rt[1], r[2] - rt[2]};
      double aR = abs3(vR), iR = 1.0 / aR;
      double tD = dot3(vR, w) * iR * iR * iR;
      const double sg = (!periodic && sd < 0) ? -1.0 : 1.0;
      F += sg * wt * tD;
    }
    F *= BQD2_IFP;
    F = wave_sum(F);
    if (lane == 0) part[((size_t)ip * (nci + 1) + ic) * BP_N + 8] = F;
  }
}

// one thread per boundary point: the chunk sums in order, the normalisation
// 1 / (1 + F) or 1 / F and, for the total field, the incident wave
// (pv_s_bd_aqd2 / pv_t_bd_aqd2, bem3_aw_qd2_field.c:394-469)
__global__ void __launch_bounds__(256) k_field_bd_fin(DevMesh M, FieldCtx X, int total, int did, int n, int nch,
                                                      const int *__restrict__ tpos, const double *__restrict__ ze,
                                                      const double *__restrict__ part, double *__restrict__ out) {
  const int ip = blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= n) return;
  const DomView &V = X.dom[did];
  cd Pa = {}, PVa[3] = {};
  double F = 0.0;
  for (int ic = 0; ic < nch; ic++) {
    const double *q = part + ((size_t)ip * nch + ic) * BP_N;
    Pa += cd{q[0], q[1]};
    for (int c2 = 0; c2 < 3; c2++) PVa[c2] += cd{q[2 + 2 * c2], q[3 + 2 * c2]};
    F += q[8];
  }
  const double iC = did == 0 ? 1.0 / (1.0 + F) : 1.0 / F;
  cd p = cd{-V.k2x, -V.k2y} * Pa * iC, pv[3];
  for (int c2 = 0; c2 < 3; c2++) pv[c2] = PVa[c2] * iC;
  if (total && did == 0) {
    ElemC Et;
    load_elem(M, abs(V.sid[tpos[ip]]), Et);
    double rt[3], wt_[3];
    bil_rw(Et, ze[2 * ip], ze[2 * ip + 1], rt, wt_);
    cd pi, vi[3];
    incident_pv(X, rt, pi, vi);
    p += pi;
    for (int c2 = 0; c2 < 3; c2++) pv[c2] += vi[c2];
  }
  out[8 * ip] = p.x;
  out[8 * ip + 1] = p.y;
  for (int c2 = 0; c2 < 3; c2++) {
    out[8 * ip + 2 + 2 * c2] = pv[c2].x;
    out[8 * ip + 3 + 2 * c2] = pv[c2].y;
  }
}

// boundary field at n points already on the device: partial sums in part
// (n x (chunks + 1) x BP_N doubles), results in out (n x 8); amp: the far
// field's amplitudes of rule type `type` (domain 0 with X.far_on), else null
static int launch_field_bd(bqd2_problem *p, hipStream_t st, FieldCtx &X, int type, int total, int did, int n,
                           const int *d_t, const double *d_ze, double *d_part, double *d_out, long long *ctr,
                           const double *amp) {
  const int nch = bd_item_chunks((long long)p->dom[did].Ne * rule_points(type, p->any_tri)) + 1;
  if (nch > 1) {
    hipLaunchKernelGGL(k_field_bd<0>, dim3(n, nch - 1), dim3(64), 0, st, p->m, p->tab, X, p->gf, type, did, n, d_t,
                       d_ze, d_part, ctr);
    hipLaunchKernelGGL(k_field_bd_F, dim3(n), dim3(64), 0, st, p->m, p->tab, X, type, did, n, d_t, d_ze, d_part);
  }
  hipLaunchKernelGGL(k_field_bd<1>, dim3(n, 1), dim3(64), 0, st, p->m, p->tab, X, p->gf, type, did, n, d_t, d_ze,
                     d_part, ctr);
  if (amp)
    hipLaunchKernelGGL(k_far_eval<1>, dim3(n), dim3(64), 0, st, p->m, X, p->gf, p->far, n, nch,
                       (const double *)nullptr, (const int *)nullptr, d_t, d_ze, did, amp, d_part, ctr);
  hipLaunchKernelGGL(k_field_bd_fin, dim3(bqd2_blocks(n, 256)), dim3(256), 0, st, p->m, X, total, did, n, nch, d_t,
                     d_ze, d_part, d_out);
  BQD2_HIP(hipGetLastError());
  return BQD2_OK;
}

static size_t bd_part_bytes(const bqd2_problem *p, int type, int did, int n) {
  const int nch = bd_item_chunks((long long)p->dom[did].Ne * rule_points(type, p->any_tri)) + 1;
  return (size_t)n * nch * BP_N * sizeof(double);
}

// ------------------------------------------------------------ force
// per open-domain element: stress tensor at 4 nodes (type 0) or 9 points
// (type 1, boundary field from k_field_bd in bdval) (bil_force_4p/9p)
__global__ void __launch_bounds__(256) k_force_elem(DevMesh M, const Tab *__restrict__ tab, FieldCtx X, int type,
                                                    const double *__restrict__ bdval, const int *__restrict__ bdoff,
                                                    double rcx, double rcy, double rcz, double *__restrict__ part) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const DomView &V = X.dom[0];
  if (s > V.Ne) return;
  const Tab &T = *tab;
  const int as = abs(V.sid[s]);
  const bool tri = M.tri[as] != 0;
  const double c1 = 0.25 / (X.pw_c0 * X.pw_c0 * X.pw_rho0), c2 = 0.25 * X.pw_rho0, c3 = 0.5 * X.pw_rho0;
  const double rc[3] = {rcx, rcy, rcz};
  ElemC E;
  load_elem(M, as, E);
  const double *bv = bdval ? bdval + 8 * (size_t)bdoff[s - 1] : nullptr;
  double F[3] = {0, 0, 0}, N[3] = {0, 0, 0};
  const int npt = type == 0 ? 4 : tri ? 7 : 9;
  for (int i = 0; i < npt; i++) {
    double r[3], w[3], wt;
    cd p, v[3];
    if (type == 0 && !(tri && i == 3)) {
      for (int d = 0; d < 3; d++) {
        r[d] = M.ren[12 * as + 3 * i + d];
        w[d] = M.wen[12 * as + 3 * i + d];
      }
      wt = tri ? T.wt_34[i] : T.wt_44[i];
      incident_pv(X, r, p, v);
      p -= ld(V.P, 4 * s + i) * cd{X.dom[0].k2x, X.dom[0].k2y};
      for (int d = 0; d < 3; d++) v[d] += ld(V.pv, 12 * s + 3 * i + d);
    } else {
      // boundary field at a Gauss point (triangles: the centroid of the 4-point
      // rule, or the 7-point rule), computed beforehand by k_field_bd
      const double z = type == 0 ? T.zt_34[3] : tri ? T.zt_37[i] : T.zt_49[i];
      const double e = type == 0 ? T.et_34[3] : tri ? T.et_37[i] : T.et_49[i];
      bil_rw(E, z, e, r, w);
      wt = type == 0 ? T.wt_34[3] : tri ? T.wt_37[i] : T.wt_49[i];
      const double *o = bv + 8 * (type == 0 ? 0 : i);
      p = cd{o[0], o[1]};
      for (int d = 0; d < 3; d++) v[d] = cd{o[2 + 2 * d], o[3 + 2 * d]};
    }
    double pp = p.x * p.x + p.y * p.y;
    double vv = v[0].x * v[0].x + v[0].y * v[0].y + v[1].x * v[1].x + v[1].y * v[1].y + v[2].x * v[2].x + v[2].y * v[2].y;
    double t1 = c1 * pp - c2 * vv, Tm[9];
    for (int l = 0; l < 3; l++)
      for (int m = 0; m < 3; m++) Tm[3 * l + m] = c3 * (v[l].x * v[m].x + v[l].y * v[m].y);
    Tm[0] += t1; Tm[4] += t1; Tm[8] += t1;
    for (int l = 0; l < 3; l++) F[l] += (Tm[3 * l] * w[0] + Tm[3 * l + 1] * w[1] + Tm[3 * l + 2] * w[2]) * wt;
    double d0 = r[0] - rc[0], d1 = r[1] - rc[1], d2 = r[2] - rc[2];
    N[0] += ((d1 * Tm[6] - d2 * Tm[3]) * w[0] + (d1 * Tm[7] - d2 * Tm[4]) * w[1] + (d1 * Tm[8] - d2 * Tm[5]) * w[2]) * wt;
    N[1] += ((d2 * Tm[0] - d0 * Tm[6]) * w[0] + (d2 * Tm[1] - d0 * Tm[7]) * w[1] + (d2 * Tm[2] - d0 * Tm[8]) * w[2]) * wt;
    N[2] += ((d0 * Tm[3] - d1 * Tm[0]) * w[0] + (d0 * Tm[4] - d1 * Tm[1]) * w[1] + (d0 * Tm[5] - d1 * Tm[2]) * w[2]) * wt;
  }
  const double sc = tri ? 0.5 : 1.0;  // reference triangle area (lit_force_4p/7p)
  for (int d = 0; d < 3; d++) {
    part[6 * (s - 1) + d] = F[d] * sc;
    part[6 * (s - 1) + 3 + d] = N[d] * sc;
  }
}

// ordered sum of the per-element parts (one wavefront, strided partials then a
// fixed reduction tree: deterministic run to run)
__global__ void __launch_bounds__(64) k_force_sum(int ne, const double *__restrict__ part, double *__restrict__ FN) {
  const int lane = threadIdx.x;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int s = lane; s < ne; s += 64)
    for (int d = 0; d < 6; d++) acc[d] += part[6 * s + d];
  for (int d = 0; d < 6; d++) {
    double v = wave_sum(acc[d]);
    if (lane == 0) FN[d] = v;
  }
}

// ------------------------------------------------------------ entry points
static FieldCtx make_ctx(const bqd2_problem *p) {
  FieldCtx X;
  memset(&X, 0, sizeof(X));
  X.MN = p->m.MN;
  X.ps = p->m.ps;
  X.any_tri = p->any_tri;
  for (int d = 0; d <= p->m.MN; d++) {
    const DevDomain &D = p->dom[d];
    X.dom[d].Ne = D.Ne;
    X.dom[d].sid = D.sid;
    X.dom[d].fperm = D.fperm;
    X.dom[d].P = D.P;
    X.dom[d].dP = D.dP;
    X.dom[d].pv = D.pv;
    X.dom[d].dpv = D.dpv;
    X.dom[d].k = p->k0[d];
    X.dom[d].k2x = p->k2[d][0];
    X.dom[d].k2y = p->k2[d][1];
  }
  memcpy(X.pw_p, p->pw_p, sizeof(X.pw_p));
  X.pw_rho0 = p->pw_rho0;
  X.pw_c0 = p->pw_c0;
  X.pw_k0 = p->pw_k0;
  memcpy(X.pw_tv, p->pw_tv, sizeof(X.pw_tv));
  memcpy(X.pw_R, p->pw_R, sizeof(X.pw_R));
  X.d1x = p->qd8[4]; X.d1y = p->qd8[5]; X.d2x = p->qd8[6]; X.d2y = p->qd8[7];
  X.idd = 1.0 / (X.d1x * X.d2y - X.d2x * X.d1y);
  X.vkx = p->qd8[1];
  X.vky = p->qd8[2];
  X.far_on = (p->m.ps == 1 && p->far.on && p->far_enabled) ? 1 : 0;
  X.far_ztop = p->far.ztop;
  X.far_zbot = p->far.zbot;
  X.far_h = p->far.h;
  X.far_bands = p->far_bands;
  return X;
}

// device bytes of the far field's amplitudes (2 sides x FAR_NB bands x G orders)
static size_t far_amp_bytes(const bqd2_problem *p) { return (size_t)2 * FAR_NB * p->far.G * 8 * sizeof(double); }

// the far field's amplitudes of rule type `type` into amp, on stream st
static int far_launch(bqd2_problem *p, const FieldCtx &X, int type, double *amp, long long *ctr, hipStream_t st) {
  const FarGeom &F = p->far;
  const long long items = (long long)p->dom[0].Ne * rule_points(type, X.any_tri);
  hipLaunchKernelGGL(k_far_amp, dim3((F.G + 63) / 64, 2, FAR_NB), dim3(64), 0, st, p->m, p->tab, X, p->gf, F, type,
                     amp);
  hipLaunchKernelGGL(k_far_fin, dim3(bqd2_blocks(2 * F.G, 256)), dim3(256), 0, st, p->gf, F, p->far_bands, amp,
                     2 * (long long)F.G * items, ctr);
  BQD2_HIP(hipGetLastError());
  return BQD2_OK;
}

// the amplitudes of the current solution for rule type `type`: made on the
// caller's stream when the solution changed since they were last made (under
// the problem's far-field lock), else the caller's stream waits for the call
// that made them
static int far_amplitudes(bqd2_problem *p, CallCtx *cx, const FieldCtx &X, int type, long long *ctr,
                          const double **amp) {
  std::lock_guard<std::mutex> g(p->far_mu);
  const unsigned long long want = p->sol_epoch.load() + 1;
  if (p->far_amp_epoch[type] == want) {
    BQD2_HIP(hipStreamWaitEvent(cx->st, p->far_ev[type], 0));
    *amp = p->far_amp[type];
    return BQD2_OK;
  }
  int rc;
  if (!p->far_amp[type] && (rc = bqd2_alloc(p, (void **)&p->far_amp[type], far_amp_bytes(p)))) return rc;
  if (!p->far_ev[type]) BQD2_HIP(hipEventCreateWithFlags(&p->far_ev[type], hipEventDisableTiming));
  if ((rc = far_launch(p, X, type, p->far_amp[type], ctr, cx->st))) return rc;
  BQD2_HIP(hipEventRecord(p->far_ev[type], cx->st));
  p->far_amp_epoch[type] = want;
  *amp = p->far_amp[type];
  return BQD2_OK;
}

extern "C" int bqd2_set_field_far(bqd2_problem *p, int on) {
  BQD2_CHECK(p && (on == 0 || on == 1), BQD2_E_ARG, "bqd2_set_field_far: bad arguments");
  p->far_enabled = on;
  return BQD2_OK;
}

extern "C" int bqd2_set_field_timing(bqd2_problem *p, int on) {
  BQD2_CHECK(p && (on == 0 || on == 1), BQD2_E_ARG, "bqd2_set_field_timing: bad arguments");
  p->field_timing = on;
  return BQD2_OK;
}

extern "C" int bqd2_field_kernel_ms(bqd2_problem *p, double *ms) {
  BQD2_CHECK(p && ms, BQD2_E_ARG, "bqd2_field_kernel_ms: bad arguments");
  *ms = p->field_ms_part;
  return BQD2_OK;
}

extern "C" int bqd2_field_far_info(bqd2_problem *p, int *on, int *orders, double *zbot, double *ztop, double *h) {
  BQD2_CHECK(p, BQD2_E_ARG, "bqd2_field_far_info: null problem");
  if (on) *on = (p->m.ps == 1 && p->far.on && p->far_enabled) ? 1 : 0;
  if (orders) *orders = p->far.G;
  if (zbot) *zbot = p->far.zbot;
  if (ztop) *ztop = p->far.ztop;
  if (h) *h = p->far.h;
  return BQD2_OK;
}

// per-call device buffers carved from the problem's workspace, which grows
// on demand and is kept between calls (single points from a caller's loop
// cost no allocation); calls on one problem are serialised by the caller
static size_t pad256(size_t b) { return (b + 255) & ~(size_t)255; }
struct Scratch {
  char *base = nullptr;
  size_t off = 0, cap = 0;
  int init(CallCtx *c, size_t total) {
    cap = total;
    return bqd2_workspace(c, total, &base);
  }
  int get(void **q, size_t bytes) {
    const size_t b = pad256(bytes ? bytes : 16);
    if (off + b > cap) {
      bqd2_set_error("workspace overflow (%zu + %zu > %zu bytes)", off, b, cap);
      return BQD2_E_NOMEM;
    }
    *q = base + off;
    off += b;
    return BQD2_OK;
  }
};

extern "C" int bqd2_field(bqd2_problem *p, int n, const double *pts, int type, int total, double *out, int *did,
                          long long *gf_evals) {
  BQD2_CHECK(p && n >= 0 && pts && out, BQD2_E_ARG, "bqd2_field: bad arguments");
  BQD2_CHECK(type >= 0 && type <= 3, BQD2_E_ARG, "bqd2_field: type must be 0..3 (4-point, 9-point, GLN, GHN)");
  if (n == 0) return BQD2_OK;
  CallCtx *cx = bqd2_call_ctx(p);
  if (!cx) return BQD2_E_HIP;
  hipStream_t st = cx->st;
  long long *const ctr = cx->cnt;
  Scratch S;
  double *d_pts, *d_r0, *d_out, *d_part;
  int *d_l, *d_did;
  int rc;
  const size_t nn = (size_t)n;
  const int npt = rule_points(type, p->any_tri);
  int nchunk = 1;
  for (int d = 0; d <= p->m.MN; d++) {
    const long long items = (long long)p->dom[d].Ne * npt, chunk = field_chunk_items(items);
    nchunk = std::max(nchunk, (int)((items + chunk - 1) / chunk));
  }
  FieldCtx X = make_ctx(p);
  const bool any_far = X.far_on != 0;  // the amplitudes serve every point of domain 0
  if ((rc = S.init(cx, pad256(nn * 24) * 2 + pad256(nn * 64) + pad256(nn * 8) + pad256(nn * 4) +
                          pad256(nn * nchunk * FP_N * 8))))
    return rc;
  if ((rc = S.get((void **)&d_pts, (size_t)n * 24)) || (rc = S.get((void **)&d_r0, (size_t)n * 24)) ||
      (rc = S.get((void **)&d_out, (size_t)n * 64)) || (rc = S.get((void **)&d_l, (size_t)n * 8)) ||
      (rc = S.get((void **)&d_did, (size_t)n * 4)) || (rc = S.get((void **)&d_part, nn * nchunk * FP_N * 8)))
    return rc;
  BQD2_HIP(hipMemcpyAsync(d_pts, pts, (size_t)n * 24, hipMemcpyHostToDevice, st));
  BQD2_HIP(hipMemsetAsync(ctr, 0, CNT_N * sizeof(long long), st));
  const double *amp = nullptr;
  if (any_far && (rc = far_amplitudes(p, cx, X, type, ctr, &amp))) return rc;
  hipLaunchKernelGGL(k_domain_id, dim3(bqd2_blocks(n, 4)), dim3(256), 0, st, p->m, X, p->rn, p->ed, n, d_pts, d_did,
                     d_r0, d_l);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (p->field_timing) {
    BQD2_HIP(hipEventCreate(&e0));
    BQD2_HIP(hipEventCreate(&e1));
    BQD2_HIP(hipEventRecord(e0, st));
  }
  hipLaunchKernelGGL(k_field_part, dim3(n, nchunk), dim3(64), 0, st, p->m, p->tab, X, p->gf, type, n, nchunk, d_r0,
                     d_did, d_part, ctr);
  if (e1) BQD2_HIP(hipEventRecord(e1, st));
  hipLaunchKernelGGL(k_field_F, dim3(n), dim3(64), 0, st, p->m, p->tab, X, type, n, nchunk, d_r0, d_did, d_part);
  if (any_far)
    hipLaunchKernelGGL(k_far_eval<0>, dim3(n), dim3(64), 0, st, p->m, X, p->gf, p->far, n, nchunk, d_r0, d_did,
                       (const int *)nullptr, (const double *)nullptr, 0, amp, d_part,
                       ctr);
  hipLaunchKernelGGL(k_field_fin, dim3(bqd2_blocks(n, 256)), dim3(256), 0, st, X, total, n, nchunk, d_part, d_l, d_pts,
                     d_did, d_out);
  BQD2_HIP(hipGetLastError());
  BQD2_HIP(hipMemcpyAsync(out, d_out, (size_t)n * 64, hipMemcpyDeviceToHost, st));
  if (did) BQD2_HIP(hipMemcpyAsync(did, d_did, (size_t)n * 4, hipMemcpyDeviceToHost, st));
  long long cnt[CNT_N];
  BQD2_HIP(hipMemcpyAsync(cnt, ctr, sizeof(cnt), hipMemcpyDeviceToHost, st));
  BQD2_HIP(hipStreamSynchronize(st));
  if (e1) {
    float ms = 0.0f;
    BQD2_HIP(hipEventElapsedTime(&ms, e0, e1));
    p->field_ms_part = ms;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  if (gf_evals) *gf_evals = cnt[CNT_FIELD + 4];
  BQD2_CHECK(cnt[CNT_ERR] == 0, BQD2_E_GF, "bqd2_field: Green's function error at %lld points", cnt[CNT_ERR]);
  return BQD2_OK;
}

// ------------------------------------------------- field without allocation
// bqd2_field_noalloc's device memory: one file-scope array per GPU, held until
// the process ends.  Host code never passes the array's name to a kernel (on
// the host it is only a stand-in for the symbol): its device address is looked
// up once per device with hipGetSymbolAddress and kept in noalloc_base[].
static __device__ __align__(256) char scratch[BQD2_NOALLOC_SCRATCH_BYTES];
static constexpr int NOALLOC_MAX_DEV = 64;
static char *noalloc_base[NOALLOC_MAX_DEV];
static std::mutex noalloc_mu;

// the current device's copy of the scratch array (set up on the first call on
// that device, with the device current)
static int noalloc_scratch(char **base) {
  int dev = 0;
  BQD2_HIP(hipGetDevice(&dev));
  BQD2_CHECK(dev >= 0 && dev < NOALLOC_MAX_DEV, BQD2_E_ARG, "bqd2_field_noalloc: device %d out of range", dev);
  std::lock_guard<std::mutex> g(noalloc_mu);
  if (!noalloc_base[dev]) {
    char *ptr = nullptr;
    BQD2_HIP(hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch)));
    noalloc_base[dev] = ptr;
  }
  *base = noalloc_base[dev];
  return BQD2_OK;
}

__global__ void k_zero_counters(long long *c) {
  if (threadIdx.x < CNT_N) c[threadIdx.x] = 0;
}

__global__ void k_copy_errors(const long long *c, long long *errors) {
  if (threadIdx.x == 0) *errors = c[CNT_ERR];
}

extern "C" int bqd2_field_noalloc(bqd2_problem *p, int n, const double *d_pts, int type, int total, double *d_out,
                                  int *d_did, long long *d_errors, void *stream_) {
  hipStream_t stream = (hipStream_t)stream_;
  BQD2_CHECK(p && n >= 0 && (n == 0 || (d_pts && d_out && d_did)), BQD2_E_ARG, "bqd2_field_noalloc: bad arguments");
  BQD2_CHECK(type >= 0 && type <= 3, BQD2_E_ARG, "bqd2_field_noalloc: type must be 0..3 (4-point, 9-point, GLN, GHN)");
  if (n == 0) return BQD2_OK;
  const size_t nn = (size_t)n;
  const int npt = rule_points(type, p->any_tri);
  int nchunk = 1;
  for (int d = 0; d <= p->m.MN; d++) {
    const long long items = (long long)p->dom[d].Ne * npt, chunk = field_chunk_items(items);
    nchunk = std::max(nchunk, (int)((items + chunk - 1) / chunk));
  }
  FieldCtx X = make_ctx(p);
  // scratch layout: counters, r0 (n x 3), lattice cell (n x 2 int), chunk
  // partials, then the far field's amplitudes (made in every call: the same
  // kernels on the same data as bqd2_field's, so the same bits)
  const size_t o_r0 = 256, o_l = o_r0 + pad256(nn * 24), o_part = o_l + pad256(nn * 8);
  const size_t o_famp = o_part + pad256(nn * nchunk * FP_N * 8);
  const size_t need = o_famp + (X.far_on ? far_amp_bytes(p) : 0);
  BQD2_CHECK(need <= (size_t)BQD2_NOALLOC_SCRATCH_BYTES, BQD2_E_ARG,
             "bqd2_field_noalloc: %d points of type %d need %zu bytes of scratch, more than its %zu bytes "
             "(at most %zu points per call on this problem)",
             n, type, need, (size_t)BQD2_NOALLOC_SCRATCH_BYTES,
             (size_t)((BQD2_NOALLOC_SCRATCH_BYTES - 256 - 768 - (need - o_famp)) / (24 + 8 + (size_t)nchunk * FP_N * 8)));
  char *base = nullptr;
  int rc;
  if ((rc = noalloc_scratch(&base))) return rc;
  long long *ctr = (long long *)base;
  double *d_r0 = (double *)(base + o_r0), *d_part = (double *)(base + o_part);
  int *d_l = (int *)(base + o_l);
  double *d_famp = (double *)(base + o_famp);
  hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, stream, ctr);
  if (X.far_on && (rc = far_launch(p, X, type, d_famp, ctr, stream))) return rc;
  hipLaunchKernelGGL(k_domain_id, dim3(bqd2_blocks(n, 4)), dim3(256), 0, stream, p->m, X, p->rn, p->ed, n, d_pts,
                     d_did, d_r0, d_l);
  hipLaunchKernelGGL(k_field_part, dim3(n, nchunk), dim3(64), 0, stream, p->m, p->tab, X, p->gf, type, n, nchunk, d_r0,
                     d_did, d_part, ctr);
  hipLaunchKernelGGL(k_field_F, dim3(n), dim3(64), 0, stream, p->m, p->tab, X, type, n, nchunk, d_r0, d_did, d_part);
  if (X.far_on)
    hipLaunchKernelGGL(k_far_eval<0>, dim3(n), dim3(64), 0, stream, p->m, X, p->gf, p->far, n, nchunk, d_r0, d_did,
                       (const int *)nullptr, (const double *)nullptr, 0, d_famp, d_part, ctr);
  hipLaunchKernelGGL(k_field_fin, dim3(bqd2_blocks(n, 256)), dim3(256), 0, stream, X, total, n, nchunk, d_part, d_l,
                     d_pts, d_did, d_out);
  if (d_errors) hipLaunchKernelGGL(k_copy_errors, dim3(1), dim3(64), 0, stream, ctr, d_errors);
  BQD2_HIP(hipGetLastError());
  return BQD2_OK;
}

extern "C" int bqd2_field_bd(bqd2_problem *p, int did, int n, const int *t, const double *ze, int type, int total,
                             double *out) {
  BQD2_CHECK(p && did >= 0 && did <= p->m.MN && n >= 0 && t && ze && out, BQD2_E_ARG, "bqd2_field_bd: bad arguments");
  BQD2_CHECK(type >= 0 && type <= 3, BQD2_E_ARG, "bqd2_field_bd: type must be 0..3");
  if (n == 0) return BQD2_OK;
  for (int i = 0; i < n; i++)
    BQD2_CHECK(t[i] >= 1 && t[i] <= p->dom[did].Ne, BQD2_E_ARG, "bqd2_field_bd: element %d out of range", t[i]);
  CallCtx *cx = bqd2_call_ctx(p);
  if (!cx) return BQD2_E_HIP;
  hipStream_t st = cx->st;
  long long *const ctr = cx->cnt;
  Scratch S;
  int *d_t;
  double *d_ze, *d_out, *d_part;
  int rc;
  const size_t bpart = bd_part_bytes(p, type, did, n);
  if ((rc = S.init(cx, pad256((size_t)n * 4) + pad256((size_t)n * 16) + pad256((size_t)n * 64) + pad256(bpart))))
    return rc;
  if ((rc = S.get((void **)&d_t, (size_t)n * 4)) || (rc = S.get((void **)&d_ze, (size_t)n * 16)) ||
      (rc = S.get((void **)&d_out, (size_t)n * 64)) || (rc = S.get((void **)&d_part, bpart)))
    return rc;
  BQD2_HIP(hipMemcpyAsync(d_t, t, (size_t)n * 4, hipMemcpyHostToDevice, st));
  BQD2_HIP(hipMemcpyAsync(d_ze, ze, (size_t)n * 16, hipMemcpyHostToDevice, st));
  FieldCtx X = make_ctx(p);
  BQD2_HIP(hipMemsetAsync(ctr, 0, CNT_N * sizeof(long long), st));
  const double *amp = nullptr;
  if (X.far_on && did == 0 && (rc = far_amplitudes(p, cx, X, type, ctr, &amp))) return rc;
  if ((rc = launch_field_bd(p, st, X, type, total, did, n, d_t, d_ze, d_part, d_out, ctr, amp))) return rc;
  BQD2_HIP(hipMemcpyAsync(out, d_out, (size_t)n * 64, hipMemcpyDeviceToHost, st));
  long long cnt[CNT_N];
  BQD2_HIP(hipMemcpyAsync(cnt, ctr, sizeof(cnt), hipMemcpyDeviceToHost, st));
  BQD2_HIP(hipStreamSynchronize(st));
  BQD2_CHECK(cnt[CNT_ERR] == 0, BQD2_E_GF, "bqd2_field_bd: Green's function error at %lld points", cnt[CNT_ERR]);
  return BQD2_OK;
}

extern "C" int bqd2_force(bqd2_problem *p, const double *rc, int type, double *F, double *N) {
  BQD2_CHECK(p && rc && F && N, BQD2_E_ARG, "bqd2_force: bad arguments");
  const DevDomain &D0 = p->dom[0];
  const int ne = D0.Ne;
  if (ne == 0) {
    for (int i = 0; i < 3; i++) F[i] = N[i] = 0.0;
    return BQD2_OK;
  }
  CallCtx *cx = bqd2_call_ctx(p);
  if (!cx) return BQD2_E_HIP;
  hipStream_t st = cx->st;
  long long *const ctr = cx->cnt;
  // boundary-field points per element: the 9-point rule of quadrilaterals and
  // the 7-point rule of triangles for type 1, the centroid of triangles for
  // type 0 (bil_force_4p/9p, lit_force_4p/7p)
  std::vector<int> off((size_t)ne + 1, 0);
  for (int s = 1; s <= ne; s++)
    off[s] = off[s - 1] + (p->tripos[0][s] ? (type == 0 ? 1 : 7) : (type == 0 ? 0 : 9));
  const size_t nbd = (size_t)off[ne];
  Scratch S;
  double *d_bd = nullptr, *d_part, *d_FN;
  int *d_off = nullptr;
  int rc_;
  if ((rc_ = S.init(cx, pad256((size_t)ne * 48) + pad256(48) +
                           (nbd ? pad256(nbd * 64) + pad256(nbd * 4) + pad256(nbd * 16) + pad256(off.size() * 4) +
                                      pad256(bd_part_bytes(p, type, 0, (int)nbd))
                                : 0))))
    return rc_;
  if ((rc_ = S.get((void **)&d_part, (size_t)ne * 48)) || (rc_ = S.get((void **)&d_FN, 48))) return rc_;
  FieldCtx X = make_ctx(p);
  if (nbd) {
    std::vector<int> tpos(nbd);
    std::vector<double> ze(2 * nbd);
    const Tab &T = p->htab;
    for (int s = 1; s <= ne; s++) {
      const bool tri = p->tripos[0][s] != 0;
      for (int i = 0; i < off[s] - off[s - 1]; i++) {
        const size_t j = (size_t)off[s - 1] + i;
        tpos[j] = s;
        ze[2 * j] = type == 0 ? T.zt_34[3] : tri ? T.zt_37[i] : T.zt_49[i];
        ze[2 * j + 1] = type == 0 ? T.et_34[3] : tri ? T.et_37[i] : T.et_49[i];
      }
    }
    int *d_t;
    double *d_ze;
    if ((rc_ = S.get((void **)&d_bd, nbd * 64)) || (rc_ = S.get((void **)&d_t, nbd * 4)) ||
        (rc_ = S.get((void **)&d_ze, nbd * 16)) || (rc_ = S.get((void **)&d_off, off.size() * 4)))
      return rc_;
    BQD2_HIP(hipMemcpyAsync(d_t, tpos.data(), nbd * 4, hipMemcpyHostToDevice, st));
    BQD2_HIP(hipMemcpyAsync(d_ze, ze.data(), nbd * 16, hipMemcpyHostToDevice, st));
    BQD2_HIP(hipMemcpyAsync(d_off, off.data(), off.size() * 4, hipMemcpyHostToDevice, st));
    BQD2_HIP(hipMemsetAsync(ctr, 0, CNT_N * sizeof(long long), st));
    double *d_bdpart;
    if ((rc_ = S.get((void **)&d_bdpart, bd_part_bytes(p, type, 0, (int)nbd)))) return rc_;
    const double *amp = nullptr;
    if (X.far_on && (rc_ = far_amplitudes(p, cx, X, type, ctr, &amp))) return rc_;
    if ((rc_ = launch_field_bd(p, st, X, type, 1, 0, (int)nbd, d_t, d_ze, d_bdpart, d_bd, ctr, amp))) return rc_;
    BQD2_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(k_force_elem, dim3(bqd2_blocks(ne, 256)), dim3(256), 0, st, p->m, p->tab, X, type, d_bd, d_off,
                     rc[0], rc[1], rc[2], d_part);
  hipLaunchKernelGGL(k_force_sum, dim3(1), dim3(64), 0, st, ne, d_part, d_FN);
  BQD2_HIP(hipGetLastError());
  double FN[6];
  long long cnt[CNT_N] = {0};
  BQD2_HIP(hipMemcpyAsync(FN, d_FN, 48, hipMemcpyDeviceToHost, st));
  if (nbd) BQD2_HIP(hipMemcpyAsync(cnt, ctr, sizeof(cnt), hipMemcpyDeviceToHost, st));
  BQD2_HIP(hipStreamSynchronize(st));
  BQD2_CHECK(cnt[CNT_ERR] == 0, BQD2_E_GF, "bqd2_force: Green's function error at %lld boundary points", cnt[CNT_ERR]);
  for (int i = 0; i < 3; i++) {
    F[i] = FN[i];
    N[i] = FN[3 + i];
  }
  return BQD2_OK;
}

extern "C" int bqd2_domain_id(bqd2_problem *p, int n, const double *pts, int *did) {
  BQD2_CHECK(p && n >= 0 && pts && did, BQD2_E_ARG, "bqd2_domain_id: bad arguments");
  if (n == 0) return BQD2_OK;
  CallCtx *cx = bqd2_call_ctx(p);
  if (!cx) return BQD2_E_HIP;
  hipStream_t st = cx->st;
  long long *const ctr = cx->cnt;
  Scratch S;
  double *d_pts, *d_r0;
  int *d_l, *d_did, rc;
  if ((rc = S.init(cx, pad256((size_t)n * 24) * 2 + pad256((size_t)n * 8) + pad256((size_t)n * 4)))) return rc;
  if ((rc = S.get((void **)&d_pts, (size_t)n * 24)) || (rc = S.get((void **)&d_r0, (size_t)n * 24)) ||
      (rc = S.get((void **)&d_l, (size_t)n * 8)) || (rc = S.get((void **)&d_did, (size_t)n * 4)))
    return rc;
  BQD2_HIP(hipMemcpyAsync(d_pts, pts, (size_t)n * 24, hipMemcpyHostToDevice, st));
  FieldCtx X = make_ctx(p);
  hipLaunchKernelGGL(k_domain_id, dim3(bqd2_blocks(n, 4)), dim3(256), 0, st, p->m, X, p->rn, p->ed, n, d_pts, d_did,
                     d_r0, d_l);
  BQD2_HIP(hipGetLastError());
  BQD2_HIP(hipMemcpyAsync(did, d_did, (size_t)n * 4, hipMemcpyDeviceToHost, st));
  BQD2_HIP(hipStreamSynchronize(st));
  return BQD2_OK;
}
