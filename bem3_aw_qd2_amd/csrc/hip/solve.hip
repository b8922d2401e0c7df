// solve.hip -- boundary-value system, its solution and the velocity recovery.
//
//   k_system_A / k_system_b   create_matrix_csr_dac (bem3_aw_qd2_solve_bieq.c:244-345),
//                             written densely and column-major for the LU
//   k_store_solution          tmatrix_bd_store (:370-427)
//   k_velocity_rows           dpdt_fcm + inv_matrix_33 per node (:493-530, :564-607, :689-729)
//   k_pv_open_domain          the open-domain velocity from the interface
//                             conditions of q_solve_pv_bv (:711-748)
//   k_dpv_system              G dpv = H pv of solve_dpv_bv (:533-653)
#include <cstring>
#include "elem_dev.h"

struct PwDev {  // incident plane wave (aw_pw.c:128-199)
  double p[2], rho0, c0, k0, tv[3], R[9];
};

__device__ __forceinline__ void pw_pv(const PwDev &W, const double *r, cd &p, cd *v) {
  const double *R = W.R;
  double rc0 = r[0] - W.tv[0], rc1 = r[1] - W.tv[1], rc2 = r[2] - W.tv[2];
  double z = -R[6] * rc0 - R[7] * rc1 + R[8] * rc2;
  cd ce = cis(W.k0 * z);
  p = cd{W.p[0], W.p[1]} * ce;
  cd vz = p * (1.0 / (W.rho0 * W.c0));
  for (int i = 0; i < 3; i++) v[i] = vz * R[3 * i + 2];
}

__device__ __forceinline__ void inv33(double *ma, const double *vn, const double *vtz, const double *vte) {
  double idet = 1.0 / (vn[0] * vtz[1] * vte[2] + vn[1] * vtz[2] * vte[0] + vn[2] * vtz[0] * vte[1] -
                       (vn[2] * vtz[1] * vte[0] + vn[1] * vtz[0] * vte[2] + vn[0] * vtz[2] * vte[1]));
  ma[0] = vtz[1] * vte[2] - vtz[2] * vte[1];
  ma[1] = -vn[1] * vte[2] + vn[2] * vte[1];
  ma[2] = vn[1] * vtz[2] - vn[2] * vtz[1];
  ma[3] = -vtz[0] * vte[2] + vtz[2] * vte[0];
  ma[4] = vn[0] * vte[2] - vn[2] * vte[0];
  ma[5] = -vn[0] * vtz[2] + vn[2] * vtz[0];
  ma[6] = vtz[0] * vte[1] - vtz[1] * vte[0];
  ma[7] = -vn[0] * vte[1] + vn[1] * vte[0];
  ma[8] = vn[0] * vtz[1] - vn[1] * vtz[0];
  for (int i = 0; i < 9; i++) ma[i] *= idet;
}

// A entries of domain D's rows: one thread per (row, s).  Main side: P column
// H, dP column -G; sub side: dP column G (its P column hk2*H is written by
// k_system_b together with the right-hand side).
// element (r, c) of the dense system: column-major for the LU, row-major for GMRES
__device__ __forceinline__ size_t sys_at(int rm, size_t r, size_t c, int na) {
  return rm ? r * na + c : c * na + r;
}

__global__ void __launch_bounds__(256) k_system_A(DevMesh M, DevDomain D, int rowoff, double *__restrict__ A, int na,
                                                  int rm) {
  long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (long long)D.N * D.Ne) return;
  const int row = (int)(gid / D.Ne), s = (int)(gid % D.Ne) + 1;
  const int arow_local = D.rowmap[row];
  if (arow_local < 0) return;
  const size_t arow = (size_t)rowoff + arow_local;
  const int asd = abs(D.sid[s]);
  const int nl = M.tri[asd] ? 3 : 4;
  const bool main_side = (D.did == M.md[asd]);
  const int nn = M.NN;
  const double *Gr = D.G + 2 * dom_gh(D, row, (size_t)(s - 1) * 4);
  const double *Hr = D.H + 2 * dom_gh(D, row, (size_t)(s - 1) * 4);
  for (int l = 0; l < nl; l++) {
    const int col = M.eni[4 * asd + l];
    double *b = A + 2 * sys_at(rm, arow, (size_t)nn + col, na);
    if (main_side) {
      double *a = A + 2 * sys_at(rm, arow, col, na);
      a[0] = Hr[2 * l];
      a[1] = Hr[2 * l + 1];
      b[0] = -Gr[2 * l];
      b[1] = -Gr[2 * l + 1];
    } else {
      b[0] = Gr[2 * l];
      b[1] = Gr[2 * l + 1];
    }
  }
}

// sub-side P columns carry hk2 = k2[md]/k2[sd]; right-hand side per row
__global__ void __launch_bounds__(64) k_system_b(DevMesh M, DevDomain D, int rowoff,
                                                              double *__restrict__ A, int na, int rm,
                                                              const double *__restrict__ hk2tab,
                                                              double *__restrict__ bvec) {
  const int row = blockIdx.x, lane = threadIdx.x;
  const int arow_local = D.rowmap[row];
  if (arow_local < 0) return;
  const size_t arow = (size_t)rowoff + arow_local;
  cd acc = cd{0.0, 0.0};
  for (int s = 1 + lane; s <= D.Ne; s += 64) {
    const int sd = D.sid[s], asd = abs(sd);
    const int mdid = M.md[asd], sdid = M.sd[asd];
    if (D.did == mdid) continue;
    const int nl = M.tri[asd] ? 3 : 4;
    cd hk2 = cd{hk2tab[2 * (mdid * 16 + sdid)], hk2tab[2 * (mdid * 16 + sdid) + 1]};
    const double *Gr = D.G + 2 * dom_gh(D, row, (size_t)(s - 1) * 4);
    const double *Hr = D.H + 2 * dom_gh(D, row, (size_t)(s - 1) * 4);
    for (int l = 0; l < nl; l++) {
      const int col = M.eni[4 * asd + l];
      cd h = hk2 * cd{Hr[2 * l], Hr[2 * l + 1]};
      if (A) {
        double *a = A + 2 * sys_at(rm, arow, col, na);
        a[0] = h.x; a[1] = h.y;
      }
      if (mdid == 0) {
        cd pi = cd{M.Pi[2 * (4 * asd + l)], M.Pi[2 * (4 * asd + l) + 1]};
        cd dpi = cd{M.dPi[2 * (4 * asd + l)], M.dPi[2 * (4 * asd + l) + 1]};
        acc -= h * pi + cd{Gr[2 * l], Gr[2 * l + 1]} * dpi;
      }
    }
  }
  acc.x = wave_sum(acc.x);
  acc.y = wave_sum(acc.y);
  if (lane == 0) {
    bvec[2 * arow] = acc.x;
    bvec[2 * arow + 1] = acc.y;
  }
}

// scatter the system solution to P, dP of every element side of domain D
__global__ void __launch_bounds__(256) k_store_solution(DevMesh M, DevDomain D, const double *__restrict__ x,
                                                        const double *__restrict__ hk2tab) {
  int s = blockIdx.x * blockDim.x + threadIdx.x + 1;
  if (s > D.Ne) return;
  const int sd = D.sid[s], asd = abs(sd);
  const int mdid = M.md[asd], sdid = M.sd[asd];
  const int nl = M.tri[asd] ? 3 : 4, nn = M.NN;
  cd hk2 = cd{hk2tab[2 * (mdid * 16 + sdid)], hk2tab[2 * (mdid * 16 + sdid) + 1]};
  for (int l = 0; l < nl; l++) {
    int e = M.eni[4 * asd + l];
    cd xp = cd{x[2 * e], x[2 * e + 1]}, xd = cd{x[2 * (nn + e)], x[2 * (nn + e) + 1]};
    cd P, dP;
    if (D.did == mdid) {
      P = xp;
      dP = xd;
    } else {
      P = hk2 * xp;
      dP = -xd;
      if (mdid == 0) {
        P += hk2 * cd{M.Pi[2 * (4 * asd + l)], M.Pi[2 * (4 * asd + l) + 1]};
        dP -= cd{M.dPi[2 * (4 * asd + l)], M.dPi[2 * (4 * asd + l) + 1]};
      }
    }
    D.P[2 * (4 * s + l)] = P.x; D.P[2 * (4 * s + l) + 1] = P.y;
    D.dP[2 * (4 * s + l)] = dP.x; D.dP[2 * (4 * s + l) + 1] = dP.y;
  }
}

// per node (t, tn) of domain D: tangential derivatives of P from the dG, dH rows
// and the node velocity pv = -[n; tz; te]^-1 (dP/dn, dP/dtz, dP/dte); with
// open_out != NULL also the open-domain velocity from the interface conditions
__global__ void __launch_bounds__(64) k_velocity_rows(DevMesh M, DevDomain D, const Tab *__restrict__ tab, PwDev W,
                                                      double rho_d, double rho_0, double *__restrict__ open_out) {
  const int row = D.r0 + blockIdx.x, lane = threadIdx.x;
  int t = row / 4 + 1, tn = row % 4;
  const int at = D.sid[t], atd = abs(at);
  if (M.tri[atd] && tn == 3) return;
  const size_t N = D.N;
  const double *g0 = D.dG + 2 * dom_dgh(D, row, 0, 0), *g1 = D.dG + 2 * dom_dgh(D, row, 1, 0);
  const double *h0 = D.dH + 2 * dom_dgh(D, row, 0, 0), *h1 = D.dH + 2 * dom_dgh(D, row, 1, 0);
  cd z = cd{0, 0}, e = cd{0, 0};
  for (size_t c = lane; c < N; c += 64) {
    // column c = (s-1)*4 + sn -> P[s][sn] at index 4 s + sn = c + 4
    cd dp = cd{D.dP[2 * (c + 4)], D.dP[2 * (c + 4) + 1]}, p = cd{D.P[2 * (c + 4)], D.P[2 * (c + 4) + 1]};
    z += cd{g0[2 * c], g0[2 * c + 1]} * dp - cd{h0[2 * c], h0[2 * c + 1]} * p;
    e += cd{g1[2 * c], g1[2 * c + 1]} * dp - cd{h1[2 * c], h1[2 * c + 1]} * p;
  }
  z = cd{wave_sum(z.x), wave_sum(z.y)};
  e = cd{wave_sum(e.x), wave_sum(e.y)};
  if (lane != 0) return;
  const double *dF = D.dF + 3 * row;
  cd Pt = cd{D.P[2 * (4 * t + tn)], D.P[2 * (4 * t + tn) + 1]};
  cd dPt = cd{D.dP[2 * (4 * t + tn)], D.dP[2 * (4 * t + tn) + 1]};
  cd dPz = (z - Pt * dF[1]) * (1.0 / dF[0]);
  cd dPe = (e - Pt * dF[2]) * (1.0 / dF[0]);
  double vtz[3], vte[3], vn[3], a[9];
  node_tangents(M, *tab, atd, tn, vtz, vte);
  const double sig = at < 0 ? -1.0 : 1.0;
  for (int i = 0; i < 3; i++) vn[i] = sig * M.wen[12 * atd + 3 * tn + i];
  double an = abs3(vn);
  if (an != 0.0) { vn[0] /= an; vn[1] /= an; vn[2] /= an; }
  inv33(a, vn, vtz, vte);
  cd pv[3];
  for (int i = 0; i < 3; i++) {
    pv[i] = -(dPt * a[3 * i] + dPz * a[3 * i + 1] + dPe * a[3 * i + 2]);
    D.pv[2 * (12 * t + 3 * tn + i)] = pv[i].x;
    D.pv[2 * (12 * t + 3 * tn + i) + 1] = pv[i].y;
  }
  if (open_out && M.md[atd] == 0) {
    const double *r = M.ren + 12 * atd + 3 * tn;
    cd pi, pvi[3];
    pw_pv(W, r, pi, pvi);
    double rtz[3], rte[3];
    for (int i = 0; i < 3; i++) {
      rtz[i] = rho_0 * vtz[i];
      rte[i] = rho_0 * vte[i];
    }
    cd rhs[3];
    rhs[0] = (pv[0] - pvi[0]) * vn[0] + (pv[1] - pvi[1]) * vn[1] + (pv[2] - pvi[2]) * vn[2];
    rhs[1] = (pv[0] * rho_d - pvi[0] * rho_0) * vtz[0] + (pv[1] * rho_d - pvi[1] * rho_0) * vtz[1] +
             (pv[2] * rho_d - pvi[2] * rho_0) * vtz[2];
    rhs[2] = (pv[0] * rho_d - pvi[0] * rho_0) * vte[0] + (pv[1] * rho_d - pvi[1] * rho_0) * vte[1] +
             (pv[2] * rho_d - pvi[2] * rho_0) * vte[2];
    inv33(a, vn, rtz, rte);
    const int en = M.eni[4 * atd + tn];
    for (int i = 0; i < 3; i++) {
      cd v = rhs[0] * a[3 * i] + rhs[1] * a[3 * i + 1] + rhs[2] * a[3 * i + 2];
      open_out[2 * (3 * en + i)] = v.x;
      open_out[2 * (3 * en + i) + 1] = v.y;
    }
  }
}

__global__ void __launch_bounds__(256) k_pv_open_domain(DevMesh M, DevDomain D0, const double *__restrict__ tpv) {
  int s = blockIdx.x * blockDim.x + threadIdx.x + 1;
  if (s > D0.Ne) return;
  const int asd = abs(D0.sid[s]);
  const int nl = M.tri[asd] ? 3 : 4;
  for (int l = 0; l < nl; l++) {
    const int en = M.eni[4 * asd + l];
    for (int i = 0; i < 3; i++) {
      D0.pv[2 * (12 * s + 3 * l + i)] = tpv[2 * (3 * en + i)];
      D0.pv[2 * (12 * s + 3 * l + i) + 1] = tpv[2 * (3 * en + i) + 1];
    }
  }
}

// dpv system of one domain: A = G (compressed rows/cols, column-major), B = H pv
__global__ void __launch_bounds__(256) k_dpv_A(DevMesh M, DevDomain D, double *__restrict__ A) {
  long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (long long)D.N * D.N) return;
  int row = (int)(gid / D.N), col = (int)(gid % D.N);
  int r = D.rowmap[row], c = D.rowmap[col];
  if (r < 0 || c < 0) return;
  const double *g = D.G + 2 * dom_gh(D, row, col);
  double *a = A + 2 * ((size_t)c * D.nn + r);
  a[0] = g[0];
  a[1] = g[1];
}

__global__ void __launch_bounds__(64) k_dpv_B(DevMesh M, DevDomain D, double *__restrict__ B) {
  const int row = blockIdx.x, lane = threadIdx.x;
  const int r = D.rowmap[row];
  if (r < 0) return;
  cd acc[3] = {};
  for (int c = lane; c < D.N; c += 64) {
    if (D.rowmap[c] < 0) continue;
    cd h = cd{D.H[2 * dom_gh(D, row, c)], D.H[2 * dom_gh(D, row, c) + 1]};
    for (int i = 0; i < 3; i++) acc[i] += h * cd{D.pv[2 * (12 + 3 * c + i)], D.pv[2 * (12 + 3 * c + i) + 1]};
  }
  for (int i = 0; i < 3; i++) {
    acc[i] = cd{wave_sum(acc[i].x), wave_sum(acc[i].y)};
    if (lane == 0) {
      B[2 * ((size_t)i * D.nn + r)] = acc[i].x;
      B[2 * ((size_t)i * D.nn + r) + 1] = acc[i].y;
    }
  }
}

// the same right-hand side on the N row slots of the domain (zero on unused
// slots), B[i N + row], for GMRES on G itself
__global__ void __launch_bounds__(64) k_dpv_B_slots(DevDomain D, double *__restrict__ B) {
  const int row = blockIdx.x, lane = threadIdx.x;
  const bool used = D.rowmap[row] >= 0;
  cd acc[3] = {};
  if (used)
    for (int c = lane; c < D.N; c += 64) {
      if (D.rowmap[c] < 0) continue;
      cd h = cd{D.H[2 * dom_gh(D, row, c)], D.H[2 * dom_gh(D, row, c) + 1]};
      for (int i = 0; i < 3; i++) acc[i] += h * cd{D.pv[2 * (12 + 3 * c + i)], D.pv[2 * (12 + 3 * c + i) + 1]};
    }
  for (int i = 0; i < 3; i++) {
    acc[i] = cd{wave_sum(acc[i].x), wave_sum(acc[i].y)};
    if (lane == 0) {
      B[2 * ((size_t)i * D.N + row)] = acc[i].x;
      B[2 * ((size_t)i * D.N + row) + 1] = acc[i].y;
    }
  }
}

__global__ void __launch_bounds__(256) k_dpv_store_slots(DevDomain D, const double *__restrict__ X) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= D.N || D.rowmap[c] < 0) return;
  for (int i = 0; i < 3; i++) {
    D.dpv[2 * (12 + 3 * c + i)] = X[2 * ((size_t)i * D.N + c)];
    D.dpv[2 * (12 + 3 * c + i) + 1] = X[2 * ((size_t)i * D.N + c) + 1];
  }
}

__global__ void __launch_bounds__(256) k_dpv_store(DevDomain D, const double *__restrict__ B) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= D.N) return;
  int r = D.rowmap[c];
  if (r < 0) return;
  for (int i = 0; i < 3; i++) {
    D.dpv[2 * (12 + 3 * c + i)] = B[2 * ((size_t)i * D.nn + r)];
    D.dpv[2 * (12 + 3 * c + i) + 1] = B[2 * ((size_t)i * D.nn + r) + 1];
  }
}

// Incident potential at the collocation nodes for the problem's plane wave:
// Pi = -p / k2, dPi = -(dp/dn) / k2 (init_boundary_data, bem3_aw_qd2.c:1104-1146,
// with calc_apw_dpdn, aw_pw.c:178-199).  One thread per (element, node slot).
__global__ void __launch_bounds__(256) k_incident(DevMesh M, PwDev W, cd k2, double *__restrict__ Pi,
                                                  double *__restrict__ dPi) {
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= 4 * M.Ne) return;
  const int s = 1 + id / 4, j = id % 4;
  const double *r = M.ren + 12 * (size_t)s + 3 * j, *w = M.wen + 12 * (size_t)s + 3 * j;
  const double wn = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  const double n[3] = {w[0] / wn, w[1] / wn, w[2] / wn};
  const double *R = W.R;
  const double rc0 = r[0] - W.tv[0], rc1 = r[1] - W.tv[1], rc2 = r[2] - W.tv[2];
  const double z = -R[6] * rc0 - R[7] * rc1 + R[8] * rc2;
  const double nz = -R[6] * n[0] - R[7] * n[1] + R[8] * n[2];
  const cd p = cd{W.p[0], W.p[1]} * cis(W.k0 * z);
  const cd dp = mul_i(p) * (W.k0 * nz);
  const cd cp = -crecip(k2);
  const cd a = cp * p, b = cp * dp;
  const size_t o = 2 * (4 * (size_t)s + j);
  Pi[o] = a.x; Pi[o + 1] = a.y;
  dPi[o] = b.x; dPi[o + 1] = b.y;
}

// No counterpart in the reference (one plane wave per model there).  A new
// incident plane wave pw7 = {Re p, Im p, theta, psi, tx, ty, tz} (the fields of
// pw.txt, aw_pw.c:3-77) on an assembled problem.  The operator depends on the
// incidence only through the in-plane wave vector, modulo the reciprocal
// lattice: the new one must equal the problem's up to a reciprocal lattice
// vector (the opposite kz, or the direction of another propagating diffraction
// order), otherwise BQD2_E_ARG.  The assembled matrices and the factors of the
// mixed-precision solver are kept; the next bqd2_solve_surface uses the new
// right-hand side.
extern "C" int bqd2_set_incidence(bqd2_problem *p, const double *pw7) {
  BQD2_CHECK(p && pw7, BQD2_E_ARG, "bqd2_set_incidence: null argument");
  const double ct = cos(pw7[2]), st = sin(pw7[2]), cp = cos(pw7[3]), sp = sin(pw7[3]);
  const double R[9] = {ct * cp * cp + sp * sp, sp * cp * (ct - 1.0), st * cp,
                       sp * cp * (ct - 1.0),   ct * sp * sp + cp * cp, st * sp,
                       -st * cp,               -st * sp,               ct};
  return bqd2_set_incidence_R(p, pw7, pw7 + 4, R);
}

int bqd2_set_incidence_R(bqd2_problem *p, const double *pw_p, const double *tv, const double *R) {
  BQD2_CHECK(R[8] != 0.0, BQD2_E_ARG, "bqd2_set_incidence: kz must not be 0");
  if (p->m.ps == 1) {
    // (kx, ky) - (kx0, ky0) = m1 g1 + m2 g2 with integer m1, m2
    const double dx = p->pw_k0 * R[2] - p->qd8[1], dy = p->pw_k0 * R[5] - p->qd8[2];
    const GfParams &g = p->gf;
    const double det = g.g1x * g.g2y - g.g1y * g.g2x;
    const double m1 = (dx * g.g2y - dy * g.g2x) / det, m2 = (g.g1x * dy - g.g1y * dx) / det;
    BQD2_CHECK(fabs(m1 - nearbyint(m1)) < 1e-9 && fabs(m2 - nearbyint(m2)) < 1e-9, BQD2_E_ARG,
               "bqd2_set_incidence: in-plane wave vector (%g, %g) differs from the operator's (%g, %g) by "
               "(%g, %g) reciprocal lattice vectors, not an integer combination",
               p->pw_k0 * R[2], p->pw_k0 * R[5], p->qd8[1], p->qd8[2], m1, m2);
  }
  p->pw_p[0] = pw_p[0];
  p->pw_p[1] = pw_p[1];
  memcpy(p->pw_tv, tv, sizeof(p->pw_tv));
  memcpy(p->pw_R, R, sizeof(p->pw_R));
  PwDev W;
  memcpy(W.p, p->pw_p, sizeof(W.p));
  W.rho0 = p->pw_rho0; W.c0 = p->pw_c0; W.k0 = p->pw_k0;
  memcpy(W.tv, p->pw_tv, sizeof(W.tv));
  memcpy(W.R, p->pw_R, sizeof(W.R));
  hipLaunchKernelGGL(k_incident, dim3(bqd2_blocks(4 * p->m.Ne, 256)), dim3(256), 0, p->stream, p->m, W,
                     cd{p->k2[0][0], p->k2[0][1]}, (double *)p->m.Pi, (double *)p->m.dPi);
  BQD2_HIP(hipGetLastError());
  BQD2_HIP(hipStreamSynchronize(p->stream));
  p->solved = false;
  return BQD2_OK;
}

void bqd2_hk2_table(const bqd2_problem *p, double *tab) {
  // tab[(m*16 + s)] = k2[m] / k2[s]
  for (int m = 0; m <= p->m.MN; m++)
    for (int s = 0; s <= p->m.MN; s++) {
      double ar = p->k2[m][0], ai = p->k2[m][1], br = p->k2[s][0], bi = p->k2[s][1];
      double den = br * br + bi * bi;
      tab[2 * (m * 16 + s)] = (ar * br + ai * bi) / den;
      tab[2 * (m * 16 + s) + 1] = (ai * br - ar * bi) / den;
    }
}

// GMRES tolerance (scaled residual) and limits for the solves below
#define GM_TOL 1.0e-15
#define GM_RESTART 100
#define GM_MAXIT 400

static int ensure_hk2(bqd2_problem *p) {
  if (p->hk2d) return BQD2_OK;
  double tab[16 * 16 * 2];
  memset(tab, 0, sizeof(tab));
  bqd2_hk2_table(p, tab);
  int rc = bqd2_alloc(p, (void **)&p->hk2d, sizeof(tab));
  if (rc) return rc;
  BQD2_HIP(hipMemcpyAsync(p->hk2d, tab, sizeof(tab), hipMemcpyHostToDevice, p->stream));
  BQD2_HIP(hipStreamSynchronize(p->stream));  // tab is released on return
  return BQD2_OK;
}

// the right-hand side of the system alone (the matrix is kept: its LU factors)
static int build_rhs(bqd2_problem *p, hipStream_t st) {
  BQD2_HIP(hipMemsetAsync(p->b, 0, (size_t)p->na * 16, st));
  int rowoff = 0;
  for (int d = 0; d <= p->m.MN; d++) {
    DevDomain &D = p->dom[d];
    if (D.Ne == 0) continue;
    hipLaunchKernelGGL(k_system_b, dim3(D.N), dim3(64), 0, st, p->m, D, rowoff, (double *)nullptr, p->na, 0, p->hk2d,
                       p->b);
    BQD2_HIP(hipGetLastError());
    rowoff += D.nn;
  }
  return BQD2_OK;
}

// the dense system of all domains, column-major (rm = 0) or row-major (rm = 1)
static int build_system(bqd2_problem *p, int rm, hipStream_t st) {
  const int na = p->na;
  p->lu64_valid = false;
  BQD2_HIP(hipMemsetAsync(p->A, 0, (size_t)na * na * 16, st));
  BQD2_HIP(hipMemsetAsync(p->b, 0, (size_t)na * 16, st));
  int rowoff = 0;
  for (int d = 0; d <= p->m.MN; d++) {
    DevDomain &D = p->dom[d];
    if (D.Ne == 0) continue;
    long long n = (long long)D.N * D.Ne;
    hipLaunchKernelGGL(k_system_A, dim3(bqd2_blocks(n, 256)), dim3(256), 0, st, p->m, D, rowoff, p->A, na, rm);
    hipLaunchKernelGGL(k_system_b, dim3(D.N), dim3(64), 0, st, p->m, D, rowoff, p->A, na, rm, p->hk2d, p->b);
    BQD2_HIP(hipGetLastError());
    rowoff += D.nn;
  }
  BQD2_CHECK(rowoff == na, BQD2_E_STATE, "bqd2_solve_surface: %d system rows for %d unknowns", rowoff, na);
  return BQD2_OK;
}

extern "C" int bqd2_set_solver(bqd2_problem *p, int mode) {
  BQD2_CHECK(p && mode >= 0 && mode <= 2, BQD2_E_ARG,
             "bqd2_set_solver: mode must be 0 (GMRES, LU fallback), 1 (LU) or 2 (complex float LU + GMRES refinement)");
  p->solver = mode;
  return BQD2_OK;
}

// GMRES steps and restart length of the refinement (each step is one FP64
// matrix-vector product and one application of the float factors)
#define IR_RESTART 30
#define IR_MAXIT 60

// Complex float LU of the row-scaled row-major system in p->A, once per
// assembly; the factors stay on the device for every later right-hand side.
static int factor_lu32(bqd2_problem *p, hipStream_t st, float bound) {
  const int na = p->na;
  int rc;
  if (!p->Af) {
    if ((rc = bqd2_alloc(p, (void **)&p->Af, (size_t)na * na * 8))) return rc;
    if ((rc = bqd2_alloc(p, &p->luws, bqd2_cgetrf_wsize(na)))) return rc;
    if ((rc = bqd2_alloc(p, (void **)&p->lutmp, (size_t)na * 6 * 8))) return rc;
    if ((rc = bqd2_alloc(p, (void **)&p->dinv, (size_t)na * 16))) return rc;
  }
  hipEvent_t e0, e1;
  BQD2_HIP(hipEventCreate(&e0));
  BQD2_HIP(hipEventCreate(&e1));
  BQD2_HIP(hipEventRecord(e0, st));
  if ((rc = bqd2_jacobi(na, p->A, na, nullptr, p->dinv, st))) return rc;
  if ((rc = bqd2_lu32_convert(na, p->A, na, nullptr, p->dinv, p->Af, st))) return rc;
  if ((rc = bqd2_cgetrf_bound(na, p->Af, p->ipiv, p->luws, st, bound, &p->lu32_growth))) return rc;
  BQD2_HIP(hipEventRecord(e1, st));
  BQD2_HIP(hipEventSynchronize(e1));
  float ms = 0.0f;
  BQD2_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  p->ms_factor = ms;
  p->lu32_valid = true;
  return BQD2_OK;
}

extern "C" int bqd2_lu32_info(bqd2_problem *p, double *ms_factor) {
  BQD2_CHECK(p && ms_factor, BQD2_E_ARG, "bqd2_lu32_info: null argument");
  *ms_factor = p->lu32_valid ? p->ms_factor : 0.0;
  return BQD2_OK;
}

extern "C" int bqd2_solver_info(bqd2_problem *p, int *iters, double *relres, int *lu) {
  BQD2_CHECK(p, BQD2_E_ARG, "bqd2_solver_info: null problem");
  for (int i = 0; i <= p->m.MN + 1 && i < 17; i++) {
    if (iters) iters[i] = p->info_its[i];
    if (relres) relres[i] = p->info_res[i];
  }
  if (lu) *lu = p->info_lu;
  return BQD2_OK;
}

int bqd2_lg_solve(bqd2_problem *p, const double **x);
static int store_solution(bqd2_problem *p, const double *x, double *x_host);
int bqd2_lg_velocity(bqd2_problem *p, const void *W, double *open_out);

// the velocity rows of domain D's rows [D.r0, D.r0 + D.nr) (its dG, dH held for them)
int bqd2_velocity_rows(bqd2_problem *p, const DevDomain &D, const void *W, double *open_out, hipStream_t st) {
  const int d = D.did;
  hipLaunchKernelGGL(k_velocity_rows, dim3(D.nr), dim3(64), 0, st, p->m, D, p->tab, *(const PwDev *)W, p->rho0[d],
                     p->rho0[0], open_out);
  BQD2_HIP(hipGetLastError());
  return BQD2_OK;
}

extern "C" int bqd2_solve_surface(bqd2_problem *p, double *x_host) {
  BQD2_CHECK(p, BQD2_E_ARG, "bqd2_solve_surface: null problem");
  p->sol_epoch++;  // the field's plane-wave amplitudes are made again
  for (int d = 0; d <= p->m.MN; d++)
    BQD2_CHECK(p->dom[d].Ne == 0 || p->dom[d].assembled, BQD2_E_STATE, "bqd2_solve_surface: domain %d not assembled", d);
  const int na = p->na;
  hipStream_t st = p->stream;
  int rc;
  if (p->large) {
    if ((rc = ensure_hk2(p))) return rc;
    p->info_lu = 0;
    for (int i = 0; i < 17; i++) { p->info_its[i] = 0; p->info_res[i] = 0.0; }
    const double *xl = nullptr;
    if ((rc = bqd2_lg_solve(p, &xl))) return rc;
    return store_solution(p, xl, x_host);
  }
  if (!p->A) {
    if ((rc = bqd2_alloc(p, (void **)&p->A, (size_t)na * na * 16))) return rc;
    if ((rc = bqd2_alloc(p, (void **)&p->b, (size_t)na * 16))) return rc;
    if ((rc = bqd2_alloc(p, (void **)&p->xg, (size_t)na * 16))) return rc;
    if ((rc = bqd2_alloc(p, (void **)&p->ipiv, (size_t)na * sizeof(int)))) return rc;
  }
  if ((rc = ensure_hk2(p))) return rc;
  p->info_lu = 0;
  for (int i = 0; i < 17; i++) { p->info_its[i] = 0; p->info_res[i] = 0.0; }
  const double *x = nullptr;
  if (p->solver == 2) {
    // complex float LU of the scaled system, then GMRES on the FP64 system
    // right-preconditioned by it; plain GMRES and the FP64 LU remain as fallbacks
    if ((rc = build_system(p, 1, st))) return rc;
    if (!p->lu32_valid) {
      // block-local pivots unchecked; a zero pivot in a diagonal block: the stable bound
      rc = factor_lu32(p, st, BQD2_LU32_RELAXED_GROWTH);
      if (rc == BQD2_E_SINGULAR) rc = factor_lu32(p, st, BQD2_LU32_STABLE_GROWTH);
      if (rc) return rc;
    }
    Precond pc{p->Af, p->luws, p->lutmp, &p->lu_graph};
    rc = bqd2_gmres(&p->main, na, 1, p->A, na, nullptr, p->b, na, p->xg, na, GM_TOL, IR_RESTART, IR_MAXIT,
                    &p->info_its[0], &p->info_res[0], &pc);
    if (rc == 1 && p->lu32_growth > BQD2_LU32_STABLE_GROWTH) {
      // the relaxed pivots made too weak a preconditioner: factor again with the stable bound
      if ((rc = factor_lu32(p, st, BQD2_LU32_STABLE_GROWTH))) return rc;
      rc = bqd2_gmres(&p->main, na, 1, p->A, na, nullptr, p->b, na, p->xg, na, GM_TOL, IR_RESTART, IR_MAXIT,
                      &p->info_its[0], &p->info_res[0], &pc);
    }
    if (rc < 0) return rc;
    if (rc == 0) x = p->xg;
  }
  if (!x && p->solver != 1) {
    // Jacobi-scaled GMRES on the row-major system
    if ((rc = build_system(p, 1, st))) return rc;
    rc = bqd2_gmres(&p->main, na, 1, p->A, na, nullptr, p->b, na, p->xg, na, GM_TOL, GM_RESTART, GM_MAXIT,
                    &p->info_its[0], &p->info_res[0]);
    if (rc < 0) return rc;
    if (rc == 0) x = p->xg;
  }
  if (!x) {
    // dense LU with partial pivoting on the column-major system; the factors
    // stay valid for later right-hand sides on the same assembly
    p->info_lu++;
    if (p->lu64_valid) {
      if ((rc = build_rhs(p, st))) return rc;
    } else {
      if ((rc = build_system(p, 0, st))) return rc;
      if ((rc = bqd2_zgetrf(na, p->A, p->ipiv, st))) return rc;
      p->lu64_valid = true;
    }
    if ((rc = bqd2_zgetrs(na, 1, p->A, p->ipiv, p->b, st))) return rc;
    x = p->b;
  }
  return store_solution(p, x, x_host);
}

// the system's solution x (device) to P, dP of every domain; x_host may be null
static int store_solution(bqd2_problem *p, const double *x, double *x_host) {
  hipStream_t st = p->stream;
  const int na = p->na;
  for (int d = 0; d <= p->m.MN; d++) {
    DevDomain &D = p->dom[d];
    if (D.Ne == 0) continue;
    hipLaunchKernelGGL(k_store_solution, dim3(bqd2_blocks(D.Ne, 256)), dim3(256), 0, st, p->m, D, x, p->hk2d);
  }
  BQD2_HIP(hipGetLastError());
  if (x_host) BQD2_HIP(hipMemcpyAsync(x_host, x, (size_t)na * 16, hipMemcpyDeviceToHost, st));
  BQD2_HIP(hipStreamSynchronize(st));
  p->solved = true;
  return BQD2_OK;
}

// complex float factors of domain D's row-scaled G (identity on unused rows),
// once per assembly: the preconditioner of its dpv systems in the mixed solver
static int factor_g32(bqd2_problem *p, DevDomain &D, hipStream_t st, float bound) {
  const size_t N = D.N;
  int rc;
  if (!D.gAf) {
    if ((rc = bqd2_alloc(p, (void **)&D.gAf, N * N * 8))) return rc;
    if ((rc = bqd2_alloc(p, &D.gws, bqd2_cgetrf_wsize(D.N)))) return rc;
    if ((rc = bqd2_alloc(p, (void **)&D.gtmp, N * 6 * 8))) return rc;
    if ((rc = bqd2_alloc(p, (void **)&D.gdinv, N * 16))) return rc;
    if ((rc = bqd2_alloc(p, (void **)&D.gipiv, N * sizeof(int)))) return rc;
  }
  if ((rc = bqd2_jacobi(D.N, D.G, D.ldm, D.rowmap, D.gdinv, st))) return rc;
  if ((rc = bqd2_lu32_convert(D.N, D.G, D.ldm, D.rowmap, D.gdinv, D.gAf, st))) return rc;
  if ((rc = bqd2_cgetrf_bound(D.N, D.gAf, D.gipiv, D.gws, st, bound, &D.g_growth))) return rc;
  D.g_lu_valid = true;
  return BQD2_OK;
}

// G dpv = H pv of one domain by the dense LU on the compressed rows
static int dpv_lu(bqd2_problem *p, DevDomain &D, hipStream_t st) {
  double *A = nullptr, *B = nullptr;
  int *ipiv = nullptr;
  int rc = bqd2_alloc(p, (void **)&A, (size_t)D.nn * D.nn * 16);
  if (!rc) rc = bqd2_alloc(p, (void **)&B, (size_t)D.nn * 3 * 16);
  if (!rc) rc = bqd2_alloc(p, (void **)&ipiv, (size_t)D.nn * sizeof(int));
  if (rc) return rc;
  BQD2_HIP(hipMemsetAsync(A, 0, (size_t)D.nn * D.nn * 16, st));
  long long n2 = (long long)D.N * D.N;
  hipLaunchKernelGGL(k_dpv_A, dim3(bqd2_blocks(n2, 256)), dim3(256), 0, st, p->m, D, A);
  hipLaunchKernelGGL(k_dpv_B, dim3(D.N), dim3(64), 0, st, p->m, D, B);
  BQD2_HIP(hipGetLastError());
  if ((rc = bqd2_zgetrf(D.nn, A, ipiv, st))) return rc;
  if ((rc = bqd2_zgetrs(D.nn, 3, A, ipiv, B, st))) return rc;
  hipLaunchKernelGGL(k_dpv_store, dim3(bqd2_blocks(D.N, 256)), dim3(256), 0, st, D, B);
  BQD2_HIP(hipGetLastError());
  BQD2_HIP(hipStreamSynchronize(st));
  // the work buffers of the fallback are released right away
  p->alloc_bytes -= (size_t)D.nn * D.nn * 16 + (size_t)D.nn * 3 * 16 + (size_t)D.nn * sizeof(int);
  for (void *q : {(void *)A, (void *)B, (void *)ipiv}) {
    (void)hipFree(q);
    for (auto it = p->owned.begin(); it != p->owned.end(); ++it)
      if (*it == q) { p->owned.erase(it); break; }
  }
  return BQD2_OK;
}

extern "C" int bqd2_solve_velocity(bqd2_problem *p) {
  BQD2_CHECK(p && p->solved, BQD2_E_STATE, "bqd2_solve_velocity: surface solution missing");
  p->sol_epoch++;
  hipStream_t st = p->stream;
  PwDev W;
  memcpy(W.p, p->pw_p, sizeof(W.p));
  W.rho0 = p->pw_rho0; W.c0 = p->pw_c0; W.k0 = p->pw_k0;
  memcpy(W.tv, p->pw_tv, sizeof(W.tv));
  memcpy(W.R, p->pw_R, sizeof(W.R));
  const bool periodic = (p->m.ps == 1);
  int rc;
  if (periodic) {
    if (!p->tpv && (rc = bqd2_alloc(p, (void **)&p->tpv, (size_t)p->m.NN * 3 * 16))) return rc;
    BQD2_HIP(hipMemsetAsync(p->tpv, 0, (size_t)p->m.NN * 3 * 16, st));
  }
  if (p->large) {
    // the sphere's rows window by window (domain 0 assembled again if an FP64 LU consumed it)
    if ((rc = bqd2_lg_velocity(p, &W, p->tpv))) return rc;
  } else {
    for (int d = periodic ? 1 : 0; d <= p->m.MN; d++) {
      DevDomain &D = p->dom[d];
      if (D.Ne == 0) continue;
      if ((rc = bqd2_velocity_rows(p, D, &W, periodic ? p->tpv : (double *)nullptr, st))) return rc;
    }
  }
  if (periodic && p->dom[0].Ne > 0)
    hipLaunchKernelGGL(k_pv_open_domain, dim3(bqd2_blocks(p->dom[0].Ne, 256)), dim3(256), 0, st, p->m, p->dom[0],
                       p->tpv);
  BQD2_HIP(hipGetLastError());
  // normal derivative of the velocity: G dpv = H pv per domain (3 right-hand sides)
  if (!p->dpvw) {
    int nmax = 0;
    for (int d = 0; d <= p->m.MN; d++) nmax = p->dom[d].N > nmax ? p->dom[d].N : nmax;
    if ((rc = bqd2_alloc(p, (void **)&p->dpvw, (size_t)2 * 3 * nmax * 16 + 16))) return rc;
  }
  for (int d = 0; d <= p->m.MN; d++) {
    DevDomain &D = p->dom[d];
    if (D.Ne == 0) continue;
    bool done = false;
    // the system layout solves the dpv systems by FP64 GMRES whatever the surface solver
    const int solver = p->large ? 0 : p->solver;
    if (solver != 1) {
      double *B = p->dpvw, *X = p->dpvw + 2 * (size_t)3 * D.N;
      hipLaunchKernelGGL(k_dpv_B_slots, dim3(D.N), dim3(64), 0, st, D, B);
      BQD2_HIP(hipGetLastError());
      rc = 1;
      if (solver == 2) {
        // complex float factors of G (once per assembly) as the preconditioner
        if (!D.g_lu_valid) {
          rc = factor_g32(p, D, st, BQD2_LU32_RELAXED_GROWTH);
          if (rc == BQD2_E_SINGULAR) rc = factor_g32(p, D, st, BQD2_LU32_STABLE_GROWTH);
          if (rc) return rc;
        }
        Precond pc{D.gAf, D.gws, D.gtmp, &D.g_graph};
        rc = bqd2_gmres(&p->main, D.N, 3, D.G, D.ldm, D.rowmap, B, D.N, X, D.N, GM_TOL, IR_RESTART, IR_MAXIT,
                        &p->info_its[1 + d], &p->info_res[1 + d], &pc);
        if (rc == 1 && D.g_growth > BQD2_LU32_STABLE_GROWTH) {
          if ((rc = factor_g32(p, D, st, BQD2_LU32_STABLE_GROWTH))) return rc;
          rc = bqd2_gmres(&p->main, D.N, 3, D.G, D.ldm, D.rowmap, B, D.N, X, D.N, GM_TOL, IR_RESTART, IR_MAXIT,
                          &p->info_its[1 + d], &p->info_res[1 + d], &pc);
        }
        if (rc < 0) return rc;
      }
      if (rc == 1)
        rc = bqd2_gmres(&p->main, D.N, 3, D.G, D.ldm, D.rowmap, B, D.N, X, D.N, GM_TOL, GM_RESTART, GM_MAXIT,
                        &p->info_its[1 + d], &p->info_res[1 + d]);
      if (rc < 0) return rc;
      if (rc == 0) {
        hipLaunchKernelGGL(k_dpv_store_slots, dim3(bqd2_blocks(D.N, 256)), dim3(256), 0, st, D, X);
        BQD2_HIP(hipGetLastError());
        done = true;
      }
    }
    if (!done) {
      p->info_lu++;
      if ((rc = dpv_lu(p, D, st))) return rc;
    }
  }
  BQD2_HIP(hipStreamSynchronize(st));
  return BQD2_OK;
}
