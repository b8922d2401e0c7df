// large.hip -- the system layout: problems whose matrices do not fit one GPU in
// the default layout (bqd2_create_large).
//
// The default layout keeps G, H of both domains, dG, dH of the sphere and the
// dense system as separate FP64 arrays; for 13 824 quadrilaterals that is
// 98 + 98 + 196 + 196 GB.  Here both domains' matrices sit in one row-major
// na x na FP64 array S = [H0 G0; H1 G1] (na = 2N), which for this mesh (one
// sphere of quadrilaterals, each with its own nodes) is the system matrix of
// create_matrix_csr_dac (bem3_aw_qd2_solve_bieq.c:244-345) up to column factors:
// the lattice domain's rows are [H0, -G0], the sphere's [hk2 H1, G1] with
// hk2 = k2[0] / k2[1], and only the sphere's rows carry a right-hand side,
// b = -(hk2 H1 Pi + G1 dPi).  The sphere's rows are free-space integrals, cheap
// next to the lattice sums of domain 0; they are assembled in windows of lwin
// rows whenever they are needed, so that no more than S is ever resident:
//   FP64 GMRES (solver 0)  the sphere's rows assembled into S; GMRES on the
//                          operator x -> dinv [S_top (xP, -xD); S_bot (hk2 xP, xD)]
//   mixed (solver 2)       the complex float factors (na^2 x 8 bytes, the size of
//                          S's lower half) take that half's memory; the sphere's
//                          rows go to window buffers kept resident as far as
//                          device memory allows, and the rest are assembled
//                          again in every product of the refinement
//   FP64 LU (solver 1)     S turned into the system in place (column factors,
//                          transposed to column-major), factored and solved; the
//                          matrices are consumed, and the velocity recovery
//                          assembles domain 0 again
//   velocity recovery      the sphere's rows assembled once more, window by
//                          window: G1 and H1 into S (for the dpv systems), dG1 and
//                          dH1 into a window buffer that the velocity rows of
//                          q_solve_pv_bv (:689-729) consume at once
#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>
#include "elem_dev.h"

int bqd2_decide_only(bqd2_problem *p, int did, bqd2_assembly_stats *st);
int bqd2_assemble_full(bqd2_problem *p, int did, bqd2_assembly_stats *st);
int bqd2_fs_rows(bqd2_problem *p, const DevDomain &D, hipStream_t st);
int bqd2_lg_assemble(bqd2_problem *p, int did, bqd2_assembly_stats *st);
void bqd2_hk2_table(const bqd2_problem *p, double *tab);

// ------------------------------------------------------------------ kernels
__device__ __forceinline__ cd lg_ld(const double *p, size_t i) { return cd{p[2 * i], p[2 * i + 1]}; }
__device__ __forceinline__ void lg_st(double *p, size_t i, cd v) { p[2 * i] = v.x; p[2 * i + 1] = v.y; }

// dinv[i] = 1 / W[i][c0 + i] for the nr rows of W (row-major, leading dimension ld)
__global__ void __launch_bounds__(256) k_lg_dinv(int nr, const double *__restrict__ W, long long ld, long long c0,
                                                 double *__restrict__ dinv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nr) return;
  lg_st(dinv, i, crecip(lg_ld(W, (size_t)i * ld + c0 + i)));
}

// y[i] = f_i sum_c W[i][c] u[c] over ncol columns, f_i = dinv[i] (dinv given) or
// alpha; one wavefront per row, WPB rows per workgroup
template <int WPB>
__global__ void __launch_bounds__(64 * WPB) k_lg_mv(int nr, int ncol, const double *__restrict__ W, long long ld,
                                                    const double *__restrict__ u, const double *__restrict__ dinv,
                                                    cd alpha, double *__restrict__ y) {
  const int r = blockIdx.x * WPB + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (r >= nr) return;
  const double *w = W + 2 * (size_t)r * ld;
  cd acc = cd{0.0, 0.0};
  for (int c = lane; c < ncol; c += 64) acc += lg_ld(w, c) * lg_ld(u, c);
  acc = cd{wave_sum(acc.x), wave_sum(acc.y)};
  if (lane == 0) lg_st(y, r, acc * (dinv ? lg_ld(dinv, r) : alpha));
}

// the column vectors the operator applies to S's two halves:
// u0 = (xP, -xD) for the lattice rows, u1 = (hk2 xP, xD) for the sphere's
__global__ void __launch_bounds__(256) k_lg_u(int N, const double *__restrict__ X, cd hk2, double *__restrict__ u0,
                                              double *__restrict__ u1) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const cd xp = lg_ld(X, i), xd = lg_ld(X, (size_t)N + i);
  lg_st(u0, i, xp);
  lg_st(u0, (size_t)N + i, -xd);
  lg_st(u1, i, hk2 * xp);
  lg_st(u1, (size_t)N + i, xd);
}

// ub = (hk2 Pi, dPi) over the node columns (c = 4 (s-1) + l <-> node l of element s)
__global__ void __launch_bounds__(256) k_lg_bu(int N, const double *__restrict__ Pi, const double *__restrict__ dPi,
                                               cd hk2, double *__restrict__ ub) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  lg_st(ub, c, hk2 * lg_ld(Pi, (size_t)c + 4));
  lg_st(ub, (size_t)N + c, lg_ld(dPi, (size_t)c + 4));
}

// rows R0 .. R0 + nr - 1 of the row-scaled system, rounded to complex float into
// the column-major Af: a(r, c) = dinv[R0 + r] f W[r][c], f = f0 for c < N and f1
// for c >= N; 32 x 32 tiles through LDS so that reads and writes coalesce
__global__ void __launch_bounds__(256) k_lg_cm32(int nr, int na, int N, const double *__restrict__ W, long long ld,
                                                 cd f0, cd f1, const double *__restrict__ dinv, int R0,
                                                 float2 *__restrict__ Af) {
  __shared__ float2 t[32][33];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int r0 = blockIdx.y * 32, c0 = blockIdx.x * 32;
  for (int i = ty; i < 32; i += 8) {
    const int r = r0 + i, c = c0 + tx;
    float2 v = make_float2(0.0f, 0.0f);
    if (r < nr && c < na) {
      const cd a = lg_ld(W, (size_t)r * ld + c) * (c < N ? f0 : f1) * lg_ld(dinv, (size_t)R0 + r);
      v = make_float2((float)a.x, (float)a.y);
    }
    t[i][tx] = v;
  }
  __syncthreads();
  for (int i = ty; i < 32; i += 8) {
    const int c = c0 + i, r = r0 + tx;
    if (r < nr && c < na) Af[(size_t)c * na + R0 + r] = t[tx][i];
  }
}

// S (row-major na x na, [H0 G0; H1 G1]) -> the system A in column-major order,
// in place: A[r][c] = S[r][c] f(r, c), f = -1 on the lattice rows' G columns, hk2
// on the sphere rows' H columns, 1 elsewhere.  Workgroup (I, J), I <= J, swaps
// the 32 x 32 tiles (I, J) and (J, I) through LDS; the others return at once.
__device__ __forceinline__ cd lg_factor(int r, int c, int N, cd hk2) {
  if (r < N) return c < N ? cd{1.0, 0.0} : cd{-1.0, 0.0};
  return c < N ? hk2 : cd{1.0, 0.0};
}

__global__ void __launch_bounds__(256) k_lg_tsys(int na, int N, cd hk2, double *__restrict__ S) {
  const int I = blockIdx.x, J = blockIdx.y;
  if (I > J) return;
  __shared__ cd t1[32][33], t2[32][33];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int rI = I * 32, rJ = J * 32;
  for (int i = ty; i < 32; i += 8) {
    // t1[i][tx] = A[rI + i][rJ + tx], t2[i][tx] = A[rJ + i][rI + tx]
    int r = rI + i, c = rJ + tx;
    if (r < na && c < na) t1[i][tx] = lg_ld(S, (size_t)r * na + c) * lg_factor(r, c, N, hk2);
    r = rJ + i; c = rI + tx;
    if (r < na && c < na) t2[i][tx] = lg_ld(S, (size_t)r * na + c) * lg_factor(r, c, N, hk2);
  }
  __syncthreads();
  for (int i = ty; i < 32; i += 8) {
    // column-major A: S'[c][r] = A[r][c]; S'[rI + i][rJ + tx] = A[rJ + tx][rI + i] = t2[tx][i]
    int r = rI + i, c = rJ + tx;
    if (r < na && c < na) lg_st(S, (size_t)r * na + c, t2[tx][i]);
    if (I == J) continue;
    r = rJ + i; c = rI + tx;
    if (r < na && c < na) lg_st(S, (size_t)r * na + c, t1[tx][i]);
  }
}

// ------------------------------------------------------------------ host side
static double lg_now() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static cd lg_hk2(const bqd2_problem *p) {
  double tab[16 * 16 * 2];
  memset(tab, 0, sizeof(tab));
  bqd2_hk2_table(p, tab);
  return cd{tab[2 * (0 * 16 + 1)], tab[2 * (0 * 16 + 1) + 1]};
}

static int lg_nwin(const bqd2_problem *p) { return (p->dom[1].N + p->lwin - 1) / p->lwin; }

// the sphere's rows [w0, w0 + nr) held at HG (rows of [H1 G1], row stride na)
static DevDomain lg_win(bqd2_problem *p, int w0, int nr, double *HG, double *dG, double *dH) {
  DevDomain D = p->dom[1];
  D.r0 = w0;
  D.nr = nr;
  D.H = HG;
  D.G = HG + 2 * (size_t)D.N;
  D.ldm = p->na;
  D.dG = dG;
  D.dH = dH;
  D.Fpair = p->lFp;
  return D;
}

// row r of the sphere's half of S
static double *lg_lower(bqd2_problem *p, int r) { return p->S + 2 * ((size_t)p->dom[1].N + r) * (size_t)p->na; }

static int lg_need(bqd2_problem *p, double **buf, size_t bytes) {
  if (*buf) return BQD2_OK;
  return bqd2_alloc(p, (void **)buf, bytes);
}

static size_t lg_win_bytes(const bqd2_problem *p) { return (size_t)p->lwin * p->na * 16; }
static size_t lg_dgh_bytes(const bqd2_problem *p) { return (size_t)2 * p->lwin * p->dom[1].N * 16; }

static void lg_free_pool(bqd2_problem *p) {
  const size_t wb = lg_win_bytes(p);
  for (double *&w : p->lpool) bqd2_release(p, (void **)&w, wb);
  p->lpool.clear();
  bqd2_release(p, (void **)&p->lscr, wb);
}

static void lg_free_velocity(bqd2_problem *p) {
  bqd2_release(p, (void **)&p->ldgw, lg_dgh_bytes(p));
  bqd2_release(p, (void **)&p->ldhw, lg_dgh_bytes(p));
}

template <int WPB>
static void lg_mv_launch(int nr, int ncol, const double *W, long long ld, const double *u, const double *dinv,
                         cd alpha, double *y, hipStream_t st) {
  hipLaunchKernelGGL((k_lg_mv<WPB>), dim3(bqd2_blocks(nr, WPB)), dim3(64 * WPB), 0, st, nr, ncol, W, ld, u, dinv,
                     alpha, y);
}

// the sphere's rows of window w: a resident window buffer, or assembled again
// into the scratch window
static int lg_rows(bqd2_problem *p, int w, double **rows, int *recomputed, hipStream_t st) {
  const int N = p->dom[1].N, w0 = w * p->lwin, nr = std::min(p->lwin, N - w0);
  if (w < (int)p->lpool.size()) {
    *rows = p->lpool[w];
    return BQD2_OK;
  }
  int rc = lg_need(p, &p->lscr, lg_win_bytes(p));
  if (rc) return rc;
  if ((rc = bqd2_fs_rows(p, lg_win(p, w0, nr, p->lscr, nullptr, nullptr), st))) return rc;
  if (recomputed) (*recomputed)++;
  *rows = p->lscr;
  return BQD2_OK;
}

// a progress report every 8 windows of a pass over the sphere's rows
static int lg_report(bqd2_problem *p, const char *stage, int done, hipStream_t st) {
  const int nw = lg_nwin(p);
  if (!bqd2_progress_on() || (done % 8 != 0 && done != nw)) return BQD2_OK;
  BQD2_HIP(hipStreamSynchronize(st));
  bqd2_progress(stage, done, nw);
  return BQD2_OK;
}

// every window of the sphere's rows assembled into S's lower half
static int lg_fill_lower(bqd2_problem *p, hipStream_t st) {
  if (p->s_lower == LG_G1H1) return BQD2_OK;
  const int N = p->dom[1].N;
  p->s_lower = LG_EMPTY;
  p->lu32_valid = false;
  int rc = lg_need(p, &p->lFp, (size_t)p->lwin * p->dom[1].Ne * 3 * 8);
  if (rc) return rc;
  for (int w0 = 0; w0 < N; w0 += p->lwin) {
    const int nr = std::min(p->lwin, N - w0);
    if ((rc = bqd2_fs_rows(p, lg_win(p, w0, nr, lg_lower(p, w0), nullptr, nullptr), st))) return rc;
    lg_report(p, "sphere rows", w0 / p->lwin + 1, st);
  }
  p->s_lower = LG_G1H1;
  return BQD2_OK;
}

// the operator of GMRES: Y = dinv A X with A = [S_top (xP, -xD); S_bot (hk2 xP, xD)]
struct LgOp {
  bqd2_problem *p;
  bool lower_resident;  // S's lower half holds G1, H1; otherwise the windows
  cd hk2;
  double *u0, *u1;
  int recomputed;       // window assemblies inside the products
};

static int lg_op_mv(void *ctx, const double *X, double *Y, hipStream_t st) {
  LgOp *o = (LgOp *)ctx;
  bqd2_problem *p = o->p;
  const int N = p->dom[0].N, na = p->na;
  hipLaunchKernelGGL(k_lg_u, dim3(bqd2_blocks(N, 256)), dim3(256), 0, st, N, X, o->hk2, o->u0, o->u1);
  lg_mv_launch<4>(N, na, p->S, na, o->u0, p->ldinv, cd{0.0, 0.0}, Y, st);
  if (o->lower_resident) {
    lg_mv_launch<4>(N, na, lg_lower(p, 0), na, o->u1, p->ldinv + 2 * (size_t)N, cd{0.0, 0.0}, Y + 2 * (size_t)N, st);
  } else {
    for (int w = 0; w < lg_nwin(p); w++) {
      const int w0 = w * p->lwin, nr = std::min(p->lwin, N - w0);
      double *rows;
      int rc = lg_rows(p, w, &rows, &o->recomputed, st);
      if (rc) return rc;
      lg_mv_launch<4>(nr, na, rows, na, o->u1, p->ldinv + 2 * ((size_t)N + w0), cd{0.0, 0.0},
                      Y + 2 * ((size_t)N + w0), st);
    }
  }
  BQD2_HIP(hipGetLastError());
  return BQD2_OK;
}

// vectors of every solver: b, x, the row scaling, pivots, and the operator's and
// right-hand side's column vectors (u0, u1, ub: 3 na complex)
static int lg_vectors(bqd2_problem *p) {
  const size_t na = (size_t)p->na;
  int rc;
  if ((rc = lg_need(p, &p->b, na * 16))) return rc;
  if ((rc = lg_need(p, &p->xg, na * 16))) return rc;
  if ((rc = lg_need(p, &p->ldinv, na * 16))) return rc;
  if ((rc = lg_need(p, &p->lvec, 3 * na * 16))) return rc;
  if (!p->ipiv && (rc = bqd2_alloc(p, (void **)&p->ipiv, na * sizeof(int)))) return rc;
  return lg_need(p, &p->lFp, (size_t)p->lwin * p->dom[1].Ne * 3 * 8);
}

// ub = (hk2 Pi, dPi) of the current incidence
static int lg_bu(bqd2_problem *p, cd hk2, hipStream_t st) {
  const int N = p->dom[1].N;
  double *ub = p->lvec + 4 * (size_t)p->na;
  hipLaunchKernelGGL(k_lg_bu, dim3(bqd2_blocks(N, 256)), dim3(256), 0, st, N, p->m.Pi, p->m.dPi, hk2, ub);
  BQD2_HIP(hipGetLastError());
  return BQD2_OK;
}

// complex float LU of the row-scaled system in the memory of S's lower half:
// the lattice rows from S, the sphere's rows window by window (kept in resident
// window buffers as far as device memory allows); also the row scaling and the
// right-hand side.  Only the right-hand side when the factors are kept.
static int lg_mixed_build(bqd2_problem *p, cd hk2, bool factor, float bound, hipStream_t st) {
  const int N = p->dom[1].N, na = p->na;
  float2 *Af = (float2 *)lg_lower(p, 0);
  double *ub = p->lvec + 4 * (size_t)na;
  int rc;
  double t0 = lg_now();
  if (factor) {
    p->s_lower = LG_EMPTY;  // from here on the lower half holds the float factors
    p->lu32_valid = false;
    if (p->lpool.empty()) {
      // resident windows: what device memory allows beyond the float LU's
      // workspace, the scratch window and 6 GiB
      size_t fr = 0, tot = 0;
      BQD2_HIP(hipMemGetInfo(&fr, &tot));
      const size_t wb = lg_win_bytes(p);
      const size_t keep = ((size_t)6 << 30) + wb + bqd2_cgetrf_wsize(na) + (size_t)na * 6 * 8 + ((size_t)1 << 30);
      int nres = fr > keep ? (int)((fr - keep) / wb) : 0;
      nres = std::min(nres, lg_nwin(p));
      if (p->lmax_resident >= 0) nres = std::min(nres, p->lmax_resident);
      for (int w = 0; w < nres; w++) {
        double *buf = nullptr;
        if ((rc = bqd2_alloc(p, (void **)&buf, wb))) return rc;
        p->lpool.push_back(buf);
      }
    }
    if (!p->luws && (rc = bqd2_alloc(p, &p->luws, bqd2_cgetrf_wsize(na)))) return rc;
    if (!p->lutmp && (rc = bqd2_alloc(p, (void **)&p->lutmp, (size_t)na * 6 * 8))) return rc;
    // the lattice rows: scaling, rounding (their right-hand side is zero)
    hipLaunchKernelGGL(k_lg_dinv, dim3(bqd2_blocks(N, 256)), dim3(256), 0, st, N, p->S, (long long)na, 0LL,
                       p->ldinv);
    hipLaunchKernelGGL(k_lg_cm32, dim3(bqd2_blocks(na, 32), bqd2_blocks(N, 32)), dim3(256), 0, st, N, na, N, p->S,
                       (long long)na, cd{1.0, 0.0}, cd{-1.0, 0.0}, p->ldinv, 0, Af);
    BQD2_HIP(hipGetLastError());
  }
  BQD2_HIP(hipMemsetAsync(p->b, 0, (size_t)N * 16, st));
  if ((rc = lg_bu(p, hk2, st))) return rc;
  double ms_rows = 0.0;
  for (int w = 0; w < lg_nwin(p); w++) {
    const int w0 = w * p->lwin, nr = std::min(p->lwin, N - w0);
    double *rows;
    const double tr = lg_now();
    if (w < (int)p->lpool.size() && factor) {
      // a resident window is assembled once per factorisation
      rows = p->lpool[w];
      if ((rc = bqd2_fs_rows(p, lg_win(p, w0, nr, rows, nullptr, nullptr), st))) return rc;
    } else {
      if ((rc = lg_rows(p, w, &rows, nullptr, st))) return rc;
    }
    BQD2_HIP(hipStreamSynchronize(st));
    ms_rows += lg_now() - tr;
    lg_report(p, "float system rows", w + 1, st);
    double *bw = p->b + 2 * ((size_t)N + w0);
    lg_mv_launch<4>(nr, na, rows, na, ub, nullptr, cd{-1.0, 0.0}, bw, st);
    if (factor) {
      double *dw = p->ldinv + 2 * ((size_t)N + w0);
      hipLaunchKernelGGL(k_lg_dinv, dim3(bqd2_blocks(nr, 256)), dim3(256), 0, st, nr, rows, (long long)na,
                         (long long)N + w0, dw);
      hipLaunchKernelGGL(k_lg_cm32, dim3(bqd2_blocks(na, 32), bqd2_blocks(nr, 32)), dim3(256), 0, st, nr, na, N,
                         rows, (long long)na, hk2, cd{1.0, 0.0}, p->ldinv, N + w0, Af);
    }
    BQD2_HIP(hipGetLastError());
  }
  BQD2_HIP(hipStreamSynchronize(st));
  p->lst.ms_rows += ms_rows;
  p->lst.ms_build += lg_now() - t0 - ms_rows;
  p->lst.resident = (int)p->lpool.size();
  if (!factor) return BQD2_OK;
  const double tf = lg_now();
  rc = bqd2_cgetrf_bound(na, (float *)Af, p->ipiv, p->luws, st, bound, &p->lu32_growth);
  BQD2_HIP(hipStreamSynchronize(st));
  p->lst.ms_factor += lg_now() - tf;
  p->ms_factor = p->lst.ms_factor;
  if (rc) return rc;
  p->s_lower = LG_AF;
  p->lu32_valid = true;
  return BQD2_OK;
}

#define LG_TOL 1.0e-15
#define LG_RESTART 30
#define LG_MAXIT 60
#define LG_ACCEPT 1.0e-12  // a refinement that stalls above LG_TOL is accepted below this

static int lg_solve_mixed(bqd2_problem *p, cd hk2, hipStream_t st, const double **x) {
  int rc;
  const bool have = p->lu32_valid && p->s_lower == LG_AF;
  if (!have) {
    rc = lg_mixed_build(p, hk2, true, BQD2_LU32_RELAXED_GROWTH, st);
    if (rc == BQD2_E_SINGULAR) rc = lg_mixed_build(p, hk2, true, BQD2_LU32_STABLE_GROWTH, st);
  } else {
    rc = lg_mixed_build(p, hk2, false, 0.0f, st);
  }
  if (rc) return rc;
  LgOp o{p, false, hk2, p->lvec, p->lvec + 2 * (size_t)p->na, 0};
  GmOp op{p->ldinv, lg_op_mv, &o};
  Precond pc{(const float *)lg_lower(p, 0), p->luws, p->lutmp, &p->lu_graph};
  double t0 = lg_now();
  rc = bqd2_gmres(&p->main, p->na, 1, nullptr, 0, nullptr, p->b, p->na, p->xg, p->na, LG_TOL, LG_RESTART, LG_MAXIT,
                  &p->info_its[0], &p->info_res[0], &pc, &op);
  if (rc == 1 && p->lu32_growth > BQD2_LU32_STABLE_GROWTH && p->info_res[0] > LG_ACCEPT) {
    // the relaxed pivots made too weak a preconditioner: factors with the stable bound
    p->lst.ms_solve += lg_now() - t0;
    if ((rc = lg_mixed_build(p, hk2, true, BQD2_LU32_STABLE_GROWTH, st))) return rc;
    t0 = lg_now();
    rc = bqd2_gmres(&p->main, p->na, 1, nullptr, 0, nullptr, p->b, p->na, p->xg, p->na, LG_TOL, LG_RESTART, LG_MAXIT,
                    &p->info_its[0], &p->info_res[0], &pc, &op);
  }
  p->lst.ms_solve += lg_now() - t0;
  p->lst.recomputed = o.recomputed;
  if (rc < 0) return rc;
  if (rc == 1 && p->info_res[0] > LG_ACCEPT) return 1;
  *x = p->xg;
  return BQD2_OK;
}

static int lg_solve_gmres(bqd2_problem *p, cd hk2, hipStream_t st, const double **x) {
  const int N = p->dom[1].N, na = p->na;
  int rc;
  double t0 = lg_now();
  if ((rc = lg_fill_lower(p, st))) return rc;
  BQD2_HIP(hipStreamSynchronize(st));
  p->lst.ms_rows += lg_now() - t0;
  t0 = lg_now();
  // every system row's diagonal is S's; only the sphere's rows carry a right-hand side
  hipLaunchKernelGGL(k_lg_dinv, dim3(bqd2_blocks(na, 256)), dim3(256), 0, st, na, p->S, (long long)na, 0LL,
                     p->ldinv);
  BQD2_HIP(hipMemsetAsync(p->b, 0, (size_t)N * 16, st));
  if ((rc = lg_bu(p, hk2, st))) return rc;
  lg_mv_launch<4>(N, na, lg_lower(p, 0), na, p->lvec + 4 * (size_t)na, nullptr, cd{-1.0, 0.0},
                  p->b + 2 * (size_t)N, st);
  BQD2_HIP(hipGetLastError());
  BQD2_HIP(hipStreamSynchronize(st));
  p->lst.ms_build += lg_now() - t0;
  LgOp o{p, true, hk2, p->lvec, p->lvec + 2 * (size_t)na, 0};
  GmOp op{p->ldinv, lg_op_mv, &o};
  t0 = lg_now();
  rc = bqd2_gmres(&p->main, na, 1, nullptr, 0, nullptr, p->b, na, p->xg, na, LG_TOL, 100, 400, &p->info_its[0],
                  &p->info_res[0], nullptr, &op);
  p->lst.ms_solve += lg_now() - t0;
  if (rc < 0) return rc;
  if (rc == 1 && p->info_res[0] > LG_ACCEPT) return 1;
  *x = p->xg;
  return BQD2_OK;
}

// FP64 LU with partial pivoting of the whole system, in place on S
static int lg_solve_lu(bqd2_problem *p, cd hk2, hipStream_t st, const double **x) {
  const int N = p->dom[1].N, na = p->na;
  int rc;
  double t0 = lg_now();
  if ((rc = lg_fill_lower(p, st))) return rc;
  BQD2_HIP(hipStreamSynchronize(st));
  p->lst.ms_rows += lg_now() - t0;
  t0 = lg_now();
  BQD2_HIP(hipMemsetAsync(p->b, 0, (size_t)N * 16, st));
  if ((rc = lg_bu(p, hk2, st))) return rc;
  lg_mv_launch<4>(N, na, lg_lower(p, 0), na, p->lvec + 4 * (size_t)na, nullptr, cd{-1.0, 0.0},
                  p->b + 2 * (size_t)N, st);
  // the matrices become the system, column-major: consumed from here on
  p->dom[0].assembled = false;
  p->s_lower = LG_EMPTY;
  p->lu32_valid = false;
  const int nt = bqd2_blocks(na, 32);
  hipLaunchKernelGGL(k_lg_tsys, dim3(nt, nt), dim3(256), 0, st, na, N, hk2, p->S);
  BQD2_HIP(hipGetLastError());
  BQD2_HIP(hipStreamSynchronize(st));
  p->lst.ms_build += lg_now() - t0;
  t0 = lg_now();
  rc = bqd2_zgetrf(na, p->S, p->ipiv, st);
  BQD2_HIP(hipStreamSynchronize(st));
  p->lst.ms_factor += lg_now() - t0;
  if (rc) return rc;
  t0 = lg_now();
  if ((rc = bqd2_zgetrs(na, 1, p->S, p->ipiv, p->b, st))) return rc;
  BQD2_HIP(hipStreamSynchronize(st));
  p->lst.ms_solve += lg_now() - t0;
  p->info_lu++;
  *x = p->b;
  return BQD2_OK;
}

// bqd2_solve_surface of the system layout; x: the solution on the device
int bqd2_lg_solve(bqd2_problem *p, const double **x) {
  BQD2_CHECK(p->dom[0].assembled && p->dom[1].assembled, BQD2_E_STATE,
             "bqd2_solve_surface: both domains must be assembled (domain 0 again after an FP64 LU)");
  hipStream_t st = p->stream;
  int rc;
  if ((rc = lg_vectors(p))) return rc;
  const cd hk2 = lg_hk2(p);
  const int windows = p->lst.windows;
  memset(&p->lst, 0, sizeof(p->lst));
  p->lst.windows = windows;
  p->lst.resident = (int)p->lpool.size();
  *x = nullptr;
  rc = 1;
  if (p->solver == 2) {
    if ((rc = lg_solve_mixed(p, hk2, st, x)) < 0) return rc;
  }
  if (rc == 1 && p->solver == 0) {
    if ((rc = lg_solve_gmres(p, hk2, st, x)) < 0) return rc;
  }
  if (rc == 1) {
    // the FP64 LU: asked for, or the fallback of an iteration that missed its tolerance
    lg_free_pool(p);
    if ((rc = lg_solve_lu(p, hk2, st, x))) return rc;
  }
  return BQD2_OK;
}

// the velocity rows of the sphere (k_velocity_rows of solve.hip), window by
// window with their dG, dH; S's lower half receives G1, H1 for the dpv systems.
// Domain 0 is assembled again first when an FP64 LU consumed it.
int bqd2_velocity_rows(bqd2_problem *p, const DevDomain &D, const void *W, double *open_out, hipStream_t st);
int bqd2_lg_velocity(bqd2_problem *p, const void *W, double *open_out) {
  hipStream_t st = p->stream;
  const int N = p->dom[1].N;
  int rc;
  p->lst.ms_velocity_rows = p->lst.ms_reassemble = 0.0;
  if (!p->dom[0].assembled) {
    const double t0 = lg_now();
    if ((rc = bqd2_lg_assemble(p, 0, nullptr))) return rc;
    p->lst.ms_reassemble = lg_now() - t0;
  }
  const double t0 = lg_now();
  // the resident windows of the mixed solver are stale once S holds G1, H1
  lg_free_pool(p);
  if ((rc = lg_need(p, &p->ldgw, lg_dgh_bytes(p)))) return rc;
  if ((rc = lg_need(p, &p->ldhw, lg_dgh_bytes(p)))) return rc;
  if ((rc = lg_need(p, &p->lFp, (size_t)p->lwin * p->dom[1].Ne * 3 * 8))) return rc;
  p->s_lower = LG_EMPTY;
  p->lu32_valid = false;
  for (int w0 = 0; w0 < N; w0 += p->lwin) {
    const int nr = std::min(p->lwin, N - w0);
    const DevDomain Dw = lg_win(p, w0, nr, lg_lower(p, w0), p->ldgw, p->ldhw);
    if ((rc = bqd2_fs_rows(p, Dw, st))) return rc;
    if ((rc = bqd2_velocity_rows(p, Dw, W, open_out, st))) return rc;
    lg_report(p, "sphere rows with dG, dH", w0 / p->lwin + 1, st);
  }
  p->s_lower = LG_G1H1;
  lg_free_velocity(p);
  BQD2_HIP(hipStreamSynchronize(st));
  p->lst.ms_velocity_rows = lg_now() - t0;
  return BQD2_OK;
}

// bqd2_assemble of the system layout
int bqd2_lg_assemble(bqd2_problem *p, int did, bqd2_assembly_stats *st) {
  DevDomain &D = p->dom[did];
  int rc;
  if (did == 1) {
    // the sphere's rows are assembled when a solve or the velocity recovery needs them
    lg_free_pool(p);
    p->s_lower = LG_EMPTY;
    p->lu32_valid = false;
    p->lst.windows = lg_nwin(p);
    if ((rc = bqd2_decide_only(p, 1, st))) return rc;
    D.assembled = true;
    return BQD2_OK;
  }
  // domain 0 into S's upper half; its per-pair F sums (and, the first time, the
  // pair order keys) in buffers released afterwards
  const size_t np = (size_t)D.N * D.Ne;
  if ((rc = bqd2_alloc(p, (void **)&D.Fpair, np * 3 * 8))) return rc;
  const bool first = D.plist_n == 0;
  if (first && (rc = bqd2_alloc(p, (void **)&D.pkey, np * sizeof(float)))) {
    bqd2_release(p, (void **)&D.Fpair, np * 3 * 8);
    return rc;
  }
  rc = bqd2_assemble_full(p, 0, st);
  (void)hipStreamSynchronize(p->stream);
  bqd2_release(p, (void **)&D.Fpair, np * 3 * 8);
  if (first) bqd2_release(p, (void **)&D.pkey, np * sizeof(float));
  if (rc) D.assembled = false;
  return rc;
}

extern "C" int bqd2_large_config(bqd2_problem *p, int rows_per_window, int max_resident) {
  BQD2_CHECK(p && p->large, BQD2_E_ARG, "bqd2_large_config: not a problem of bqd2_create_large");
  BQD2_CHECK(rows_per_window >= 64 && rows_per_window <= 65535 && max_resident >= -1, BQD2_E_ARG,
             "bqd2_large_config: rows per window must lie in [64, 65535], max_resident >= -1");
  lg_free_pool(p);
  lg_free_velocity(p);
  bqd2_release(p, (void **)&p->lFp, (size_t)p->lwin * p->dom[1].Ne * 3 * 8);
  p->lwin = rows_per_window;
  p->lmax_resident = max_resident;
  p->lst.windows = lg_nwin(p);
  p->lu32_valid = false;
  if (p->s_lower == LG_AF) p->s_lower = LG_EMPTY;
  return BQD2_OK;
}

extern "C" int bqd2_large_info(bqd2_problem *p, bqd2_large_stats *st) {
  BQD2_CHECK(p && p->large && st, BQD2_E_ARG, "bqd2_large_info: not a problem of bqd2_create_large");
  *st = p->lst;
  st->windows = lg_nwin(p);
  return BQD2_OK;
}
