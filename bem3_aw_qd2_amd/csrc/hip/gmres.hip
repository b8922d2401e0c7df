// gmres.hip -- restarted GMRES for the dense complex systems of the solve.
//
// The reference solves its boundary-value system with MKL PARDISO
// (d3qd2_mkl_solver.c:18-148) and the per-domain dpv systems G dpv = H pv with
// LAPACKE_zgesv (bem3_aw_qd2_solve_bieq.c:617), both direct.  A dense LU costs
// one pivot step per unknown on the critical path; these systems are of the
// second kind after row scaling and converge in some tens of Krylov steps, each
// one pass over the matrix in HBM.  GMRES here: left Jacobi scaling (rows
// divided by their diagonal), classical Gram-Schmidt applied twice, Givens
// rotations and the small least-squares solve on the host, restarts, and a
// final check of the true scaled residual.  Several right-hand sides that share
// the matrix run in lock step (one pass over the matrix per step for all).  The
// caller falls back to the LU when the tolerance is not met.
//
// With a preconditioner (the complex float LU of the row-scaled matrix,
// lu32.hip) the iteration is right-preconditioned: the Krylov space is built on
// D A M^-1 and the update is X += M^-1 V y, so the residual monitored is still
// the true scaled residual D (b - A X) of the FP64 system.  This is GMRES-based
// iterative refinement: the float factors set the convergence rate, the FP64
// matrix-vector products set the attainable accuracy.
//
// Matrices are row-major, lda >= n.  An optional row map marks unused rows
// (map[r] < 0, the fourth slot of a triangle): they are solved as identity rows
// with a zero right-hand side.
#include <complex>
#include <cstring>
#include <vector>
#include "elem_dev.h"
#include "tune.h"

typedef std::complex<double> zc;

__device__ __forceinline__ cd ldc(const double *p, size_t i) { return cd{p[2 * i], p[2 * i + 1]}; }
__device__ __forceinline__ void stc(double *p, size_t i, cd v) { p[2 * i] = v.x; p[2 * i + 1] = v.y; }

__device__ __forceinline__ double block_sum(double v, double *sh) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < nw; i++) s += sh[i];
  return s;
}

// dinv[r] = 1 / A[r][r] (1 on unused rows)
__global__ void __launch_bounds__(256) k_gm_dinv(int n, const double *__restrict__ A, long long lda,
                                                 const int *__restrict__ map, double *__restrict__ dinv) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  cd d = cd{1.0, 0.0};
  if (!(map && map[r] < 0)) d = crecip(ldc(A, (size_t)r * lda + r));
  stc(dinv, r, d);
}

// Y_s = dinv (A X_s) for s < NR (identity on unused rows); one wavefront per row
template <int NR, int WPB>
__global__ void __launch_bounds__(64 * WPB) k_gm_mv(int n, const double *__restrict__ A, long long lda,
                                               const int *__restrict__ map, const double *__restrict__ dinv,
                                               const double *__restrict__ X, long long ldx,
                                               double *__restrict__ Y, long long ldy) {
  const int r = blockIdx.x * WPB + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (r >= n) return;
  if (map && map[r] < 0) {
    if (lane < NR) stc(Y + 2 * (size_t)lane * ldy, r, ldc(X + 2 * (size_t)lane * ldx, r));
    return;
  }
  const double *a = A + 2 * (size_t)r * lda;
  cd acc[NR];
#pragma unroll
  for (int s = 0; s < NR; s++) acc[s] = cd{0.0, 0.0};
  for (int c = lane; c < n; c += 64) {
    const cd v = ldc(a, c);
#pragma unroll
    for (int s = 0; s < NR; s++) acc[s] += v * ldc(X + 2 * (size_t)s * ldx, c);
  }
  const cd di = ldc(dinv, r);
#pragma unroll
  for (int s = 0; s < NR; s++) {
    cd v = cd{wave_sum(acc[s].x), wave_sum(acc[s].y)} * di;
    if (lane == 0) stc(Y + 2 * (size_t)s * ldy, r, v);
  }
}

// h[s][i] = <V_{s,i}, W_s> for i <= j; grid (j + 1, nrhs)
__global__ void __launch_bounds__(256) k_gm_dots(int n, const double *__restrict__ V, long long ldv, long long vstride,
                                                 const double *__restrict__ W, long long ldw,
                                                 double *__restrict__ h, int hstride) {
  __shared__ double sh[8];
  const int i = blockIdx.x, s = blockIdx.y;
  const double *v = V + 2 * ((size_t)s * vstride + (size_t)i * ldv);
  const double *w = W + 2 * (size_t)s * ldw;
  double re = 0.0, im = 0.0;
  for (int r = threadIdx.x; r < n; r += blockDim.x) {
    const cd a = ldc(v, r), b = ldc(w, r);
    re += a.x * b.x + a.y * b.y;
    im += a.x * b.y - a.y * b.x;
  }
  re = block_sum(re, sh);
  im = block_sum(im, sh);
  if (threadIdx.x == 0) stc(h, (size_t)s * hstride + i, cd{re, im});
}

// W_s -= sum_{i <= j} V_{s,i} h[s][i]
__global__ void __launch_bounds__(256) k_gm_axpy(int n, int j, const double *__restrict__ V, long long ldv,
                                                 long long vstride, const double *__restrict__ h, int hstride,
                                                 double *__restrict__ W, long long ldw) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x, s = blockIdx.y;
  if (r >= n) return;
  const double *v = V + 2 * (size_t)s * vstride;
  cd acc = cd{0.0, 0.0};
  for (int i = 0; i <= j; i++) acc += ldc(v, (size_t)i * ldv + r) * ldc(h, (size_t)s * hstride + i);
  double *w = W + 2 * (size_t)s * ldw;
  stc(w, r, ldc(w, r) - acc);
}

// out[s] = |W_s|
__global__ void __launch_bounds__(256) k_gm_norm(int n, const double *__restrict__ W, long long ldw,
                                                 double *__restrict__ out) {
  __shared__ double sh[8];
  const int s = blockIdx.x;
  const double *w = W + 2 * (size_t)s * ldw;
  double a = 0.0;
  for (int r = threadIdx.x; r < n; r += blockDim.x) {
    const cd v = ldc(w, r);
    a += v.x * v.x + v.y * v.y;
  }
  a = block_sum(a, sh);
  if (threadIdx.x == 0) out[s] = sqrt(a);
}

// V_{s,k} = W_s / nrm[s] (left as is when the norm vanishes)
__global__ void __launch_bounds__(256) k_gm_scale(int n, const double *__restrict__ W, long long ldw,
                                                  const double *__restrict__ nrm, double *__restrict__ V,
                                                  long long ldv, long long vstride, int k) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x, s = blockIdx.y;
  if (r >= n) return;
  const double a = nrm[s], f = a > 0.0 ? 1.0 / a : 1.0;
  stc(V + 2 * ((size_t)s * vstride + (size_t)k * ldv), r, ldc(W + 2 * (size_t)s * ldw, r) * f);
}

// R_s = Db_s - Y_s
__global__ void __launch_bounds__(256) k_gm_sub(int n, const double *__restrict__ Db, const double *__restrict__ Y,
                                                double *__restrict__ R, long long ld) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x, s = blockIdx.y;
  if (r >= n) return;
  const size_t o = (size_t)s * ld + r;
  stc(R, o, ldc(Db, o) - ldc(Y, o));
}

// Db_s = dinv B_s
__global__ void __launch_bounds__(256) k_gm_scale_rhs(int n, const double *__restrict__ dinv,
                                                      const double *__restrict__ B, long long ldb,
                                                      double *__restrict__ Db, long long ld) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x, s = blockIdx.y;
  if (r >= n) return;
  stc(Db, (size_t)s * ld + r, ldc(dinv, r) * ldc(B + 2 * (size_t)s * ldb, r));
}

// X_s += sum_{i < k} V_{s,i} y[s][i]
__global__ void __launch_bounds__(256) k_gm_update(int n, int k, const double *__restrict__ V, long long ldv,
                                                   long long vstride, const double *__restrict__ y, int ystride,
                                                   double *__restrict__ X, long long ldx) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x, s = blockIdx.y;
  if (r >= n) return;
  const double *v = V + 2 * (size_t)s * vstride;
  cd acc = cd{0.0, 0.0};
  for (int i = 0; i < k; i++) acc += ldc(v, (size_t)i * ldv + r) * ldc(y, (size_t)s * ystride + i);
  double *x = X + 2 * (size_t)s * ldx;
  stc(x, r, ldc(x, r) + acc);
}

// WPB rows per workgroup (one wavefront each), as the tuner picks: each row's
// sum runs in its own wavefront whatever the workgroup size
template <int WPB>
static void launch_mv_wpb(int nrhs, int n, const double *A, long long lda, const int *map, const double *dinv,
                          const double *X, long long ldx, double *Y, long long ldy, hipStream_t st) {
  dim3 g(bqd2_blocks(n, WPB)), b(64 * WPB);
  if (nrhs == 1) hipLaunchKernelGGL((k_gm_mv<1, WPB>), g, b, 0, st, n, A, lda, map, dinv, X, ldx, Y, ldy);
  else if (nrhs == 2) hipLaunchKernelGGL((k_gm_mv<2, WPB>), g, b, 0, st, n, A, lda, map, dinv, X, ldx, Y, ldy);
  else hipLaunchKernelGGL((k_gm_mv<3, WPB>), g, b, 0, st, n, A, lda, map, dinv, X, ldx, Y, ldy);
}

static void launch_mv(int nrhs, int n, const double *A, long long lda, const int *map, const double *dinv,
                      const double *X, long long ldx, double *Y, long long ldy, hipStream_t st) {
  const TuneTicket tk = tune_begin(TUNE_GMV, 8LL * tune_class(n) + nrhs, 16.0 * n * n, st);
  switch (tk.cand) {
    case 1: launch_mv_wpb<2>(nrhs, n, A, lda, map, dinv, X, ldx, Y, ldy, st); break;
    case 2: launch_mv_wpb<8>(nrhs, n, A, lda, map, dinv, X, ldx, Y, ldy, st); break;
    case 3: launch_mv_wpb<1>(nrhs, n, A, lda, map, dinv, X, ldx, Y, ldy, st); break;
    default: launch_mv_wpb<4>(nrhs, n, A, lda, map, dinv, X, ldx, Y, ldy, st);
  }
  tune_end(tk, st);
}

// X_s += Z_s
__global__ void __launch_bounds__(256) k_gm_add(int n, const double *__restrict__ Z, long long ldz,
                                                double *__restrict__ X, long long ldx) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x, s = blockIdx.y;
  if (r >= n) return;
  double *x = X + 2 * (size_t)s * ldx;
  stc(x, r, ldc(x, r) + ldc(Z + 2 * (size_t)s * ldz, r));
}

int bqd2_jacobi(int n, const double *A, long long lda, const int *map, double *dinv, hipStream_t st) {
  hipLaunchKernelGGL(k_gm_dinv, dim3(bqd2_blocks(n, 256)), dim3(256), 0, st, n, A, lda, map, dinv);
  BQD2_HIP(hipGetLastError());
  return BQD2_OK;
}

int bqd2_gmres(CallCtx *cx, int n, int nrhs, const double *A, long long lda, const int *map, const double *B,
               long long ldb, double *X, long long ldx, double tol, int m, int maxit, int *its, double *relres,
               const Precond *pc, const GmOp *op) {
  BQD2_CHECK(n > 0 && nrhs >= 1 && nrhs <= 3 && m >= 2 && (!op || nrhs == 1), BQD2_E_ARG, "bqd2_gmres: bad arguments");
  hipStream_t st = cx->st;
  // workspace: V (nrhs x (m+1) x n), W, R, Db (nrhs x n), dinv (n), h (nrhs x (2(m+1)+1)), y (nrhs x m)
  const long long vstride = (long long)(m + 1) * n;
  const int hstride = 2 * (m + 1) + 2;
  const size_t bV = (size_t)nrhs * vstride * 16, bN = (size_t)nrhs * n * 16, bD = (size_t)n * 16;
  const size_t bH = (size_t)nrhs * hstride * 16, bY = (size_t)nrhs * m * 16;
  const size_t total = bV + 4 * bN + bD + bH + bY + 8 * 256;
  char *base;
  int rc = bqd2_workspace(cx, total, &base);
  if (rc) return rc;
  auto take = [&](size_t bytes) { double *q = (double *)base; base += (bytes + 255) / 256 * 256; return q; };
  double *V = take(bV), *W = take(bN), *R = take(bN), *Db = take(bN), *dinv = take(bD), *H = take(bH), *Y = take(bY);
  double *Z = take(bN);  // M^-1 of the basis vector / of the update (preconditioned runs)
  const dim3 gv(bqd2_blocks(n, 256), nrhs), b256(256);
  // Y = dinv A X: the dense row-major matrix, or the caller's operator
  auto mv = [&](const double *Xv, long long ldxv, double *Yv) -> int {
    if (op) return op->mv(op->ctx, Xv, Yv, st);
    launch_mv(nrhs, n, A, lda, map, dinv, Xv, ldxv, Yv, (long long)n, st);
    return BQD2_OK;
  };

  if (op) BQD2_HIP(hipMemcpyAsync(dinv, op->dinv, bD, hipMemcpyDeviceToDevice, st));
  else hipLaunchKernelGGL(k_gm_dinv, dim3(bqd2_blocks(n, 256)), b256, 0, st, n, A, lda, map, dinv);
  hipLaunchKernelGGL(k_gm_scale_rhs, gv, b256, 0, st, n, dinv, B, ldb, Db, (long long)n);
  for (int s = 0; s < nrhs; s++) BQD2_HIP(hipMemsetAsync(X + 2 * (size_t)s * ldx, 0, (size_t)n * 16, st));
  BQD2_HIP(hipGetLastError());
  // |Db_s| for the relative residual
  std::vector<double> hb((size_t)nrhs * hstride * 2), bnorm(nrhs), res(nrhs, 1.0);
  hipLaunchKernelGGL(k_gm_norm, dim3(nrhs), b256, 0, st, n, Db, (long long)n, H);
  BQD2_HIP(hipMemcpyAsync(bnorm.data(), H, nrhs * sizeof(double), hipMemcpyDeviceToHost, st));
  BQD2_HIP(hipStreamSynchronize(st));
  for (int s = 0; s < nrhs; s++)
    if (bnorm[s] == 0.0) bnorm[s] = 1.0;

  int total_its = 0;
  bool first = true;
  std::vector<zc> Hh((size_t)nrhs * (m + 1) * m), g((size_t)nrhs * (m + 1)), cs((size_t)nrhs * m),
      sn((size_t)nrhs * m), yv((size_t)nrhs * m);
  std::vector<double> beta(nrhs);
  for (;;) {
    // residual R = Db - dinv A X (= Db while X is zero)
    if (first) {
      BQD2_HIP(hipMemcpyAsync(R, Db, bN, hipMemcpyDeviceToDevice, st));
    } else {
      if ((rc = mv(X, ldx, W))) return rc;
      hipLaunchKernelGGL(k_gm_sub, gv, b256, 0, st, n, Db, W, R, (long long)n);
    }
    hipLaunchKernelGGL(k_gm_norm, dim3(nrhs), b256, 0, st, n, R, (long long)n, H);
    BQD2_HIP(hipGetLastError());
    BQD2_HIP(hipMemcpyAsync(beta.data(), H, nrhs * sizeof(double), hipMemcpyDeviceToHost, st));
    BQD2_HIP(hipStreamSynchronize(st));
    bool done = true;
    for (int s = 0; s < nrhs; s++) {
      res[s] = beta[s] / bnorm[s];
      done = done && res[s] <= tol;
    }
    if (done || total_its >= maxit) break;
    first = false;
    hipLaunchKernelGGL(k_gm_scale, gv, b256, 0, st, n, R, (long long)n, H, V, (long long)n, vstride, 0);
    std::fill(g.begin(), g.end(), zc(0.0));
    for (int s = 0; s < nrhs; s++) g[(size_t)s * (m + 1)] = beta[s];
    int k = 0;  // basis vectors of this cycle used in the update
    for (int j = 0; j < m && total_its < maxit; j++) {
      if (pc) {
        int rc2 = bqd2_lu32_apply(n, nrhs, pc->Af, pc->ws, V + 2 * (size_t)j * n, vstride, Z, (long long)n, pc->tmp,
                                  st, pc->graph);
        if (rc2) return rc2;
        if ((rc = mv(Z, (long long)n, W))) return rc;
      } else {
        if ((rc = mv(V + 2 * (size_t)j * n, vstride, W))) return rc;
      }
      // classical Gram-Schmidt, twice; h of pass p at H[s][p (m+1) + i], the norm at H[s][2 (m+1)]
      for (int pass = 0; pass < 2; pass++) {
        double *hp = H + 2 * (size_t)pass * (m + 1);
        hipLaunchKernelGGL(k_gm_dots, dim3(j + 1, nrhs), b256, 0, st, n, V, (long long)n, vstride, W, (long long)n,
                           hp, hstride);
        hipLaunchKernelGGL(k_gm_axpy, gv, b256, 0, st, n, j, V, (long long)n, vstride, hp, hstride, W, (long long)n);
      }
      // the norms go after the two passes of every system: H[s][2 (m+1)] (real part)
      double *hn = H + 2 * (size_t)2 * (m + 1);
      hipLaunchKernelGGL(k_gm_norm, dim3(nrhs), b256, 0, st, n, W, (long long)n, hn);
      BQD2_HIP(hipGetLastError());
      BQD2_HIP(hipMemcpyAsync(hb.data(), H, bH, hipMemcpyDeviceToHost, st));
      BQD2_HIP(hipStreamSynchronize(st));
      total_its++;
      k = j + 1;
      bool all = true;
      std::vector<double> nrm(nrhs);
      for (int s = 0; s < nrhs; s++) {
        const double *hs = hb.data() + 2 * (size_t)s * hstride;
        // the norm of system s sits at element 2 (m+1) of system 0's row block
        // (k_gm_norm writes out[s] contiguously), so read it from there
        nrm[s] = (hb.data() + 2 * (size_t)2 * (m + 1))[s];
        zc *Hc = Hh.data() + (size_t)s * (m + 1) * m;  // column-major (m+1) x m
        for (int i = 0; i <= j; i++)
          Hc[(size_t)j * (m + 1) + i] = zc(hs[2 * i], hs[2 * i + 1]) +
                                        zc(hs[2 * (m + 1 + i)], hs[2 * (m + 1 + i) + 1]);
        Hc[(size_t)j * (m + 1) + j + 1] = nrm[s];
        for (int i = 0; i < j; i++) {
          zc a = Hc[(size_t)j * (m + 1) + i], b = Hc[(size_t)j * (m + 1) + i + 1];
          zc c = cs[(size_t)s * m + i], sv = sn[(size_t)s * m + i];
          Hc[(size_t)j * (m + 1) + i] = std::conj(c) * a + std::conj(sv) * b;
          Hc[(size_t)j * (m + 1) + i + 1] = -sv * a + c * b;
        }
        zc a = Hc[(size_t)j * (m + 1) + j], b = Hc[(size_t)j * (m + 1) + j + 1];
        double den = std::sqrt(std::norm(a) + std::norm(b));
        zc c = den > 0.0 ? a / den : zc(1.0), sv = den > 0.0 ? b / den : zc(0.0);
        cs[(size_t)s * m + j] = c;
        sn[(size_t)s * m + j] = sv;
        Hc[(size_t)j * (m + 1) + j] = std::conj(c) * a + std::conj(sv) * b;
        Hc[(size_t)j * (m + 1) + j + 1] = 0.0;
        zc g0 = g[(size_t)s * (m + 1) + j];
        g[(size_t)s * (m + 1) + j] = std::conj(c) * g0;
        g[(size_t)s * (m + 1) + j + 1] = -sv * g0;
        res[s] = std::abs(g[(size_t)s * (m + 1) + j + 1]) / bnorm[s];
        all = all && (res[s] <= tol || nrm[s] == 0.0);
      }
      if (all || j == m - 1 || total_its >= maxit) break;
      hipLaunchKernelGGL(k_gm_scale, gv, b256, 0, st, n, W, (long long)n, hn, V, (long long)n, vstride, j + 1);
    }
    // y = H^-1 g (upper triangular k x k), X += V y
    for (int s = 0; s < nrhs; s++) {
      const zc *Hc = Hh.data() + (size_t)s * (m + 1) * m;
      for (int i = k - 1; i >= 0; i--) {
        zc acc = g[(size_t)s * (m + 1) + i];
        for (int q = i + 1; q < k; q++) acc -= Hc[(size_t)q * (m + 1) + i] * yv[(size_t)s * m + q];
        zc d = Hc[(size_t)i * (m + 1) + i];
        yv[(size_t)s * m + i] = (d != zc(0.0)) ? acc / d : zc(0.0);
      }
    }
    BQD2_HIP(hipMemcpyAsync(Y, yv.data(), bY, hipMemcpyHostToDevice, st));
    if (pc) {
      // X += M^-1 (V y)
      BQD2_HIP(hipMemsetAsync(Z, 0, bN, st));
      hipLaunchKernelGGL(k_gm_update, gv, b256, 0, st, n, k, V, (long long)n, vstride, Y, m, Z, (long long)n);
      int rc2 = bqd2_lu32_apply(n, nrhs, pc->Af, pc->ws, Z, (long long)n, Z, (long long)n, pc->tmp, st, pc->graph);
      if (rc2) return rc2;
      hipLaunchKernelGGL(k_gm_add, gv, b256, 0, st, n, Z, (long long)n, X, ldx);
    } else {
      hipLaunchKernelGGL(k_gm_update, gv, b256, 0, st, n, k, V, (long long)n, vstride, Y, m, X, ldx);
    }
    BQD2_HIP(hipGetLastError());
    BQD2_HIP(hipStreamSynchronize(st));
  }
  if (its) *its = total_its;
  double worst = 0.0;
  for (int s = 0; s < nrhs; s++) worst = res[s] > worst ? res[s] : worst;
  if (relres) *relres = worst;
  return worst <= tol ? 0 : 1;
}
