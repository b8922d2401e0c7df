// lu.hip -- dense complex double LU with partial pivoting and the matching
// triangular solves, on column-major device matrices.
//
// Replaces the reference's direct solves: MKL PARDISO on the 2NN system
// (d3qd2_mkl_solver.c:18-148) and LAPACKE_zgesv for the per-domain 3-RHS
// systems (bem3_aw_qd2_solve_bieq.c:617).  Right-looking blocked algorithm with
// panel width NB: per column a pivot search and rank-1 panel update, then the
// panel's row interchanges on the other columns, a unit-lower triangular solve
// for the U12 block and the trailing update A22 -= L21 U12 as a tiled complex
// GEMM (the one dense contraction, O(n^3)).  Pivot choice is max |Re|+|Im|
// (LAPACK izamax convention).
#include "common.h"
#include "cplx.h"

#define NB 64

// ---------------------------------------------------------------- panel
// pivot of column j over rows j..n-1, interchange inside the panel columns
// [c0, c1), scale the column below the diagonal
__global__ void __launch_bounds__(1024) k_pivot(int n, double *__restrict__ A, int j, int c0, int c1,
                                                int *__restrict__ ipiv) {
  __shared__ double sv[1024];
  __shared__ int si[1024];
  const int tid = threadIdx.x;
  double best = -1.0;
  int bi = j;
  for (int i = j + tid; i < n; i += blockDim.x) {
    const double *a = A + 2 * ((size_t)j * n + i);
    double v = fabs(a[0]) + fabs(a[1]);
    if (v > best) { best = v; bi = i; }
  }
  sv[tid] = best;
  si[tid] = bi;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (tid < s) {
      double o = sv[tid + s];
      int oi = si[tid + s];
      if (o > sv[tid] || (o == sv[tid] && oi < si[tid])) { sv[tid] = o; si[tid] = oi; }
    }
    __syncthreads();
  }
  const int p = si[0];
  if (tid == 0) ipiv[j] = p;
  if (p != j) {
    for (int c = c0 + tid; c < c1; c += blockDim.x) {
      double *x = A + 2 * ((size_t)c * n + j), *y = A + 2 * ((size_t)c * n + p);
      double t0 = x[0], t1 = x[1];
      x[0] = y[0]; x[1] = y[1];
      y[0] = t0; y[1] = t1;
    }
  }
  __syncthreads();
  const double *d = A + 2 * ((size_t)j * n + j);
  cd piv = cd{d[0], d[1]};
  if (piv.x == 0.0 && piv.y == 0.0) return;  // singular: leave the column, the host checks
  cd ip = crecip(piv);
  for (int i = j + 1 + tid; i < n; i += blockDim.x) {
    double *a = A + 2 * ((size_t)j * n + i);
    cd v = cd{a[0], a[1]} * ip;
    a[0] = v.x; a[1] = v.y;
  }
}

// rank-1 update of the rest of the panel: A[i, c] -= A[i, j] A[j, c], i > j, j < c < c1
__global__ void __launch_bounds__(256) k_panel_rank1(int n, double *__restrict__ A, int j, int c1) {
  const int c = j + 1 + blockIdx.y;
  if (c >= c1) return;
  const int i = j + 1 + blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double *u = A + 2 * ((size_t)c * n + j);
  const double *l = A + 2 * ((size_t)j * n + i);
  double *a = A + 2 * ((size_t)c * n + i);
  cd v = cd{a[0], a[1]} - cd{l[0], l[1]} * cd{u[0], u[1]};
  a[0] = v.x; a[1] = v.y;
}

// the panel's interchanges applied to the columns outside [k0, k0+kb)
__global__ void __launch_bounds__(256) k_swap_rows(int n, int ncols, double *__restrict__ A, int ld, int k0, int kb,
                                                   const int *__restrict__ ipiv, int skip0, int skip1) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncols || (c >= skip0 && c < skip1)) return;
  for (int j = k0; j < k0 + kb; j++) {
    int p = ipiv[j];
    if (p == j) continue;
    double *x = A + 2 * ((size_t)c * ld + j), *y = A + 2 * ((size_t)c * ld + p);
    double t0 = x[0], t1 = x[1];
    x[0] = y[0]; x[1] = y[1];
    y[0] = t0; y[1] = t1;
  }
}

// Triangular solves with the kb x kb diagonal block T = A[k0.., k0..] of a
// factored panel: X[k0:k0+kb, c] <- T^{-1} X[k0:k0+kb, c] for the columns c of
// X, T unit lower (forward) or upper (backward).  One lane per column of X, 64
// columns per workgroup; T and the 64 columns are staged in LDS (the lanes of a
// wave read consecutive column entries, T by broadcast).
template <bool UPPER>
__global__ void __launch_bounds__(64) k_trsm(int n, const double *__restrict__ T, int k0, int kb,
                                             double *__restrict__ X, int ldx, int c0, int ncols) {
  __shared__ cd sT[NB * NB];  // column-major: (r, q) at q * NB + r
  __shared__ cd sX[NB * 64];  // (r, lane) at r * 64 + lane
  const int lane = threadIdx.x;
  for (int e = lane; e < kb * kb; e += 64) {
    const int r = e % kb, q = e / kb;
    const double *a = T + 2 * ((size_t)(k0 + q) * n + k0 + r);
    sT[q * NB + r] = cd{a[0], a[1]};
  }
  const int c = c0 + blockIdx.x * 64 + lane;
  const bool on = c < c0 + ncols;
  double *x = X + 2 * ((size_t)c * ldx + k0);
  if (on)
    for (int r = 0; r < kb; r++) sX[r * 64 + lane] = cd{x[2 * r], x[2 * r + 1]};
  __syncthreads();
  if (!on) return;
  if (!UPPER) {
    for (int q = 0; q < kb; q++) {
      const cd vq = sX[q * 64 + lane];
      for (int r = q + 1; r < kb; r++) sX[r * 64 + lane] -= sT[q * NB + r] * vq;
    }
  } else {
    for (int q = kb - 1; q >= 0; q--) {
      const cd vq = sX[q * 64 + lane] * crecip(sT[q * NB + q]);
      sX[q * 64 + lane] = vq;
      for (int r = 0; r < q; r++) sX[r * 64 + lane] -= sT[q * NB + r] * vq;
    }
  }
  for (int r = 0; r < kb; r++) {
    const cd v = sX[r * 64 + lane];
    x[2 * r] = v.x;
    x[2 * r + 1] = v.y;
  }
}

static void trsm(bool upper, int n, const double *T, int k0, int kb, double *X, int ldx, int c0, int ncols,
                 hipStream_t st) {
  if (ncols <= 0) return;
  if (upper)
    hipLaunchKernelGGL(k_trsm<true>, dim3(bqd2_blocks(ncols, 64)), dim3(64), 0, st, n, T, k0, kb, X, ldx, c0, ncols);
  else
    hipLaunchKernelGGL(k_trsm<false>, dim3(bqd2_blocks(ncols, 64)), dim3(64), 0, st, n, T, k0, kb, X, ldx, c0, ncols);
}

// ---------------------------------------------------------------- trailing update
// C[m x nc] -= Am[m x kk] * B[kk x nc], all column-major complex, tiles of 64 x 64
// with a 4 x 4 complex block per lane, K staged through LDS in slices of 16.
#define TM 64
#define TN 64
#define TK 16
__global__ void __launch_bounds__(256) k_zgemm_sub(int m, int nc, int kk, const double *__restrict__ Am, int lda,
                                                   const double *__restrict__ B, int ldb, double *__restrict__ C,
                                                   int ldc) {
  __shared__ double sA[TK][TM][2];
  __shared__ double sB[TK][TN][2];
  const int tid = threadIdx.x;
  const int tr = tid % 16, tc = tid / 16;  // 16 x 16 lanes, 4 x 4 outputs each
  const int row0 = blockIdx.x * TM, col0 = blockIdx.y * TN;
  cd acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 4; b++) acc[a][b] = cd{0.0, 0.0};
  for (int k0 = 0; k0 < kk; k0 += TK) {
    // A slice: TM rows x TK cols; B slice: TK rows x TN cols
    for (int e = tid; e < TM * TK; e += 256) {
      int r = e % TM, c = e / TM;
      int gr = row0 + r, gc = k0 + c;
      double x = 0.0, y = 0.0;
      if (gr < m && gc < kk) {
        const double *p = Am + 2 * ((size_t)gc * lda + gr);
        x = p[0]; y = p[1];
      }
      sA[c][r][0] = x; sA[c][r][1] = y;
    }
    for (int e = tid; e < TK * TN; e += 256) {
      int r = e % TK, c = e / TK;
      int gr = k0 + r, gc = col0 + c;
      double x = 0.0, y = 0.0;
      if (gr < kk && gc < nc) {
        const double *p = B + 2 * ((size_t)gc * ldb + gr);
        x = p[0]; y = p[1];
      }
      sB[r][c][0] = x; sB[r][c][1] = y;
    }
    __syncthreads();
#pragma unroll 4
    for (int q = 0; q < TK; q++) {
      cd av[4], bv[4];
#pragma unroll
      for (int a = 0; a < 4; a++) av[a] = cd{sA[q][tr + 16 * a][0], sA[q][tr + 16 * a][1]};
#pragma unroll
      for (int b = 0; b < 4; b++) bv[b] = cd{sB[q][tc + 16 * b][0], sB[q][tc + 16 * b][1]};
#pragma unroll
      for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) acc[a][b] += av[a] * bv[b];
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 4; b++) {
      int gr = row0 + tr + 16 * a, gc = col0 + tc + 16 * b;
      if (gr < m && gc < nc) {
        double *p = C + 2 * ((size_t)gc * ldc + gr);
        p[0] -= acc[a][b].x;
        p[1] -= acc[a][b].y;
      }
    }
}

static int gemm_sub(int m, int nc, int kk, const double *Am, int lda, const double *B, int ldb, double *C, int ldc,
                    hipStream_t st) {
  if (m <= 0 || nc <= 0 || kk <= 0) return BQD2_OK;
  dim3 grid(bqd2_blocks(m, TM), bqd2_blocks(nc, TN));
  hipLaunchKernelGGL(k_zgemm_sub, grid, dim3(256), 0, st, m, nc, kk, Am, lda, B, ldb, C, ldc);
  BQD2_HIP(hipGetLastError());
  return BQD2_OK;
}

// zero pivot detection after the factorisation
__global__ void k_check_diag(int n, const double *__restrict__ A, int *__restrict__ flag) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double *d = A + 2 * ((size_t)i * n + i);
  if (d[0] == 0.0 && d[1] == 0.0) atomicMin(flag, i);
}

extern "C" int bqd2_zgetrf(int n, double *A, int *ipiv, void *stream) {
  BQD2_CHECK(n > 0 && A && ipiv, BQD2_E_ARG, "bqd2_zgetrf: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  for (int k0 = 0; k0 < n; k0 += NB) {
    const int kb = (n - k0 < NB) ? n - k0 : NB;
    const int c1 = k0 + kb;
    for (int j = k0; j < c1; j++) {
      hipLaunchKernelGGL(k_pivot, dim3(1), dim3(1024), 0, st, n, A, j, k0, c1, ipiv);
      if (j + 1 < c1) {
        dim3 g(bqd2_blocks(n - j - 1, 256), c1 - j - 1);
        if (n - j - 1 > 0) hipLaunchKernelGGL(k_panel_rank1, g, dim3(256), 0, st, n, A, j, c1);
      }
    }
    BQD2_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_swap_rows, dim3(bqd2_blocks(n, 256)), dim3(256), 0, st, n, n, A, n, k0, kb, ipiv, k0, c1);
    const int rest = n - c1;
    if (rest > 0) {
      trsm(false, n, A, k0, kb, A, n, c1, rest, st);
      int rc = gemm_sub(rest, rest, kb, A + 2 * ((size_t)k0 * n + c1), n, A + 2 * ((size_t)c1 * n + k0), n,
                        A + 2 * ((size_t)c1 * n + c1), n, st);
      if (rc) return rc;
    }
    BQD2_HIP(hipGetLastError());
  }
  int *d_flag = nullptr, h_flag = n;
  BQD2_HIP(hipMalloc(&d_flag, sizeof(int)));
  BQD2_HIP(hipMemcpyAsync(d_flag, &h_flag, sizeof(int), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_check_diag, dim3(bqd2_blocks(n, 256)), dim3(256), 0, st, n, A, d_flag);
  BQD2_HIP(hipMemcpyAsync(&h_flag, d_flag, sizeof(int), hipMemcpyDeviceToHost, st));
  BQD2_HIP(hipStreamSynchronize(st));
  (void)hipFree(d_flag);
  BQD2_CHECK(h_flag == n, BQD2_E_SINGULAR, "bqd2_zgetrf: zero pivot at column %d", h_flag);
  return BQD2_OK;
}

extern "C" int bqd2_zgetrs(int n, int nrhs, const double *A, const int *ipiv, double *B, void *stream) {
  BQD2_CHECK(n > 0 && nrhs > 0 && A && ipiv && B, BQD2_E_ARG, "bqd2_zgetrs: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  // row interchanges on B (all pivots, in order)
  for (int k0 = 0; k0 < n; k0 += NB) {
    const int kb = (n - k0 < NB) ? n - k0 : NB;
    hipLaunchKernelGGL(k_swap_rows, dim3(bqd2_blocks(nrhs, 256)), dim3(256), 0, st, n, nrhs, B, n, k0, kb, ipiv, -1, -1);
  }
  // forward: L y = b
  for (int k0 = 0; k0 < n; k0 += NB) {
    const int kb = (n - k0 < NB) ? n - k0 : NB;
    trsm(false, n, A, k0, kb, B, n, 0, nrhs, st);
    int rest = n - k0 - kb;
    int rc = gemm_sub(rest, nrhs, kb, A + 2 * ((size_t)k0 * n + k0 + kb), n, B + 2 * (size_t)k0, n,
                      B + 2 * (size_t)(k0 + kb), n, st);
    if (rc) return rc;
  }
  // backward: U x = y
  int last = ((n - 1) / NB) * NB;
  for (int k0 = last; k0 >= 0; k0 -= NB) {
    const int kb = (n - k0 < NB) ? n - k0 : NB;
    trsm(true, n, A, k0, kb, B, n, 0, nrhs, st);
    int rc = gemm_sub(k0, nrhs, kb, A + 2 * ((size_t)k0 * n), n, B + 2 * (size_t)k0, n, B, n, st);
    if (rc) return rc;
  }
  BQD2_HIP(hipGetLastError());
  return BQD2_OK;
}
