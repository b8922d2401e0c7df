// lu.hip -- dense complex double LU with partial pivoting and the matching
// triangular solves, on column-major device matrices.
//
// Replaces the reference's direct solves: MKL PARDISO on the 2NN system
// (d3qd2_mkl_solver.c:18-148) and LAPACKE_zgesv for the per-domain 3-RHS
// systems (bem3_aw_qd2_solve_bieq.c:617).  Right-looking blocked algorithm with
// panel width NB; pivot choice max |Re|+|Im| (LAPACK's izamax convention),
// ties to the lowest row, so the interchanges are LAPACK's zgetrf's.
//   panel     one launch per column (k_zstep): every workgroup owns a slice of
//             the panel's rows; the launch takes the pivot from the candidates
//             the previous launch left (one per workgroup), interchanges, scales
//             and applies the rank-1 update to its rows, then leaves its
//             candidate for the next column.  Launch boundaries are the only
//             grid-wide synchronisation.
//   swaps     the panel's interchanges on the other columns (k_swap_rows)
//   U12       = L11^-1 A12 (k_trsm, one lane per column)
//   A22      -= L21 U12, the O(n^3) part: k_zgemm on v_mfma_f64_16x16x4_f64
// The candidates and the zero-pivot flag live in a device workspace allocated
// once per device and kept (calls on one process are serialised on it).
#include <algorithm>
#include <mutex>
#include "common.h"
#include "cplx.h"

#define NB 64       // panel width = K of the trailing update
#define GMAXW 1024  // panel workgroups at most
#define PROWS 32    // panel rows per workgroup (more when GMAXW workgroups do not cover the rows)
#define PEL (NB / 8)

typedef double doublex4 __attribute__((ext_vector_type(4)));

// device workspace (size independent of n):
//   slot  2 x GMAXW pivot candidates (value, row) of the panel workgroups
//   cand  2 x GMAXW candidate rows of NB entries
//   rowj  2 x NB entries: row j of the panel before the interchange
//   flag  first zero pivot (n when none)
struct ZWs {
  cd *slot;
  cd *cand;
  cd *rowj;
  int *flag;
};
#define ZWS_BYTES ((size_t)2 * GMAXW * 16 + (size_t)2 * GMAXW * NB * 16 + (size_t)2 * NB * 16 + 256)

static std::mutex g_zws_mu;
static void *g_zws[64];

// the device's workspace, allocated at the first factorisation on it; call with g_zws_mu held
static int zws_get(ZWs *w) {
  int dev = 0;
  BQD2_HIP(hipGetDevice(&dev));
  BQD2_CHECK(dev >= 0 && dev < 64, BQD2_E_ARG, "bqd2_zgetrf: device %d", dev);
  if (!g_zws[dev]) BQD2_HIP(hipMalloc(&g_zws[dev], ZWS_BYTES));
  char *q = (char *)g_zws[dev];
  w->slot = (cd *)q;
  q += (size_t)2 * GMAXW * 16;
  w->cand = (cd *)q;
  q += (size_t)2 * GMAXW * NB * 16;
  w->rowj = (cd *)q;
  q += (size_t)2 * NB * 16;
  w->flag = (int *)q;
  return BQD2_OK;
}

// ---------------------------------------------------------------- panel
// Launch for column j of the panel [k0, k0 + kb) (j = k0 - 1: the first
// candidates only).  Workgroup w owns rows [k0 + w chunk, k0 + (w+1) chunk).
// Lane (i, g) = (tid % 32, tid / 32) holds, for each block of 32 of those rows,
// row i of the block in the columns cf + g + 8 u, u < PEL (cf = j): the whole
// remaining panel row in registers.  The first block is loaded before the pivot
// is known, so that its latency overlaps the reduction of the candidates.
// Reads the candidates of parity (j - k0) & 1, writes those of the other.  A
// column whose candidates are all NaN keeps its row (p = j) and is flagged.
__device__ __forceinline__ bool zbetter(double v, int r, double bv, int br) {
  return v > bv || (v == bv && r < br);
}

__device__ __forceinline__ void zblock_best(double &bv, int &br, double *sv, int *sr) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int off = 32; off >= 1; off >>= 1) {
    const double ov = __shfl_xor(bv, off, 64);
    const int orr = __shfl_xor(br, off, 64);
    if (zbetter(ov, orr, bv, br)) { bv = ov; br = orr; }
  }
  __syncthreads();
  if (lane == 0) { sv[wv] = bv; sr[wv] = br; }
  __syncthreads();
  bv = sv[0]; br = sr[0];
  for (int q = 1; q < 4; q++)
    if (zbetter(sv[q], sr[q], bv, br)) { bv = sv[q]; br = sr[q]; }
}

__global__ void __launch_bounds__(256) k_zstep(int n, cd *__restrict__ A, int k0, int kb, int j, int chunk, int nw,
                                               int *__restrict__ ipiv, ZWs W) {
  __shared__ cd su[NB];  // pivot row over the panel's columns
  __shared__ cd sj[NB];  // row j before the interchange
  __shared__ cd sl[32];  // multipliers of the current block of rows
  __shared__ double sv[4];
  __shared__ int sr[4];
  const int tid = threadIdx.x, w = blockIdx.x;
  const int li = tid & 31, lg = tid >> 5;
  const int r0 = k0 + w * chunk, r1 = min(n, r0 + chunk);
  const size_t ld = (size_t)n;
  const int kend = k0 + kb;
  const bool step = j >= k0;
  // rows [rs, r1) in blocks of 32, columns [cf, cend)
  const int rs = step ? max(r0, j + 1) : r0;
  const int cf = step ? j : k0;
  const int cend = step ? kend : k0 + 1;
  const int jn = j + 1;
  cd v[PEL];
  auto load = [&](int rb) {
    const int r = rb + li;
#pragma unroll
    for (int u = 0; u < PEL; u++) {
      const int c = cf + lg + 8 * u;
      if (r < r1 && c < cend) v[u] = A[(size_t)c * ld + r];
    }
  };
  load(rs);
  int p = j;
  cd ip = cd{0.0, 0.0};
  if (step) {
    const int par = (j - k0) & 1;
    // pivot: the best of the workgroups' candidates (ties: the lowest row)
    double bv = -1.0;
    int br = 0x7fffffff;
    for (int q = tid; q < nw; q += 256) {
      const cd sc = W.slot[par * GMAXW + q];
      const int r = (int)sc.y;
      if (zbetter(sc.x, r, bv, br)) { bv = sc.x; br = r; }
    }
    zblock_best(bv, br, sv, sr);
    const bool none = (br == 0x7fffffff);
    p = none ? j : br;
    const cd *rj = W.rowj + par * NB;
    const cd *prow = none ? rj : W.cand + ((size_t)par * GMAXW + (p - k0) / chunk) * NB;
    for (int c = tid; c < kb; c += 256) { su[c] = prow[c]; sj[c] = rj[c]; }
    if (w == 0 && tid == 0) ipiv[j] = p;
    // interchange rows j and p over the panel's columns
    if (p != j) {
      if (p >= r0 && p < r1)
        for (int c = tid; c < kb; c += 256) A[(size_t)(k0 + c) * ld + p] = rj[c];
      if (j >= r0 && j < r1)
        for (int c = tid; c < kb; c += 256) A[(size_t)(k0 + c) * ld + j] = prow[c];
    }
    __syncthreads();
    const cd piv = su[j - k0];
    const bool zero = none || (piv.x == 0.0 && piv.y == 0.0);
    if (zero && w == 0 && tid == 0) atomicMin(W.flag, j);
    ip = zero ? cd{0.0, 0.0} : crecip(piv);
  }
  double bv = -1.0;
  int br = 0x7fffffff;
  for (int rb = rs; rb < r1; rb += 32) {
    if (rb != rs) load(rb);
    const int r = rb + li;
    const bool on = r < r1;
    if (step) {
      // row p now holds row j
      if (on && r == p) {
#pragma unroll
        for (int u = 0; u < PEL; u++) {
          const int c = cf + lg + 8 * u;
          if (c < cend) v[u] = sj[c - k0];
        }
      }
      // multiplier of row r (column j: lane group 0, u = 0)
      if (lg == 0 && on) {
        v[0] = v[0] * ip;
        sl[li] = v[0];
        A[(size_t)j * ld + r] = v[0];
      }
      __syncthreads();
      const cd l = sl[li];
#pragma unroll
      for (int u = 0; u < PEL; u++) {
        const int c = cf + lg + 8 * u;
        if (!on || c == j || c >= cend) continue;
        v[u] -= l * su[c - k0];
        A[(size_t)c * ld + r] = v[u];
      }
      __syncthreads();  // sl is rewritten by the next block
    }
    // candidate for column jn: lane group 1 (step) or 0 (first launch), u = 0
    if (on && lg == (step ? 1 : 0) && jn < kend) {
      const double a = fabs(v[0].x) + fabs(v[0].y);
      if (zbetter(a, r, bv, br)) { bv = a; br = r; }
    }
  }
  if (jn >= kend) return;
  const int parn = (jn - k0) & 1;
  zblock_best(bv, br, sv, sr);  // its barriers also order this workgroup's stores above
  if (tid == 0) W.slot[parn * GMAXW + w] = cd{bv, (double)br};
  cd *cr = W.cand + ((size_t)parn * GMAXW + w) * NB;
  if (br != 0x7fffffff)
    for (int c = tid; c < kb; c += 256) cr[c] = A[(size_t)(k0 + c) * ld + br];
  if (jn >= r0 && jn < r1)
    for (int c = tid; c < kb; c += 256) W.rowj[parn * NB + c] = A[(size_t)(k0 + c) * ld + jn];
}

__global__ void k_zinit_flag(int *flag, int n) {
  if (threadIdx.x == 0) *flag = n;
}

// the panel's interchanges applied to the columns outside [skip0, skip1)
__global__ void __launch_bounds__(256) k_swap_rows(int n, int ncols, cd *__restrict__ A, int ld, int k0, int kb,
                                                   const int *__restrict__ ipiv, int skip0, int skip1) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncols || (c >= skip0 && c < skip1)) return;
  for (int j = k0; j < k0 + kb; j++) {
    int p = ipiv[j];
    if (p == j) continue;
    cd *x = A + (size_t)c * ld + j, *y = A + (size_t)c * ld + p;
    const cd t = *x;
    *x = *y;
    *y = t;
  }
}

// Triangular solves with the kb x kb diagonal block T = A[k0.., k0..] of a
// factored panel: X[k0:k0+kb, c] <- T^{-1} X[k0:k0+kb, c] for the columns c of
// X, T unit lower (forward) or upper (backward).  One lane per column of X, 64
// columns per workgroup; T and the 64 columns are staged in LDS (the lanes of a
// wave read consecutive column entries, T by broadcast).
template <bool UPPER>
__global__ void __launch_bounds__(64) k_trsm(int n, const cd *__restrict__ T, int k0, int kb, cd *__restrict__ X,
                                             int ldx, int c0, int ncols) {
  __shared__ cd sT[NB * NB];  // column-major: (r, q) at q * NB + r
  __shared__ cd sX[NB * 64];  // (r, lane) at r * 64 + lane
  const int lane = threadIdx.x;
  for (int e = lane; e < kb * kb; e += 64) {
    const int r = e % kb, q = e / kb;
    sT[q * NB + r] = T[(size_t)(k0 + q) * n + k0 + r];
  }
  const int c = c0 + blockIdx.x * 64 + lane;
  const bool on = c < c0 + ncols;
  cd *x = X + (size_t)c * ldx + k0;
  if (on)
    for (int r = 0; r < kb; r++) sX[r * 64 + lane] = x[r];
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
  for (int r = 0; r < kb; r++) x[r] = sX[r * 64 + lane];
}

static void trsm(bool upper, int n, const cd *T, int k0, int kb, cd *X, int ldx, int c0, int ncols, hipStream_t st) {
  if (ncols <= 0) return;
  if (upper)
    hipLaunchKernelGGL(k_trsm<true>, dim3(bqd2_blocks(ncols, 64)), dim3(64), 0, st, n, T, k0, kb, X, ldx, c0, ncols);
  else
    hipLaunchKernelGGL(k_trsm<false>, dim3(bqd2_blocks(ncols, 64)), dim3(64), 0, st, n, T, k0, kb, X, ldx, c0, ncols);
}

// ---------------------------------------------------------------- MFMA GEMM
// C[m x nc] -= A[m x kk] B[kk x nc], column-major complex double, on
// v_mfma_f64_16x16x4_f64: D[i][j] += sum_{k<4} X[i][k] Y[k][j], lane l holding
// X[l%16][l/16] and Y[l/16][l%16], and D[i][j] in element r of lane l for
// j = l%16, i = l/16 + 4 r.  The product is formed as C^T = B^T A^T (j runs
// along the rows of C: 16 lanes store 16 consecutive rows of one column), and
// the four k slots carry two complex k as (re, im) pairs: X = (Br, Bi) for
// both parts of C, Y = (Ar, -Ai) for Re C and (Ai, Ar) for Im C.  Four waves
// (2 x 2) of WA x WB blocks of 16 x 16: a workgroup tile of 32 WA x 32 WB;
// K staged in LDS in slices of 16 complex columns.
#define ZGK 16
template <int WA, int WB>
__global__ void __launch_bounds__(256, 2) k_zgemm(int m, int nc, int kk, const cd *__restrict__ A, long long lda,
                                                  const cd *__restrict__ B, long long ldb, cd *__restrict__ C,
                                                  long long ldc) {
  constexpr int TR = 32 * WA, TC = 32 * WB;
  __shared__ cd sA[ZGK][TR + 1];  // (k, row)
  __shared__ cd sB[ZGK][TC + 1];  // (k, col)
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave & 1, wc = wave >> 1;
  const int row0 = blockIdx.x * TR, col0 = blockIdx.y * TC;
  const int l16 = lane & 15, ks = lane >> 4;  // k slot: complex k = 2 q + (ks >> 1), part ks & 1
  const int kh = ks >> 1;
  const bool im = (ks & 1) != 0;
  doublex4 accR[WB][WA], accI[WB][WA];
#pragma unroll
  for (int b = 0; b < WB; b++)
#pragma unroll
    for (int a = 0; a < WA; a++)
#pragma unroll
      for (int r = 0; r < 4; r++) { accR[b][a][r] = 0.0; accI[b][a][r] = 0.0; }
  for (int k0 = 0; k0 < kk; k0 += ZGK) {
    __syncthreads();
    for (int e = tid; e < TR * ZGK; e += 256) {
      const int r = e % TR, k = e / TR;
      const int gr = row0 + r, gk = k0 + k;
      sA[k][r] = (gr < m && gk < kk) ? A[(size_t)gk * lda + gr] : cd{0.0, 0.0};
    }
    for (int e = tid; e < TC * ZGK; e += 256) {
      const int k = e % ZGK, c = e / ZGK;
      const int gc = col0 + c, gk = k0 + k;
      sB[k][c] = (gc < nc && gk < kk) ? B[(size_t)gc * ldb + gk] : cd{0.0, 0.0};
    }
    __syncthreads();
#pragma unroll 2
    for (int q = 0; q < ZGK; q += 2) {
      double x[WB], yr[WA], yi[WA];
#pragma unroll
      for (int b = 0; b < WB; b++) {
        const cd v = sB[q + kh][wc * 16 * WB + b * 16 + l16];
        x[b] = im ? v.y : v.x;
      }
#pragma unroll
      for (int a = 0; a < WA; a++) {
        const cd v = sA[q + kh][wr * 16 * WA + a * 16 + l16];
        yr[a] = im ? -v.y : v.x;
        yi[a] = im ? v.x : v.y;
      }
#pragma unroll
      for (int b = 0; b < WB; b++)
#pragma unroll
        for (int a = 0; a < WA; a++) {
          accR[b][a] = __builtin_amdgcn_mfma_f64_16x16x4f64(x[b], yr[a], accR[b][a], 0, 0, 0);
          accI[b][a] = __builtin_amdgcn_mfma_f64_16x16x4f64(x[b], yi[a], accI[b][a], 0, 0, 0);
        }
    }
  }
#pragma unroll
  for (int b = 0; b < WB; b++)
#pragma unroll
    for (int a = 0; a < WA; a++) {
      const int gr = row0 + wr * 16 * WA + a * 16 + l16;
      if (gr >= m) continue;
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int gc = col0 + wc * 16 * WB + b * 16 + ks + 4 * r;
        if (gc >= nc) continue;
        cd *p = C + (size_t)gc * ldc + gr;
        cd o = *p;
        o.x -= accR[b][a][r];
        o.y -= accI[b][a][r];
        *p = o;
      }
    }
}

static int gemm_sub(int m, int nc, int kk, const cd *A, long long lda, const cd *B, long long ldb, cd *C,
                    long long ldc, hipStream_t st) {
  if (m <= 0 || nc <= 0 || kk <= 0) return BQD2_OK;
  if (nc <= 16) {
    // few right-hand sides: 128 x 32 tiles
    dim3 grid(bqd2_blocks(m, 128), bqd2_blocks(nc, 32));
    hipLaunchKernelGGL((k_zgemm<4, 1>), grid, dim3(256), 0, st, m, nc, kk, A, lda, B, ldb, C, ldc);
  } else {
    dim3 grid(bqd2_blocks(m, 64), bqd2_blocks(nc, 64));
    hipLaunchKernelGGL((k_zgemm<2, 2>), grid, dim3(256), 0, st, m, nc, kk, A, lda, B, ldb, C, ldc);
  }
  BQD2_HIP(hipGetLastError());
  return BQD2_OK;
}

extern "C" int bqd2_zgetrf(int n, double *d_A, int *ipiv, void *stream) {
  BQD2_CHECK(n > 0 && d_A && ipiv, BQD2_E_ARG, "bqd2_zgetrf: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  cd *A = (cd *)d_A;
  const size_t ld = (size_t)n;
  std::lock_guard<std::mutex> lk(g_zws_mu);
  ZWs W;
  int rc = zws_get(&W);
  if (rc) return rc;
  hipLaunchKernelGGL(k_zinit_flag, dim3(1), dim3(64), 0, st, W.flag, n);
  // with a progress callback: an event every 64 panels, the one before waited
  // for and reported (two parts in flight)
  const int npan = (n + NB - 1) / NB, every = 64;
  const bool prog = bqd2_progress_on() && npan >= 2 * every;
  hipEvent_t pe[2] = {nullptr, nullptr};
  int pk = 0;
  for (int k0 = 0; k0 < n; k0 += NB) {
    const int kb = std::min(NB, n - k0);
    const int rows = n - k0, rest = rows - kb;
    int nw = bqd2_blocks(rows, PROWS);
    if (nw > GMAXW) nw = GMAXW;
    const int chunk = bqd2_blocks(rows, nw);
    nw = bqd2_blocks(rows, chunk);
    for (int j = k0 - 1; j < k0 + kb; j++)
      hipLaunchKernelGGL(k_zstep, dim3(nw), dim3(256), 0, st, n, A, k0, kb, j, chunk, nw, ipiv, W);
    BQD2_HIP(hipGetLastError());
    if (n > kb)
      hipLaunchKernelGGL(k_swap_rows, dim3(bqd2_blocks(n, 256)), dim3(256), 0, st, n, n, A, n, k0, kb, ipiv, k0,
                         k0 + kb);
    if (rest > 0) {
      trsm(false, n, A, k0, kb, A, n, k0 + kb, rest, st);
      rc = gemm_sub(rest, rest, kb, A + (size_t)k0 * ld + k0 + kb, n, A + (size_t)(k0 + kb) * ld + k0, n,
                    A + (size_t)(k0 + kb) * ld + k0 + kb, n, st);
      if (rc) return rc;
    }
    BQD2_HIP(hipGetLastError());
    const int pan = k0 / NB + 1;
    if (prog && pan % every == 0) {
      hipEvent_t &e = pe[pk & 1];
      if (!e) BQD2_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      BQD2_HIP(hipEventRecord(e, st));
      pk++;
      if (pk >= 2) {
        BQD2_HIP(hipEventSynchronize(pe[pk & 1]));
        bqd2_progress("FP64 LU panels", (long long)(pan - every), npan);
      }
    }
  }
  for (hipEvent_t e : pe)
    if (e) (void)hipEventDestroy(e);
  int h_flag = n;
  BQD2_HIP(hipMemcpyAsync(&h_flag, W.flag, sizeof(int), hipMemcpyDeviceToHost, st));
  BQD2_HIP(hipStreamSynchronize(st));
  BQD2_CHECK(h_flag == n, BQD2_E_SINGULAR, "bqd2_zgetrf: zero pivot at column %d", h_flag);
  return BQD2_OK;
}

extern "C" int bqd2_zgetrs(int n, int nrhs, const double *d_A, const int *ipiv, double *d_B, void *stream) {
  BQD2_CHECK(n > 0 && nrhs > 0 && d_A && ipiv && d_B, BQD2_E_ARG, "bqd2_zgetrs: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  const cd *A = (const cd *)d_A;
  cd *B = (cd *)d_B;
  const size_t ld = (size_t)n;
  // row interchanges on B (all pivots, in order)
  for (int k0 = 0; k0 < n; k0 += NB) {
    const int kb = std::min(NB, n - k0);
    hipLaunchKernelGGL(k_swap_rows, dim3(bqd2_blocks(nrhs, 256)), dim3(256), 0, st, n, nrhs, B, n, k0, kb, ipiv, -1,
                       -1);
  }
  // forward: L y = b
  for (int k0 = 0; k0 < n; k0 += NB) {
    const int kb = std::min(NB, n - k0);
    trsm(false, n, A, k0, kb, B, n, 0, nrhs, st);
    const int rest = n - k0 - kb;
    int rc = gemm_sub(rest, nrhs, kb, A + (size_t)k0 * ld + k0 + kb, n, B + k0, n, B + k0 + kb, n, st);
    if (rc) return rc;
  }
  // backward: U x = y
  const int last = ((n - 1) / NB) * NB;
  for (int k0 = last; k0 >= 0; k0 -= NB) {
    const int kb = std::min(NB, n - k0);
    trsm(true, n, A, k0, kb, B, n, 0, nrhs, st);
    int rc = gemm_sub(k0, nrhs, kb, A + (size_t)k0 * ld, n, B + k0, n, B, n, st);
    if (rc) return rc;
  }
  BQD2_HIP(hipGetLastError());
  return BQD2_OK;
}
