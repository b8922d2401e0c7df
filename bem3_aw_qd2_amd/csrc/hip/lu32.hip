// lu32.hip -- dense LU in complex FP32 with the trailing updates on the MFMA
// units, and the solves with its factors.
//
// The reference solves its boundary-value system directly in complex double
// (MKL PARDISO, d3qd2_mkl_solver.c:18-148).  This factorisation is the
// mixed-precision form of that solve: the row-scaled system is rounded to
// complex float, factored here with partial pivoting (max |Re|+|Im|, LAPACK's
// izamax convention), and applied as the preconditioner of GMRES on the FP64
// system (gmres.hip), which restores the FP64 result.
//
// Layout: column-major, leading dimension n, complex numbers as float2 (re, im).
// Right-looking blocked algorithm with panel width NB.  Panels:
//   block-local (default)  the 128 x 128 diagonal block factored by one
//             workgroup with pivots from its own rows (k_cpanel_local), which
//             also forms L11^-1 and U11^-1; L21 = A21 U11^-1 and
//             U12 = L11^-1 A12 are MFMA GEMMs.  Checked (the C ABI): a panel
//             whose max |L21| exceeds the growth bound is restored and redone
//             over all rows.  Unchecked (the preconditioners, whose FP64
//             refinement validates the factors): no host decision per panel,
//             and the next panel runs on its own CUs beside the trailing update.
//   all rows  one launch per column (k_cstep): every workgroup owns a slice of
//             the panel's rows; the launch takes the pivot from the candidates
//             the previous launch left (one per workgroup), interchanges, scales
//             and applies the rank-1 update to its rows, then leaves its
//             candidate for the next column.  Launch boundaries are the only
//             grid-wide synchronisation.  U12 by k_ctrsm_l.
//   swaps     the panel's interchanges on the other columns (k_cswap_cols)
//   A22      -= L21 U12, the O(n^3) part: k_cgemm on v_mfma_f32_32x32x2_f32,
//             three real products per complex product, FP32 accumulation
// After the last panel: inverses of the diagonal blocks' unit-lower (k_clinv)
// and upper (k_ctrtri_u) factors, kept for the solves.
// Solves (bqd2_cgetrs): permutation, then per diagonal block x_k = T_k b_k and
// the update of the remaining rows, forward with L and backward with U; GEMV
// kernels for up to 4 right-hand sides, the MFMA GEMM beyond.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>
#include "common.h"
#include "tune.h"

#define NB 128      // panel width = K of the trailing update
#define GMAXW 1024  // panel workgroups at most
#define PROWS 32    // panel rows per workgroup (more when GMAXW workgroups do not cover the rows)

typedef float floatx16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ float2 cmulf(float2 a, float2 b) {
  return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ float2 crecipf(float2 a) {
  const float d = 1.0f / fmaf(a.x, a.x, a.y * a.y);
  return make_float2(a.x * d, -a.y * d);
}

// device workspace of a factorisation of order n (bqd2_cgetrf_wsize bytes):
//   T     ceil(n / NB) diagonal blocks x 2 (L^-1, U^-1), NB x NB column-major each
//   perm  n ints: row i of P A is row perm[i] of A
//   slot  2 x GMAXW pivot candidates (value, row) of the panel workgroups
//   cand  2 x GMAXW candidate rows of NB entries
//   rowj  2 x NB entries: row j of the panel before the interchange
//   flag  first zero pivot (n when none)
//   U, stat, bk  the block-local panel's inverses, status and backup
struct LuWs {
  float2 *T;
  int *perm;
  float2 *slot;  // (value, row as float bits)
  float2 *cand;
  float2 *rowj;
  int *flag;
  float2 *U;      // 2 x NB x NB: L11^-1 and U11^-1 of the current panel (block-local panel)
  float2 *bk;     // n x NB: the panel before its block-local factorisation
  int *stat;      // [0] zero pivot in the block, [1] max |L21| as float bits
};

static size_t align256(size_t b) { return (b + 255) / 256 * 256; }

static LuWs lu_ws(int n, void *base) {
  char *q = (char *)base;
  LuWs w;
  const size_t nbk = (size_t)(n + NB - 1) / NB;
  w.T = (float2 *)q; q += align256(nbk * 2 * NB * NB * sizeof(float2));
  w.perm = (int *)q; q += align256((size_t)n * sizeof(int));
  w.slot = (float2 *)q; q += align256(2 * GMAXW * sizeof(float2));
  w.cand = (float2 *)q; q += align256((size_t)2 * GMAXW * NB * sizeof(float2));
  w.rowj = (float2 *)q; q += align256(2 * NB * sizeof(float2));
  w.flag = (int *)q; q += 256;
  w.U = (float2 *)q; q += align256((size_t)2 * NB * NB * sizeof(float2));
  w.stat = (int *)q; q += 256;
  w.bk = (float2 *)q;
  return w;
}

extern "C" size_t bqd2_cgetrf_wsize(int n) {
  if (n <= 0) return 0;
  const size_t nbk = (size_t)(n + NB - 1) / NB;
  return align256(nbk * 2 * NB * NB * sizeof(float2)) + align256((size_t)n * sizeof(int)) +
         align256(2 * GMAXW * sizeof(float2)) + align256((size_t)2 * GMAXW * NB * sizeof(float2)) +
         align256(2 * NB * sizeof(float2)) + 256 + align256((size_t)2 * NB * NB * sizeof(float2)) + 256 +
         align256((size_t)n * NB * sizeof(float2));
}

// ---------------------------------------------------------------- panel
// Launch for column j of the panel [k0, k0 + kb) (j = k0 - 1: the first
// candidates only).  Workgroup w owns rows [k0 + w chunk, k0 + (w+1) chunk).
// Lane (i, g) = (tid % 32, tid / 32) holds, for each block of 32 of those rows,
// row i of the block in the columns cf + g + 8 u, u < 16 (cf = j): the whole
// remaining panel row in registers.  The first block is loaded before the pivot
// is known, so that its latency overlaps the reduction of the candidates.
// Reads the candidates of parity (j - k0) & 1, writes those of the other.  A
// column whose candidates are all NaN keeps its row (p = j) and is flagged.
#define PEL (NB / 8)
__device__ __forceinline__ bool better(float v, int r, float bv, int br) { return v > bv || (v == bv && r < br); }

__device__ __forceinline__ void block_best(float &bv, int &br, float *sv, int *sr) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int off = 32; off >= 1; off >>= 1) {
    const float ov = __shfl_xor(bv, off, 64);
    const int orr = __shfl_xor(br, off, 64);
    if (better(ov, orr, bv, br)) { bv = ov; br = orr; }
  }
  __syncthreads();
  if (lane == 0) { sv[wv] = bv; sr[wv] = br; }
  __syncthreads();
  bv = sv[0]; br = sr[0];
  for (int q = 1; q < 4; q++)
    if (better(sv[q], sr[q], bv, br)) { bv = sv[q]; br = sr[q]; }
}

__global__ void __launch_bounds__(256) k_cstep(int n, float2 *__restrict__ A, int k0, int kb, int j, int chunk,
                                               int nw, int *__restrict__ ipiv, LuWs W) {
  __shared__ float2 su[NB];     // pivot row over the panel's columns
  __shared__ float2 sj[NB];     // row j before the interchange
  __shared__ float2 sl[32];     // multipliers of the current block of rows
  __shared__ float sv[4];
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
  float2 v[PEL];
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
  float2 ip = make_float2(0.0f, 0.0f);
  if (step) {
    const int par = (j - k0) & 1;
    // pivot: the best of the workgroups' candidates (ties: the lowest row)
    float bv = -1.0f;
    int br = 0x7fffffff;
    for (int q = tid; q < nw; q += 256) {
      const float2 sc = W.slot[par * GMAXW + q];
      const int r = __float_as_int(sc.y);
      if (better(sc.x, r, bv, br)) { bv = sc.x; br = r; }
    }
    block_best(bv, br, sv, sr);
    const bool none = (br == 0x7fffffff);
    p = none ? j : br;
    const float2 *rj = W.rowj + par * NB;
    const float2 *prow = none ? rj : W.cand + ((size_t)par * GMAXW + (p - k0) / chunk) * NB;
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
    const float2 piv = su[j - k0];
    const bool zero = none || (piv.x == 0.0f && piv.y == 0.0f);
    if (zero && w == 0 && tid == 0) atomicMin(W.flag, j);
    ip = zero ? make_float2(0.0f, 0.0f) : crecipf(piv);
  }
  float bv = -1.0f;
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
        v[0] = cmulf(v[0], ip);
        sl[li] = v[0];
        A[(size_t)j * ld + r] = v[0];
      }
      __syncthreads();
      const float2 l = sl[li];
#pragma unroll
      for (int u = 0; u < PEL; u++) {
        const int c = cf + lg + 8 * u;
        if (!on || c == j || c >= cend) continue;
        const float2 uc = su[c - k0];
        v[u].x -= fmaf(l.x, uc.x, -l.y * uc.y);
        v[u].y -= fmaf(l.x, uc.y, l.y * uc.x);
        A[(size_t)c * ld + r] = v[u];
      }
      __syncthreads();  // sl is rewritten by the next block
    }
    // candidate for column jn: lane group 1 (step) or 0 (first launch), u = 0
    if (on && lg == (step ? 1 : 0) && jn < kend) {
      const float a = fabsf(v[0].x) + fabsf(v[0].y);
      if (better(a, r, bv, br)) { bv = a; br = r; }
    }
  }
  if (jn >= kend) return;
  const int parn = (jn - k0) & 1;
  block_best(bv, br, sv, sr);  // its barriers also order this workgroup's stores above
  if (tid == 0) W.slot[parn * GMAXW + w] = make_float2(bv, __int_as_float(br));
  float2 *cd = W.cand + ((size_t)parn * GMAXW + w) * NB;
  if (br != 0x7fffffff)
    for (int c = tid; c < kb; c += 256) cd[c] = A[(size_t)(k0 + c) * ld + br];
  if (jn >= r0 && jn < r1)
    for (int c = tid; c < kb; c += 256) W.rowj[parn * NB + c] = A[(size_t)(k0 + c) * ld + jn];
}

// ---------------------------------------------------------------- block-local panel
// The panel with its pivots restricted to the rows of its diagonal block: the
// kb x kb block is factored by one workgroup (partial pivoting within the
// block, max |Re|+|Im|, ties to the lowest row) with the block in registers,
// and the inverses L11^-1 (unit lower) and U11^-1 are formed for
// L21 = A21 U11^-1 and U12 = L11^-1 A12 (MFMA GEMMs); the largest |L21| bounds
// the growth.  Within the growth bound (16 unless set) the panel is kept (one launch and two GEMMs instead
// of one launch per column and a triangular solve); otherwise, or at a zero
// pivot in the block, the panel is restored and factored with pivoting over
// all its rows (k_cstep).  Either way P A = L U holds: the factors
// precondition GMRES on the FP64 system, which restores the FP64 result
// whatever the pivots.

// Thread t holds row r = t % 128 of the block in the columns c = g + 8 u,
// u < 16, g = t / 128, as a 32-float vector (re, im of column g + 8 u at
// 2u, 2u + 1): the step's column is read with a uniform index, which stays
// in registers (indexed moves), and the loop is not unrolled, so that the
// code of one column fits the instruction cache.  Column j's pivot search
// runs in the two waves of g = j % 8: the largest |Re|+|Im| (as float bits,
// which order as integers for non-negative floats) by a butterfly, then the
// lowest row holding it by a ballot; rows j and p are exchanged through LDS,
// the multipliers of column j and the pivot row meet in LDS, and every thread
// updates its own entries: two barriers per column.  The inverses then take
// thread t as column c = t % 128 of X, rows 8 u + g, with the factors in LDS:
// X = U^-1 by back substitution (row q of X final at step q, the rows above
// updated), X = L^-1 by forward substitution, one barrier per step.
// stat[0] = 1 on a zero (or NaN) pivot in the block; stat[1] (the growth
// maximum of k_cmaxabs) is cleared here.  With sticky set, stat[0] only
// accumulates the zero pivots of all panels and stat[1] is left alone.  Linv, Uinv: NB x NB column-major.
typedef float floatx32 __attribute__((ext_vector_type(32)));

__device__ __forceinline__ float2 vget(const floatx32 &v, int u) { return make_float2(v[2 * u], v[2 * u + 1]); }
__device__ __forceinline__ void vset(floatx32 &v, int u, float2 x) {
  v[2 * u] = x.x;
  v[2 * u + 1] = x.y;
}

__global__ void __launch_bounds__(1024) k_cpanel_local(int n, float2 *__restrict__ A, int k0, int kb,
                                                       int *__restrict__ ipiv, float2 *__restrict__ Linv,
                                                       float2 *__restrict__ Uinv, int *__restrict__ stat,
                                                       int sticky) {
  __shared__ float2 sD[NB * NB];  // after the factorisation: L\U column-major, (r, c) at c * NB + r
  __shared__ float2 sP[NB], sJ[NB], sl[NB], sx[2][NB];
  __shared__ unsigned wk[2];
  __shared__ int wr[2];
  __shared__ float2 wz[2], sjj;
  const int tid = threadIdx.x, r = tid & (NB - 1), g = tid >> 7, lane = tid & 63;
  const size_t ld = (size_t)n;
  floatx32 a;
#pragma unroll
  for (int u = 0; u < 16; u++) {
    const int c = g + 8 * u;
    vset(a, u, (r < kb && c < kb) ? A[(size_t)(k0 + c) * ld + k0 + r] : make_float2(0.0f, 0.0f));
  }
  bool bad = false;
  for (int j = 0; j < kb; j++) {
    const int gj = j & 7, uj = j >> 3;
    if (g == gj) {
      const float2 x = vget(a, uj);
      const float v = fabsf(x.x) + fabsf(x.y);
      const bool cand = r >= j && r < kb && v == v;
      const unsigned key = cand ? __float_as_uint(v) : 0u;
      unsigned mx = key;
      for (int off = 32; off >= 1; off >>= 1) mx = max(mx, (unsigned)__shfl_xor((int)mx, off, 64));
      const unsigned long long hit = __ballot(cand && key == mx);
      if (hit && lane == __ffsll((long long)hit) - 1) {
        wk[(tid >> 6) & 1] = mx;
        wr[(tid >> 6) & 1] = r;
        wz[(tid >> 6) & 1] = x;
      } else if (!hit && lane == 0) {
        wk[(tid >> 6) & 1] = 0u;
        wr[(tid >> 6) & 1] = -1;
      }
      if (r == j) sjj = x;
    }
    __syncthreads();
    // wave 0 holds the lower rows: it wins ties
    const int w = (wr[1] >= 0 && (wr[0] < 0 || wk[1] > wk[0])) ? 1 : 0;
    int p = wr[w];
    float2 piv = wz[w];
    const bool zero = p < 0 || wk[w] == 0u;  // no finite candidate, or a zero column
    if (zero) {
      p = j;
      piv = make_float2(0.0f, 0.0f);
      bad = true;
    }
    const float2 ip = zero ? make_float2(0.0f, 0.0f) : crecipf(piv);
    if (tid == 0) ipiv[k0 + j] = k0 + p;
    if (r == p) {
#pragma unroll
      for (int u = 0; u < 16; u++) sP[g + 8 * u] = vget(a, u);
    }
    if (r == j && p != j) {
#pragma unroll
      for (int u = 0; u < 16; u++) sJ[g + 8 * u] = vget(a, u);
    }
    if (g == gj && r > j) sl[r] = cmulf((r == p) ? sjj : vget(a, uj), ip);
    __syncthreads();
    if (p != j) {
      if (r == j) {
#pragma unroll
        for (int u = 0; u < 16; u++) vset(a, u, sP[g + 8 * u]);
      } else if (r == p) {
#pragma unroll
        for (int u = 0; u < 16; u++) vset(a, u, sJ[g + 8 * u]);
      }
    }
    if (r > j && r < kb) {
      const float2 l = sl[r];
#pragma unroll
      for (int u = 0; u < 16; u++) {
        const int c = g + 8 * u;
        if (c == j) {
          vset(a, u, l);
        } else if (c > j) {
          const float2 uc = sP[c];
          float2 e = vget(a, u);
          e.x -= fmaf(l.x, uc.x, -l.y * uc.y);
          e.y -= fmaf(l.x, uc.y, l.y * uc.x);
          vset(a, u, e);
        }
      }
    }
  }
  if (tid == 0) {
    if (!sticky) {
      stat[0] = bad ? 1 : 0;
      stat[1] = 0;
    } else if (bad) {
      stat[0] = 1;
    }
  }
#pragma unroll
  for (int u = 0; u < 16; u++) {
    const int c = g + 8 * u;
    sD[c * NB + r] = vget(a, u);
    if (r < kb && c < kb) A[(size_t)(k0 + c) * ld + k0 + r] = vget(a, u);
  }
  __syncthreads();
  // inverses: thread (column c, rows 8 u + g); rows and columns beyond kb are the identity's
  const int c = r;
  floatx32 x;
#pragma unroll
  for (int u = 0; u < 16; u++) vset(x, u, make_float2((8 * u + g == c) ? 1.0f : 0.0f, 0.0f));
  for (int q = kb - 1; q >= 0; q--) {
    if ((q & 7) == g) {
      const float2 v = cmulf(vget(x, q >> 3), crecipf(sD[q * NB + q]));
      vset(x, q >> 3, v);
      sx[q & 1][c] = v;
    }
    __syncthreads();
    const float2 xq = sx[q & 1][c];
#pragma unroll
    for (int u = 0; u < 16; u++) {
      const int rr = 8 * u + g;
      if (rr < q) {
        const float2 uq = sD[q * NB + rr];
        float2 e = vget(x, u);
        e.x -= fmaf(uq.x, xq.x, -uq.y * xq.y);
        e.y -= fmaf(uq.x, xq.y, uq.y * xq.x);
        vset(x, u, e);
      }
    }
  }
#pragma unroll
  for (int u = 0; u < 16; u++) {
    Uinv[(size_t)c * NB + 8 * u + g] = vget(x, u);
    vset(x, u, make_float2((8 * u + g == c) ? 1.0f : 0.0f, 0.0f));
  }
  __syncthreads();  // the last step's sx reads are done
  for (int q = 0; q < kb; q++) {
    if ((q & 7) == g) sx[q & 1][c] = vget(x, q >> 3);
    __syncthreads();
    const float2 xq = sx[q & 1][c];
#pragma unroll
    for (int u = 0; u < 16; u++) {
      const int rr = 8 * u + g;
      if (rr > q && rr < kb) {
        const float2 lq = sD[q * NB + rr];
        float2 e = vget(x, u);
        e.x -= fmaf(lq.x, xq.x, -lq.y * xq.y);
        e.y -= fmaf(lq.x, xq.y, lq.y * xq.x);
        vset(x, u, e);
      }
    }
  }
#pragma unroll
  for (int u = 0; u < 16; u++) Linv[(size_t)c * NB + 8 * u + g] = vget(x, u);
}

// max |Re| + |Im| over rows x cols of A (leading dimension lda) into stat[1] as
// float bits (non-negative floats order as integers; NaN above every finite):
// blockIdx.y is the column, the x blocks stride its rows
__global__ void __launch_bounds__(256) k_cmaxabs(int rows, const float2 *__restrict__ A, long long lda,
                                                 int *__restrict__ stat) {
  const float2 *col = A + (size_t)blockIdx.y * lda;
  float mx = 0.0f;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < rows; i += gridDim.x * 256) {
    const float2 v = col[i];
    const float a = fabsf(v.x) + fabsf(v.y);
    mx = (a > mx || a != a) ? a : mx;
  }
  for (int off = 32; off >= 1; off >>= 1) {
    const float o = __shfl_xor(mx, off, 64);
    mx = (o > mx || o != o) ? o : mx;
  }
  if ((threadIdx.x & 63) == 0) atomicMax(stat + 1, __float_as_int(mx));
}

// a growth bound at or above this keeps every block-local panel unchecked
#define BQD2_LU32_UNCHECKED 1.0e30f
static std::atomic<int> g_local_panel{1};
static std::atomic<float> g_growth_bound{16.0f};
static std::atomic<long long> g_panels_local{0}, g_panels_full{0};
// growth of each block-local panel of the most recent factorisation (any thread)
static std::mutex g_growth_mu;
static std::vector<float> g_growth_log;

extern "C" int bqd2_set_lu32_pivoting(int mode) {
  BQD2_CHECK(mode == 0 || mode == 1, BQD2_E_ARG, "bqd2_set_lu32_pivoting: mode is 0 (all rows) or 1 (block-local)");
  g_local_panel = mode;
  return BQD2_OK;
}

static bool g_lookahead = true;
extern "C" int bqd2_set_lu32_lookahead(int on) {
  g_lookahead = on != 0;
  return BQD2_OK;
}

extern "C" int bqd2_set_lu32_growth_bound(double bound) {
  BQD2_CHECK(bound >= 1.0, BQD2_E_ARG, "bqd2_set_lu32_growth_bound: the bound is at least 1");
  g_growth_bound = (float)bound;
  return BQD2_OK;
}

extern "C" int bqd2_lu32_panel_growth(float *out, int cap) {
  BQD2_CHECK(out || cap == 0, BQD2_E_ARG, "bqd2_lu32_panel_growth: null argument");
  std::lock_guard<std::mutex> lk(g_growth_mu);
  const int cnt = (int)g_growth_log.size();
  for (int i = 0; i < cnt && i < cap; i++) out[i] = g_growth_log[i];
  return cnt;
}

extern "C" int bqd2_lu32_panel_counts(long long *out) {
  BQD2_CHECK(out, BQD2_E_ARG, "bqd2_lu32_panel_counts: null argument");
  out[0] = g_panels_local.load();
  out[1] = g_panels_full.load();
  return BQD2_OK;
}

// the interchanges of the panel [k0, k0 + kb) on the columns outside it
__global__ void __launch_bounds__(256) k_cswap_cols(int n, float2 *__restrict__ A, int k0, int kb,
                                                    const int *__restrict__ ipiv) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n - kb) return;
  if (c >= k0) c += kb;
  float2 *a = A + (size_t)c * n;
  for (int j = k0; j < k0 + kb; j++) {
    const int p = ipiv[j];
    if (p == j) continue;
    const float2 t = a[j];
    a[j] = a[p];
    a[p] = t;
  }
}

// X <- L^-1 X for the ncols columns of X (column c at X + c ldx, rows 0..kb-1),
// L the unit lower part of the kb x kb block at Ld (leading dimension n): the
// inverse of a panel's L11 (X the identity).  One lane per column, 256 columns
// per workgroup.  The strictly lower part of L is packed row by row in LDS
// ((r, q), q < r, at r (r-1)/2 + q; every lane reads the same entry, a
// broadcast).  The rows go in blocks of 32 held in registers: the block's
// triangle is solved in registers, then each later row of the lane's column
// takes one 32-term dot product with it.
#define TB 32
// the strictly lower triangle of the kb x kb block into sL, packed by rows: a
// flat walk (consecutive lanes on consecutive rows of a column) with eight
// loads in flight per lane, since a walk column by column waits for one load
// per column
__device__ __forceinline__ void stage_l(int n, const float2 *__restrict__ Ld, int kb, float2 *sL) {
  const int tid = threadIdx.x, nt = blockDim.x, tot = kb * kb;
  for (int e0 = tid; e0 < tot; e0 += 8 * nt) {
    float2 v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int e = e0 + u * nt;
      v[u] = (e < tot) ? Ld[(size_t)(e / kb) * n + e % kb] : make_float2(0.0f, 0.0f);
    }
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int e = e0 + u * nt, q = e / kb, r = e % kb;
      if (e < tot && r > q) sL[r * (r - 1) / 2 + q] = v[u];
    }
  }
  __syncthreads();
}

// column c of X (ncols columns, ldx apart) := L^-1 X for the unit lower kb x kb
// factor staged in sL by stage_l, by forward substitution in 32-row blocks held
// in registers: the block's triangle row by row, then x_r -= L[r][block]
// x_block for the rows below
__device__ __forceinline__ void trsm_cols(int kb, float2 *Xall, long long ldx, int ncols, int c, const float2 *sL) {
  if (c >= ncols) return;
  float2 *X = Xall + (size_t)c * ldx;
  for (int b0 = 0; b0 < kb; b0 += TB) {
    float2 x[TB];
#pragma unroll
    for (int t = 0; t < TB; t++) x[t] = (b0 + t < kb) ? X[b0 + t] : make_float2(0.0f, 0.0f);
    // the block's triangle
#pragma unroll
    for (int t = 1; t < TB; t++) {
      if (b0 + t >= kb) break;
      float2 acc = make_float2(0.0f, 0.0f);
#pragma unroll
      for (int q = 0; q < t; q++) {
        const float2 l = sL[(b0 + t) * (b0 + t - 1) / 2 + b0 + q];
        acc.x += fmaf(l.x, x[q].x, -l.y * x[q].y);
        acc.y += fmaf(l.x, x[q].y, l.y * x[q].x);
      }
      x[t].x -= acc.x;
      x[t].y -= acc.y;
    }
#pragma unroll
    for (int t = 0; t < TB; t++)
      if (b0 + t < kb) X[b0 + t] = x[t];
    // later rows: x_r -= L[r][b0:b0+TB] x_block, RB rows of X loaded at once
    // (a lane's column is strided in memory: one load per row would wait for
    // each in turn)
    constexpr int RB = 16;
    for (int r0 = b0 + TB; r0 < kb; r0 += RB) {
      float2 y[RB];
#pragma unroll
      for (int u = 0; u < RB; u++) y[u] = (r0 + u < kb) ? X[r0 + u] : make_float2(0.0f, 0.0f);
#pragma unroll
      for (int u = 0; u < RB; u++) {
        const int r = r0 + u;
        if (r >= kb) break;
        float2 acc = make_float2(0.0f, 0.0f);
#pragma unroll
        for (int q = 0; q < TB; q++) {
          const float2 l = sL[r * (r - 1) / 2 + b0 + q];
          acc.x += fmaf(l.x, x[q].x, -l.y * x[q].y);
          acc.y += fmaf(l.x, x[q].y, l.y * x[q].x);
        }
        y[u].x -= acc.x;
        y[u].y -= acc.y;
        X[r] = y[u];
      }
    }
  }
}

// 256 columns per workgroup, one lane each
__global__ void __launch_bounds__(256) k_ctrsm_l(int n, const float2 *__restrict__ Ld, int kb, float2 *X,
                                                 long long ldx, int ncols) {
  __shared__ float2 sL[NB * (NB - 1) / 2];
  stage_l(n, Ld, kb, sL);
  trsm_cols(kb, X, ldx, ncols, blockIdx.x * 256 + threadIdx.x, sL);
}

// L11^-1 of every diagonal block (block blockIdx.x) into T + 2 NB NB blockIdx.x,
// which holds the identity: the solve's forward-sweep blocks
__global__ void __launch_bounds__(256) k_clinv(int n, const float2 *__restrict__ A, float2 *T) {
  __shared__ float2 sL[NB * (NB - 1) / 2];
  const int k0 = blockIdx.x * NB, kb = min(NB, n - k0);
  stage_l(n, A + (size_t)k0 * n + k0, kb, sL);
  trsm_cols(kb, T + (size_t)blockIdx.x * 2 * NB * NB, (long long)NB, kb, threadIdx.x, sL);
}

// Inverse of the upper factor of the diagonal block k0 = NB blockIdx.x (kb x kb),
// after the factorisation (a block is final once its panel is done: later
// interchanges touch only rows below it), written to T + (2 blockIdx.x + 1) NB NB
// (ldT = NB).  One lane per column of the inverse; the block staged in LDS.
__global__ void __launch_bounds__(NB) k_ctrtri_u(int n, const float2 *__restrict__ A, float2 *__restrict__ Tall) {
  __shared__ float2 sD[NB * NB];  // column-major (r, q) at q * NB + r
  const int c = threadIdx.x;
  const int k0 = blockIdx.x * NB, kb = min(NB, n - k0);
#pragma unroll 8
  for (int e = c; e < kb * kb; e += NB) {
    const int r = e % kb, q = e / kb;
    sD[q * NB + r] = A[(size_t)(k0 + q) * n + k0 + r];
  }
  __syncthreads();
  if (c >= kb) return;
  float2 *X = Tall + (size_t)(2 * blockIdx.x + 1) * NB * NB + (size_t)c * NB;  // column c of the inverse
  // U X = I: X[c][c] = 1 / U[c][c], X[r][c] = -(sum_{r < q <= c} U[r][q] X[q][c]) / U[r][r]
  for (int r = c + 1; r < kb; r++) X[r] = make_float2(0.0f, 0.0f);
  X[c] = crecipf(sD[c * NB + c]);
  for (int r = c - 1; r >= 0; r--) {
    float2 acc = make_float2(0.0f, 0.0f);
    for (int q = r + 1; q <= c; q++) {
      const float2 u = sD[q * NB + r], x = X[q];
      acc.x += fmaf(u.x, x.x, -u.y * x.y);
      acc.y += fmaf(u.x, x.y, u.y * x.x);
    }
    const float2 v = cmulf(acc, crecipf(sD[r * NB + r]));
    X[r] = make_float2(-v.x, -v.y);
  }
}

// the L blocks of T (block k at T + 2 NB NB k) set to the identity
__global__ void __launch_bounds__(256) k_ceye(int nbk, float2 *__restrict__ T) {
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= (size_t)nbk * NB * NB) return;
  const int r = e % NB, c = (e / NB) % NB;
  T[(e / (NB * NB)) * 2 * NB * NB + e % (NB * NB)] = make_float2(r == c ? 1.0f : 0.0f, 0.0f);
}

// ---------------------------------------------------------------- MFMA GEMM
// C[m x nc] -= A[m x kk] B[kk x nc] (SUB) or C = A B (SET), column-major complex
// float.  Four waves (2 x 2) of WA x WB blocks of 32 x 32: a workgroup tile of
// 64 WA rows x 64 WB columns; K staged in LDS in slices of 16 complex columns.
// v_mfma_f32_32x32x2_f32 computes D[i][j] += sum_{k<2} X[i][k] Y[k][j], with
// lane l holding X[l%32][l/32] and Y[l/32][l%32] and D[i][j] in register r of
// lane l for j = l%32, i = 8 (r/4) + 4 (l/32) + r%4.  The product is formed as
// C^T = B^T A^T, so j runs along the rows of C and the 32 lanes of a register
// store 32 consecutive rows of one column.
// Complex products, two forms:
//   4M  one instruction per complex k for each of Re C and Im C: the two k
//       slots carry the real and imaginary parts, X = (Br, Bi) for both,
//       Y = (Ar, -Ai) for Re C and (Ai, Ar) for Im C;
//   3M  (the trailing update) the two k slots carry two consecutive complex k
//       and three real products are accumulated, P1 = Ar Br, P2 = Ai Bi and
//       P3 = (Ar + Ai)(Br + Bi); Re = P1 - P2, Im = P3 - P1 - P2 at the end:
//       three instructions per two complex k instead of four (the 3M method,
//       A. Karatsuba's identity applied to matrices; N.J. Higham, "Stability
//       of a method for multiplying complex matrices with three real matrix
//       multiplications", SIAM J. Matrix Anal. Appl. 13 (1992) 681-687: the
//       error bound of Im C grows by a small constant factor).
// In-place forms: B may be C when one workgroup covers all m rows, or A may be
// C when one covers all nc columns: a workgroup then reads only the columns
// (rows) it writes, and writes after its last read.
// Every tile shape computes each element of C from the same MFMA chain over k,
// so the tile shape changes no bit of the result (the form does).
#define GK 16
#define GPAD 1
template <int SET, int WA, int WB, int M3>
__global__ void __launch_bounds__(256, 2) k_cgemm(int m, int nc, int kk, const float2 *A, long long lda,
                                                  const float2 *B, long long ldb, float2 *C, long long ldc) {
  constexpr int TR = 64 * WA, TC = 64 * WB;  // tile rows, columns
  constexpr int AKL = 256 / TR;              // lanes along k when staging A
  constexpr int NACC = M3 ? 3 : 2;
  __shared__ float2 sA[GK][TR + GPAD];  // (k, row)
  __shared__ float2 sB[GK][TC + GPAD];  // (k, col)
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave & 1, wc = wave >> 1;
  const int row0 = blockIdx.x * TR, col0 = blockIdx.y * TC;
  const int l32 = lane & 31;
  const int hi = lane >= 32 ? 1 : 0;
  // global -> register staging of one K slice: A TR rows x 16 k (consecutive
  // lanes on consecutive rows), B 16 k x TC columns (16 lanes per column)
  const int ar = tid % TR, ak = tid / TR;
  const int bk = tid & 15, bc = tid >> 4;
  float2 ra[GK / AKL], rb[TC / 16];
  auto load = [&](int k0) {
    const int gr = row0 + ar;
#pragma unroll
    for (int i = 0; i < GK / AKL; i++) {
      const int gk = k0 + ak + AKL * i;
      ra[i] = (gr < m && gk < kk) ? A[(size_t)gk * lda + gr] : make_float2(0.0f, 0.0f);
    }
    const int gk = k0 + bk;
#pragma unroll
    for (int i = 0; i < TC / 16; i++) {
      const int gc = col0 + bc + 16 * i;
      rb[i] = (gc < nc && gk < kk) ? B[(size_t)gc * ldb + gk] : make_float2(0.0f, 0.0f);
    }
  };
  floatx16 acc[NACC][WB][WA];
#pragma unroll
  for (int t = 0; t < NACC; t++)
#pragma unroll
    for (int b = 0; b < WB; b++)
#pragma unroll
      for (int a = 0; a < WA; a++)
#pragma unroll
        for (int r = 0; r < 16; r++) acc[t][b][a][r] = 0.0f;
  load(0);
  for (int k0 = 0; k0 < kk; k0 += GK) {
    __syncthreads();
#pragma unroll
    for (int i = 0; i < GK / AKL; i++) sA[ak + AKL * i][ar] = ra[i];
#pragma unroll
    for (int i = 0; i < TC / 16; i++) sB[bk][bc + 16 * i] = rb[i];
    __syncthreads();
    if (k0 + GK < kk) load(k0 + GK);  // next slice in flight during this one's MFMAs
    if (M3) {
#pragma unroll 4
      for (int q = 0; q < GK; q += 2) {
        float xr[WB], xi[WB], xs[WB], yr[WA], yi[WA], ys[WA];
#pragma unroll
        for (int b = 0; b < WB; b++) {
          const float2 v = sB[q + hi][wc * 32 * WB + b * 32 + l32];
          xr[b] = v.x;
          xi[b] = v.y;
          xs[b] = v.x + v.y;
        }
#pragma unroll
        for (int a = 0; a < WA; a++) {
          const float2 v = sA[q + hi][wr * 32 * WA + a * 32 + l32];
          yr[a] = v.x;
          yi[a] = v.y;
          ys[a] = v.x + v.y;
        }
#pragma unroll
        for (int b = 0; b < WB; b++)
#pragma unroll
          for (int a = 0; a < WA; a++) {
            acc[0][b][a] = __builtin_amdgcn_mfma_f32_32x32x2f32(xr[b], yr[a], acc[0][b][a], 0, 0, 0);
            acc[1][b][a] = __builtin_amdgcn_mfma_f32_32x32x2f32(xi[b], yi[a], acc[1][b][a], 0, 0, 0);
            acc[NACC - 1][b][a] = __builtin_amdgcn_mfma_f32_32x32x2f32(xs[b], ys[a], acc[NACC - 1][b][a], 0, 0, 0);
          }
      }
    } else {
#pragma unroll 4
      for (int q = 0; q < GK; q++) {
        float xb[WB], yr[WA], yi[WA];
#pragma unroll
        for (int b = 0; b < WB; b++) {
          const float2 v = sB[q][wc * 32 * WB + b * 32 + l32];
          xb[b] = hi ? v.y : v.x;
        }
#pragma unroll
        for (int a = 0; a < WA; a++) {
          const float2 v = sA[q][wr * 32 * WA + a * 32 + l32];
          yr[a] = hi ? -v.y : v.x;
          yi[a] = hi ? v.x : v.y;
        }
#pragma unroll
        for (int b = 0; b < WB; b++)
#pragma unroll
          for (int a = 0; a < WA; a++) {
            acc[0][b][a] = __builtin_amdgcn_mfma_f32_32x32x2f32(xb[b], yr[a], acc[0][b][a], 0, 0, 0);
            acc[1][b][a] = __builtin_amdgcn_mfma_f32_32x32x2f32(xb[b], yi[a], acc[1][b][a], 0, 0, 0);
          }
      }
    }
  }
  // acc[.][b][a] register r: row l32 of row block a, column 8 (r/4) + 4 hi + r%4 of column block b
#pragma unroll
  for (int b = 0; b < WB; b++)
#pragma unroll
    for (int a = 0; a < WA; a++) {
      const int gr = row0 + wr * 32 * WA + a * 32 + l32;
      if (gr >= m) continue;
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const int gc = col0 + wc * 32 * WB + b * 32 + 8 * (r >> 2) + 4 * hi + (r & 3);
        if (gc >= nc) continue;
        float re = acc[0][b][a][r], im = acc[1][b][a][r];
        if (M3) {
          const float p1 = re, p2 = im;
          re = p1 - p2;
          im = acc[NACC - 1][b][a][r] - p1 - p2;
        }
        float2 *p = C + (size_t)gc * ldc + gr;
        if (SET) {
          *p = make_float2(re, im);
        } else {
          float2 o = *p;
          o.x -= re;
          o.y -= im;
          *p = o;
        }
      }
    }
}

// Launch shapes.  GEMM_SQ128: 128 x 128 tiles, 4M (the in-place forms need a
// workgroup to cover all of C's rows or columns, as stated above);
// GEMM_SQ64: 64 x 64 tiles, 4M; GEMM_3M64 / GEMM_3M128X64: the trailing
// update, 3M, on 64 x 64 or 128 x 64 tiles (distinct operands only).
enum GemmShape { GEMM_SQ128 = 0, GEMM_SQ64 = 1, GEMM_3M64 = 2, GEMM_3M128X64 = 3 };
static int cgemm(bool set, int m, int nc, int kk, const float2 *A, long long lda, const float2 *B, long long ldb,
                 float2 *C, long long ldc, hipStream_t st, GemmShape shape = GEMM_SQ128) {
  if (m <= 0 || nc <= 0 || kk <= 0) return BQD2_OK;
  if (set && (shape == GEMM_3M64 || shape == GEMM_3M128X64)) shape = GEMM_SQ64;
  const int tr = (shape == GEMM_SQ128 || shape == GEMM_3M128X64) ? 128 : 64;
  const int tc = shape == GEMM_SQ128 ? 128 : 64;
  dim3 grid(bqd2_blocks(m, tr), bqd2_blocks(nc, tc));
#define BQD2_GEMM_LAUNCH(S, WA, WB, M3) \
  hipLaunchKernelGGL((k_cgemm<S, WA, WB, M3>), grid, dim3(256), 0, st, m, nc, kk, A, lda, B, ldb, C, ldc)
  switch (shape) {
    case GEMM_SQ128:
      if (set) BQD2_GEMM_LAUNCH(1, 2, 2, 0); else BQD2_GEMM_LAUNCH(0, 2, 2, 0);
      break;
    case GEMM_SQ64:
      if (set) BQD2_GEMM_LAUNCH(1, 1, 1, 0); else BQD2_GEMM_LAUNCH(0, 1, 1, 0);
      break;
    case GEMM_3M64: BQD2_GEMM_LAUNCH(0, 1, 1, 1); break;
    case GEMM_3M128X64: BQD2_GEMM_LAUNCH(0, 2, 1, 1); break;
  }
#undef BQD2_GEMM_LAUNCH
  BQD2_HIP(hipGetLastError());
  return BQD2_OK;
}

// ---------------------------------------------------------------- GEMV (few right-hand sides)
// Y[m x NR] -= A[m x kk] X[kk x NR] (SUB) or Y = A X (SET), kk <= NB; X staged
// in LDS before any write, so Y may alias X when one workgroup covers m.
template <int NR, int SET>
__global__ void __launch_bounds__(256) k_cgemv(int m, int kk, const float2 *__restrict__ A, long long lda,
                                               const float2 *X, long long ldx, float2 *Y, long long ldy) {
  __shared__ float2 sx[NR][NB];
  for (int e = threadIdx.x; e < NR * kk; e += blockDim.x) sx[e / kk][e % kk] = X[(size_t)(e / kk) * ldx + e % kk];
  __syncthreads();
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= m) return;
  float2 acc[NR];
#pragma unroll
  for (int s = 0; s < NR; s++) acc[s] = make_float2(0.0f, 0.0f);
  const float2 *a = A + r;
  for (int k = 0; k < kk; k++) {
    const float2 v = a[(size_t)k * lda];
#pragma unroll
    for (int s = 0; s < NR; s++) {
      const float2 x = sx[s][k];
      acc[s].x += fmaf(v.x, x.x, -v.y * x.y);
      acc[s].y += fmaf(v.x, x.y, v.y * x.x);
    }
  }
#pragma unroll
  for (int s = 0; s < NR; s++) {
    float2 *y = Y + (size_t)s * ldy + r;
    if (SET) *y = acc[s];
    else { float2 o = *y; o.x -= acc[s].x; o.y -= acc[s].y; *y = o; }
  }
}

template <int NR>
static void gemv_nr(bool set, int m, int kk, const float2 *A, long long lda, const float2 *X, long long ldx, float2 *Y,
                    long long ldy, hipStream_t st) {
  const int bs = set ? NB : 256;  // SET is used in place on one diagonal block: one workgroup
  dim3 g(bqd2_blocks(m, bs));
  if (set) hipLaunchKernelGGL((k_cgemv<NR, 1>), g, dim3(bs), 0, st, m, kk, A, lda, X, ldx, Y, ldy);
  else hipLaunchKernelGGL((k_cgemv<NR, 0>), g, dim3(bs), 0, st, m, kk, A, lda, X, ldx, Y, ldy);
}

// Y -= A X or Y = A X for nrhs right-hand sides (SET only in place on <= NB rows)
static int small_update(bool set, int m, int kk, int nrhs, const float2 *A, long long lda, const float2 *X,
                        long long ldx, float2 *Y, long long ldy, hipStream_t st) {
  if (m <= 0 || kk <= 0) return BQD2_OK;
  switch (nrhs) {
    case 1: gemv_nr<1>(set, m, kk, A, lda, X, ldx, Y, ldy, st); break;
    case 2: gemv_nr<2>(set, m, kk, A, lda, X, ldx, Y, ldy, st); break;
    case 3: gemv_nr<3>(set, m, kk, A, lda, X, ldx, Y, ldy, st); break;
    case 4: gemv_nr<4>(set, m, kk, A, lda, X, ldx, Y, ldy, st); break;
    default: return cgemm(set, m, nrhs, kk, A, lda, X, ldx, Y, ldy, st);
  }
  BQD2_HIP(hipGetLastError());
  return BQD2_OK;
}

// perm from the interchanges: row i of P A is row perm[i] of A (one lane: the
// interchanges are sequential)
// the row permutation from the interchanges: a sequential chain, run in LDS
// when the permutation fits (PERM_LDS ints), else in memory
#define PERM_LDS 32768
__global__ void __launch_bounds__(256) k_perm(int n, const int *__restrict__ ipiv, int *__restrict__ perm,
                                              int in_lds) {
  __shared__ int sperm[PERM_LDS];
  int *w = in_lds ? sperm : perm;
  for (int i = threadIdx.x; i < n; i += 256) w[i] = i;
  __syncthreads();
  if (threadIdx.x == 0)
    for (int i = 0; i < n; i++) {
      const int p = ipiv[i];
      const int t = w[i];
      w[i] = w[p];
      w[p] = t;
    }
  __syncthreads();
  if (in_lds)
    for (int i = threadIdx.x; i < n; i += 256) perm[i] = w[i];
}

__global__ void k_init_flag(int *flag, int n, int *stat) {
  if (threadIdx.x == 0) *flag = n;
  if (threadIdx.x < 2) stat[threadIdx.x] = 0;
}

// After the panels: inverses of the diagonal blocks' factors for the solves,
// the row permutation, the solves' form of the off-diagonal blocks, and the
// zero-pivot check.
static int cgetrf_finish(int n, float2 *A, int *d_ipiv, const LuWs &W, hipStream_t st) {
  const int nbk = bqd2_blocks(n, NB);
  hipLaunchKernelGGL(k_ceye, dim3(bqd2_blocks((long long)nbk * NB * NB, 256)), dim3(256), 0, st, nbk, W.T);
  hipLaunchKernelGGL(k_clinv, dim3(nbk), dim3(256), 0, st, n, A, W.T);
  hipLaunchKernelGGL(k_ctrtri_u, dim3(nbk), dim3(NB), 0, st, n, A, W.T);
  {
    const bool in_lds = n <= PERM_LDS;
    hipLaunchKernelGGL(k_perm, dim3(1), dim3(256), 0, st, n, d_ipiv, W.perm, in_lds ? 1 : 0);
  }
  BQD2_HIP(hipGetLastError());
  // the solves' form of the factors: each off-diagonal block times the inverse of
  // its column's diagonal block (L_rk L_kk^-1 below, U_rk U_kk^-1 above), in place
  for (int k0 = 0; k0 < n; k0 += NB) {
    const int kb = (n - k0 < NB) ? n - k0 : NB, rest = n - k0 - kb;
    const float2 *TL = W.T + (size_t)(k0 / NB) * 2 * NB * NB;
    float2 *Lb = A + (size_t)k0 * n + k0 + kb, *Ub = A + (size_t)k0 * n;
    int rc = cgemm(true, rest, kb, kb, Lb, n, TL, NB, Lb, n, st);
    if (!rc) rc = cgemm(true, k0, kb, kb, Ub, n, TL + NB * NB, NB, Ub, n, st);
    if (rc) return rc;
  }
  BQD2_HIP(hipGetLastError());
  int h_flag = n;
  BQD2_HIP(hipMemcpyAsync(&h_flag, W.flag, sizeof(int), hipMemcpyDeviceToHost, st));
  BQD2_HIP(hipStreamSynchronize(st));
  BQD2_CHECK(h_flag == n, BQD2_E_SINGULAR, "bqd2_cgetrf: zero pivot at column %d", h_flag);
  return BQD2_OK;
}

// Streams of the relaxed factorisation: the trailing work and the panels on
// disjoint sets of CUs (LA_PANEL_CUS for the panels), so that a panel (one
// workgroup of 1024 threads and 136 KB of LDS) starts beside the running
// update instead of waiting for a CU to drain.  Created once per host thread
// and device, default priority, kept for the process.
#define LA_PANEL_CUS 8
static bool lookahead_streams(hipStream_t *m, hipStream_t *pn) {
  struct Pair {
    int dev = -1;
    hipStream_t main = nullptr, panel = nullptr;
  };
  static thread_local Pair S;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return false;
  if (S.main && S.dev == dev) {
    *m = S.main;
    *pn = S.panel;
    return true;
  }
  int ncu = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 8 * LA_PANEL_CUS)
    return false;
  const int words = (ncu + 31) / 32;
  std::vector<uint32_t> mm(words, 0u), pm(words, 0u);
  for (int c = 0; c < ncu; c++) (c < LA_PANEL_CUS ? pm : mm)[c / 32] |= 1u << (c % 32);
  hipStream_t a = nullptr, b = nullptr;
  if (hipExtStreamCreateWithCUMask(&a, (uint32_t)words, mm.data()) != hipSuccess) return false;
  if (hipExtStreamCreateWithCUMask(&b, (uint32_t)words, pm.data()) != hipSuccess) {
    (void)hipStreamDestroy(a);
    return false;
  }
  S.dev = dev;
  S.main = a;
  S.panel = b;
  *m = a;
  *pn = b;
  return true;
}

// The relaxed factorisation (block-local pivots whatever the growth; the FP64
// refinement that these factors precondition validates them): no host
// decision per panel, so every launch is queued at once.  Order per panel k
// on the main stream: L21 = A21 U11^-1, swaps, U12 = L11^-1 A12, A22 -= L21 U12
// on the next panel's columns, then on the rest; the next panel's kernel runs
// on the panel stream between the two updates (events order it after the
// first and the next L21 after it; the rest of the update touches none of its
// columns).  Returns BQD2_E_SINGULAR when a diagonal block met a zero pivot.
static int cgetrf_relaxed(int n, float2 *A, int *d_ipiv, const LuWs &W, hipStream_t st) {
  hipStream_t M = st, P = st;
  const bool two = g_lookahead && lookahead_streams(&M, &P);
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};  // start/end, next panel's columns updated, panel done
  struct EvGuard {
    hipEvent_t *e;
    ~EvGuard() {
      for (int i = 0; i < 3; i++)
        if (e[i]) (void)hipEventDestroy(e[i]);
    }
  } guard{ev};
  if (two) {
    for (int i = 0; i < 3; i++) BQD2_HIP(hipEventCreateWithFlags(&ev[i], hipEventDisableTiming));
    BQD2_HIP(hipEventRecord(ev[0], st));
    BQD2_HIP(hipStreamWaitEvent(M, ev[0], 0));
    BQD2_HIP(hipStreamWaitEvent(P, ev[0], 0));
  }
  const size_t ld = (size_t)n;
  hipLaunchKernelGGL(k_init_flag, dim3(1), dim3(64), 0, M, W.flag, n, W.stat);
  auto panel = [&](hipStream_t s, int k0) {
    hipLaunchKernelGGL(k_cpanel_local, dim3(1), dim3(1024), 0, s, n, A, k0, std::min(NB, n - k0), d_ipiv, W.U,
                       W.U + NB * NB, W.stat, 1);
  };
  panel(M, 0);
  BQD2_HIP(hipGetLastError());
  for (int k0 = 0; k0 < n; k0 += NB) {
    const int kb = std::min(NB, n - k0);
    const int rest = n - k0 - kb;
    if (k0 > 0 && two) BQD2_HIP(hipStreamWaitEvent(M, ev[2], 0));
    float2 *A21 = A + (size_t)k0 * ld + k0 + kb;
    if (rest > 0) {
      int rc = cgemm(true, rest, kb, kb, A21, n, W.U + NB * NB, NB, A21, n, M, GEMM_SQ128);
      if (rc) return rc;
    }
    if (n > kb)
      hipLaunchKernelGGL(k_cswap_cols, dim3(bqd2_blocks(n - kb, 256)), dim3(256), 0, M, n, A, k0, kb, d_ipiv);
    BQD2_HIP(hipGetLastError());
    if (rest <= 0) continue;
    float2 *A12 = A + (size_t)(k0 + kb) * ld + k0;
    float2 *A22 = A + (size_t)(k0 + kb) * ld + k0 + kb;
    int rc = cgemm(true, kb, rest, kb, W.U, NB, A12, n, A12, n, M, GEMM_SQ128);
    const int c0 = std::min(NB, rest);
    if (!rc) rc = cgemm(false, rest, c0, kb, A21, n, A12, n, A22, n, M, GEMM_3M64);
    if (rc) return rc;
    if (two) {
      BQD2_HIP(hipEventRecord(ev[1], M));
      BQD2_HIP(hipStreamWaitEvent(P, ev[1], 0));
      panel(P, k0 + kb);
      BQD2_HIP(hipEventRecord(ev[2], P));
    } else {
      panel(M, k0 + kb);
    }
    BQD2_HIP(hipGetLastError());
    if (rest > c0) {
      const int nc = rest - c0;
      const TuneTicket tk = tune_begin(TUNE_CGEMM, tune_class(rest), 8.0 * rest * nc * kb, M);
      rc = cgemm(false, rest, nc, kb, A21, n, A12 + (size_t)c0 * ld, n, A22 + (size_t)c0 * ld, n, M,
                 tk.cand == 1 ? GEMM_3M128X64 : GEMM_3M64);
      tune_end(tk, M);
      if (rc) return rc;
    }
  }
  int rc = cgetrf_finish(n, A, d_ipiv, W, M);
  int zero = 0;
  if (!rc) {
    BQD2_HIP(hipMemcpyAsync(&zero, W.stat, sizeof(int), hipMemcpyDeviceToHost, M));
    BQD2_HIP(hipStreamSynchronize(M));
  }
  if (two) {
    BQD2_HIP(hipEventRecord(ev[0], M));
    BQD2_HIP(hipStreamWaitEvent(st, ev[0], 0));
  }
  if (rc) return rc;
  g_panels_local += bqd2_blocks(n, NB);
  BQD2_CHECK(!zero, BQD2_E_SINGULAR, "bqd2_cgetrf: zero pivot in a diagonal block (block-local pivots)");
  return BQD2_OK;
}

// Host-visible status of the block-local panel: pinned, one per host thread
// (a factorisation waits for its own panels), allocated on first use and kept.
static int *panel_status_buffer() {
  static thread_local int *h = nullptr;
  if (!h && hipHostMalloc((void **)&h, 64, hipHostMallocDefault) != hipSuccess) h = nullptr;
  return h;
}

// Order of the work per panel k with block-local pivoting (one stream):
//   swaps_k, U12_k = L11^-1 A12 (GEMM), A22 -= L21 U12 on the next panel's
//   columns, the next panel's attempt (backup, k_cpanel_local, L21 GEMM,
//   growth), then A22 -= L21 U12 on the remaining columns.
// The host waits for the attempt's status (an event) while the GPU runs the
// big update behind it, so the decision costs no idle time; a rejected panel
// is restored and factored with k_cstep launches queued after that update
// (which does not touch the panel's columns).
int bqd2_cgetrf_bound(int n, float *d_A, int *d_ipiv, void *d_ws, void *stream, float bound, float *max_growth) {
  BQD2_CHECK(n > 0 && d_A && d_ipiv && d_ws, BQD2_E_ARG, "bqd2_cgetrf: bad arguments");
  if (max_growth) *max_growth = 0.0f;
  hipStream_t st = (hipStream_t)stream;
  float2 *A = (float2 *)d_A;
  LuWs W = lu_ws(n, d_ws);
  if (g_local_panel.load() && bound >= BQD2_LU32_UNCHECKED) {
    if (max_growth) *max_growth = INFINITY;  // not measured
    return cgetrf_relaxed(n, A, d_ipiv, W, st);
  }
  const size_t ld = (size_t)n;
  int *h_stat = nullptr;
  hipEvent_t ev = nullptr;
  bool local_on = g_local_panel.load() != 0;
  std::vector<float> growths;
  if (local_on) {
    h_stat = panel_status_buffer();
    BQD2_CHECK(h_stat, BQD2_E_HIP, "bqd2_cgetrf: pinned status buffer");
    BQD2_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  }
  struct EvGuard {
    hipEvent_t e;
    ~EvGuard() { if (e) (void)hipEventDestroy(e); }
  } guard{ev};
  hipLaunchKernelGGL(k_init_flag, dim3(1), dim3(64), 0, st, W.flag, n, W.stat);
  // attempt of the block-local panel at k0: backup, factor, L21, growth, status to the host
  auto attempt = [&](int k0) -> int {
    const int kb = std::min(NB, n - k0), rows = n - k0, rest = rows - kb;
    BQD2_HIP(hipMemcpy2DAsync(W.bk, (size_t)rows * sizeof(float2), A + (size_t)k0 * ld + k0, ld * sizeof(float2),
                              (size_t)rows * sizeof(float2), kb, hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(k_cpanel_local, dim3(1), dim3(1024), 0, st, n, A, k0, kb, d_ipiv, W.U, W.U + NB * NB, W.stat,
                       0);
    BQD2_HIP(hipGetLastError());
    if (rest > 0) {
      float2 *A21 = A + (size_t)k0 * ld + k0 + kb;
      int rc = cgemm(true, rest, kb, kb, A21, n, W.U + NB * NB, NB, A21, n, st, GEMM_SQ128);
      if (rc) return rc;
      hipLaunchKernelGGL(k_cmaxabs, dim3(std::min(bqd2_blocks(rest, 256), 8), kb), dim3(256), 0, st, rest, A21,
                         (long long)n, W.stat);
      BQD2_HIP(hipGetLastError());
    }
    BQD2_HIP(hipMemcpyAsync(h_stat, W.stat, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
    BQD2_HIP(hipEventRecord(ev, st));
    return BQD2_OK;
  };
  if (local_on) {
    int rc = attempt(0);
    if (rc) return rc;
  }
  for (int k0 = 0; k0 < n; k0 += NB) {
    const int kb = std::min(NB, n - k0);
    const int rows = n - k0;
    const int rest = rows - kb;
    bool local = false;
    if (local_on) {
      BQD2_HIP(hipEventSynchronize(ev));
      float growth;
      memcpy(&growth, &h_stat[1], sizeof(float));
      if (h_stat[0]) growth = INFINITY;
      growths.push_back(growth);
      local = growth <= bound;
      if (local && max_growth && growth > *max_growth) *max_growth = growth;
      if (!local)
        BQD2_HIP(hipMemcpy2DAsync(A + (size_t)k0 * ld + k0, ld * sizeof(float2), W.bk, (size_t)rows * sizeof(float2),
                                  (size_t)rows * sizeof(float2), kb, hipMemcpyDeviceToDevice, st));
    }
    if (local) {
      g_panels_local++;
    } else {
      g_panels_full++;
      int nw = bqd2_blocks(rows, PROWS);
      if (nw > GMAXW) nw = GMAXW;
      const int chunk = bqd2_blocks(rows, nw);
      nw = bqd2_blocks(rows, chunk);
      for (int j = k0 - 1; j < k0 + kb; j++)
        hipLaunchKernelGGL(k_cstep, dim3(nw), dim3(256), 0, st, n, A, k0, kb, j, chunk, nw, d_ipiv, W);
    }
    if (n > kb)
      hipLaunchKernelGGL(k_cswap_cols, dim3(bqd2_blocks(n - kb, 256)), dim3(256), 0, st, n, A, k0, kb, d_ipiv);
    BQD2_HIP(hipGetLastError());
    if (rest > 0) {
      // U12 = L11^-1 A12 in place
      float2 *A12 = A + (size_t)(k0 + kb) * ld + k0;
      if (local) {
        int rc = cgemm(true, kb, rest, kb, W.U, NB, A12, n, A12, n, st, GEMM_SQ128);
        if (rc) return rc;
      } else {
        hipLaunchKernelGGL(k_ctrsm_l, dim3(bqd2_blocks(rest, 256)), dim3(256), 0, st, n, A + (size_t)k0 * ld + k0,
                           kb, A12, (long long)n, rest);
        BQD2_HIP(hipGetLastError());
      }
      // A22 -= L21 U12 on the MFMA units (3M form); with block-local panels the
      // next panel's columns first, then its attempt, then the other columns
      const float2 *L21 = A + (size_t)k0 * ld + k0 + kb;
      float2 *A22 = A + (size_t)(k0 + kb) * ld + k0 + kb;
      int c0 = 0;
      if (local_on) {
        c0 = std::min(NB, rest);
        int rc = cgemm(false, rest, c0, kb, L21, n, A12, n, A22, n, st, GEMM_3M64);
        if (!rc) rc = attempt(k0 + kb);
        if (rc) return rc;
      }
      if (rest > c0) {
        // operands distinct from A22: either tile shape, as the tuner picks
        const int nc = rest - c0;
        const TuneTicket tk = tune_begin(TUNE_CGEMM, tune_class(rest), 8.0 * rest * nc * kb, st);
        int rc = cgemm(false, rest, nc, kb, L21, n, A12 + (size_t)c0 * ld, n, A22 + (size_t)c0 * ld, n, st,
                       tk.cand == 1 ? GEMM_3M128X64 : GEMM_3M64);
        tune_end(tk, st);
        if (rc) return rc;
      }
    }
  }
  if (local_on) {
    std::lock_guard<std::mutex> lk(g_growth_mu);
    g_growth_log.swap(growths);
  }
  return cgetrf_finish(n, A, d_ipiv, W, st);
}

extern "C" int bqd2_cgetrf(int n, float *d_A, int *d_ipiv, void *d_ws, void *stream) {
  return bqd2_cgetrf_bound(n, d_A, d_ipiv, d_ws, stream, g_growth_bound.load(), nullptr);
}

// One block step of a triangular sweep.  The factorisation leaves the
// off-diagonal blocks multiplied by the inverse of their column's diagonal
// block (M_rk = L_rk L_kk^-1 below the diagonal, U_rk U_kk^-1 above), so the
// step is y_r -= M_rk y_k for the rows still to be eliminated ([k0 + kb, n)
// forward, [0, k0) backward) and x_k = T_k y_k, written to X by the last
// workgroup.  Nobody writes rows k0.. of Y in this launch.  A workgroup takes
// 64 rows; its wave w sums the columns 32 w .. 32 w + 31 of the block (lanes on
// consecutive rows), and the four partial sums meet in LDS.
template <int NR, bool UPPER>
__global__ void __launch_bounds__(256) k_csweep(int n, const float2 *__restrict__ A, const float2 *__restrict__ Tk,
                                                int k0, int kb, float2 *Y, float2 *X, long long ld) {
  __shared__ float2 sy[NR][NB];
  __shared__ float2 part[4][NR][64];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  for (int e = tid; e < NR * kb; e += 256) sy[e / kb][e % kb] = Y[(size_t)(e / kb) * ld + k0 + e % kb];
  __syncthreads();
  if (blockIdx.x == gridDim.x - 1) {
    // x_k = T_k y_k
    if (tid >= kb) return;
    float2 acc[NR];
#pragma unroll
    for (int s = 0; s < NR; s++) acc[s] = make_float2(0.0f, 0.0f);
#pragma unroll 8
    for (int q = 0; q < kb; q++) {
      const float2 t = Tk[(size_t)q * NB + tid];
#pragma unroll
      for (int s = 0; s < NR; s++) {
        const float2 y = sy[s][q];
        acc[s].x += fmaf(t.x, y.x, -t.y * y.y);
        acc[s].y += fmaf(t.x, y.y, t.y * y.x);
      }
    }
#pragma unroll
    for (int s = 0; s < NR; s++) X[(size_t)s * ld + k0 + tid] = acc[s];
    return;
  }
  const int r0 = UPPER ? 0 : k0 + kb, r1 = UPPER ? k0 : n;
  const int r = r0 + blockIdx.x * 64 + lane;
  float2 acc[NR];
#pragma unroll
  for (int s = 0; s < NR; s++) acc[s] = make_float2(0.0f, 0.0f);
  if (r < r1) {
    const int q0 = 32 * w, q1 = min(kb, q0 + 32);
    const float2 *a = A + (size_t)k0 * n + r;
#pragma unroll 8
    for (int q = q0; q < q1; q++) {
      const float2 v = a[(size_t)q * n];
#pragma unroll
      for (int s = 0; s < NR; s++) {
        const float2 x = sy[s][q];
        acc[s].x += fmaf(v.x, x.x, -v.y * x.y);
        acc[s].y += fmaf(v.x, x.y, v.y * x.x);
      }
    }
  }
#pragma unroll
  for (int s = 0; s < NR; s++) part[w][s][lane] = acc[s];
  __syncthreads();
  if (w != 0 || r >= r1) return;
#pragma unroll
  for (int s = 0; s < NR; s++) {
    float2 t = part[0][s][lane];
    for (int v = 1; v < 4; v++) {
      t.x += part[v][s][lane].x;
      t.y += part[v][s][lane].y;
    }
    float2 *y = Y + (size_t)s * ld + r;
    float2 o = *y;
    o.x -= t.x;
    o.y -= t.y;
    *y = o;
  }
}

// One block step: launched on st, or (gr != nullptr) added to the graph gr
// after the node *prev, which it then replaces.
template <int NR>
static int sweep_nr(bool upper, int n, const float2 *A, const float2 *Tk, int k0, int kb, float2 *Y, float2 *X,
                    long long ld, hipStream_t st, hipGraph_t gr, hipGraphNode_t *prev) {
  const int rows = upper ? k0 : n - k0 - kb;
  dim3 g(bqd2_blocks(rows, 64) + 1);  // row workgroups + the one that forms x_k
  void *fn = upper ? (void *)k_csweep<NR, true> : (void *)k_csweep<NR, false>;
  if (!gr) {
    if (upper) hipLaunchKernelGGL((k_csweep<NR, true>), g, dim3(256), 0, st, n, A, Tk, k0, kb, Y, X, ld);
    else hipLaunchKernelGGL((k_csweep<NR, false>), g, dim3(256), 0, st, n, A, Tk, k0, kb, Y, X, ld);
    return BQD2_OK;
  }
  void *args[] = {(void *)&n, (void *)&A, (void *)&Tk, (void *)&k0, (void *)&kb, (void *)&Y, (void *)&X, (void *)&ld};
  hipKernelNodeParams kp = {};
  kp.func = fn;
  kp.gridDim = g;
  kp.blockDim = dim3(256);
  kp.sharedMemBytes = 0;
  kp.kernelParams = args;
  kp.extra = nullptr;
  hipGraphNode_t node;
  BQD2_HIP(hipGraphAddKernelNode(&node, gr, *prev ? prev : nullptr, *prev ? 1 : 0, &kp));
  *prev = node;
  return BQD2_OK;
}

// Y (n x nrhs, column-major, leading dimension ld, already permuted) <- U^-1 L^-1 Y;
// X is a second n x nrhs buffer: the forward sweep writes its result there and
// the backward sweep writes back into Y.  More than 4 right-hand sides: the
// block steps as MFMA GEMMs, in place on Y.
static int cgetrs_perm(int n, int nrhs, const float2 *A, const LuWs &W, float2 *Y, float2 *X, long long ld,
                       hipStream_t st, hipGraph_t gr = nullptr) {
  const int nbk = bqd2_blocks(n, NB);
  if (nrhs <= 4) {
    hipGraphNode_t prev = nullptr;
    for (int pass = 0; pass < 2; pass++) {
      const bool upper = pass == 1;
      float2 *in = upper ? X : Y, *out = upper ? Y : X;
      for (int b = 0; b < nbk; b++) {
        const int k0 = (upper ? nbk - 1 - b : b) * NB, kb = (n - k0 < NB) ? n - k0 : NB;
        const float2 *Tk = W.T + (size_t)(k0 / NB) * 2 * NB * NB + (upper ? NB * NB : 0);
        int rc;
        switch (nrhs) {
          case 1: rc = sweep_nr<1>(upper, n, A, Tk, k0, kb, in, out, ld, st, gr, &prev); break;
          case 2: rc = sweep_nr<2>(upper, n, A, Tk, k0, kb, in, out, ld, st, gr, &prev); break;
          case 3: rc = sweep_nr<3>(upper, n, A, Tk, k0, kb, in, out, ld, st, gr, &prev); break;
          default: rc = sweep_nr<4>(upper, n, A, Tk, k0, kb, in, out, ld, st, gr, &prev); break;
        }
        if (rc) return rc;
      }
    }
    BQD2_HIP(hipGetLastError());
    return BQD2_OK;
  }
  BQD2_CHECK(!gr, BQD2_E_ARG, "bqd2 sweeps: graphs take at most 4 right-hand sides");
  // forward: y_r -= M_rk y_k below, then y_k = T_k y_k (in place), by diagonal blocks
  for (int k0 = 0; k0 < n; k0 += NB) {
    const int kb = (n - k0 < NB) ? n - k0 : NB;
    const float2 *Tk = W.T + (size_t)(k0 / NB) * 2 * NB * NB;
    int rc = small_update(false, n - k0 - kb, kb, nrhs, A + (size_t)k0 * n + k0 + kb, n, Y + k0, ld, Y + k0 + kb,
                          ld, st);
    if (!rc) rc = small_update(true, kb, kb, nrhs, Tk, NB, Y + k0, ld, Y + k0, ld, st);
    if (rc) return rc;
  }
  // backward: the same with the upper factor, last block first
  for (int k0 = ((n - 1) / NB) * NB; k0 >= 0; k0 -= NB) {
    const int kb = (n - k0 < NB) ? n - k0 : NB;
    const float2 *Tk = W.T + (size_t)(k0 / NB) * 2 * NB * NB + NB * NB;
    int rc = small_update(false, k0, kb, nrhs, A + (size_t)k0 * n, n, Y + k0, ld, Y, ld, st);
    if (!rc) rc = small_update(true, kb, kb, nrhs, Tk, NB, Y + k0, ld, Y + k0, ld, st);
    if (rc) return rc;
  }
  return BQD2_OK;
}

// complex float right-hand sides: out[i] = in[perm[i]]
__global__ void __launch_bounds__(256) k_gather_f(int n, const int *__restrict__ perm, const float2 *__restrict__ in,
                                                  long long ldi, float2 *__restrict__ out, long long ldo) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, s = blockIdx.y;
  if (i >= n) return;
  out[(size_t)s * ldo + i] = in[(size_t)s * ldi + perm[i]];
}

extern "C" int bqd2_cgetrs(int n, int nrhs, const float *d_A, const void *d_ws, float *d_B, float *d_tmp,
                           void *stream) {
  BQD2_CHECK(n > 0 && nrhs > 0 && d_A && d_ws && d_B && d_tmp, BQD2_E_ARG, "bqd2_cgetrs: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  LuWs W = lu_ws(n, (void *)d_ws);
  float2 *B = (float2 *)d_B, *T = (float2 *)d_tmp;
  hipLaunchKernelGGL(k_gather_f, dim3(bqd2_blocks(n, 256), nrhs), dim3(256), 0, st, n, W.perm, B, (long long)n, T,
                     (long long)n);
  int rc = cgetrs_perm(n, nrhs, (const float2 *)d_A, W, T, B, n, st);
  if (rc) return rc;
  BQD2_HIP(hipMemcpyAsync(B, T, (size_t)n * nrhs * sizeof(float2), hipMemcpyDeviceToDevice, st));
  return BQD2_OK;
}

// ---------------------------------------------------------------- FP64 <-> FP32 for the refinement
// Af (column-major complex float) = rows of the row-major complex double A, each
// scaled by dinv[r] (rows with map[r] < 0: identity rows); 32 x 32 tiles through
// LDS so that reads and writes coalesce.
__global__ void __launch_bounds__(256) k_rm64_to_cm32(int n, const double *__restrict__ A, long long lda,
                                                      const int *__restrict__ map, const double *__restrict__ dinv,
                                                      float2 *__restrict__ Af) {
  __shared__ float2 t[32][33];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  const int r0 = blockIdx.y * 32, c0 = blockIdx.x * 32;
  for (int i = ty; i < 32; i += 8) {
    const int r = r0 + i, c = c0 + tx;
    float2 v = make_float2(0.0f, 0.0f);
    if (r < n && c < n) {
      if (map && map[r] < 0) {
        v = make_float2(r == c ? 1.0f : 0.0f, 0.0f);  // unused row: identity, as GMRES treats it
      } else {
        const double *a = A + 2 * ((size_t)r * lda + c);
        const double dr = dinv ? dinv[2 * r] : 1.0, di = dinv ? dinv[2 * r + 1] : 0.0;
        v = make_float2((float)(a[0] * dr - a[1] * di), (float)(a[0] * di + a[1] * dr));
      }
    }
    t[i][tx] = v;
  }
  __syncthreads();
  for (int i = ty; i < 32; i += 8) {
    const int c = c0 + i, r = r0 + tx;
    if (r < n && c < n) Af[(size_t)c * n + r] = t[tx][i];
  }
}

// x (complex double, nrhs columns of leading dimension ld) -> P x in complex float
__global__ void __launch_bounds__(256) k_gather_d2f(int n, const int *__restrict__ perm, const double *__restrict__ x,
                                                    long long ldx, float2 *__restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, s = blockIdx.y;
  if (i >= n) return;
  const double *v = x + 2 * ((size_t)s * ldx + perm[i]);
  y[(size_t)s * n + i] = make_float2((float)v[0], (float)v[1]);
}

__global__ void __launch_bounds__(256) k_f2d(int n, const float2 *__restrict__ y, double *__restrict__ x,
                                             long long ldx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, s = blockIdx.y;
  if (i >= n) return;
  const float2 v = y[(size_t)s * n + i];
  double *o = x + 2 * ((size_t)s * ldx + i);
  o[0] = v.x;
  o[1] = v.y;
}

// round the row-scaled system to complex float, column-major
int bqd2_lu32_convert(int n, const double *A, long long lda, const int *map, const double *dinv, float *Af,
                      hipStream_t st) {
  dim3 g(bqd2_blocks(n, 32), bqd2_blocks(n, 32));
  hipLaunchKernelGGL(k_rm64_to_cm32, g, dim3(256), 0, st, n, A, lda, map, dinv, (float2 *)Af);
  BQD2_HIP(hipGetLastError());
  return BQD2_OK;
}

// out = M^-1 in for complex double vectors (nrhs columns, leading dimension ld),
// M = P^T L U the complex float factors; tmp holds 2 n nrhs complex floats
// Sweeps as a HIP graph (0: launched one by one; for tests that compare the two)
static bool g_use_graphs = true;
extern "C" int bqd2_set_graphs(int on) {
  g_use_graphs = on != 0;
  return BQD2_OK;
}

int bqd2_lu32_apply(int n, int nrhs, const float *Af, const void *ws, const double *in, long long ldi, double *out,
                    long long ldo, float *tmp, hipStream_t st, hipGraphExec_t *graph) {
  LuWs W = lu_ws(n, (void *)ws);
  float2 *T = (float2 *)tmp;
  dim3 g(bqd2_blocks(n, 256), nrhs);
  hipLaunchKernelGGL(k_gather_d2f, g, dim3(256), 0, st, n, W.perm, in, ldi, T);
  if (graph && g_use_graphs) {
    // the block steps of both sweeps read and write fixed buffers (the factors,
    // the workspace and tmp): made a graph once per preconditioner and replayed
    if (!*graph) {
      // built node by node (no stream capture: a capture in progress would
      // conflict with other threads' synchronous work on the null stream)
      hipGraph_t gr = nullptr;
      BQD2_HIP(hipGraphCreate(&gr, 0));
      int rc = cgetrs_perm(n, nrhs, (const float2 *)Af, W, T, T + (size_t)n * nrhs, n, st, gr);
      hipError_t e = rc ? hipSuccess : hipGraphInstantiate(graph, gr, nullptr, nullptr, 0);
      (void)hipGraphDestroy(gr);
      if (rc) return rc;
      BQD2_HIP(e);
    }
    BQD2_HIP(hipGraphLaunch(*graph, st));
  } else {
    int rc = cgetrs_perm(n, nrhs, (const float2 *)Af, W, T, T + (size_t)n * nrhs, n, st);
    if (rc) return rc;
  }
  hipLaunchKernelGGL(k_f2d, g, dim3(256), 0, st, n, T, out, ldo);
  BQD2_HIP(hipGetLastError());
  return BQD2_OK;
}

extern "C" int bqd2_cgemm(int set, int m, int nc, int kk, const float *d_A, long long lda, const float *d_B,
                          long long ldb, float *d_C, long long ldc, void *stream) {
  BQD2_CHECK(m >= 0 && nc >= 0 && kk >= 0 && d_A && d_B && d_C, BQD2_E_ARG, "bqd2_cgemm: bad arguments");
  // C = A B: 128 x 128 tiles (the in-place forms); C -= A B: the trailing update's 3M form
  return cgemm(set != 0, m, nc, kk, (const float2 *)d_A, lda, (const float2 *)d_B, ldb, (float2 *)d_C, ldc,
               (hipStream_t)stream, set ? GEMM_SQ128 : GEMM_3M64);
}
