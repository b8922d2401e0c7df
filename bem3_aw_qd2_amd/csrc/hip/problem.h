// problem.h -- device-resident state of one boundary-value problem.
#pragma once
#include <atomic>
#include <mutex>
#include <thread>
#include <vector>
#include "common.h"
#include "gf_dev.h"

// quadrature and collocation tables, one device copy per problem
struct Tab {
  double xli[12], wli[12], xhi[64], whi[64];
  double zt_44[4], et_44[4], wt_44[4], zt_34[4], et_34[4], wt_34[4];
  double zt_49[9], et_49[9], wt_49[9], zt_37[7], et_37[7], wt_37[7];
  // Gauss-Legendre rules of 4, 6 and 8 points for the smooth remainder of the
  // lattice sum (bqd2_gauss_legendre)
  double xs4[4], ws4[4], xs6[6], ws6[6], xs8[8], ws8[8];
};

// element data on the device, indexed by element id (row 0 unused)
struct DevMesh {
  int Ne, NN, MN, ps;
  const double *cr;   // (Ne+1) x 12
  const double *cw;   // (Ne+1) x 9
  const double *ren;  // (Ne+1) x 12
  const double *wen;  // (Ne+1) x 12
  const int *tri;     // (Ne+1): 1 for a linear triangle
  const int *eni;     // (Ne+1) x 4
  const int *md, *sd; // (Ne+1)
  const double *Pi, *dPi;  // (Ne+1) x 4 complex
};

// per-domain device buffers
struct DevDomain {
  int did, Ne, N;      // N = 4 Ne
  int nn;              // collocation nodes of the domain (3 or 4 per element)
  int has_tri;         // 1 when the domain has linear triangles
  int *sid;            // Ne+1 signed element ids
  int *rowmap;         // N: compressed node index of row (t-1)*4+tn, -1 for unused triangle slot
  int *fperm;          // Ne: element positions in the field kernels' order (latitude bands, then azimuth)
  double k;            // real wave number of the domain
  double *G, *H;       // N x N complex: row `row` at (row - r0) ldm
  double *dG, *dH;     // 2N x N complex (interior / isolated domains): rows 2 (row - r0) + j, leading dimension N
  long long ldm;       // row stride of G and H in complex entries: N, or 2N in the system layout
  int r0, nr;          // rows held by G, H, dG, dH and Fpair: [r0, r0 + nr) (all N rows unless streamed in windows)
  double *dF;          // N x 3
  double *Fpair;       // nr x Ne x 3 : per pair F, dF/dtz, dF/dte
  unsigned char *rule; // N x Ne : 1 = GH64 rule
  int *plist;          // N x Ne + PLIST_PAD : (row, s) pairs sorted by rule and element kind
  float *pkey;         // N x Ne : |z| offset of the element centre from the target (pair order)
  int plist_n;         // entries of plist, 0 until the first assembly (rules and list are cached)
  long long n_gh, n_split;  // GH64 pairs, split ones among them
  long long n_smooth;       // pairs whose lattice remainder takes a 4-, 6- or 8-point rule
  long long n_free_pts;     // free-space kernel points summed by the split pairs
  double *P, *dP;      // (Ne+1) x 4 complex
  double *pv, *dpv;    // (Ne+1) x 4 x 3 complex
  bool assembled;
  // mixed-precision solver: complex float factors of the row-scaled G for the
  // dpv systems (N x N, identity on unused rows), valid until the next assembly
  float *gAf, *gtmp;   // N x N; 2 x N x 3
  void *gws;           // bqd2_cgetrf_wsize(N)
  double *gdinv;       // N complex
  int *gipiv;          // N
  bool g_lu_valid;
  float g_growth;  // largest block-local panel growth kept in its factors
  hipGraphExec_t g_graph;  // its triangular sweeps (3 right-hand sides) as a graph
};

// complex-entry offsets of G/H entry (row, col), of dG/dH row 2 row + j, column
// col, and the double offset of the Fpair triple (row, source position s)
__host__ __device__ __forceinline__ size_t dom_gh(const DevDomain &D, int row, size_t col) {
  return (size_t)(row - D.r0) * (size_t)D.ldm + col;
}
__host__ __device__ __forceinline__ size_t dom_dgh(const DevDomain &D, int row, int j, size_t col) {
  return (size_t)(2 * (row - D.r0) + j) * (size_t)D.N + col;
}
__host__ __device__ __forceinline__ size_t dom_fp(const DevDomain &D, int row, int s) {
  return ((size_t)(row - D.r0) * (size_t)D.Ne + (size_t)(s - 1)) * 3;
}

// execution context of one calling host thread for the field and force entry
// points: its own stream, grow-only workspace and counters, so that callers
// that evaluate points from several threads overlap on the device
struct CallCtx {
  hipStream_t st;
  char *ws;
  size_t ws_bytes;
  long long *cnt;
};

// Field points of the periodic domain beyond the boundary's z extent by h or
// more: there the lattice sum is the spectral series in which the source and
// target factors separate (field.hip, k_far_*).  Orders (n1, n2) on the grid
// n1lo .. n1lo + n1n - 1, n2lo .. n2lo + n2n - 1 (G = n1n n2n of them); set at
// creation, on = 0 where the grid would exceed FAR_MAX_ORDERS.
struct FarGeom {
  int on;
  int n1lo, n1n, n2lo, n2n, G;
  double ztop, zbot, h;
};
// The field kernels visit the periodic domain's elements in FAR_NB bands of the
// element centre's z (DevDomain.fperm).  Band b holds the elements at fperm
// positions [es, ee); zpre is the highest node of bands 0..b, zsuf the lowest
// node of bands b..FAR_NB-1.  A field point at height z takes the bands with
// zpre <= z - h (below it) and those with zsuf >= z + h (above it) from
// plane-wave amplitudes, the rest by the lattice sums.
enum { FAR_NB = 32 };
struct FarBand {
  int es, ee;
  double zpre, zsuf;
};
enum { FAR_MAX_ORDERS = 32768 };
// entries of the pair list beyond N x Ne: each of the 18 classes (9 rules, two
// element kinds) is padded to a multiple of the pairs per wavefront (up to 16)
enum { PLIST_PAD = 512 };
// the order grid of a periodic model (QPD2 layout qd8): G orders, (2 n1 + 1) x
// (2 n2 + 1), and the height h; 0 when it would exceed FAR_MAX_ORDERS
int far_grid(const double *qd8, int *n1, int *n2, double *h);

struct bqd2_problem {
  DevMesh m;
  Tab *tab;            // device
  Tab htab;            // host copy
  GfParams gf;
  double qd8[8];
  double *rn;         // (Nn+1) x 3 mesh nodes (device)
  int *ed;            // (Ne+1) x 4 corner node ids (device)
  double k2[16][2];    // host copies of the media constants (MN+1 <= 16)
  double k0[16], rho0[16];
  double pw_p[2], pw_rho0, pw_c0, pw_k0, pw_tv[3], pw_R[9];
  std::vector<DevDomain> dom;
  int any_tri;                // the mesh has linear triangles
  std::vector<std::vector<char>> tripos;  // host, per domain: 1 where element position t is a triangle
  std::vector<void *> owned;  // device allocations released by bqd2_destroy
  size_t alloc_bytes;         // their size in bytes
  hipStream_t stream;
  long long *d_counters;      // device counters (GF terms)
  // system
  double *A, *b;
  int *ipiv;
  int na;
  bool solved;
  int solver;              // 0: GMRES, LU when it misses the tolerance; 1: LU only;
                           // 2: complex float LU (MFMA) + FP64 GMRES refinement
  // mixed-precision factorisation of the row-scaled system (solver 2); valid
  // until the next assembly
  float *Af;               // na x na complex float, column-major
  void *luws;              // bqd2_cgetrf_wsize(na) bytes
  float *lutmp;            // 2 x na x 3 complex float
  double *dinv;            // na complex: inverse diagonal of the system
  hipGraphExec_t lu_graph; // the triangular sweeps of its factors as a graph
  bool lu32_valid;
  bool lu64_valid;         // p->A holds the FP64 LU factors of the current assembly (with ipiv)
  float lu32_growth;       // largest block-local panel growth kept in the factors
  double ms_factor;        // time of the last factorisation (conversion + LU)
  double *hk2d;            // k2[m] / k2[s] table (16 x 16 complex), set once
  double *xg;              // GMRES solution of the system (na)
  double *tpv;             // open-domain node velocities (NN x 3)
  double *dpvw;            // dpv right-hand sides and solutions on the row slots (2 x 3 x N)
  int info_its[17];        // Krylov steps of the last solve: system, then dpv of domain d at 1 + d
  double info_res[17];     // final scaled residuals
  int info_lu;             // systems of the last solve that fell back to the LU
  // the creating thread uses main (the problem's stream and counters); other
  // threads get a context of their own on first use
  CallCtx main;
  std::thread::id owner;
  std::mutex ctx_mu;
  std::vector<std::pair<std::thread::id, CallCtx *>> ctxs;
  // system layout (bqd2_create_large): both domains' matrices in one row-major
  // na x na array S = [H0 G0; H1 G1] (na = 2N, the system's own column order);
  // the interior domain's rows are assembled in windows of lwin rows when needed
  int large;
  double *S;
  int s_lower;               // what rows [N, na) of S hold: LG_EMPTY, LG_G1H1 or LG_AF
  int lwin;                  // rows per window of the interior domain
  int lmax_resident;         // windows kept resident by the mixed solver (-1: as many as fit)
  std::vector<double *> lpool;  // resident windows of [H1 G1] rows, windows 0, 1, ...
  double *lscr;              // one window of [H1 G1] rows for the others
  double *lFp;               // Fpair of one window
  double *ldgw, *ldhw;       // dG, dH of one window (velocity recovery)
  double *ldinv;             // na complex: inverse diagonal of the system
  double *lvec;              // 3 na complex: the operator's two column vectors, the right-hand side's
  bqd2_large_stats lst;      // timings of the last solve and velocity recovery
  // far field: geometry, and the plane-wave amplitudes of the current solution
  // per field rule type, made at the first field call that needs them after
  // the solution changed (sol_epoch counts the changes)
  FarGeom far;
  int far_enabled;           // bqd2_set_field_far (1 by default)
  int smooth_mode;           // bqd2_set_remainder_rules (1 by default)
  std::atomic<unsigned long long> sol_epoch;
  std::mutex far_mu;
  FarBand *far_bands;        // FAR_NB bands (device)
  std::vector<int> fband0;   // host: band of each fperm position of domain 0
  std::vector<int> fperm0;   // host: domain 0's fperm
  double *far_amp[4];        // 2 sides x FAR_NB bands x G orders x 8 doubles
  unsigned long long far_amp_epoch[4];  // sol_epoch + 1 they were made for (0: none)
  hipEvent_t far_ev[4];      // recorded after they were made
  // bqd2_set_field_timing: HIP events around the field's point kernel
  int field_timing;
  double field_ms_part;      // k_field_part of the last bqd2_field call (ms)
};
enum { LG_EMPTY = 0, LG_G1H1 = 1, LG_AF = 2 };

int bqd2_alloc(bqd2_problem *p, void **ptr, size_t bytes);
// release one allocation of bqd2_alloc before the problem ends (*ptr = nullptr)
void bqd2_release(bqd2_problem *p, void **ptr, size_t bytes);

// Right preconditioner of GMRES: the complex float LU factors (lu32.hip) of the
// row-scaled matrix D A (D = the inverse diagonal), so that D A M^-1 is close to
// the identity.  tmp: 2 n nrhs complex floats.
struct Precond {
  const float *Af;
  const void *ws;
  float *tmp;
  hipGraphExec_t *graph;  // the sweeps, captured at the first application (owned by the factors' holder)
};
// A matrix given as an operator: mv(ctx, X, Y, st) sets Y = dinv A X for one
// complex vector of n entries (the caller's row scaling dinv, n complex)
struct GmOp {
  const double *dinv;
  int (*mv)(void *ctx, const double *X, double *Y, hipStream_t st);
  void *ctx;
};
// restarted GMRES on nrhs <= 3 right-hand sides sharing the row-major matrix A
// (gmres.hip), right-preconditioned by pc when it is not null; returns 0 when
// the scaled residual is below tol, 1 when not (the caller falls back to
// another solver), < 0 on errors
int bqd2_gmres(CallCtx *cx, int n, int nrhs, const double *A, long long lda, const int *map, const double *B,
               long long ldb, double *X, long long ldx, double tol, int m, int maxit, int *its, double *relres,
               const Precond *pc = nullptr, const GmOp *op = nullptr);
// lu32.hip: the row-scaled row-major complex double matrix rounded to column-major
// complex float; out = M^-1 in with the factors (in may be out)
// bqd2_cgetrf with the growth bound of block-local panels given (the C ABI
// uses the process-wide one); *max_growth (may be null) receives the largest
// growth of a kept block-local panel, 0 when none was kept.
int bqd2_cgetrf_bound(int n, float *d_A, int *d_ipiv, void *d_ws, void *stream, float bound, float *max_growth);
// Growth bounds: the stable one of the C ABI's default (threshold pivoting
// with |L21| <= 16), and the relaxed one of the preconditioners (block-local
// pivots unchecked: refinement on the FP64 system restores the accuracy, and
// the factors are redone with the stable bound when refinement does not
// converge or a diagonal block meets a zero pivot, BQD2_E_SINGULAR).
#define BQD2_LU32_STABLE_GROWTH 16.0f
#define BQD2_LU32_RELAXED_GROWTH 3.0e38f
int bqd2_lu32_convert(int n, const double *A, long long lda, const int *map, const double *dinv, float *Af,
                      hipStream_t st);
int bqd2_lu32_apply(int n, int nrhs, const float *Af, const void *ws, const double *in, long long ldi, double *out,
                    long long ldo, float *tmp, hipStream_t st, hipGraphExec_t *graph);
// gmres.hip: dinv[r] = 1 / A[r][r] (complex), 1 on rows with map[r] < 0
int bqd2_jacobi(int n, const double *A, long long lda, const int *map, double *dinv, hipStream_t st);
// the calling thread's context (nullptr on failure, error set)
CallCtx *bqd2_call_ctx(bqd2_problem *p);
// make the calling thread the problem's owner (its main context): for a caller
// that hands a problem from thread to thread and never uses it from two at once
void bqd2_rebind(bqd2_problem *p);
// device bytes held by the problem: allocations and every context's workspace
size_t bqd2_bytes(bqd2_problem *p);
// new incident plane wave given by amplitude, translation and rotation (solve.hip)
int bqd2_set_incidence_R(bqd2_problem *p, const double *pw_p, const double *tv, const double *R);
// at least `bytes` of device workspace at *base (previous contents are lost)
int bqd2_workspace(CallCtx *c, size_t bytes, char **base);

// device counter slots: errors, then per kernel family the summed Green's
// function terms by kind (spatial, propagating, evanescent, pure spectral) and
// the number of evaluations
// and, from CNT_FAR, the far field's points, their point-order products and
// the item-order products of its amplitudes
enum { CNT_ERR = 0, CNT_PAIRS = 1, CNT_SELF = 6, CNT_FIELD = 11, CNT_FAR = 16, CNT_N = 19 };

__device__ __forceinline__ void tally_add(long long *c, long long sp, long long pr, long long ev, long long fd,
                                          long long evals) {
  atomicAdd((unsigned long long *)&c[0], (unsigned long long)sp);
  atomicAdd((unsigned long long *)&c[1], (unsigned long long)pr);
  atomicAdd((unsigned long long *)&c[2], (unsigned long long)ev);
  atomicAdd((unsigned long long *)&c[3], (unsigned long long)fd);
  atomicAdd((unsigned long long *)&c[4], (unsigned long long)evals);
}
