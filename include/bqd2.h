/*
 * bqd2.h -- C ABI of the HIP library libbqd2_hip.so (MI355X / gfx950).
 *
 * Plain pointers and sizes only.  Complex numbers are interleaved (re, im)
 * doubles.  Pointers named d_* are device pointers; every other pointer is host
 * memory.  Every entry point returns 0 on success or a negative BQD2_E* code;
 * bqd2_last_error() gives a message for the last failure on the calling thread.
 *
 * The host C library (libbem3_aw_qd2.so, the bem3_aw_qd2.h ABI) is the caller;
 * tests and bench.py call the same entry points through ctypes.
 */
#ifndef BQD2_H_
#define BQD2_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  BQD2_OK = 0,
  BQD2_E_ARG = -1,       /* invalid argument / shape                         */
  BQD2_E_HIP = -2,       /* HIP runtime error                                 */
  BQD2_E_GF = -3,        /* a Green's function evaluation returned an error   */
  BQD2_E_SINGULAR = -4,  /* zero pivot in the dense LU                        */
  BQD2_E_NOMEM = -5,     /* device allocation failed                          */
  BQD2_E_STATE = -6      /* call out of order (e.g. solve before factor)      */
};

const char *bqd2_last_error(void);

/* ------------------------------------------------------------------------
 * Quasi-periodic Green's function
 * ------------------------------------------------------------------------ */

/* Instance constants of the 2-D periodic Green's function, derived once from the
 * reference's QPD2 (qgf2_src/d3_qpgf_d2.h:32-36) by bqd2_gf_params(). */
typedef struct {
  double k, vkx, vky;          /* wave number, in-plane incident wave vector   */
  double d1x, d1y, d2x, d2y;   /* lattice vectors                              */
  double idd, aidd;            /* 1/det(d1,d2), |1/det|                         */
  double g1x, g1y, g2x, g2y;   /* reciprocal lattice vectors                   */
  double d2n, g2n, g12;        /* |d2|^2, |g2|^2, g1.g2                          */
  double E0;                   /* Ewald parameter before the per-point cap     */
  double bN, bNp, bNm;         /* spectral-decay probe at orders +-S_LIMIT     */
  int nc;                      /* centre row n1 of the reciprocal sums         */
  int ewald_only;              /* 1: probe orders propagate, Ewald everywhere  */
  double fd_lift;              /* the pure spectral sum is taken at heights    */
                               /* |z| / fd_lift that pass the reference's probe */
                               /* (1: the reference's choice; default 2.2)     */
} GfParams;

/* qd8 = {k, vk[0], vk[1], vk[2], vd1[0], vd1[1], vd2[0], vd2[1]} (QPD2 layout).
 * Replaces the per-call set-up at the top of d3hm_qpgf_d2_v1/_v2/_qG/_dqG
 * (qgf2_src/d3_qpgf_d2.c:10-230) and d3hm_qpgf_d2v1_ew_sd (:1671-1693). */
int bqd2_gf_params(const double *qd8, GfParams *p);

/* Batched Green's function at n offsets r (d_r: n x 3, source minus target).
 * mode 0: value (d3hm_qpgf_d2_qG, d3_qpgf_d2.h:125), 1: gradient (_dqG, :126),
 * 2: value+gradient (_v1, :112), 3: +Hessian (_v2, :124).
 * Outputs (device, may be NULL when the mode does not produce them):
 * d_G n x 2, d_dG n x 6 (x,y,z), d_d2G n x 12 (xx,yy,zz,xy,yz,zx), d_terms n:
 * lattice terms summed (> 0) or the reference's negative error code. */
int bqd2_gf_batch(const GfParams *p, int mode, int n, const double *d_r, double eps,
                  double *d_G, double *d_dG, double *d_d2G, int *d_terms, void *stream);

/* ------------------------------------------------------------------------
 * Boundary-value problem on the device
 * ------------------------------------------------------------------------ */

/* Host description of one model, filled by the host library from an AQD2
 * (include/bem3_aw_qd2.h).  Element arrays are indexed by element id 1..Ne
 * (row 0 unused), as in AQD2; complex values are (re, im) pairs. */
typedef struct {
  int Ne, Nn, NN, MN, ps;
  const double *rn;      /* (Nn+1) x 3  : mesh nodes                      */
  const double *cr;      /* (Ne+1) x 12 : cr[s][d][j] at 4d+j            */
  const double *cw;      /* (Ne+1) x 9  : cw[s][d][j] at 3d+j            */
  const double *ren;     /* (Ne+1) x 12 : node n, coordinate d at 3n+d   */
  const double *wen;     /* (Ne+1) x 12                                  */
  const int *ed;         /* (Ne+1) x 4  : corner node ids (ed[s][3]==0: triangle) */
  const int *eni;        /* (Ne+1) x 4  : collocation node index          */
  const int *md, *sd;    /* (Ne+1)      : main / sub domain               */
  const double *Pi, *dPi;/* (Ne+1) x 4 complex: incident potential at nodes */
  const int *dom_ne;     /* MN+1        : elements per domain             */
  const int *dom_sid;    /* sum(dom_ne)+MN+1: per domain 1-based signed ids, entry 0 unused */
  const double *k0;      /* MN+1 real wave numbers                        */
  const double *k2;      /* MN+1 complex i omega rho                      */
  const double *rho0;    /* MN+1 densities                                */
  double qd[8];          /* QPD2 of the host medium                       */
  double xli[12], wli[12], xhi[64], whi[64];
  double zt_44[4], et_44[4], wt_44[4], zt_34[4], et_34[4], wt_34[4];
  double zt_49[9], et_49[9], wt_49[9], zt_37[7], et_37[7], wt_37[7];
  /* incident plane wave: amplitude (re, im), rho0, c0, k0, translation, rotation */
  double pw_p[2], pw_rho0, pw_c0, pw_k0, pw_tv[3], pw_R[9];
} bqd2_model;

typedef struct bqd2_problem bqd2_problem;

/* counters of the last assembly of one domain */
typedef struct {
  long long pairs_gl, pairs_gh, self_rows;  /* off-diagonal pairs per rule, self blocks */
  long long gf_evals;                       /* distinct Green's function evaluations    */
  long long gf_terms;                       /* lattice terms summed over them           */
  double ms_decide, ms_pairs, ms_self;      /* kernel times (HIP events on the problem's stream):
                                               rule decision, off-diagonal pairs, self blocks + row sums */
  /* periodic domain only: terms by kind (spatial images, propagating and evanescent
   * spectral orders, pure spectral orders) and evaluations, for the off-diagonal
   * pair kernel and for the self-element kernel separately */
  long long pair_terms[4], pair_evals;
  long long self_terms[4], self_evals;
  /* GH64 pairs of the periodic domain evaluated as the free-space kernel on
   * GH64 plus the smooth remainder of the lattice sum on a Gauss rule */
  long long pairs_split;
  /* pairs of the periodic domain whose lattice remainder takes a 4-, 6- or
   * 8-point Gauss rule (GL12 pairs and split GH64 pairs), and the free-space
   * kernel points summed by all split pairs */
  long long pairs_smooth, free_points;
} bqd2_assembly_stats;

/* The n-point Gauss-Legendre rule on [-1, 1] (nodes ascending), 1 <= n <= 64:
 * the rules of the smooth lattice remainder.  Host only, no device needed.  No
 * counterpart in the reference, whose rules are GL12 and GH64 only. */
int bqd2_gauss_legendre(int n, double *x, double *w);

int bqd2_create(const bqd2_model *m, bqd2_problem **out);
int bqd2_destroy(bqd2_problem *p);

/* Coefficient matrices of domain did on the device (create_cmatrix_domain,
 * d3qd2_src/bem3_aw_qd2_solve_bieq.c:133-218, cm.type = 2):
 *   G, H [N x N], dG, dH [2N x N] (interior and isolated domains only), dF [N x 3],
 * N = 4 Ne_d, row (t-1)*4+tn, column (s-1)*4+n. st may be NULL. */
int bqd2_assemble(bqd2_problem *p, int did, bqd2_assembly_stats *st);

/* No counterpart in the reference.  In the periodic domain the lattice sum of a
 * pair is split into the free-space kernel, integrated on the reference's rule
 * (GL12, or GH64 for near pairs), and the remainder, which is analytic over the
 * element up to the other lattice images of the target and is integrated on a
 * Gauss rule of 4, 6 or 8 points a side chosen per pair from the distance to
 * those images (error bound below 1e-14 of the remainder); pairs with a point
 * level with the target keep the reference's evaluation.  on = 0 keeps the
 * lattice sum on every GL12 point (and GL12 for the remainder of split GH64
 * pairs); on = 4, 6 or 8 takes that order wherever a remainder rule is allowed
 * (convergence checks); the rules are decided again at the next assembly. */
int bqd2_set_remainder_rules(bqd2_problem *p, int on);

/* No counterpart in the reference.  bqd2_assemble of domain did for `count`
 * problems, with the off-diagonal pairs of their periodic domains computed in
 * one persistent launch (uneven instances share the GPU; each pair is computed
 * exactly as bqd2_assemble computes it).  st: count entries or NULL; the pair
 * times are those of the whole batch. */
int bqd2_assemble_batch(bqd2_problem **ps, int count, int did, bqd2_assembly_stats *st);

/* Device counters of the last assembly, field or force call: [0] failed Green's
 * function evaluations; then for the off-diagonal pair kernel [1..5], the
 * self-element kernel [6..10] and the field kernels [11..15]: terms summed by
 * kind (spatial, propagating spectral, evanescent spectral, pure spectral) and
 * evaluations.  n <= 16 values are copied. */
int bqd2_counters(bqd2_problem *p, long long *out, int n);

/* copy domain did's matrices to host buffers (any may be NULL) */
int bqd2_download_domain(bqd2_problem *p, int did, double *G, double *H, double *dG, double *dH, double *dF);
/* the same for nrows rows (indices (t-1)*4 + tn): G, H nrows x N complex, dG,
 * dH nrows x 2N (the row's two tangential derivatives), dF nrows x 3 */
int bqd2_download_rows(bqd2_problem *p, int did, int nrows, const int *rows, double *G, double *H, double *dG,
                       double *dH, double *dF);

/* Build the dense 2NN system (create_matrix_csr_dac, :244-345) and solve it; the
 * reference uses MKL PARDISO (d3qd2_mkl_solver.c:18-148).  Default: GMRES with
 * Jacobi row scaling to a scaled residual of 1e-15, and the dense LU with partial
 * pivoting when that is not reached; bqd2_set_solver(p, 1) selects the LU alone.
 * Scatters the solution to the per-domain P, dP (tmatrix_bd_store, :370-427).
 * x (host, 2NN complex) may be NULL. */
int bqd2_solve_surface(bqd2_problem *p, double *x);

/* Particle velocity and its normal derivative on the boundary: q_solve_pv_bv /
 * solve_pv_bv (:429-491, :655-750) and solve_dpv_bv (:533-653; LAPACKE_zgesv
 * there, GMRES with the LU fallback here, as for the system). */
int bqd2_solve_velocity(bqd2_problem *p);

/* No counterpart in the reference.  Solver of bqd2_solve_surface and
 * bqd2_solve_velocity: 0 GMRES with the LU fallback (default), 1 LU only,
 * 2 mixed precision: the row-scaled system factored once per assembly in
 * complex float on the MFMA units (bqd2_cgetrf) and used as the right
 * preconditioner of GMRES on the FP64 system, which refines the solution to the
 * same scaled residual as mode 0 (the dpv systems run as in mode 0). */
int bqd2_set_solver(bqd2_problem *p, int mode);

/* No counterpart in the reference.  Krylov steps and final scaled residuals of
 * the last solves: entry 0 the system, entry 1 + d the dpv system of domain d
 * (MN + 2 entries, 0 where the LU ran); lu: systems that fell back to the LU. */
int bqd2_solver_info(bqd2_problem *p, int *iters, double *relres, int *lu);

/* No counterpart in the reference (one plane wave per model there).  A new
 * incident plane wave on an assembled problem, pw7 = {Re p, Im p, theta, psi,
 * tx, ty, tz} as in pw.txt (aw_pw.c:3-77).  Its in-plane wave vector must equal
 * the problem's modulo the reciprocal lattice (the mirrored direction pi - theta,
 * or another propagating diffraction order), since the periodic operator depends
 * on it (BQD2_E_ARG otherwise).  The assembled matrices and the mixed-precision
 * factors are kept: the next bqd2_solve_surface only forms and solves the new
 * right-hand side. */
int bqd2_set_incidence(bqd2_problem *p, const double *pw7);

/* per-domain solution arrays, (Ne_d+1) x 4 complex for P, dP and
 * (Ne_d+1) x 4 x 3 complex for pv, dpv, entry 0 unused, as in SUBD */
int bqd2_download_solution(bqd2_problem *p, int did, double *P, double *dP, double *pv, double *dpv);
int bqd2_upload_solution(bqd2_problem *p, int did, const double *P, const double *dP, const double *pv,
                         const double *dpv);

/* Pressure and particle velocity at n points (pts host n x 3): the batched form
 * of pv_t_aqd2 (total = 1) / pv_s_aqd2 (total = 0), bem3_aw_qd2_field.c:103-297,
 * with rule type 0: 4-point, 1: 9-point, 2: GLN^2, 3: GHN^2.  out (host n x 8):
 * p (re, im), pv x3 (re, im); did (host n, may be NULL): the domain id the
 * reference returns.  gf_evals (may be NULL): Green's function evaluations. */
int bqd2_field(bqd2_problem *p, int n, const double *pts, int type, int total, double *out, int *did,
               long long *gf_evals);

/* Field points of the periodic domain and the boundary elements at least h
 * below or above them (h 1/12 to 1/4 of the shorter lattice vector): between them
 * the lattice sum is the pure spectral series of d3_qpgf_d2.c:484-533, whose
 * terms factor into a source and a target part, so bqd2_field sums the boundary
 * once per order into plane-wave amplitudes of 32 bands of elements by height
 * (kept on the device until the solution changes) and each point over the
 * orders for the bands that clear h, the other bands by the lattice sums.  No
 * counterpart in the reference, which sums every (point, source) pair.  on = 0
 * takes the per-pair sums at every point (for comparisons);
 * bqd2_field_far_info reports whether the problem uses the amplitudes, the
 * number of orders, the boundary's z extent and h. */
int bqd2_set_field_far(bqd2_problem *p, int on);
int bqd2_field_far_info(bqd2_problem *p, int *on, int *orders, double *zbot, double *ztop, double *h);
/* No counterpart in the reference (measurement).  With timing on, bqd2_field
 * records HIP events around its point kernel on the call's stream;
 * bqd2_field_kernel_ms gives that kernel's time in the last timed call. */
int bqd2_set_field_timing(bqd2_problem *p, int on);
int bqd2_field_kernel_ms(bqd2_problem *p, double *ms);

/* bqd2_field for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  It calls no hipMalloc, hipFree or allocating
 * library routine and does not synchronise: the kernels of bqd2_field run on
 * `stream`, with their partial sums in a fixed device array of
 * BQD2_NOALLOC_SCRATCH_BYTES per GPU (set up at the first call on a device and
 * held until the process ends).  Inputs and outputs are the caller's device
 * buffers: d_pts (n x 3), d_out (n x 8 as in bqd2_field), d_did (n), d_errors
 * (may be NULL: receives the number of points whose Green's function
 * evaluation failed).  Results are the bits bqd2_field returns.  A call needs
 * at most 1024 + n (32 + 144 c) + 4096 G bytes, c <= 64 the element chunks per
 * point of the problem and rule, G <= 32768 the far field's orders (0 where it
 * is not used; bqd2_field_far_info): the far field's amplitudes are made in
 * every call (256 MiB: about 14 800 points at c = 64 and G = 32041); a larger n is
 * rejected with BQD2_E_ARG before anything runs.  Every
 * call on one GPU shares that GPU's array: calls that overlap on one GPU (two
 * streams, two threads, two problems) race and must be ordered by the caller. */
#define BQD2_NOALLOC_SCRATCH_BYTES (256u << 20)
int bqd2_field_noalloc(bqd2_problem *p, int n, const double *d_pts, int type, int total, double *d_out,
                       int *d_did, long long *d_errors, void *stream);

/* Domain of each point (domain_id / domain_id_aqd2, bem3_aw_qd2.c:111-156, :1148-1194):
 * solid-angle classification of the point reduced to the unit cell. */
int bqd2_domain_id(bqd2_problem *p, int n, const double *pts, int *did);

/* Same at boundary points (zeta, eta) of elements t (1-based position in
 * domain did's list): pv_s_bd_aqd2 / pv_t_bd_aqd2 (:394-469). ze host n x 2. */
int bqd2_field_bd(bqd2_problem *p, int did, int n, const int *t, const double *ze, int type, int total, double *out);

/* Radiation force F and torque N about rc on the open-domain boundary:
 * force_FN_aqd2 (bem3_aw_qd2_force.c:4-42); type 0 node values, otherwise the
 * 9-point rule with boundary field evaluations. */
int bqd2_force(bqd2_problem *p, const double *rc, int type, double *F, double *N);

/* dense complex LU helpers on device buffers (column-major, leading dimension n),
 * used by the solves above and exposed for tests: A is overwritten by its
 * factors, ipiv (device, n ints) receives the pivots; B (n x nrhs) by X. */
int bqd2_zgetrf(int n, double *d_A, int *d_ipiv, void *stream);
int bqd2_zgetrs(int n, int nrhs, const double *d_A, const int *d_ipiv, double *d_B, void *stream);

/* Complex float LU with partial pivoting on the MFMA units, the factorisation of
 * the mixed-precision solver (bqd2_set_solver(p, 2)); replaces the same MKL
 * solves as bqd2_zgetrf.  Column-major, leading dimension n, (re, im) floats.
 * d_ws: device workspace of bqd2_cgetrf_wsize(n) bytes that keeps the inverted
 * diagonal blocks and the row permutation for bqd2_cgetrs.  d_A receives the
 * factors in the form the solves use: the 128 x 128 diagonal blocks hold L and
 * U, every off-diagonal block is multiplied by the inverse of its column's
 * diagonal block.  bqd2_cgetrs solves in place on d_B (n x nrhs); d_tmp holds
 * n x nrhs complex floats. */
size_t bqd2_cgetrf_wsize(int n);
int bqd2_cgetrf(int n, float *d_A, int *d_ipiv, void *d_ws, void *stream);
/* Pivoting of bqd2_cgetrf's panels (process-wide; no counterpart in the
 * reference): 1 (default) restricts a panel's pivots to the rows of its
 * 128 x 128 diagonal block and keeps that when max |L21| is within the growth
 * bound, else redoes the panel with pivots over all its rows; 0 always pivots
 * over all rows.  Counts of panels factored each way since the process
 * started: out[0] block-local, out[1] all rows. */
int bqd2_set_lu32_pivoting(int mode);
int bqd2_lu32_panel_counts(long long *out);
/* No counterpart in the reference.  The growth bound of block-local panels
 * (max |L21|, default 16, at least 1; process-wide; 1e30 or more: unchecked,
 * the form the mixed-precision solver uses, which reports a zero pivot inside
 * a diagonal block as BQD2_E_SINGULAR), and the growth of each checked panel
 * of the most recent checked factorisation (inf for a zero pivot in the
 * block): up to cap values into out, returns their number. */
int bqd2_set_lu32_growth_bound(double bound);
int bqd2_lu32_panel_growth(float *out, int cap);
int bqd2_cgetrs(int n, int nrhs, const float *d_A, const void *d_ws, float *d_B, float *d_tmp, void *stream);
/* No counterpart in the reference.  Whether the unchecked factorisation of the
 * mixed-precision solver runs each next panel on a stream of its own restricted
 * to 8 CUs beside the trailing update (1, default) or everything on the
 * caller's stream (0); process-wide. */
int bqd2_set_lu32_lookahead(int on);
/* The trailing-update kernel of bqd2_cgetrf, exposed for tests: C -= A B
 * (set = 0, three real products per complex product, as the trailing update)
 * or C = A B (set = 1, four real products; B may be C when m <= 128, A may be
 * C when nc <= 128), column-major complex float, A m x kk, B kk x nc,
 * C m x nc. */
int bqd2_cgemm(int set, int m, int nc, int kk, const float *d_A, long long lda, const float *d_B, long long ldb,
               float *d_C, long long ldc, void *stream);
/* No counterpart in the reference.  The mixed-precision solves replay their
 * triangular sweeps as HIP graphs (built once per factorisation); 0 launches
 * them one by one (process-wide; for tests that compare the two). */
int bqd2_set_graphs(int on);

/* No counterpart in the reference (the batching form's autotuner).  Two
 * launches have several tile or workgroup sizes that give the same bits:
 *   site 0  trailing update of bqd2_cgetrf: 64 x 64 (0) or 128 x 128 (1) tiles;
 *   site 1  GMRES matrix-vector product: 4 (0), 2 (1), 8 (2) or 1 (3) rows per workgroup.
 * With mode 1 (the default) the library times every candidate on its own calls
 * of each size class (half octaves of the dimension; for site 1 also the number
 * of right-hand sides), three times each, as a rate (the call's work per
 * millisecond: flops for site 0, bytes of the matrix for site 1), and then
 * keeps the fastest for that class and device; mode 0 uses candidate 0
 * everywhere.  bqd2_autotune_force pins a site to one candidate (-1 releases
 * it).  bqd2_autotune_info returns the number of classes of a site on the
 * current device and fills up to max of them: key (the class), chosen (-1
 * while exploring) and best_rate[4 max] (highest rate per candidate, -1 where
 * none yet); wait != 0 first waits for timings in flight.
 * bqd2_autotune_reset forgets every class. */
int bqd2_set_autotune(int mode);
int bqd2_autotune_force(int site, int cand);
int bqd2_autotune_info(int site, int wait, int max, long long *keys, int *chosen, float *best_rate, int *ncand);
int bqd2_autotune_reset(void);

/* No counterpart in the reference.  Milliseconds of the last complex float
 * factorisation of the system (row scaling, rounding and bqd2_cgetrf), 0 when
 * none has run since the last assembly. */
int bqd2_lu32_info(bqd2_problem *p, double *ms_factor);

/* ------------------------------------------------------------------------
 * The system layout (no counterpart in the reference, which writes every
 * domain matrix to temporary files and reads it back to build the system,
 * bem3_aw_qd2_solve_bieq.c:55-86, :202-217, then hands it to PARDISO)
 * ------------------------------------------------------------------------ */

/* A problem whose matrices do not fit one GPU in the default layout: one sphere
 * (MN = 1) of quadrilaterals, each with its own four nodes, in a periodic lattice
 * (BQD2_E_ARG otherwise).  Both domains' matrices live in one row-major
 * 2N x 2N complex double array in the system's own column order, [H0 G0; H1 G1];
 * the sphere's rows are assembled in windows whenever they are needed, and dG,
 * dH of the sphere never exist whole.  The calls are those of the default
 * layout (bqd2_assemble of both domains, bqd2_solve_surface, bqd2_solve_velocity,
 * bqd2_field, bqd2_force) with these differences:
 *   bqd2_assemble(p, 1) only decides the sphere's quadrature rules;
 *   solver 2 keeps its complex float factors in the memory of the sphere's half
 *     of the array, and the sphere's rows in window buffers as far as device
 *     memory allows (the others are assembled again in every product of the
 *     refinement);
 *   solver 1 (FP64 LU) factors the array in place: the next bqd2_solve_velocity
 *     assembles domain 0 again;
 *   the dpv systems of bqd2_solve_velocity use FP64 GMRES whatever the solver. */
int bqd2_create_large(const bqd2_model *m, bqd2_problem **out);
/* rows per window of the sphere's rows (default 2048) and the most windows the
 * mixed solver keeps resident (-1: as many as device memory allows, keeping
 * 6 GiB free); takes effect at the next bqd2_assemble of domain 1 */
int bqd2_large_config(bqd2_problem *p, int rows_per_window, int max_resident);
typedef struct {
  int windows;             /* windows of the sphere's rows */
  int resident;            /* of them kept resident by the last mixed solve */
  int recomputed;          /* window assemblies inside the last solve's products */
  int pad_;
  double ms_rows;          /* last solve: assembling the sphere's rows before the solve proper */
  double ms_build;         /* ... the row scaling, right-hand side and float rounding of the system */
  double ms_factor;        /* ... the factorisation (complex float, or FP64 LU with its transposition) */
  double ms_solve;         /* ... GMRES (refinement or FP64) or the FP64 triangular solves */
  double ms_velocity_rows; /* last velocity recovery: the sphere's rows with dG, dH and the velocity rows */
  double ms_reassemble;    /* ... assembling domain 0 again after an FP64 LU */
} bqd2_large_stats;
int bqd2_large_info(bqd2_problem *p, bqd2_large_stats *st);

/* No counterpart in the reference.  A process-wide progress callback for calls
 * that run for minutes at large sizes: fn(stage, done, total, ctx) is called on
 * the calling thread as parts of the periodic pair assembly, of the FP64 LU and
 * of the system layout's window passes complete.  NULL (the default): none, and
 * the work is launched without the waits that the reports need. */
typedef void (*bqd2_progress_fn)(const char *stage, long long done, long long total, void *ctx);
int bqd2_set_progress(bqd2_progress_fn fn, void *ctx);

/* ------------------------------------------------------------------------
 * Serving (no counterpart in the reference, which solves one model per process:
 * aw_d3qd2_bv_solver.c:5-11, then example1.c:10-48 for fields and forces)
 * ------------------------------------------------------------------------ */

/* Device bytes that one request's problem may hold: matrices, system, solver
 * factors and workspaces, field and force workspaces (an upper estimate; the
 * server admits requests by it). */
size_t bqd2_footprint(const bqd2_model *m, int solver, int npts, int force_type);
/* device bytes a problem holds now (allocations and workspaces) */
int bqd2_device_bytes(bqd2_problem *p, size_t *bytes);

typedef struct bqd2_server bqd2_server;

/* One request: the whole solve of a model (bqd2_create, bqd2_assemble of every
 * domain, bqd2_solve_surface, bqd2_solve_velocity), then optionally the total
 * field at npts points (rule type 0) and the radiation force about rc.  The
 * model and the points are copied at submission.  op_key != 0 promises that
 * requests with the same key share mesh, media, frequency, lattice and in-plane
 * wave vector modulo the reciprocal lattice (the incidence may differ): the
 * operator of the first stays resident and the others reuse it (checked). */
typedef struct {
  const bqd2_model *model;
  unsigned long long op_key;
  int solver;          /* bqd2_set_solver mode: 0 FP64 GMRES, 1 FP64 LU, 2 mixed precision */
  int npts;            /* field points (host npts x 3), 0 for none */
  const double *pts;
  int force_type;      /* -1 no force, else the type of bqd2_force */
  double rc[3];        /* torque centre */
} bqd2_request;

typedef struct {
  int status;          /* 0 or the BQD2_E code of the failure (error holds the message) */
  int reused;          /* 1: the operator was resident (no assembly, no factorisation) */
  int krylov_steps;    /* GMRES steps of the surface solve */
  int pad_;
  double relres;       /* its final scaled residual */
  double F[3], N[3];   /* force and torque (force_type >= 0) */
  double t_queue;      /* seconds from submission to the start of execution */
  double t_run;        /* seconds of execution */
  double t_done;       /* completion time on the server's clock */
  size_t device_bytes; /* device bytes held by the request's problem at completion */
  char error[256];
} bqd2_result;

/* A server with `slots` concurrent requests (worker threads, each request on its
 * problem's own stream) and a device-memory budget of hbm_bytes for resident
 * and running problems. */
int bqd2_server_create(size_t hbm_bytes, int slots, bqd2_server **out);
/* Queue a request; field_out (host npts x 8: p, v as in bqd2_field) must stay
 * valid until bqd2_server_wait returns for the ticket. */
int bqd2_server_submit(bqd2_server *s, const bqd2_request *rq, double *field_out, long long *ticket);
/* Block until the request is done; returns its status. */
int bqd2_server_wait(bqd2_server *s, long long ticket, bqd2_result *res);
/* counts: hits, misses, evictions, failed, resident operators, pending requests;
 * bytes: accounted device bytes in use, budget (either may be NULL) */
int bqd2_server_stats(bqd2_server *s, long long *counts, size_t *bytes);
/* Finish the queued requests, release every resident problem. */
int bqd2_server_destroy(bqd2_server *s);

#ifdef __cplusplus
}
#endif

#endif /* BQD2_H_ */
