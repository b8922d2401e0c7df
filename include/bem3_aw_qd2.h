/*
 * bem3_aw_qd2.h -- public C ABI of the MI355X implementation of the periodic
 * acoustic multi-domain BEM (bilinear quadrilateral / linear triangular
 * elements, discontinuous collocation nodes, 2-D quasi-periodic lattice).
 *
 * Drop-in replacement for the reference's d3qd2_src/bem3_aw_qd2.h:13-96 and
 * the types it pulls in from d3qd2_elem.h:13-74, aw_pw.h:16-34,
 * d3_qpgf_d2.h:32-36, my_utils.h and gauleg.h.  The struct layouts below are
 * byte-identical to the reference (sizeof(AQD2) == 2720 on LP64, double complex
 * is 16 bytes with 8-byte alignment), because ex.dat starts with a raw image of
 * AQD2 and the example drivers allocate AQD2 on their stack.
 *
 * Every array reachable from AQD2 is 1-based (index 0 unused) and allocated as
 * pointer-to-pointer, exactly as the reference lays them out, so callers that
 * index qd->bd.sb[d].P[s][n] keep working.
 */
#ifndef BEM3_AW_QD2_H_
#define BEM3_AW_QD2_H_

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <complex.h>
#include <time.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- numerical settings (reference d3qd2_const.h:31-92, d3_qpgf_d2.h:19-29) ----
 * These fix the discretisation; changing any of them changes matrix entries. */
#define MEPS    1.0e-15  /* relative accuracy requested from the lattice sums   */
#define GEPS    1.0e-15  /* geometric tolerance (plane-element test)            */
#define DEPS    0.1      /* radius of the principal-value disc D_eps            */
#define CDFH    1.0e-3   /* central-difference step for dF/dv on the self element */
#define GLN     12       /* Gauss-Legendre order of the regular element rule    */
#define GHN     64       /* Gauss-Legendre order of the fallback element rule   */
#define CBD_CDF 10.0     /* |grad F| above this: field point is near the boundary */
#define CBD_SDF 0.5      /* |dF/dv| above this: switch the element rule GLN->GHN */
#define OPENDID 99       /* gmsh physical id of the open (lattice) domain        */
#define ELT3    2        /* gmsh element type: 3-node triangle                  */
#define ELT4    3        /* gmsh element type: 4-node quadrangle                */

#define S_LIMIT 15       /* reciprocal-order probe used to pick Ewald vs spectral */
#define FD_LIMIT 100     /* row/column limit of the pure spectral sum            */
#define EW_LIMIT 100     /* row/column limit of the Ewald sums                   */
#define EW_EXPM 3.0      /* cap on k^2 E^2 - r^2/(4E^2) when choosing E          */
#define EFS     0.1      /* per-term stop factor: |term| <= EFS*eps*|sum|        */

/* ---- quasi-periodic Green's function parameters (d3_qpgf_d2.h:32-36) ---- */
typedef struct qp_data_v2 {
  double k;             /* wave number of the periodic (host) medium            */
  double vk[3];         /* incident wave vector; vk[0..1] set the Bloch phase   */
  double vd1[2], vd2[2];/* lattice vectors in the xy-plane                      */
} QPD2;

/* ---- incident plane wave (aw_pw.h:16-34) ---- */
typedef struct aw_planewave_adata {
  double omega;          /* angular frequency                                   */
  double K0;             /* bulk modulus rho0*c0^2                              */
  double complex k1, k2; /* i*omega/K0, i*omega*rho0                            */
  double k0, kx, ky, kz; /* wave number and its components                      */
  double lambda0;        /* wavelength                                          */
  double R[9];           /* rotation of the propagation axis                    */
} ApwD;

typedef struct aw_planewave {
  double rho0, c0, f;    /* host density, sound speed, frequency                */
  double complex p;      /* complex amplitude                                   */
  double tv[3];          /* translation of the phase origin                     */
  double theta, psi;     /* propagation direction parameters                    */
  ApwD wd;
} Apw;

/* ---- per-domain boundary data (d3qd2_elem.h:13-19) ---- */
typedef struct sub_domain_data {
  int Ne;                    /* elements bounding this domain                   */
  int *sid;                  /* signed element ids (+: main side, -: sub side)  */
  double complex **P, **dP;  /* potential and its normal derivative, [s][node]  */
  double complex ***pv, ***dpv; /* particle velocity and its normal derivative, [s][node][xyz] */
} SUBD;

/* ---- mesh, element constants and quadrature tables (d3qd2_elem.h:21-46) ---- */
typedef struct boundary_data {
  int ps;                    /* 1: quasi-periodic, 0: isolated                  */
  QPD2 qd;
  int Nn;                    /* mesh nodes                                      */
  int Ne;                    /* surface elements (each listed twice in the msh) */
  double **rn;               /* rn[node][xyz]                                   */
  int **ed;                  /* ed[element][corner] = node id (corner 3 == 0 for triangles) */
  int NN;                    /* collocation nodes                               */
  int **eni;                 /* eni[element][corner] = collocation node index   */
  int *md, *sd, *gd;         /* main-side domain, sub-side domain, gmsh entity  */
  double ***cr, ***cw;       /* r(z,e) = cr.(1,z,e,ze); w(z,e) = cw.(1,z,e)     */
  double ***ren, ***wen;     /* collocation node position / unnormalised normal */
  double complex **Pi, **dPi;/* incident potential and normal derivative at nodes */
  double zt_44[4], et_44[4], wt_44[4], zt_49[9], et_49[9], wt_49[9];
  double zt_34[4], et_34[4], wt_34[4], zt_37[7], et_37[7], wt_37[7];
  double xli[GLN], wli[GLN];
  double xhi[GHN], whi[GHN];
  SUBD *sb;                  /* sb[0..MN]                                       */
} BOUD;

typedef struct domain_data {
  char ped_fn[128], med_fn[128], msh_fn[128];
  double rv[3], th;          /* Rodrigues rotation of the object                */
  double tv[3];              /* translation of the object                       */
  int MN;                    /* media besides the host                          */
  double *rho0, *c0;         /* [0..MN]                                         */
  double *k0, *K0;
  double complex *k1, *k2;
  Apw pw;
  BOUD bd;
} AQD2;

/* ---- model, IO (bem3_aw_qd2.h:13-20) ---- */
void read_aqd2(int argc, char **argv, AQD2 *qd);
void print_aqd2(AQD2 *qd);
void initialize_aqd2(AQD2 *qd);
void finalize_aqd2(AQD2 *qd);
int  domain_id_aqd2(double *rt, AQD2 *qd);
void domain_id_l_aqd2(int *l1, int *l2, double *rt, AQD2 *qd);
void dat_write_aqd2(char *fname, AQD2 *qd);
void dat_read_aqd2(char *fname, AQD2 *qd);

/* ---- boundary-value solve (bem3_aw_qd2.h:24) ---- */
void solve_bieq_aqd2(AQD2 *qd);

/* ---- radiation force and torque, surface area (bem3_aw_qd2.h:28-44) ---- */
void force_FN_aqd2(double *F, double *N, double *rc, int type, AQD2 *qd);
void surface_area_aqd2(double *S, int did, int type, AQD2 *qd);

/* ---- fields at a point; return the domain id (bem3_aw_qd2.h:49-96) ---- */
int phi_s_aqd2(double complex *phi, double *rt, int type, AQD2 *qd);
int phi_t_aqd2(double complex *phi, double *rt, int type, AQD2 *qd);
int phi_i_aqd2(double complex *phi, double *rt, int type, AQD2 *qd);
int p_s_aqd2(double complex *p, double *rt, int type, AQD2 *qd);
int p_t_aqd2(double complex *p, double *rt, int type, AQD2 *qd);
int p_i_aqd2(double complex *p, double *rt, int type, AQD2 *qd);
int pv_s_aqd2(double complex *p, double complex *pv, double *rt, int type, AQD2 *qd);
int pv_t_aqd2(double complex *p, double complex *pv, double *rt, int type, AQD2 *qd);
int pv_i_aqd2(double complex *p, double complex *pv, double *rt, int type, AQD2 *qd);

/* ---- fields at a boundary point (zeta_t, eta_t) of element t of domain did ---- */
void phi_s_bd_aqd2(double complex *phi, int did, int t, double zeta_t, double eta_t, int type, AQD2 *qd);
void phi_t_bd_aqd2(double complex *phi, int did, int t, double zeta_t, double eta_t, int type, AQD2 *qd);
void phi_i_bd_aqd2(double complex *phi, int did, int t, double zeta_t, double eta_t, int type, AQD2 *qd);
void pv_s_bd_aqd2(double complex *p, double complex *pv, int did, int t, double zeta_t, double eta_t, int type, AQD2 *qd);
void pv_t_bd_aqd2(double complex *p, double complex *pv, int did, int t, double zeta_t, double eta_t, int type, AQD2 *qd);
void pv_i_bd_aqd2(double complex *p, double complex *pv, int did, int t, double zeta_t, double eta_t, int type, AQD2 *qd);

/* ---- batched field evaluation (no counterpart in the reference: the reference
 * callers loop over pv_t_aqd2 point by point, example2.c:28-31).  Evaluates
 * pv_t_aqd2 (total = 1) or pv_s_aqd2 (total = 0) at n points in one GPU pass;
 * did[i] receives what pv_*_aqd2 would return for point i.  rt is n x 3. ---- */
void pv_batch_aqd2(int n, const double *rt, int type, int total,
                   double complex *p, double complex *pv, int *did, AQD2 *qd);

/* ---- device description of an initialised model (no counterpart in the
 * reference): the bqd2_model (include/bqd2.h) of qd in newly allocated packed
 * arrays, for bqd2_server_submit and bqd2_footprint; release it with
 * aqd2_model_free.  Returned as void * so that this header needs no bqd2.h. ---- */
void *aqd2_model_new(AQD2 *qd);
void aqd2_model_free(void *model);

/* ---- plane wave (aw_pw.h:44-58) ---- */
void read_data_apw(char *fname, Apw *pw);
void print_data_apw(Apw *pw);
void setup_apw(Apw *pw);
void calc_apw_pv(double complex *p, double complex *v, double *r, Apw *pw);
void calc_apw_dpdn(double complex *p, double complex *dpdn, double *r, double *n, Apw *pw);

/* ---- small utilities the reference exports to its callers (my_utils.h, gauleg.h) ---- */
void vadd_d(double *r, double *v0, double *v1);
void vsub_d(double *r, double *v0, double *v1);
int  vuni_d(double *r);
double vabs_d(double *v);
double vabs_2dm(double *v0, double *v1);
double vdot_d(double *v0, double *v1);
void vcrs_d(double *r, double *v0, double *v1);
void prt_z_en(double complex z);
void prt_z_fen(char *txt, int n, double complex z);
void *m_alloc2(size_t num, size_t size, char *txt);
void continue_message(void);
void gauleg(double x1, double x2, double x[], double w[], int n);

/* ---- layout probe (no counterpart in the reference) ----
 * sizeof(AQD2), offsetof(AQD2, bd), offsetof(BOUD, sb), sizeof(Apw),
 * sizeof(QPD2), offsetof(AQD2, MN) as compiled into the library; returns the
 * number of values written (at most n, at most 6).  For binding checks. */
int aqd2_layout(long *out, int n);

#ifdef __cplusplus
}
#endif

#endif /* BEM3_AW_QD2_H_ */
