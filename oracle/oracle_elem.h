/* oracle_elem.h -- element-level oracle routines (elem_cpu.c), test infrastructure only. */
#ifndef BQD2_ORACLE_ELEM_H_
#define BQD2_ORACLE_ELEM_H_
#include <math.h>
#include "oracle.h"

double orc_bil_N(int n, double z, double e);
/* normalise a 3-vector in place (leaves the zero vector unchanged) */
static inline void orc_unit(double *v)
{
  double a = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  if (a != 0.0) { v[0] /= a; v[1] /= a; v[2] /= a; }
}
void orc_elem_copy(const BOUD *bd, int s, double cr[3][4], double cw[3][3]);
void orc_node_tangents(const BOUD *bd, int s, int nid, double *vtz, double *vte);
int orc_bil_is_plane(double cr[3][4], double cw[3][3]);
void orc_bil_dF_GL(const BOUD *bd, const double *rt, const double *v0, const double *v1, int s, double *dF0, double *dF1);
long orc_q0_bil_pair(const BOUD *bd, const double *rt, const double *v0, const double *v1, int s, int n,
                     const double *x, const double *w, double complex *CC, double complex *dC0, double complex *dC1);
void orc_bil_pair(const BOUD *bd, const double *rt, const double *v0, const double *v1, int s, double k, int n,
                  const double *x, const double *w, double complex *CC, double complex *dC0, double complex *dC1);
long orc_q0_bil_self(const BOUD *bd, int s, double zt, double et, double complex *CC);
void orc_q0_bil_self_dF(const BOUD *bd, int s, double zt, double et, double *dFz, double *dFe);
void orc_bil_self(const BOUD *bd, int s_signed, double zt, double et, double k, double complex *CC);
void orc_bil_self_dPV(const BOUD *bd, int s_signed, double zt, double et, double k, int dir, double complex *dCC);

/* linear triangles (lit_cpu.c) */
double orc_lit_N(int n, double z, double e);
void orc_lit_part(int part, double a, double b, double *z, double *e);
void orc_lit_corners(double zt, double et, double *r, double *th);
void orc_lit_dF_GL(const BOUD *bd, const double *rt, const double *v0, const double *v1, int s, double *dF0,
                   double *dF1);
long orc_q0_lit_pair(const BOUD *bd, const double *rt, const double *v0, const double *v1, int s, int n,
                     const double *x, const double *w, double complex *CC, double complex *dC0,
                     double complex *dC1);
void orc_lit_pair(const BOUD *bd, const double *rt, const double *v0, const double *v1, int s, double k, int n,
                  const double *x, const double *w, double complex *CC, double complex *dC0, double complex *dC1);
long orc_q0_lit_self(const BOUD *bd, int s, double zt, double et, double complex *CC);
void orc_lit_self(const BOUD *bd, int s_signed, double zt, double et, double k, double complex *CC);
void orc_lit_self_dPV(const BOUD *bd, int s_signed, double zt, double et, double k, int dir, double complex *dCC);

/* assembly level (assemble_cpu.c) */
long orc_coef_block(const AQD2 *qd, int did, int td, int tn, int sd, const double *vtz, const double *vte,
                    double complex *CC, double complex *dCz, double complex *dCe);
long orc_assemble_row(const AQD2 *qd, int did, int t, int tn, double complex *G, double complex *H,
                      double complex *dG, double complex *dH, double *dF);
long orc_assemble_domain(const AQD2 *qd, int did, double complex *G, double complex *H, double complex *dG,
                         double complex *dH, double *dF);
long orc_pair_sample(const AQD2 *qd, int did, int n, const int *row, const int *src);
int orc_threads(int n);
void orc_system_row(const AQD2 *qd, int did, const double complex *Grow, const double complex *Hrow,
                    double complex *A, double complex *b);
void orc_solution_vector(const AQD2 *qd, double complex *x);
void orc_store_solution(AQD2 *qd, const double complex *X);
void orc_inv33(double *ma, const double *vn, const double *vtz, const double *vte);
void orc_dpdt(const AQD2 *qd, int d, int t, int tn, const double complex *dGrow, const double complex *dHrow,
              const double *dF, double complex *dPz, double complex *dPe);
void orc_node_pv(const AQD2 *qd, int d, int t, int tn, double complex dPz, double complex dPe, double complex *pv);
/* scattered field at a point off the boundary (field_cpu.c) */
int orc_field_s(const AQD2 *qd, int did, const double *rt, int type, double complex *p, double complex *pv,
                double *gradF);
/* scattered field at a boundary point (field_cpu.c) */
int orc_field_bd(const AQD2 *qd, int did, int t, double zt, double et, int type, double complex *p,
                 double complex *pv);

/* point rules of field types 0 and 1, incident wave, boundary field and
 * radiation force (force_cpu.c) */
long orc_coef_pts(const AQD2 *ad, int did, const double *rt, int sd, int type, double complex *CC);
void orc_apw_pv(const Apw *pw, const double *r, double complex *p, double complex *v);
int orc_field_bd_total(const AQD2 *ad, int t, double zt, double et, int type, double complex *p,
                       double complex *pv);
int orc_force_elem(const AQD2 *ad, const double *rc, int type, int t, double *F, double *N);
int orc_force(const AQD2 *ad, const double *rc, int type, double *F, double *N);

#endif
