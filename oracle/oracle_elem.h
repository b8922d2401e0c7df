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
#endif
