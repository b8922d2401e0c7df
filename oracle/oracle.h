/*
 * oracle.h -- CPU oracle of the bem3_aw_qd2 hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference's
 * algorithms, used by tests/ as the checker of the HIP implementation and by
 * bench.py as the timed CPU baseline.  Nothing in the product library links,
 * loads or calls it.
 *
 * Each function cites the reference file:line it restates.  Pins (what checks
 * each part against the reference or against mathematics) are listed in the
 * header comment of each source file.
 */
#ifndef BQD2_ORACLE_H_
#define BQD2_ORACLE_H_

#include <complex.h>
#include "bem3_aw_qd2.h"

/* Faddeeva w(z) = exp(-z^2) erfc(-iz) (faddeeva_cpu.c). */
double complex orc_faddeeva_w(double complex z);
/* exp(x^2 - y^2) * w(x + i y), the product the Ewald sums use. */
double complex orc_scaled_w(double x, double y);
void orc_faddeeva_w_batch(int n, const double *z, double *w);

/* The Ewald spectral sum's stopping rules where the offset is level with the
 * source: mode 0 (default) evaluates them as the reference does with this
 * libm; mode 1 with correctly rounded exponentials at z == 0 and exact
 * first-order differences at 0 < |z| < 2e-8 E (qpgf_cpu.c).  orc_exp_z0: the
 * exponential of the z == 0 residue in the current mode. */
void orc_set_cr_exp(int on);
int orc_cr_mode(void);
double orc_exp_z0(double x);

/* Quasi-periodic Green's function (qpgf_cpu.c).  mode: 0 value only
 * (d3hm_qpgf_d2_qG), 1 gradient only (_dqG), 2 value+gradient (_v1),
 * 3 value+gradient+Hessian (_v2).  Returns the number of lattice terms summed
 * (> 0) or the reference's negative error code. */
int orc_qpgf(int mode, const QPD2 *qd, const double *r, double eps,
             double complex *G, double complex *dG, double complex *d2G);

/* batched driver: r is n x 3 row-major, outputs G[n], dG[n*3], d2G[n*6] (may be
 * NULL for modes that do not produce them), ret[n]. OpenMP over points. */
void orc_qpgf_batch(int mode, const QPD2 *qd, int n, const double *r, double eps,
                    double complex *G, double complex *dG, double complex *d2G, int *ret);

#endif
