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

#ifdef __cplusplus
}
#endif

#endif /* BQD2_H_ */
