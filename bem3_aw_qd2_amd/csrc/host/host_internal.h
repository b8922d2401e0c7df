/* host_internal.h -- helpers shared by the host C sources of libbem3_aw_qd2.so. */
#ifndef BQD2_HOST_INTERNAL_H_
#define BQD2_HOST_INTERNAL_H_
#include "bem3_aw_qd2.h"

/* element constants cr, cw and quadrature tables (model.c) */
void aqd2_element_constants(BOUD *bd);
void aqd2_quadrature_tables(BOUD *bd);
/* position and unnormalised normal at (z, e) of element s > 0 */
void aqd2_rw(const BOUD *bd, int s, double z, double e, double *r, double *w);
/* GPU-side state attached to an AQD2, kept in a side table keyed by its address
 * (AQD2 has no spare field: its layout is fixed by ex.dat).  solve.c */
struct bqd2_problem;
struct bqd2_problem *aqd2_problem(AQD2 *qd, int *has_solution);
void aqd2_push_solution(AQD2 *qd, struct bqd2_problem *p);
/* the device solution of every domain into the AQD2 arrays (exported for tools
 * that drive the device problem directly and then read the solution) */
void aqd2_pull_solution(AQD2 *qd, struct bqd2_problem *p);
/* batches and points of the coalesced single-point field calls (field.c) */
void aqd2_point_batches(long long *batches, long long *points);
void aqd2_mark_solution(AQD2 *qd, int on);
void aqd2_release_device(AQD2 *qd);

#endif
