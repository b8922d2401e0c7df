/*
 * field.c -- the reference's field API (bem3_aw_qd2.h:49-96) on the GPU.
 *
 * Each call evaluates through libbqd2_hip.so (bqd2_field, bqd2_field_bd,
 * bqd2_domain_id); pv_batch_aqd2 is the batched entry point for grids.  The
 * single-point calls (pv_t_aqd2 and the like, which example2.c:28-31 drives
 * from an OpenMP loop) are coalesced: concurrent callers' points join one
 * batch, evaluated by one bqd2_field call (a point's result does not depend
 * on the batch it travels in).  Batches of different threads may run at once:
 * the device library gives each calling thread its own stream and workspace.
 * One lock guards the lazy upload of a solution read from ex.dat.
 */
#include <errno.h>
#include <pthread.h>
#include <string.h>
#include <time.h>
#include "host_internal.h"
#include "bqd2.h"

static pthread_mutex_t g_field_lock = PTHREAD_MUTEX_INITIALIZER;

static void fail(const char *what, int rc)
{
  printf("libbem3_aw_qd2: %s failed (%d): %s. Exit...\n", what, rc, bqd2_last_error());
  exit(1);
}

/* device problem with the current surface solution of qd */
static struct bqd2_problem *ready(AQD2 *qd)
{
  int has = 0;
  pthread_mutex_lock(&g_field_lock);
  struct bqd2_problem *p = aqd2_problem(qd, &has);
  if (!has) aqd2_push_solution(qd, p);
  pthread_mutex_unlock(&g_field_lock);
  return p;
}

void pv_batch_aqd2(int n, const double *rt, int type, int total, double complex *p, double complex *pv, int *did,
                   AQD2 *qd)
{
  if (n <= 0) return;
  double *out = malloc((size_t)n * 8 * sizeof(double));
  int *d = did ? did : malloc((size_t)n * sizeof(int));
  struct bqd2_problem *pr = ready(qd);
  int rc = bqd2_field(pr, n, rt, type, total, out, d, NULL);
  if (rc) fail("bqd2_field", rc);
  for (int i = 0; i < n; i++) {
    if (p) p[i] = out[8 * i] + I * out[8 * i + 1];
    if (pv)
      for (int c = 0; c < 3; c++) pv[3 * i + c] = out[8 * i + 2 + 2 * c] + I * out[8 * i + 3 + 2 * c];
  }
  if (!did) free(d);
  free(out);
}

/* ---------------------------------------------------------------- coalescing
 * The first caller of a batch leads it: it waits until as many points have
 * joined as there were callers inside at the recent closes (the threads
 * calling in a loop), PB_CAP points or PB_WAIT_US microseconds, closes the
 * batch, evaluates it and wakes the others, who copy their results.  A lone
 * caller leads a batch of one without waiting. */
#define PB_CAP 512
#define PB_WAIT_US 2000
typedef struct PtBatch {
  AQD2 *qd;
  int type, total;
  int n, done, users;
  double rt[3 * PB_CAP];
  double out[8 * PB_CAP];
  int did[PB_CAP];
} PtBatch;

static pthread_mutex_t g_pb_mu = PTHREAD_MUTEX_INITIALIZER;
static pthread_cond_t g_pb_cv = PTHREAD_COND_INITIALIZER;
static PtBatch *g_open;  /* the batch accepting points */
static int g_inside;     /* callers inside one_point */
static int g_expect = 1; /* points the leader waits for: the most callers inside at the last closes */
static int g_recent[4];
static long long g_batches, g_points;

/* at a close: the callers inside then (this batch's and those of batches in
 * flight, who come back for their next point) set the next leaders' wait */
static void pb_closed(int n)
{
  memmove(g_recent + 1, g_recent, 3 * sizeof(int));
  g_recent[0] = g_inside;
  int m = 1;
  for (int i = 0; i < 4; i++) m = g_recent[i] > m ? g_recent[i] : m;
  g_expect = m < PB_CAP ? m : PB_CAP;
  g_batches++;
  g_points += n;
}

static int one_point(double complex *p, double complex *pv, double *rt, int type, int total, AQD2 *qd)
{
  pthread_mutex_lock(&g_pb_mu);
  g_inside++;
  PtBatch *b = g_open;
  int lead = 0;
  if (!b || b->qd != qd || b->type != type || b->total != total || b->n == PB_CAP) {
    b = calloc(1, sizeof(*b));
    if (!b) {
      pthread_mutex_unlock(&g_pb_mu);
      fail("point batch allocation", -1);
    }
    b->qd = qd;
    b->type = type;
    b->total = total;
    g_open = b;
    lead = 1;
  }
  const int i = b->n++;
  b->users++;
  memcpy(b->rt + 3 * i, rt, 3 * sizeof(double));
  if (lead) {
    struct timespec dl;
    clock_gettime(CLOCK_REALTIME, &dl);
    dl.tv_nsec += PB_WAIT_US * 1000L;
    dl.tv_sec += dl.tv_nsec / 1000000000L;
    dl.tv_nsec %= 1000000000L;
    while (b->n < g_expect && b->n < PB_CAP)
      if (pthread_cond_timedwait(&g_pb_cv, &g_pb_mu, &dl) == ETIMEDOUT) break;
    if (g_open == b) g_open = NULL;
    const int n = b->n;
    pb_closed(n);
    pthread_mutex_unlock(&g_pb_mu);
    struct bqd2_problem *pr = ready(qd);
    int rc = bqd2_field(pr, n, b->rt, type, total, b->out, b->did, NULL);
    if (rc) fail("bqd2_field", rc);
    pthread_mutex_lock(&g_pb_mu);
    b->done = 1;
    pthread_cond_broadcast(&g_pb_cv);
  } else {
    pthread_cond_broadcast(&g_pb_cv); /* the leader counts the points */
    while (!b->done) pthread_cond_wait(&g_pb_cv, &g_pb_mu);
  }
  const double *o = b->out + 8 * i;
  if (p) *p = o[0] + I * o[1];
  if (pv)
    for (int c = 0; c < 3; c++) pv[c] = o[2 + 2 * c] + I * o[3 + 2 * c];
  const int did = b->did[i];
  if (--b->users == 0) free(b);
  g_inside--;
  pthread_mutex_unlock(&g_pb_mu);
  return did;
}

/* No counterpart in the reference: batches and points of the coalesced
 * single-point calls since the library was loaded. */
void aqd2_point_batches(long long *batches, long long *points)
{
  pthread_mutex_lock(&g_pb_mu);
  *batches = g_batches;
  *points = g_points;
  pthread_mutex_unlock(&g_pb_mu);
}

int pv_s_aqd2(double complex *p, double complex *pv, double *rt, int type, AQD2 *qd)
{
  return one_point(p, pv, rt, type, 0, qd);
}

int pv_t_aqd2(double complex *p, double complex *pv, double *rt, int type, AQD2 *qd)
{
  return one_point(p, pv, rt, type, 1, qd);
}

int p_s_aqd2(double complex *p, double *rt, int type, AQD2 *qd)
{
  return one_point(p, NULL, rt, type, 0, qd);
}

int p_t_aqd2(double complex *p, double *rt, int type, AQD2 *qd)
{
  return one_point(p, NULL, rt, type, 1, qd);
}

/* velocity potential: p = -k2 phi in every domain (bem3_aw_qd2_field.c:3-79) */
int phi_s_aqd2(double complex *phi, double *rt, int type, AQD2 *qd)
{
  double complex p;
  int did = one_point(&p, NULL, rt, type, 0, qd);
  *phi = (did >= 0) ? -p / qd->k2[did] : 0.0;
  return did;
}

int phi_t_aqd2(double complex *phi, double *rt, int type, AQD2 *qd)
{
  double complex p;
  int did = one_point(&p, NULL, rt, type, 1, qd);
  *phi = (did >= 0) ? -p / qd->k2[did] : 0.0;
  return did;
}

static int classify(double *rt, AQD2 *qd)
{
  int did = -1;
  struct bqd2_problem *pr = aqd2_problem(qd, NULL);
  int rc = bqd2_domain_id(pr, 1, rt, &did);
  if (rc) fail("bqd2_domain_id", rc);
  return did;
}

int domain_id_aqd2(double *rt, AQD2 *qd)
{
  return classify(rt, qd);
}

int phi_i_aqd2(double complex *phi, double *rt, int type, AQD2 *qd)
{
  double complex p, v[3];
  (void)type;
  calc_apw_pv(&p, v, rt, &qd->pw);
  *phi = -p / qd->pw.wd.k2;
  return classify(rt, qd);
}

int p_i_aqd2(double complex *p, double *rt, int type, AQD2 *qd)
{
  double complex v[3];
  (void)type;
  calc_apw_pv(p, v, rt, &qd->pw);
  return classify(rt, qd);
}

int pv_i_aqd2(double complex *p, double complex *pv, double *rt, int type, AQD2 *qd)
{
  (void)type;
  calc_apw_pv(p, pv, rt, &qd->pw);
  return classify(rt, qd);
}

/* ---------------------------------------------------------------- boundary */
static void bd_point(double complex *p, double complex *pv, int did, int t, double zt, double et, int type, int total,
                     AQD2 *qd)
{
  double out[8], ze[2] = {zt, et};
  struct bqd2_problem *pr = ready(qd);
  int rc = bqd2_field_bd(pr, did, 1, &t, ze, type, total, out);
  if (rc) fail("bqd2_field_bd", rc);
  if (p) *p = out[0] + I * out[1];
  if (pv)
    for (int c = 0; c < 3; c++) pv[c] = out[2 + 2 * c] + I * out[3 + 2 * c];
}

void pv_s_bd_aqd2(double complex *p, double complex *pv, int did, int t, double zeta_t, double eta_t, int type, AQD2 *qd)
{
  bd_point(p, pv, did, t, zeta_t, eta_t, type, 0, qd);
}

void pv_t_bd_aqd2(double complex *p, double complex *pv, int did, int t, double zeta_t, double eta_t, int type, AQD2 *qd)
{
  bd_point(p, pv, did, t, zeta_t, eta_t, type, 1, qd);
}

void phi_s_bd_aqd2(double complex *phi, int did, int t, double zeta_t, double eta_t, int type, AQD2 *qd)
{
  double complex p;
  bd_point(&p, NULL, did, t, zeta_t, eta_t, type, 0, qd);
  *phi = -p / qd->k2[did];
}

void phi_t_bd_aqd2(double complex *phi, int did, int t, double zeta_t, double eta_t, int type, AQD2 *qd)
{
  double complex p;
  bd_point(&p, NULL, did, t, zeta_t, eta_t, type, 1, qd);
  *phi = -p / qd->k2[did];
}

/* incident field at a boundary point (zero inside the objects), :382-392, :471-487 */
static void bd_position(double *r, int did, int t, double zt, double et, AQD2 *qd)
{
  double w[3];
  aqd2_rw(&qd->bd, abs(qd->bd.sb[did].sid[t]), zt, et, r, w);
}

void phi_i_bd_aqd2(double complex *phi, int did, int t, double zeta_t, double eta_t, int type, AQD2 *qd)
{
  (void)type;
  if (did != 0) {
    *phi = 0.0;
    return;
  }
  double r[3];
  double complex v[3];
  bd_position(r, did, t, zeta_t, eta_t, qd);
  calc_apw_pv(phi, v, r, &qd->pw);
  *phi /= -qd->pw.wd.k2;
}

void pv_i_bd_aqd2(double complex *p, double complex *pv, int did, int t, double zeta_t, double eta_t, int type, AQD2 *qd)
{
  (void)type;
  if (did != 0) {
    *p = 0.0;
    for (int i = 0; i < 3; i++) pv[i] = 0.0;
    return;
  }
  double r[3];
  bd_position(r, did, t, zeta_t, eta_t, qd);
  calc_apw_pv(p, pv, r, &qd->pw);
}

/* ---------------------------------------------------------------- force */
void force_FN_aqd2(double *F, double *N, double *rc, int type, AQD2 *qd)
{
  struct bqd2_problem *pr = ready(qd);
  int rcode = bqd2_force(pr, rc, type, F, N);
  if (rcode) fail("bqd2_force", rcode);
}

/* surface area of domain did by the 4- or 9-point rule (bem3_aw_qd2_force.c:44-66, :279-298) */
void surface_area_aqd2(double *S, int did, int type, AQD2 *qd)
{
  BOUD *bd = &qd->bd;
  double sa = 0.0;
  for (int t = 1; t <= bd->sb[did].Ne; t++) {
    int s = abs(bd->sb[did].sid[t]);
    double ts = 0.0;
    if (bd->ed[s][3] != 0) {
      if (type == 0) {
        for (int i = 0; i < 4; i++) ts += vabs_d(bd->wen[s][i]) * bd->wt_44[i];
      } else {
        for (int i = 0; i < 9; i++) {
          double r[3], w[3];
          aqd2_rw(bd, s, bd->zt_49[i], bd->et_49[i], r, w);
          ts += vabs_d(w) * bd->wt_49[i];
        }
      }
    } else {
      int n = (type == 0) ? 4 : 7;
      for (int i = 0; i < n; i++) {
        double r[3], w[3];
        if (type == 0 && i < 3) {
          ts += vabs_d(bd->wen[s][i]) * bd->wt_34[i];
          continue;
        }
        aqd2_rw(bd, s, type == 0 ? bd->zt_34[i] : bd->zt_37[i], type == 0 ? bd->et_34[i] : bd->et_37[i], r, w);
        ts += vabs_d(w) * (type == 0 ? bd->wt_34[i] : bd->wt_37[i]);
      }
      ts *= 0.5;
    }
    sa += ts;
  }
  *S = sa;
}
