/*
 * field.c -- the reference's field API (bem3_aw_qd2.h:49-96) on the GPU.
 *
 * Each call evaluates through libbqd2_hip.so (bqd2_field, bqd2_field_bd,
 * bqd2_domain_id); pv_batch_aqd2 is the batched entry point for grids.  Calls
 * from several host threads (example2.c drives pv_t_aqd2 from an OpenMP loop)
 * run concurrently: the device library gives each calling thread its own
 * stream and workspace.  One lock only guards the lazy upload of a solution
 * read from ex.dat.
 */
#include <pthread.h>
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

static int one_point(double complex *p, double complex *pv, double *rt, int type, int total, AQD2 *qd)
{
  double complex tp, tv[3];
  int did;
  pv_batch_aqd2(1, rt, type, total, &tp, tv, &did, qd);
  if (p) *p = tp;
  if (pv)
    for (int c = 0; c < 3; c++) pv[c] = tv[c];
  return did;
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
