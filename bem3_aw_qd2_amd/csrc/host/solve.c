/*
 * solve.c -- solve_bieq_aqd2 on the GPU and the AQD2 <-> device bridge.
 *
 * Same phases and console report as the reference's solve_bieq_aqd2
 * (d3qd2_src/bem3_aw_qd2_solve_bieq.c:3-52): coefficient matrices, surface
 * potential, particle velocity.  All arithmetic runs in libbqd2_hip.so; this
 * file only packs the model, drives the phases and unpacks the results into the
 * caller's AQD2.  There is no CPU path: a failing GPU step ends the process
 * with a message, as the reference does on its own errors.
 */
#include <sys/time.h>
#include <pthread.h>
#include "host_internal.h"
#include "bqd2.h"

typedef struct {
  const AQD2 *key;
  bqd2_problem *prob;
  int has_solution;   /* P, dP, pv, dpv on the device match the AQD2 */
} side_entry;

#define MAX_SIDE 64
static side_entry g_side[MAX_SIDE];
static pthread_mutex_t g_side_lock = PTHREAD_MUTEX_INITIALIZER;

static void fail(const char *what, int rc)
{
  printf("libbem3_aw_qd2: %s failed (%d): %s. Exit...\n", what, rc, bqd2_last_error());
  exit(1);
}

static double wall(void)
{
  struct timeval tv;
  gettimeofday(&tv, NULL);
  return tv.tv_sec + 1e-6 * tv.tv_usec;
}

/* flat copies of the AQD2 arrays handed to bqd2_create */
typedef struct {
  double *rn, *cr, *cw, *ren, *wen, *Pi, *dPi, *k0, *k2, *rho0;
  int *ed, *eni, *md, *sd, *dom_ne, *dom_sid;
} packed;

static void pack_model(const AQD2 *qd, bqd2_model *m, packed *pk)
{
  const BOUD *bd = &qd->bd;
  size_t ne1 = (size_t)bd->Ne + 1;
  memset(m, 0, sizeof(*m));
  pk->rn = calloc(((size_t)bd->Nn + 1) * 3, sizeof(double));
  for (int i = 0; i <= bd->Nn; i++)
    for (int d = 0; d < 3; d++) pk->rn[3 * i + d] = bd->rn[i][d];
  pk->cr = calloc(ne1 * 12, sizeof(double));
  pk->cw = calloc(ne1 * 9, sizeof(double));
  pk->ren = calloc(ne1 * 12, sizeof(double));
  pk->wen = calloc(ne1 * 12, sizeof(double));
  pk->Pi = calloc(ne1 * 8, sizeof(double));
  pk->dPi = calloc(ne1 * 8, sizeof(double));
  pk->ed = calloc(ne1 * 4, sizeof(int));
  pk->eni = calloc(ne1 * 4, sizeof(int));
  pk->md = calloc(ne1, sizeof(int));
  pk->sd = calloc(ne1, sizeof(int));
  for (size_t s = 1; s < ne1; s++) {
    for (int d = 0; d < 3; d++) {
      for (int j = 0; j < 4; j++) pk->cr[12 * s + 4 * d + j] = bd->cr[s][d][j];
      for (int j = 0; j < 3; j++) pk->cw[9 * s + 3 * d + j] = bd->cw[s][d][j];
    }
    for (int n = 0; n < 4; n++) {
      for (int d = 0; d < 3; d++) {
        pk->ren[12 * s + 3 * n + d] = bd->ren[s][n][d];
        pk->wen[12 * s + 3 * n + d] = bd->wen[s][n][d];
      }
      pk->Pi[8 * s + 2 * n] = creal(bd->Pi[s][n]);
      pk->Pi[8 * s + 2 * n + 1] = cimag(bd->Pi[s][n]);
      pk->dPi[8 * s + 2 * n] = creal(bd->dPi[s][n]);
      pk->dPi[8 * s + 2 * n + 1] = cimag(bd->dPi[s][n]);
      pk->ed[4 * s + n] = bd->ed[s][n];
      pk->eni[4 * s + n] = bd->eni[s][n];
    }
    pk->md[s] = bd->md[s];
    pk->sd[s] = bd->sd[s];
  }
  int nd = qd->MN + 1;
  pk->dom_ne = calloc(nd, sizeof(int));
  size_t tot = 0;
  for (int d = 0; d < nd; d++) tot += bd->sb[d].Ne + 1;
  pk->dom_sid = calloc(tot, sizeof(int));
  size_t off = 0;
  for (int d = 0; d < nd; d++) {
    pk->dom_ne[d] = bd->sb[d].Ne;
    for (int i = 0; i <= bd->sb[d].Ne; i++) pk->dom_sid[off + i] = bd->sb[d].sid[i];
    off += bd->sb[d].Ne + 1;
  }
  pk->k0 = calloc(nd, sizeof(double));
  pk->k2 = calloc(2 * nd, sizeof(double));
  pk->rho0 = calloc(nd, sizeof(double));
  for (int d = 0; d < nd; d++) {
    pk->k0[d] = qd->k0[d];
    pk->k2[2 * d] = creal(qd->k2[d]);
    pk->k2[2 * d + 1] = cimag(qd->k2[d]);
    pk->rho0[d] = qd->rho0[d];
  }
  m->Ne = bd->Ne;
  m->Nn = bd->Nn;
  m->NN = bd->NN;
  m->MN = qd->MN;
  m->ps = bd->ps;
  m->rn = pk->rn;
  m->cr = pk->cr; m->cw = pk->cw; m->ren = pk->ren; m->wen = pk->wen;
  m->ed = pk->ed; m->eni = pk->eni; m->md = pk->md; m->sd = pk->sd;
  m->Pi = pk->Pi; m->dPi = pk->dPi;
  m->dom_ne = pk->dom_ne; m->dom_sid = pk->dom_sid;
  m->k0 = pk->k0; m->k2 = pk->k2; m->rho0 = pk->rho0;
  m->qd[0] = bd->qd.k;
  for (int i = 0; i < 3; i++) m->qd[1 + i] = bd->qd.vk[i];
  m->qd[4] = bd->qd.vd1[0]; m->qd[5] = bd->qd.vd1[1];
  m->qd[6] = bd->qd.vd2[0]; m->qd[7] = bd->qd.vd2[1];
  memcpy(m->xli, bd->xli, sizeof(m->xli)); memcpy(m->wli, bd->wli, sizeof(m->wli));
  memcpy(m->xhi, bd->xhi, sizeof(m->xhi)); memcpy(m->whi, bd->whi, sizeof(m->whi));
  memcpy(m->zt_44, bd->zt_44, sizeof(m->zt_44)); memcpy(m->et_44, bd->et_44, sizeof(m->et_44));
  memcpy(m->wt_44, bd->wt_44, sizeof(m->wt_44)); memcpy(m->zt_34, bd->zt_34, sizeof(m->zt_34));
  memcpy(m->et_34, bd->et_34, sizeof(m->et_34)); memcpy(m->wt_34, bd->wt_34, sizeof(m->wt_34));
  memcpy(m->zt_49, bd->zt_49, sizeof(m->zt_49)); memcpy(m->et_49, bd->et_49, sizeof(m->et_49));
  memcpy(m->wt_49, bd->wt_49, sizeof(m->wt_49)); memcpy(m->zt_37, bd->zt_37, sizeof(m->zt_37));
  memcpy(m->et_37, bd->et_37, sizeof(m->et_37)); memcpy(m->wt_37, bd->wt_37, sizeof(m->wt_37));
  m->pw_p[0] = creal(qd->pw.p);
  m->pw_p[1] = cimag(qd->pw.p);
  m->pw_rho0 = qd->pw.rho0;
  m->pw_c0 = qd->pw.c0;
  m->pw_k0 = qd->pw.wd.k0;
  memcpy(m->pw_tv, qd->pw.tv, sizeof(m->pw_tv));
  memcpy(m->pw_R, qd->pw.wd.R, sizeof(m->pw_R));
}

static void free_packed(packed *pk)
{
  free(pk->rn); free(pk->cr); free(pk->cw); free(pk->ren); free(pk->wen); free(pk->Pi); free(pk->dPi);
  free(pk->ed); free(pk->eni); free(pk->md); free(pk->sd); free(pk->dom_ne); free(pk->dom_sid);
  free(pk->k0); free(pk->k2); free(pk->rho0);
}

/* the packed model and its arrays in one block (the model first) */
typedef struct {
  bqd2_model m;
  packed pk;
} model_box;

void *aqd2_model_new(AQD2 *qd)
{
  model_box *b = calloc(1, sizeof(*b));
  if (!b) {
    printf("libbem3_aw_qd2: out of memory in aqd2_model_new. Exit...\n");
    exit(1);
  }
  pack_model(qd, &b->m, &b->pk);
  return &b->m;
}

void aqd2_model_free(void *model)
{
  if (!model) return;
  model_box *b = (model_box *)model;
  free_packed(&b->pk);
  free(b);
}

/* the device problem of qd, created on first use */
bqd2_problem *aqd2_problem(AQD2 *qd, int *has_solution)
{
  pthread_mutex_lock(&g_side_lock);
  int free_slot = -1;
  for (int i = 0; i < MAX_SIDE; i++) {
    if (g_side[i].key == qd) {
      bqd2_problem *p = g_side[i].prob;
      if (has_solution) *has_solution = g_side[i].has_solution;
      pthread_mutex_unlock(&g_side_lock);
      return p;
    }
    if (free_slot < 0 && g_side[i].key == NULL) free_slot = i;
  }
  if (free_slot < 0) {
    pthread_mutex_unlock(&g_side_lock);
    printf("libbem3_aw_qd2: more than %d live AQD2 models. Exit...\n", MAX_SIDE);
    exit(1);
  }
  bqd2_model m;
  packed pk;
  pack_model(qd, &m, &pk);
  bqd2_problem *p = NULL;
  int rc = bqd2_create(&m, &p);
  if (rc == BQD2_E_NOMEM) {
    /* the default layout does not fit the device: the system layout, for the
     * meshes it holds (one sphere of quadrilaterals in a periodic lattice) */
    int rc2 = bqd2_create_large(&m, &p);
    if (rc2 == 0) rc = 0;
  }
  free_packed(&pk);
  if (rc) {
    pthread_mutex_unlock(&g_side_lock);
    fail("bqd2_create", rc);
  }
  g_side[free_slot].key = qd;
  g_side[free_slot].prob = p;
  g_side[free_slot].has_solution = 0;
  if (has_solution) *has_solution = 0;
  pthread_mutex_unlock(&g_side_lock);
  return p;
}

void aqd2_mark_solution(AQD2 *qd, int on)
{
  pthread_mutex_lock(&g_side_lock);
  for (int i = 0; i < MAX_SIDE; i++)
    if (g_side[i].key == qd) g_side[i].has_solution = on;
  pthread_mutex_unlock(&g_side_lock);
}

void aqd2_release_device(AQD2 *qd)
{
  pthread_mutex_lock(&g_side_lock);
  for (int i = 0; i < MAX_SIDE; i++) {
    if (g_side[i].key == qd) {
      bqd2_destroy(g_side[i].prob);
      g_side[i].key = NULL;
      g_side[i].prob = NULL;
      g_side[i].has_solution = 0;
    }
  }
  pthread_mutex_unlock(&g_side_lock);
}

/* copy the device solution of every domain into the AQD2 arrays */
void aqd2_pull_solution(AQD2 *qd, bqd2_problem *p)
{
  for (int d = 0; d <= qd->MN; d++) {
    SUBD *s = &qd->bd.sb[d];
    size_t n1 = (size_t)s->Ne + 1;
    double *P = malloc(n1 * 4 * 16), *dP = malloc(n1 * 4 * 16), *pv = malloc(n1 * 12 * 16), *dpv = malloc(n1 * 12 * 16);
    int rc = bqd2_download_solution(p, d, P, dP, pv, dpv);
    if (rc) fail("bqd2_download_solution", rc);
    for (size_t i = 1; i < n1; i++)
      for (int n = 0; n < 4; n++) {
        s->P[i][n] = P[8 * i + 2 * n] + I * P[8 * i + 2 * n + 1];
        s->dP[i][n] = dP[8 * i + 2 * n] + I * dP[8 * i + 2 * n + 1];
        for (int c = 0; c < 3; c++) {
          size_t k = 2 * (12 * i + 3 * n + c);
          s->pv[i][n][c] = pv[k] + I * pv[k + 1];
          s->dpv[i][n][c] = dpv[k] + I * dpv[k + 1];
        }
      }
    free(P); free(dP); free(pv); free(dpv);
  }
}

/* push the AQD2 solution arrays to the device (models read from ex.dat) */
void aqd2_push_solution(AQD2 *qd, bqd2_problem *p)
{
  for (int d = 0; d <= qd->MN; d++) {
    SUBD *s = &qd->bd.sb[d];
    size_t n1 = (size_t)s->Ne + 1;
    double *P = calloc(n1 * 8, sizeof(double)), *dP = calloc(n1 * 8, sizeof(double));
    double *pv = calloc(n1 * 24, sizeof(double)), *dpv = calloc(n1 * 24, sizeof(double));
    for (size_t i = 1; i < n1; i++)
      for (int n = 0; n < 4; n++) {
        P[8 * i + 2 * n] = creal(s->P[i][n]);
        P[8 * i + 2 * n + 1] = cimag(s->P[i][n]);
        dP[8 * i + 2 * n] = creal(s->dP[i][n]);
        dP[8 * i + 2 * n + 1] = cimag(s->dP[i][n]);
        for (int c = 0; c < 3; c++) {
          size_t k = 2 * (12 * i + 3 * n + c);
          pv[k] = creal(s->pv[i][n][c]);
          pv[k + 1] = cimag(s->pv[i][n][c]);
          dpv[k] = creal(s->dpv[i][n][c]);
          dpv[k + 1] = cimag(s->dpv[i][n][c]);
        }
      }
    int rc = bqd2_upload_solution(p, d, P, dP, pv, dpv);
    free(P); free(dP); free(pv); free(dpv);
    if (rc) fail("bqd2_upload_solution", rc);
  }
  aqd2_mark_solution(qd, 1);
}

void solve_bieq_aqd2(AQD2 *qd)
{
  printf("\nsolve acoustic wave boundary value \n");
  double t0 = wall(), ts;
  bqd2_problem *p = aqd2_problem(qd, NULL);

  printf("  coefficient matrix          ");
  fflush(stdout);
  ts = wall();
  for (int d = 0; d <= qd->MN; d++) {
    int rc = bqd2_assemble(p, d, NULL);
    if (rc) fail("bqd2_assemble", rc);
  }
  printf("finished. Elapsed time : %5g (sec)\n", wall() - ts);

  printf("  solve VP boundary value     ");
  fflush(stdout);
  ts = wall();
  int rc = bqd2_solve_surface(p, NULL);
  if (rc) fail("bqd2_solve_surface", rc);
  printf("finished. Elapsed time : %5g (sec)\n", wall() - ts);

  printf("  solve PV boundary value     ");
  fflush(stdout);
  ts = wall();
  rc = bqd2_solve_velocity(p);
  if (rc) fail("bqd2_solve_velocity", rc);
  aqd2_pull_solution(qd, p);
  aqd2_mark_solution(qd, 1);
  printf("finished. Elapsed time : %5g (sec)\n", wall() - ts);
  printf("Total elapsed time : %g (sec)\n", wall() - t0);
}
