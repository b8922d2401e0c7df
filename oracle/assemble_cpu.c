/*
 * assemble_cpu.c -- oracle assembly, system build and boundary-value solve.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates d3qd2_src/bem3_aw_qd2_solve_bieq.c: the per-domain coefficient
 * matrices of create_cmatrix_domain (:133-218), the system of
 * create_matrix_csr_dac (:244-345, dense here), the solution scatter
 * tmatrix_bd_store (:370-427), the particle-velocity recovery q_solve_pv_bv /
 * solve_pv_bv with dpdt_fcm and inv_matrix_33 (:429-530, :655-750) and the
 * normal-derivative solve solve_dpv_bv (:533-653).  The dense complex solves
 * themselves (MKL PARDISO and LAPACKE_zgesv in the reference) are left to the
 * caller: the tests use numpy's LAPACK.
 *
 * Layouts (row-major, complex interleaved):
 *   G, H   [N x N]    row (t-1)*4+tn, column (s-1)*4+n, N = 4 Ne_d
 *   dG, dH [2N x N]   row 2*((t-1)*4+tn) + {0: tz, 1: te}
 *   dF     [N x 3]    (F or 1+F, dF/dtz, dF/dte)
 */
#include <math.h>
#include <omp.h>
#include <complex.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"
#include "oracle_elem.h"

static int is_tri(const BOUD *bd, int s) { return bd->ed[abs(s)][3] == 0; }

/* coefficient block of target node (t, tn) against source element s in domain
 * did: the dispatch of d3qd2_elem.c:515-572 (interior / isolated) and :830-886
 * (periodic domain 0), cm.type == 2 (GL12 with the GH64 switch). */
long orc_coef_block(const AQD2 *qd, int did, int td, int tn, int sd, const double *vtz, const double *vte,
                    double complex *CC, double complex *dCz, double complex *dCe)
{
  const BOUD *bd = &qd->bd;
  double k = qd->k0[did];
  int at = abs(td);
  const double *rt = bd->ren[at][tn];
  long terms = 0;
  if (is_tri(bd, sd)) {
    /* linear triangles (d3qd2_elem.c:541-570, :855-884) */
    if (did == 0 && bd->ps == 1) {
      if (td != sd) {
        double dF0, dF1;
        orc_lit_dF_GL(bd, rt, vtz, vte, sd, &dF0, &dF1);
        if (fabs(dF0) > CBD_SDF || fabs(dF1) > CBD_SDF)
          terms = orc_q0_lit_pair(bd, rt, vtz, vte, sd, GHN, bd->xhi, bd->whi, CC, dCz, dCe);
        else
          terms = orc_q0_lit_pair(bd, rt, vtz, vte, sd, GLN, bd->xli, bd->wli, CC, dCz, dCe);
      } else {
        terms = orc_q0_lit_self(bd, sd, bd->zt_34[tn], bd->et_34[tn], CC);
        for (int l = 0; l < 9; l++) dCz[l] = dCe[l] = 0.0;
      }
    } else {
      if (td != sd) {
        double dF0, dF1;
        orc_lit_dF_GL(bd, rt, vtz, vte, sd, &dF0, &dF1);
        if (fabs(dF0) > CBD_SDF || fabs(dF1) > CBD_SDF)
          orc_lit_pair(bd, rt, vtz, vte, sd, k, GHN, bd->xhi, bd->whi, CC, dCz, dCe);
        else
          orc_lit_pair(bd, rt, vtz, vte, sd, k, GLN, bd->xli, bd->wli, CC, dCz, dCe);
      } else {
        orc_lit_self(bd, sd, bd->zt_34[tn], bd->et_34[tn], k, CC);
        orc_lit_self_dPV(bd, sd, bd->zt_34[tn], bd->et_34[tn], k, 0, dCz);
        orc_lit_self_dPV(bd, sd, bd->zt_34[tn], bd->et_34[tn], k, 1, dCe);
      }
    }
    return terms;
  }
  if (did == 0 && bd->ps == 1) {
    if (td != sd) {
      double dF0, dF1;
      orc_bil_dF_GL(bd, rt, vtz, vte, sd, &dF0, &dF1);
      if (fabs(dF0) > CBD_SDF || fabs(dF1) > CBD_SDF)
        terms = orc_q0_bil_pair(bd, rt, vtz, vte, sd, GHN, bd->xhi, bd->whi, CC, dCz, dCe);
      else
        terms = orc_q0_bil_pair(bd, rt, vtz, vte, sd, GLN, bd->xli, bd->wli, CC, dCz, dCe);
    } else {
      double dFz, dFe;
      terms = orc_q0_bil_self(bd, sd, bd->zt_44[tn], bd->et_44[tn], CC);
      orc_q0_bil_self_dF(bd, sd, bd->zt_44[tn], bd->et_44[tn], &dFz, &dFe);
      for (int l = 0; l < 9; l++) dCz[l] = dCe[l] = 0.0;
      dCz[8] = dFz;
      dCe[8] = dFe;
    }
  } else {
    if (td != sd) {
      double dF0, dF1;
      orc_bil_dF_GL(bd, rt, vtz, vte, sd, &dF0, &dF1);
      if (fabs(dF0) > CBD_SDF || fabs(dF1) > CBD_SDF)
        orc_bil_pair(bd, rt, vtz, vte, sd, k, GHN, bd->xhi, bd->whi, CC, dCz, dCe);
      else
        orc_bil_pair(bd, rt, vtz, vte, sd, k, GLN, bd->xli, bd->wli, CC, dCz, dCe);
    } else {
      orc_bil_self(bd, sd, bd->zt_44[tn], bd->et_44[tn], k, CC);
      orc_bil_self_dPV(bd, sd, bd->zt_44[tn], bd->et_44[tn], k, 0, dCz);
      orc_bil_self_dPV(bd, sd, bd->zt_44[tn], bd->et_44[tn], k, 1, dCe);
    }
  }
  return terms;
}

/* one row (t, tn) of domain did (bem3_aw_qd2_solve_bieq.c:157-199); t is the
 * 1-based position in the domain list. dG, dH may be NULL. Returns GF terms. */
long orc_assemble_row(const AQD2 *qd, int did, int t, int tn, double complex *G, double complex *H,
                      double complex *dG, double complex *dH, double *dF)
{
  const BOUD *bd = &qd->bd;
  int Ne = bd->sb[did].Ne, N = 4 * Ne;
  int td = bd->sb[did].sid[t];
  double vtz[3], vte[3];
  orc_node_tangents(bd, td, tn, vtz, vte);
  double F = 0.0, Fz = 0.0, Fe = 0.0;
  long terms = 0, err = 0;
  /* the elements in parallel (inside an enclosing parallel region this runs
   * on one thread); F and its derivatives summed afterwards in element order */
  double *fs = malloc(sizeof(double) * 3 * (Ne + 1));
#pragma omp parallel for schedule(dynamic) reduction(+ : terms)
  for (int s = 1; s <= Ne; s++) {
    int sd = bd->sb[did].sid[s];
    double complex CC[9], dCz[9], dCe[9];
    long c = orc_coef_block(qd, did, td, tn, sd, vtz, vte, CC, dCz, dCe);
    if (c < 0) {
#pragma omp critical
      err = c;
      continue;
    }
    terms += c;
    for (int i = 0; i < 4; i++) {
      G[(s - 1) * 4 + i] = CC[i];
      H[(s - 1) * 4 + i] = CC[i + 4];
      if (dG) {
        dG[(s - 1) * 4 + i] = dCz[i];
        dG[N + (s - 1) * 4 + i] = dCe[i];
      }
      if (dH) {
        dH[(s - 1) * 4 + i] = dCz[i + 4];
        dH[N + (s - 1) * 4 + i] = dCe[i + 4];
      }
    }
    fs[3 * s] = creal(CC[8]);
    fs[3 * s + 1] = creal(dCz[8]);
    fs[3 * s + 2] = creal(dCe[8]);
  }
  for (int s = 1; s <= Ne && !err; s++) {
    F += fs[3 * s];
    Fz += fs[3 * s + 1];
    Fe += fs[3 * s + 2];
  }
  free(fs);
  if (err) return err;
  double diag = (did == 0) ? 1.0 + F : F;
  H[(t - 1) * 4 + tn] += diag;
  dF[0] = diag;
  dF[1] = Fz;
  dF[2] = Fe;
  return terms;
}

long orc_assemble_domain(const AQD2 *qd, int did, double complex *G, double complex *H, double complex *dG,
                         double complex *dH, double *dF)
{
  const BOUD *bd = &qd->bd;
  int Ne = bd->sb[did].Ne;
  size_t N = 4 * (size_t)Ne;
  long total = 0, err = 0;
#pragma omp parallel for schedule(dynamic) reduction(+ : total)
  for (int t = 1; t <= Ne; t++) {
    int tl = is_tri(bd, bd->sb[did].sid[t]) ? 3 : 4;
    for (int tn = 0; tn < tl; tn++) {
      size_t row = (size_t)(t - 1) * 4 + tn;
      long c = orc_assemble_row(qd, did, t, tn, G + row * N, H + row * N, dG ? dG + 2 * row * N : NULL,
                                dH ? dH + 2 * row * N : NULL, dF + 3 * row);
      if (c < 0) err = c;
      else total += c;
    }
  }
  return err < 0 ? err : total;
}

/* coefficient blocks of n (row, source element) pairs of domain did, computed
 * in parallel and discarded: the timed sample of the CPU baseline.  row is
 * (t-1)*4+tn, src the 1-based element position.  Returns GF terms or <0. */
long orc_pair_sample(const AQD2 *qd, int did, int n, const int *row, const int *src)
{
  const BOUD *bd = &qd->bd;
  long total = 0, err = 0;
#pragma omp parallel for schedule(dynamic) reduction(+ : total)
  for (int i = 0; i < n; i++) {
    int t = row[i] / 4 + 1, tn = row[i] % 4;
    int td = bd->sb[did].sid[t];
    double vtz[3], vte[3];
    orc_node_tangents(bd, td, tn, vtz, vte);
    double complex CC[9], dCz[9], dCe[9];
    long c = orc_coef_block(qd, did, td, tn, bd->sb[did].sid[src[i]], vtz, vte, CC, dCz, dCe);
    if (c < 0) err = c;
    else total += c;
  }
  return err < 0 ? err : total;
}

/* system row of node (t, tn) of domain did from that domain's G, H rows
 * (bem3_aw_qd2_solve_bieq.c:279-319). A has na = 2 NN columns. */
void orc_system_row(const AQD2 *qd, int did, const double complex *Grow, const double complex *Hrow,
                    double complex *A, double complex *b)
{
  const BOUD *bd = &qd->bd;
  size_t nn = bd->NN, na = 2 * nn;
  memset(A, 0, na * sizeof(double complex));
  double complex tb = 0.0;
  for (int s = 1; s <= bd->sb[did].Ne; s++) {
    int sd = bd->sb[did].sid[s], asd = abs(sd);
    int mdid = bd->md[asd], sdid = bd->sd[asd];
    int nl = is_tri(bd, sd) ? 3 : 4;
    if (did == mdid) {
      for (int l = 0; l < nl; l++) {
        A[bd->eni[asd][l]] = Hrow[(s - 1) * 4 + l];
        A[nn + bd->eni[asd][l]] = -Grow[(s - 1) * 4 + l];
      }
    } else {
      double complex hk2 = qd->k2[mdid] / qd->k2[sdid];
      for (int l = 0; l < nl; l++) {
        A[bd->eni[asd][l]] = hk2 * Hrow[(s - 1) * 4 + l];
        A[nn + bd->eni[asd][l]] = Grow[(s - 1) * 4 + l];
        if (mdid == 0) tb += -hk2 * Hrow[(s - 1) * 4 + l] * bd->Pi[asd][l] - Grow[(s - 1) * 4 + l] * bd->dPi[asd][l];
      }
    }
  }
  *b = tb;
}

/* system unknown vector from a stored solution (inverse of tmatrix_bd_store,
 * bem3_aw_qd2_solve_bieq.c:370-427): x[eni] = P, x[nn+eni] = dP of the main side */
void orc_solution_vector(const AQD2 *qd, double complex *x)
{
  const BOUD *bd = &qd->bd;
  size_t nn = bd->NN;
  for (int d = 0; d <= qd->MN; d++) {
    for (int s = 1; s <= bd->sb[d].Ne; s++) {
      int sd = bd->sb[d].sid[s], asd = abs(sd);
      if (d != bd->md[asd]) continue;
      int nl = is_tri(bd, sd) ? 3 : 4;
      for (int l = 0; l < nl; l++) {
        x[bd->eni[asd][l]] = bd->sb[d].P[s][l];
        x[nn + bd->eni[asd][l]] = bd->sb[d].dP[s][l];
      }
    }
  }
}

/* tmatrix_bd_store (bem3_aw_qd2_solve_bieq.c:370-427) */
void orc_store_solution(AQD2 *qd, const double complex *X)
{
  BOUD *bd = &qd->bd;
  size_t nn = bd->NN;
  for (int d = 0; d <= qd->MN; d++) {
    for (int s = 1; s <= bd->sb[d].Ne; s++) {
      int sd = bd->sb[d].sid[s], asd = abs(sd);
      int mdid = bd->md[asd], sdid = bd->sd[asd];
      int nl = is_tri(bd, sd) ? 3 : 4;
      if (d == mdid) {
        for (int l = 0; l < nl; l++) {
          bd->sb[d].P[s][l] = X[bd->eni[asd][l]];
          bd->sb[d].dP[s][l] = X[nn + bd->eni[asd][l]];
        }
      } else {
        double complex hk2 = qd->k2[mdid] / qd->k2[sdid];
        for (int l = 0; l < nl; l++) {
          bd->sb[d].P[s][l] = hk2 * X[bd->eni[asd][l]];
          bd->sb[d].dP[s][l] = -1.0 * X[nn + bd->eni[asd][l]];
          if (mdid == 0) {
            bd->sb[d].P[s][l] += hk2 * bd->Pi[asd][l];
            bd->sb[d].dP[s][l] += -1.0 * bd->dPi[asd][l];
          }
        }
      }
    }
  }
}

/* inverse of the matrix with rows n, vtz, vte (bem3_aw_qd2_solve_bieq.c:493-514) */
void orc_inv33(double *ma, const double *vn, const double *vtz, const double *vte)
{
  double idet = 1.0 / (vn[0] * vtz[1] * vte[2] + vn[1] * vtz[2] * vte[0] + vn[2] * vtz[0] * vte[1] -
                       (vn[2] * vtz[1] * vte[0] + vn[1] * vtz[0] * vte[2] + vn[0] * vtz[2] * vte[1]));
  ma[0] = vtz[1] * vte[2] - vtz[2] * vte[1];
  ma[1] = -vn[1] * vte[2] + vn[2] * vte[1];
  ma[2] = vn[1] * vtz[2] - vn[2] * vtz[1];
  ma[3] = -vtz[0] * vte[2] + vtz[2] * vte[0];
  ma[4] = vn[0] * vte[2] - vn[2] * vte[0];
  ma[5] = -vn[0] * vtz[2] + vn[2] * vtz[0];
  ma[6] = vtz[0] * vte[1] - vtz[1] * vte[0];
  ma[7] = -vn[0] * vte[1] + vn[1] * vte[0];
  ma[8] = vn[0] * vtz[1] - vn[1] * vtz[0];
  for (int i = 0; i < 9; i++) ma[i] *= idet;
}

/* tangential derivatives of P at node (t, tn) from its dG, dH rows and dF
 * (dpdt_fcm, bem3_aw_qd2_solve_bieq.c:516-530) */
void orc_dpdt(const AQD2 *qd, int d, int t, int tn, const double complex *dGrow, const double complex *dHrow,
              const double *dF, double complex *dPz, double complex *dPe)
{
  const SUBD *sb = &qd->bd.sb[d];
  int Ne = sb->Ne;
  double complex z = 0.0, e = 0.0;
  for (int s = 1; s <= Ne; s++)
    for (int sn = 0; sn < 4; sn++) {
      z += dGrow[4 * (s - 1) + sn] * sb->dP[s][sn] - dHrow[4 * (s - 1) + sn] * sb->P[s][sn];
      e += dGrow[Ne * 4 + 4 * (s - 1) + sn] * sb->dP[s][sn] - dHrow[Ne * 4 + 4 * (s - 1) + sn] * sb->P[s][sn];
    }
  *dPz = (z - sb->P[t][tn] * dF[1]) / dF[0];
  *dPe = (e - sb->P[t][tn] * dF[2]) / dF[0];
}

/* particle velocity at node (t, tn) of domain d from dP/dn and the tangential
 * derivatives: pv = -[n; tz; te]^-1 (dP, dPz, dPe) (:602-607 / :700-707) */
void orc_node_pv(const AQD2 *qd, int d, int t, int tn, double complex dPz, double complex dPe, double complex *pv)
{
  const BOUD *bd = &qd->bd;
  int at = bd->sb[d].sid[t], atd = abs(at);
  double sig = at < 0 ? -1.0 : 1.0, vtz[3], vte[3], vn[3], a[9];
  orc_node_tangents(bd, atd, tn, vtz, vte);
  for (int i = 0; i < 3; i++) vn[i] = sig * bd->wen[atd][tn][i];
  orc_unit(vn);
  orc_inv33(a, vn, vtz, vte);
  double complex dP = bd->sb[d].dP[t][tn];
  for (int i = 0; i < 3; i++) pv[i] = -(dP * a[i * 3 + 0] + dPz * a[i * 3 + 1] + dPe * a[i * 3 + 2]);
}

/* ---------------------------------------------------------------- test helpers */
int orc_model_info(const AQD2 *qd, int *info)
{
  info[0] = qd->MN;
  info[1] = qd->bd.Ne;
  info[2] = qd->bd.NN;
  info[3] = qd->bd.ps;
  for (int d = 0; d <= qd->MN; d++) info[4 + d] = qd->bd.sb[d].Ne;
  return 0;
}

/* residual of one system row against the solution stored in qd:
 * |A_row x - b| / sum_j |A_row,j x_j| (x from the stored P, dP) */
double orc_row_residual(const AQD2 *qd, int did, int t, int tn, long *terms)
{
  const BOUD *bd = &qd->bd;
  size_t N = 4 * (size_t)bd->sb[did].Ne, na = 2 * (size_t)bd->NN;
  double complex *G = calloc(N, sizeof(double complex)), *H = calloc(N, sizeof(double complex));
  double complex *A = calloc(na, sizeof(double complex)), *x = calloc(na, sizeof(double complex)), b;
  double dF[3];
  long c = orc_assemble_row(qd, did, t, tn, G, H, NULL, NULL, dF);
  if (terms) *terms = c;
  orc_system_row(qd, did, G, H, A, &b);
  orc_solution_vector(qd, x);
  double complex acc = 0.0;
  double scale = 0.0;
  for (size_t j = 0; j < na; j++) {
    acc += A[j] * x[j];
    scale += cabs(A[j] * x[j]);
  }
  double r = cabs(acc - b) / (scale + cabs(b));
  free(G); free(H); free(A); free(x);
  return c < 0 ? -1.0 : r;
}

/* particle velocity of node (t, tn) of interior domain d recomputed from the
 * stored P, dP (the per-node step of q_solve_pv_bv / solve_pv_bv) */
int orc_node_pv_recompute(const AQD2 *qd, int d, int t, int tn, double complex *pv)
{
  const BOUD *bd = &qd->bd;
  size_t N = 4 * (size_t)bd->sb[d].Ne;
  double complex *G = calloc(N, sizeof(double complex)), *H = calloc(N, sizeof(double complex));
  double complex *dG = calloc(2 * N, sizeof(double complex)), *dH = calloc(2 * N, sizeof(double complex));
  double dF[3], complex_dummy = 0;
  (void)complex_dummy;
  long c = orc_assemble_row(qd, d, t, tn, G, H, dG, dH, dF);
  double complex dPz, dPe;
  orc_dpdt(qd, d, t, tn, dG, dH, dF, &dPz, &dPe);
  orc_node_pv(qd, d, t, tn, dPz, dPe, pv);
  free(G); free(H); free(dG); free(dH);
  return c < 0 ? (int)c : 0;
}

/* residual of one row of the dpv system G dpv = H pv of domain d (solve_dpv_bv,
 * bem3_aw_qd2_solve_bieq.c:560-617), relative, max over the 3 components */
double orc_dpv_row_residual(const AQD2 *qd, int d, int t, int tn)
{
  const BOUD *bd = &qd->bd;
  const SUBD *sb = &bd->sb[d];
  size_t N = 4 * (size_t)sb->Ne;
  double complex *G = calloc(N, sizeof(double complex)), *H = calloc(N, sizeof(double complex));
  double dF[3];
  orc_assemble_row(qd, d, t, tn, G, H, NULL, NULL, dF);
  double worst = 0.0;
  for (int i = 0; i < 3; i++) {
    double complex acc = 0.0;
    double scale = 0.0;
    for (int s = 1; s <= sb->Ne; s++) {
      int nl = is_tri(bd, sb->sid[s]) ? 3 : 4;
      for (int n = 0; n < nl; n++) {
        double complex a = G[4 * (s - 1) + n] * sb->dpv[s][n][i], b = H[4 * (s - 1) + n] * sb->pv[s][n][i];
        acc += a - b;
        scale += cabs(a) + cabs(b);
      }
    }
    double r = cabs(acc) / scale;
    if (r > worst) worst = r;
  }
  free(G); free(H);
  return worst;
}

/* per-source-element F, dF/dtz, dF/dte of one row (diagnostics for tests) */
int orc_row_fparts(const AQD2 *qd, int did, int t, int tn, double *out)
{
  const BOUD *bd = &qd->bd;
  int Ne = bd->sb[did].Ne, td = bd->sb[did].sid[t];
  double vtz[3], vte[3];
  orc_node_tangents(bd, td, tn, vtz, vte);
  for (int s = 1; s <= Ne; s++) {
    double complex CC[9], dCz[9], dCe[9];
    long c = orc_coef_block(qd, did, td, tn, bd->sb[did].sid[s], vtz, vte, CC, dCz, dCe);
    if (c < 0) return (int)c;
    out[3 * (s - 1)] = creal(CC[8]);
    out[3 * (s - 1) + 1] = creal(dCz[8]);
    out[3 * (s - 1) + 2] = creal(dCe[8]);
  }
  return 0;
}

/* OpenMP team of the oracle's parallel loops: set it (n > 0) and return the
 * team size the next parallel region will use */
int orc_threads(int n)
{
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
}
