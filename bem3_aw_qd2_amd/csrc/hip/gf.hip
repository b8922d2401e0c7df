// gf.hip -- instance set-up and the batched Green's function entry point.
#include <cmath>
#include <cstdarg>
#include <cstring>
#include "common.h"
#include "gf_dev.h"

static thread_local char g_errbuf[512];

void bqd2_set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_errbuf, sizeof(g_errbuf), fmt, ap);
  va_end(ap);
}

extern "C" const char *bqd2_last_error(void) { return g_errbuf; }

// Host part of the method choice and Ewald parameter of d3_qpgf_d2.c:124-176 and
// :1671-1693: everything that does not depend on the evaluation point.
extern "C" int bqd2_gf_params(const double *qd8, GfParams *p) {
  BQD2_CHECK(qd8 && p, BQD2_E_ARG, "bqd2_gf_params: null argument");
  const double k = qd8[0], vkx = qd8[1], vky = qd8[2];
  const double d1x = qd8[4], d1y = qd8[5], d2x = qd8[6], d2y = qd8[7];
  const double det = d1x * d2y - d2x * d1y;
  BQD2_CHECK(det != 0.0, BQD2_E_ARG, "bqd2_gf_params: degenerate lattice (det == 0)");
  BQD2_CHECK(k > 0.0, BQD2_E_ARG, "bqd2_gf_params: wave number must be positive");
  memset(p, 0, sizeof(*p));
  p->k = k; p->vkx = vkx; p->vky = vky;
  p->d1x = d1x; p->d1y = d1y; p->d2x = d2x; p->d2y = d2y;
  p->idd = 1.0 / det;
  p->aidd = fabs(p->idd);
  const double tp = 2.0 * M_PI;
  p->g1x = tp * p->idd * d2y;
  p->g1y = -tp * p->idd * d2x;
  p->g2x = -tp * p->idd * d1y;
  p->g2y = tp * p->idd * d1x;
  p->d2n = d2x * d2x + d2y * d2y;
  p->g2n = p->g2x * p->g2x + p->g2y * p->g2y;
  p->g12 = p->g1x * p->g2x + p->g1y * p->g2y;
  {
    double sx = d1x + d2x, sy = d1y + d2y, dx = d1x - d2x, dy = d1y - d2y;
    double adp = sqrt(sx * sx + sy * sy), adm = sqrt(dx * dx + dy * dy);
    p->E0 = sqrt(fabs(det) / (4.0 * M_PI) * adp / adm);
  }
  p->nc = (int)round((vkx * d1x + vky * d1y) / tp);
  {
    const double k2 = k * k, cpd = tp * p->idd;
    const double ad12 = d1x * d1x + d1y * d1y, ad22 = d2x * d2x + d2y * d2y;
    int n1 = (int)round((vkx * d1x + vky * d1y) / tp);
    int n2 = (int)round((vkx * d2x + vky * d2y) / tp);
    double aNx = -vkx + n1 * cpd * d2y - n2 * cpd * d1y;
    double aNy = -vky - n1 * cpd * d2x + n2 * cpd * d1x;
    int p1 = n1, m1 = n1, p2 = n2, m2 = n2;
    if (ad22 < ad12) { p1 = n1 + 15; m1 = n1 - 15; }
    else { p2 = n2 + 15; m2 = n2 - 15; }
    double apx = -vkx + p1 * cpd * d2y - p2 * cpd * d1y;
    double apy = -vky - p1 * cpd * d2x + p2 * cpd * d1x;
    double amx = -vkx + m1 * cpd * d2y - m2 * cpd * d1y;
    double amy = -vky - m1 * cpd * d2x + m2 * cpd * d1x;
    double bNp = k2 - apx * apx - apy * apy;
    double bNm = k2 - amx * amx - amy * amy;
    p->bN = sqrt(k2 - aNx * aNx - aNy * aNy);
    p->ewald_only = (bNp > 0.0 || bNm > 0.0);
    p->bNp = p->ewald_only ? 1.0 : sqrt(-bNp);
    p->bNm = p->ewald_only ? 1.0 : sqrt(-bNm);
  }
  // Just above the reference's switch height the pure spectral sum needs some
  // 900 orders where the Ewald sum needs ~110 terms of about twice the cost;
  // both converge to the same value (1e-15), so the pure spectral sum is taken
  // from 2.2 times that height on, where the two costs meet.
  p->fd_lift = 2.2;
  return BQD2_OK;
}

template <int M>
__global__ void __launch_bounds__(256) k_gf_batch(GfParams P, int n, const double *__restrict__ r, double eps,
                                                  double *__restrict__ G, double *__restrict__ dG,
                                                  double *__restrict__ d2G, int *__restrict__ terms) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Sum10 s;
  int c = qpgf_eval<M>(P, r[3 * i], r[3 * i + 1], r[3 * i + 2], eps, s);
  if (G) { G[2 * i] = s.g.x; G[2 * i + 1] = s.g.y; }
  if (dG && M != GF_QG) {
    for (int j = 0; j < 3; j++) { dG[6 * i + 2 * j] = s.dg[j].x; dG[6 * i + 2 * j + 1] = s.dg[j].y; }
  }
  if (d2G && M == GF_V2) {
    for (int j = 0; j < 6; j++) { d2G[12 * i + 2 * j] = s.d2g[j].x; d2G[12 * i + 2 * j + 1] = s.d2g[j].y; }
  }
  if (terms) terms[i] = c;
}

extern "C" int bqd2_gf_batch(const GfParams *p, int mode, int n, const double *d_r, double eps, double *d_G,
                             double *d_dG, double *d_d2G, int *d_terms, void *stream) {
  BQD2_CHECK(p && d_r && n >= 0, BQD2_E_ARG, "bqd2_gf_batch: bad arguments");
  BQD2_CHECK(mode >= 0 && mode <= 3, BQD2_E_ARG, "bqd2_gf_batch: mode must be 0..3");
  if (n == 0) return BQD2_OK;
  hipStream_t st = (hipStream_t)stream;
  dim3 grid(bqd2_blocks(n, 256)), blk(256);
  switch (mode) {
    case 0: hipLaunchKernelGGL(k_gf_batch<GF_QG>, grid, blk, 0, st, *p, n, d_r, eps, d_G, d_dG, d_d2G, d_terms); break;
    case 1: hipLaunchKernelGGL(k_gf_batch<GF_DQG>, grid, blk, 0, st, *p, n, d_r, eps, d_G, d_dG, d_d2G, d_terms); break;
    case 2: hipLaunchKernelGGL(k_gf_batch<GF_V1>, grid, blk, 0, st, *p, n, d_r, eps, d_G, d_dG, d_d2G, d_terms); break;
    default: hipLaunchKernelGGL(k_gf_batch<GF_V2>, grid, blk, 0, st, *p, n, d_r, eps, d_G, d_dG, d_d2G, d_terms); break;
  }
  BQD2_HIP(hipGetLastError());
  return BQD2_OK;
}
