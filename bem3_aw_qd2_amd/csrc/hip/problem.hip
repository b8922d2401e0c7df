// problem.hip -- creation, transfer and release of a device-resident problem.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <utility>
#include <vector>
#include "problem.h"

int bqd2_alloc(bqd2_problem *p, void **ptr, size_t bytes) {
  if (bytes == 0) bytes = 16;
  hipError_t e = hipMalloc(ptr, bytes);
  if (e != hipSuccess) {
    bqd2_set_error("hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
    return BQD2_E_NOMEM;
  }
  p->owned.push_back(*ptr);
  p->alloc_bytes += bytes;
  return BQD2_OK;
}

void bqd2_release(bqd2_problem *p, void **ptr, size_t bytes) {
  if (!*ptr) return;
  for (auto it = p->owned.begin(); it != p->owned.end(); ++it)
    if (*it == *ptr) {
      p->owned.erase(it);
      p->alloc_bytes -= bytes ? bytes : 16;
      break;
    }
  (void)hipFree(*ptr);
  *ptr = nullptr;
}

static bqd2_progress_fn g_progress = nullptr;
static void *g_progress_ctx = nullptr;
extern "C" int bqd2_set_progress(bqd2_progress_fn fn, void *ctx) {
  g_progress = fn;
  g_progress_ctx = ctx;
  return BQD2_OK;
}
bool bqd2_progress_on() { return g_progress != nullptr; }
void bqd2_progress(const char *stage, long long done, long long total) {
  bqd2_progress_fn fn = g_progress;
  if (fn) fn(stage, done, total, g_progress_ctx);
}

void bqd2_rebind(bqd2_problem *p) { p->owner = std::this_thread::get_id(); }

size_t bqd2_bytes(bqd2_problem *p) {
  std::lock_guard<std::mutex> g(p->ctx_mu);
  size_t b = p->alloc_bytes + p->main.ws_bytes;
  for (auto &tc : p->ctxs) b += tc.second->ws_bytes + CNT_N * sizeof(long long);
  return b;
}

extern "C" int bqd2_device_bytes(bqd2_problem *p, size_t *bytes) {
  BQD2_CHECK(p && bytes, BQD2_E_ARG, "bqd2_device_bytes: null argument");
  *bytes = bqd2_bytes(p);
  return BQD2_OK;
}

template <class T>
static int upload(bqd2_problem *p, T **dst, const T *src, size_t n) {
  int rc = bqd2_alloc(p, (void **)dst, n * sizeof(T));
  if (rc) return rc;
  BQD2_HIP(hipMemcpyAsync(*dst, src, n * sizeof(T), hipMemcpyHostToDevice, p->stream));
  BQD2_HIP(hipStreamSynchronize(p->stream));  // the host array may be released on return
  return BQD2_OK;
}

#define TRY(x)                \
  do {                        \
    int rc_ = (x);            \
    if (rc_) {                \
      bqd2_destroy(p);        \
      return rc_;             \
    }                         \
  } while (0)

// the system layout needs one sphere of quadrilaterals in a periodic lattice,
// each element with its own four nodes in element order, domain 0 the lattice
// side of every element and domain 1 the sphere side
static int large_layout_ok(const bqd2_model *m) {
  BQD2_CHECK(m->MN == 1 && m->ps == 1, BQD2_E_ARG,
             "bqd2_create_large: one sphere (MN = 1) in a periodic lattice (ps = 1) required, got MN=%d ps=%d", m->MN,
             m->ps);
  BQD2_CHECK(m->NN == 4 * m->Ne && m->dom_ne[0] == m->Ne && m->dom_ne[1] == m->Ne, BQD2_E_ARG,
             "bqd2_create_large: every element must bound both domains and carry four nodes");
  const int *sid0 = m->dom_sid, *sid1 = m->dom_sid + m->Ne + 1;
  for (int s = 1; s <= m->Ne; s++) {
    BQD2_CHECK(m->ed[4 * s + 3] != 0, BQD2_E_ARG, "bqd2_create_large: element %d is a triangle", s);
    BQD2_CHECK(m->md[s] == 0 && m->sd[s] == 1, BQD2_E_ARG, "bqd2_create_large: element %d does not face domain 0", s);
    BQD2_CHECK(sid0[s] == s && sid1[s] == -s, BQD2_E_ARG, "bqd2_create_large: domain element lists out of order");
    for (int l = 0; l < 4; l++)
      BQD2_CHECK(m->eni[4 * s + l] == 4 * (s - 1) + l, BQD2_E_ARG, "bqd2_create_large: node %d of element %d", l, s);
  }
  return BQD2_OK;
}

static int create_impl(const bqd2_model *m, int large, bqd2_problem **out);

// Gauss-Legendre nodes and weights by Newton's iteration on the three-term
// recurrence of the Legendre polynomials (Abramowitz and Stegun 25.4.29), from
// the starting values cos(pi (i + 3/4) / (n + 1/2))
extern "C" int bqd2_gauss_legendre(int n, double *x, double *w) {
  BQD2_CHECK(n >= 1 && n <= 64 && x && w, BQD2_E_ARG, "bqd2_gauss_legendre: bad arguments");
  for (int i = 0; i < (n + 1) / 2; i++) {
    double z = cos(M_PI * (i + 0.75) / (n + 0.5)), dp = 1.0;
    for (int it = 0; it < 100; it++) {
      double p0 = 1.0, p1 = 0.0;
      for (int j = 1; j <= n; j++) {
        const double pm = p1;
        p1 = p0;
        p0 = ((2.0 * j - 1.0) * z * p1 - (j - 1.0) * pm) / j;
      }
      dp = n * (z * p0 - p1) / (z * z - 1.0);
      const double dz = p0 / dp;
      z -= dz;
      if (fabs(dz) <= 1e-17) break;
    }
    if (2 * i + 1 == n) z = 0.0;
    const double wi = 2.0 / ((1.0 - z * z) * dp * dp);
    x[i] = -z;
    x[n - 1 - i] = z;
    w[i] = w[n - 1 - i] = wi;
  }
  return BQD2_OK;
}

// Far-field geometry of the periodic domain.  Above z_top + h (below z_bot - h)
// every source point lies at least h lower (higher), so the lattice sum's
// spectral series converges for every (point, source) pair of the field and
// its order-n term is a product of a target factor and a source factor; the
// orders kept are those with |a_n| <= sqrt(k^2 + (FAR_DECAY / h)^2): the ones
// beyond decay by exp(-FAR_DECAY) = 1e-20 or more over the height h.  h is the
// smallest of 1/12, 1/8, 1/6 and 1/4 of the shorter lattice vector whose grid
// stays within FAR_MAX_ORDERS (1/12: about 32 000 orders on lattices of equal
// sides).
// A smaller h lets more of the element bands of points level with the boundary
// take the amplitudes; each point then sums more orders, which costs far less
// than the lattice sums it replaces.
int far_grid(const double *qd8, int *n1, int *n2, double *hp) {
  const double k = qd8[0], vk = hypot(qd8[1], qd8[2]);
  const double l1 = hypot(qd8[4], qd8[5]), l2 = hypot(qd8[6], qd8[7]);
  const double FAR_DECAY = 46.0;
  for (const double f : {1.0 / 12.0, 0.125, 1.0 / 6.0, 0.25}) {
    const double h = f * std::min(l1, l2);
    if (!(h > 0.0)) return 0;
    const double R = sqrt(k * k + (FAR_DECAY / h) * (FAR_DECAY / h)) + vk;
    // a_n + vk = n1 g1 + n2 g2 with g_i . d_j = 2 pi delta_ij: |n_i| <= R |d_i| / 2 pi
    const double N1 = ceil(R * l1 / (2.0 * M_PI)), N2 = ceil(R * l2 / (2.0 * M_PI));
    const double G = (2.0 * N1 + 1.0) * (2.0 * N2 + 1.0);
    if (G > FAR_MAX_ORDERS) continue;
    if (n1) *n1 = (int)N1;
    if (n2) *n2 = (int)N2;
    if (hp) *hp = h;
    return (int)G;
  }
  return 0;
}

static void far_setup(bqd2_problem *p, const bqd2_model *m) {
  FarGeom &F = p->far;
  memset(&F, 0, sizeof(F));
  const int ne0 = m->dom_ne[0];
  const int *sid = m->dom_sid;  // domain 0's list comes first
  double zt = -1e300, zb = 1e300;
  for (int t = 1; t <= ne0; t++) {
    const int s = abs(sid[t]);
    for (int j = 0; j < 4; j++) {
      const int nd = m->ed[4 * (size_t)s + j];
      if (j == 3 && nd == 0) continue;  // linear triangle
      const double z = m->rn[3 * (size_t)nd + 2];
      zt = std::max(zt, z);
      zb = std::min(zb, z);
    }
  }
  int N1, N2;
  double h;
  const int G = far_grid(m->qd, &N1, &N2, &h);
  if (ne0 <= 0 || !(zt >= zb) || G == 0) return;
  // the bands of the field order: element ranges, and the extreme node heights
  // of the bands below and above each
  FarBand B[FAR_NB];
  double bmax[FAR_NB], bmin[FAR_NB];
  for (int b = 0; b < FAR_NB; b++) {
    bmax[b] = -1e300;
    bmin[b] = 1e300;
    B[b].es = B[b].ee = 0;
  }
  int pos = 0;
  for (int b = 0; b < FAR_NB; b++) {
    B[b].es = pos;
    while (pos < ne0 && p->fband0[pos] == b) {
      const int s = abs(sid[p->fperm0[pos]]);
      for (int j = 0; j < 4; j++) {
        const int nd = m->ed[4 * (size_t)s + j];
        if (j == 3 && nd == 0) continue;
        const double z = m->rn[3 * (size_t)nd + 2];
        bmax[b] = std::max(bmax[b], z);
        bmin[b] = std::min(bmin[b], z);
      }
      pos++;
    }
    B[b].ee = pos;
  }
  if (pos != ne0) return;  // the band order does not cover the domain
  for (int b = 0; b < FAR_NB; b++) B[b].zpre = std::max(bmax[b], b ? B[b - 1].zpre : -1e300);
  for (int b = FAR_NB - 1; b >= 0; b--) B[b].zsuf = std::min(bmin[b], b + 1 < FAR_NB ? B[b + 1].zsuf : 1e300);
  if (upload(p, &p->far_bands, B, FAR_NB)) return;
  F.n1lo = -N1;
  F.n1n = 2 * N1 + 1;
  F.n2lo = -N2;
  F.n2n = 2 * N2 + 1;
  F.G = (int)G;
  F.ztop = zt;
  F.zbot = zb;
  F.h = h;
  F.on = 1;
}

extern "C" int bqd2_create(const bqd2_model *m, bqd2_problem **out) { return create_impl(m, 0, out); }

extern "C" int bqd2_create_large(const bqd2_model *m, bqd2_problem **out) {
  BQD2_CHECK(m && out, BQD2_E_ARG, "bqd2_create_large: null argument");
  int rc = large_layout_ok(m);
  if (rc) return rc;
  return create_impl(m, 1, out);
}

static int create_impl(const bqd2_model *m, int large, bqd2_problem **out) {
  BQD2_CHECK(m && out, BQD2_E_ARG, "bqd2_create: null argument");
  BQD2_CHECK(m->Ne > 0 && m->MN >= 0 && m->MN < 16, BQD2_E_ARG, "bqd2_create: bad sizes Ne=%d MN=%d", m->Ne, m->MN);
  bqd2_problem *p = new bqd2_problem();
  *out = nullptr;
  p->large = large;
  p->lwin = 2048;
  p->lmax_resident = -1;
  // the problem's stream does not wait for the null stream, nor it for the
  // problem: requests of a serving loop on other problems never wait for it
  // (every transfer of the problem is ordered on this stream)
  BQD2_HIP(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
  p->m.Ne = m->Ne;
  p->m.NN = m->NN;
  p->m.MN = m->MN;
  p->m.ps = m->ps;
  const size_t ne1 = (size_t)m->Ne + 1;
  std::vector<int> tri(ne1, 0);
  for (size_t s = 1; s < ne1; s++) tri[s] = (m->ed[4 * s + 3] == 0);
  TRY(upload(p, (double **)&p->m.cr, m->cr, ne1 * 12));
  TRY(upload(p, (double **)&p->m.cw, m->cw, ne1 * 9));
  TRY(upload(p, (double **)&p->m.ren, m->ren, ne1 * 12));
  TRY(upload(p, (double **)&p->m.wen, m->wen, ne1 * 12));
  TRY(upload(p, (int **)&p->m.tri, tri.data(), ne1));
  TRY(upload(p, &p->rn, m->rn, ((size_t)m->Nn + 1) * 3));
  TRY(upload(p, &p->ed, m->ed, ne1 * 4));
  TRY(upload(p, (int **)&p->m.eni, m->eni, ne1 * 4));
  TRY(upload(p, (int **)&p->m.md, m->md, ne1));
  TRY(upload(p, (int **)&p->m.sd, m->sd, ne1));
  TRY(upload(p, (double **)&p->m.Pi, m->Pi, ne1 * 8));
  TRY(upload(p, (double **)&p->m.dPi, m->dPi, ne1 * 8));

  Tab &T = p->htab;
  memcpy(T.xli, m->xli, sizeof(T.xli)); memcpy(T.wli, m->wli, sizeof(T.wli));
  memcpy(T.xhi, m->xhi, sizeof(T.xhi)); memcpy(T.whi, m->whi, sizeof(T.whi));
  memcpy(T.zt_44, m->zt_44, sizeof(T.zt_44)); memcpy(T.et_44, m->et_44, sizeof(T.et_44));
  memcpy(T.wt_44, m->wt_44, sizeof(T.wt_44)); memcpy(T.zt_34, m->zt_34, sizeof(T.zt_34));
  memcpy(T.et_34, m->et_34, sizeof(T.et_34)); memcpy(T.wt_34, m->wt_34, sizeof(T.wt_34));
  memcpy(T.zt_49, m->zt_49, sizeof(T.zt_49)); memcpy(T.et_49, m->et_49, sizeof(T.et_49));
  memcpy(T.wt_49, m->wt_49, sizeof(T.wt_49)); memcpy(T.zt_37, m->zt_37, sizeof(T.zt_37));
  memcpy(T.et_37, m->et_37, sizeof(T.et_37)); memcpy(T.wt_37, m->wt_37, sizeof(T.wt_37));
  TRY(bqd2_gauss_legendre(4, T.xs4, T.ws4));
  TRY(bqd2_gauss_legendre(6, T.xs6, T.ws6));
  TRY(bqd2_gauss_legendre(8, T.xs8, T.ws8));
  TRY(upload(p, &p->tab, &T, 1));

  for (int d = 0; d <= m->MN; d++) {
    p->k0[d] = m->k0[d];
    p->rho0[d] = m->rho0[d];
    p->k2[d][0] = m->k2[2 * d];
    p->k2[d][1] = m->k2[2 * d + 1];
  }
  memcpy(p->pw_p, m->pw_p, sizeof(p->pw_p));
  p->pw_rho0 = m->pw_rho0; p->pw_c0 = m->pw_c0; p->pw_k0 = m->pw_k0;
  memcpy(p->pw_tv, m->pw_tv, sizeof(p->pw_tv));
  memcpy(p->pw_R, m->pw_R, sizeof(p->pw_R));
  memcpy(p->qd8, m->qd, sizeof(p->qd8));
  if (m->ps == 1) TRY(bqd2_gf_params(m->qd, &p->gf));
  else memset(&p->gf, 0, sizeof(p->gf));

  TRY(bqd2_alloc(p, (void **)&p->d_counters, CNT_N * sizeof(long long)));

  size_t off = 0;
  p->dom.resize(m->MN + 1);
  for (int d = 0; d <= m->MN; d++) {
    DevDomain &D = p->dom[d];
    memset(&D, 0, sizeof(D));
    D.did = d;
    D.Ne = m->dom_ne[d];
    D.N = 4 * D.Ne;
    D.k = m->k0[d];
    const int *sid = m->dom_sid + off;
    off += D.Ne + 1;
    int nn = 0;
    std::vector<int> rowmap(D.N > 0 ? D.N : 1, -1);
    p->tripos.resize(m->MN + 1);
    p->tripos[d].assign((size_t)D.Ne + 1, 0);
    for (int t = 1; t <= D.Ne; t++) p->tripos[d][t] = tri[abs(sid[t])] ? 1 : 0;
    for (int t = 1; t <= D.Ne; t++) {
      int nl = tri[abs(sid[t])] ? 3 : 4;
      if (nl == 3) D.has_tri = p->any_tri = 1;
      for (int tn = 0; tn < nl; tn++) rowmap[(t - 1) * 4 + tn] = nn++;
    }
    D.nn = nn;
    TRY(upload(p, &D.sid, sid, (size_t)D.Ne + 1));
    TRY(upload(p, &D.rowmap, rowmap.data(), rowmap.size()));
    {
      // the field kernels visit a point's elements in bands of the element
      // centre's z (32 over the domain's extent), by azimuth within a band, so
      // that the lanes of a wavefront hold neighbouring elements and their
      // lattice sums need about as many terms
      std::vector<int> ord(D.Ne > 0 ? D.Ne : 1);
      double z0 = 1e300, z1 = -1e300;
      for (int t = 1; t <= D.Ne; t++) {
        const double z = m->cr[12 * (size_t)abs(sid[t]) + 8];
        z0 = std::min(z0, z);
        z1 = std::max(z1, z);
      }
      const double zs = (z1 > z0) ? 32.0 / (z1 - z0) : 0.0;
      std::vector<std::pair<std::pair<int, double>, int>> key((size_t)D.Ne);
      for (int t = 1; t <= D.Ne; t++) {
        const double *c = m->cr + 12 * (size_t)abs(sid[t]);
        const int band = std::min(31, (int)((c[8] - z0) * zs));
        key[t - 1] = {{band, atan2(c[4], c[0])}, t};
      }
      std::stable_sort(key.begin(), key.end());
      for (int i = 0; i < D.Ne; i++) ord[i] = key[i].second;
      TRY(upload(p, &D.fperm, ord.data(), ord.size()));
      if (d == 0) {
        p->fband0.assign((size_t)D.Ne, 0);
        for (int i = 0; i < D.Ne; i++) p->fband0[i] = key[i].first.first;
        p->fperm0.assign(ord.begin(), ord.begin() + D.Ne);
      }
    }
    const size_t N = D.N;
    const bool periodic0 = (d == 0 && m->ps == 1);
    D.ldm = D.N;
    D.r0 = 0;
    D.nr = D.N;
    if (!large) {
      TRY(bqd2_alloc(p, (void **)&D.G, N * N * 16));
      TRY(bqd2_alloc(p, (void **)&D.H, N * N * 16));
      if (!periodic0) {
        TRY(bqd2_alloc(p, (void **)&D.dG, 2 * N * N * 16));
        TRY(bqd2_alloc(p, (void **)&D.dH, 2 * N * N * 16));
      }
      TRY(bqd2_alloc(p, (void **)&D.Fpair, N * D.Ne * 3 * 8));
      TRY(bqd2_alloc(p, (void **)&D.pkey, N * D.Ne * sizeof(float)));
    }
    TRY(bqd2_alloc(p, (void **)&D.dF, N * 3 * 8));
    TRY(bqd2_alloc(p, (void **)&D.rule, N * D.Ne));
    if (!large || periodic0) TRY(bqd2_alloc(p, (void **)&D.plist, (N * D.Ne + PLIST_PAD) * sizeof(int)));
    TRY(bqd2_alloc(p, (void **)&D.P, ((size_t)D.Ne + 1) * 4 * 16));
    TRY(bqd2_alloc(p, (void **)&D.dP, ((size_t)D.Ne + 1) * 4 * 16));
    TRY(bqd2_alloc(p, (void **)&D.pv, ((size_t)D.Ne + 1) * 12 * 16));
    TRY(bqd2_alloc(p, (void **)&D.dpv, ((size_t)D.Ne + 1) * 12 * 16));
    BQD2_HIP(hipMemsetAsync(D.P, 0, ((size_t)D.Ne + 1) * 4 * 16, p->stream));
    BQD2_HIP(hipMemsetAsync(D.dP, 0, ((size_t)D.Ne + 1) * 4 * 16, p->stream));
    BQD2_HIP(hipMemsetAsync(D.pv, 0, ((size_t)D.Ne + 1) * 12 * 16, p->stream));
    BQD2_HIP(hipMemsetAsync(D.dpv, 0, ((size_t)D.Ne + 1) * 12 * 16, p->stream));
  }
  p->far_enabled = 1;
  p->smooth_mode = 1;
  if (m->ps == 1) far_setup(p, m);
  p->na = 2 * m->NN;
  if (large) {
    // S = [H0 G0; H1 G1], row-major, na = 2N columns per row; the pair kernels
    // write straight into it (domain matrices with row stride na)
    const size_t na = (size_t)p->na, N = (size_t)p->dom[0].N;
    TRY(bqd2_alloc(p, (void **)&p->S, na * na * 16));
    for (int d = 0; d < 2; d++) {
      DevDomain &D = p->dom[d];
      D.H = p->S + 2 * (size_t)d * N * na;
      D.G = D.H + 2 * N;
      D.ldm = (long long)na;
    }
    p->s_lower = LG_EMPTY;
  }
  p->A = p->b = nullptr;
  p->ipiv = nullptr;
  p->solved = false;
  p->main.st = p->stream;
  p->main.ws = nullptr;
  p->main.ws_bytes = 0;
  p->main.cnt = p->d_counters;
  p->owner = std::this_thread::get_id();
  // the host arrays of the model are the caller's: every upload done on return
  BQD2_HIP(hipStreamSynchronize(p->stream));
  *out = p;
  return BQD2_OK;
}

extern "C" int bqd2_destroy(bqd2_problem *p) {
  if (!p) return BQD2_OK;
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  for (auto &tc : p->ctxs) {
    CallCtx *c = tc.second;
    (void)hipStreamSynchronize(c->st);
    if (c->ws) (void)hipFree(c->ws);
    (void)hipFree(c->cnt);
    (void)hipStreamDestroy(c->st);
    delete c;
  }
  if (p->lu_graph) (void)hipGraphExecDestroy(p->lu_graph);
  for (int t = 0; t < 4; t++)
    if (p->far_ev[t]) (void)hipEventDestroy(p->far_ev[t]);
  for (auto &D : p->dom)
    if (D.g_graph) (void)hipGraphExecDestroy(D.g_graph);
  for (void *ptr : p->owned) (void)hipFree(ptr);
  if (p->main.ws) (void)hipFree(p->main.ws);
  if (p->stream) (void)hipStreamDestroy(p->stream);
  delete p;
  return BQD2_OK;
}

extern "C" int bqd2_download_domain(bqd2_problem *p, int did, double *G, double *H, double *dG, double *dH, double *dF) {
  BQD2_CHECK(p && did >= 0 && did <= p->m.MN, BQD2_E_ARG, "bqd2_download_domain: bad domain");
  DevDomain &D = p->dom[did];
  BQD2_CHECK(D.assembled, BQD2_E_STATE, "bqd2_download_domain: domain %d not assembled", did);
  const size_t N = D.N;
  BQD2_HIP(hipStreamSynchronize(p->stream));
  BQD2_CHECK(D.r0 == 0 && D.nr == D.N && !(p->large && did == 1 && p->s_lower != LG_G1H1), BQD2_E_STATE,
             "bqd2_download_domain: domain %d's rows are not held", did);
  if (G) BQD2_HIP(hipMemcpy2DAsync(G, N * 16, D.G, (size_t)D.ldm * 16, N * 16, N, hipMemcpyDeviceToHost, p->stream));
  if (H) BQD2_HIP(hipMemcpy2DAsync(H, N * 16, D.H, (size_t)D.ldm * 16, N * 16, N, hipMemcpyDeviceToHost, p->stream));
  if (dG && D.dG) BQD2_HIP(hipMemcpyAsync(dG, D.dG, 2 * N * N * 16, hipMemcpyDeviceToHost, p->stream));
  if (dH && D.dH) BQD2_HIP(hipMemcpyAsync(dH, D.dH, 2 * N * N * 16, hipMemcpyDeviceToHost, p->stream));
  if (dF) BQD2_HIP(hipMemcpyAsync(dF, D.dF, N * 3 * 8, hipMemcpyDeviceToHost, p->stream));
  BQD2_HIP(hipStreamSynchronize(p->stream));  // the caller reads or reuses the host buffers
  return BQD2_OK;
}

extern "C" int bqd2_download_rows(bqd2_problem *p, int did, int nrows, const int *rows, double *G, double *H,
                                  double *dG, double *dH, double *dF) {
  BQD2_CHECK(p && did >= 0 && did <= p->m.MN && nrows >= 0 && (nrows == 0 || rows), BQD2_E_ARG,
             "bqd2_download_rows: bad arguments");
  DevDomain &D = p->dom[did];
  BQD2_CHECK(D.assembled, BQD2_E_STATE, "bqd2_download_rows: domain %d not assembled", did);
  const size_t N = D.N;
  BQD2_CHECK(D.r0 == 0 && D.nr == D.N && !(p->large && did == 1 && p->s_lower != LG_G1H1), BQD2_E_STATE,
             "bqd2_download_rows: domain %d's rows are not held", did);
  for (int i = 0; i < nrows; i++)
    BQD2_CHECK(rows[i] >= 0 && (size_t)rows[i] < N, BQD2_E_ARG, "bqd2_download_rows: row %d out of range", rows[i]);
  BQD2_HIP(hipStreamSynchronize(p->stream));
  for (int i = 0; i < nrows; i++) {
    const size_t r = (size_t)rows[i];
    if (G) BQD2_HIP(hipMemcpyAsync(G + 2 * N * i, D.G + 2 * dom_gh(D, (int)r, 0), N * 16, hipMemcpyDeviceToHost, p->stream));
    if (H) BQD2_HIP(hipMemcpyAsync(H + 2 * N * i, D.H + 2 * dom_gh(D, (int)r, 0), N * 16, hipMemcpyDeviceToHost, p->stream));
    if (dG && D.dG) BQD2_HIP(hipMemcpyAsync(dG + 4 * N * i, D.dG + 4 * N * r, 2 * N * 16, hipMemcpyDeviceToHost, p->stream));
    if (dH && D.dH) BQD2_HIP(hipMemcpyAsync(dH + 4 * N * i, D.dH + 4 * N * r, 2 * N * 16, hipMemcpyDeviceToHost, p->stream));
    if (dF) BQD2_HIP(hipMemcpyAsync(dF + 3 * i, D.dF + 3 * r, 3 * 8, hipMemcpyDeviceToHost, p->stream));
  }
  BQD2_HIP(hipStreamSynchronize(p->stream));  // the caller reads or reuses the host buffers
  return BQD2_OK;
}

extern "C" int bqd2_download_solution(bqd2_problem *p, int did, double *P, double *dP, double *pv, double *dpv) {
  BQD2_CHECK(p && did >= 0 && did <= p->m.MN, BQD2_E_ARG, "bqd2_download_solution: bad domain");
  DevDomain &D = p->dom[did];
  const size_t n1 = (size_t)D.Ne + 1;
  BQD2_HIP(hipStreamSynchronize(p->stream));
  if (P) BQD2_HIP(hipMemcpyAsync(P, D.P, n1 * 4 * 16, hipMemcpyDeviceToHost, p->stream));
  if (dP) BQD2_HIP(hipMemcpyAsync(dP, D.dP, n1 * 4 * 16, hipMemcpyDeviceToHost, p->stream));
  if (pv) BQD2_HIP(hipMemcpyAsync(pv, D.pv, n1 * 12 * 16, hipMemcpyDeviceToHost, p->stream));
  if (dpv) BQD2_HIP(hipMemcpyAsync(dpv, D.dpv, n1 * 12 * 16, hipMemcpyDeviceToHost, p->stream));
  BQD2_HIP(hipStreamSynchronize(p->stream));  // the caller reads or reuses the host buffers
  return BQD2_OK;
}

extern "C" int bqd2_upload_solution(bqd2_problem *p, int did, const double *P, const double *dP, const double *pv,
                                    const double *dpv) {
  BQD2_CHECK(p && did >= 0 && did <= p->m.MN, BQD2_E_ARG, "bqd2_upload_solution: bad domain");
  p->sol_epoch++;
  DevDomain &D = p->dom[did];
  const size_t n1 = (size_t)D.Ne + 1;
  BQD2_HIP(hipStreamSynchronize(p->stream));
  if (P) BQD2_HIP(hipMemcpyAsync(D.P, P, n1 * 4 * 16, hipMemcpyHostToDevice, p->stream));
  if (dP) BQD2_HIP(hipMemcpyAsync(D.dP, dP, n1 * 4 * 16, hipMemcpyHostToDevice, p->stream));
  if (pv) BQD2_HIP(hipMemcpyAsync(D.pv, pv, n1 * 12 * 16, hipMemcpyHostToDevice, p->stream));
  if (dpv) BQD2_HIP(hipMemcpyAsync(D.dpv, dpv, n1 * 12 * 16, hipMemcpyHostToDevice, p->stream));
  BQD2_HIP(hipStreamSynchronize(p->stream));  // the caller reads or reuses the host buffers
  return BQD2_OK;
}

extern "C" int bqd2_counters(bqd2_problem *p, long long *out, int n) {
  BQD2_CHECK(p && out && n > 0 && n <= CNT_N, BQD2_E_ARG, "bqd2_counters: bad arguments");
  BQD2_HIP(hipStreamSynchronize(p->stream));
  BQD2_HIP(hipMemcpyAsync(out, p->d_counters, (size_t)n * sizeof(long long), hipMemcpyDeviceToHost, p->stream));
  BQD2_HIP(hipStreamSynchronize(p->stream));  // the caller reads or reuses the host buffers
  return BQD2_OK;
}

int bqd2_workspace(CallCtx *c, size_t bytes, char **base) {
  if (bytes > c->ws_bytes) {
    BQD2_HIP(hipStreamSynchronize(c->st));
    if (c->ws) BQD2_HIP(hipFree(c->ws));
    c->ws = nullptr;
    c->ws_bytes = 0;
    size_t grow = bytes + bytes / 2;
    hipError_t e = hipMalloc((void **)&c->ws, grow);
    if (e != hipSuccess) {
      bqd2_set_error("hipMalloc(%zu bytes) of the workspace failed: %s", grow, hipGetErrorString(e));
      return BQD2_E_NOMEM;
    }
    c->ws_bytes = grow;
  }
  *base = c->ws;
  return BQD2_OK;
}

CallCtx *bqd2_call_ctx(bqd2_problem *p) {
  const std::thread::id me = std::this_thread::get_id();
  if (me == p->owner) return &p->main;
  std::lock_guard<std::mutex> g(p->ctx_mu);
  for (auto &tc : p->ctxs)
    if (tc.first == me) return tc.second;
  CallCtx *c = new CallCtx();
  c->ws = nullptr;
  c->ws_bytes = 0;
  if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) {
    bqd2_set_error("hipStreamCreate failed for a calling thread");
    delete c;
    return nullptr;
  }
  if (hipMalloc((void **)&c->cnt, CNT_N * sizeof(long long)) != hipSuccess) {
    bqd2_set_error("hipMalloc of the thread's counters failed");
    (void)hipStreamDestroy(c->st);
    delete c;
    return nullptr;
  }
  p->ctxs.push_back({me, c});
  return c;
}
