// serve.hip -- serving loop for boundary-value solve requests on one device.
//
// The reference solves one model per process (aw_d3qd2_bv_solver.c:5-11:
// read, initialise, solve_bieq_aqd2, write), then evaluates fields and forces
// from the stored solution (example1.c:10-48).  A request here is that whole
// sequence for one model and incident plane wave: assembly, surface solve,
// velocity recovery, optional field points and radiation force.
//
// Continuous admission: worker threads ("slots") take requests from a FIFO as
// slots and device memory free up.  A request whose operator is not resident
// needs its footprint (bqd2_footprint) within the device-memory budget; idle
// resident operators are evicted, least recently used first, to make room.
// Misses are admitted in arrival order (a miss that does not fit holds back the
// later misses, so large requests are not starved); requests whose operator is
// resident and idle run as soon as a slot is free.  Each problem has its own
// HIP stream, so the kernels of concurrent requests overlap on the device.
//
// Resident operators: a request with a nonzero operator key promises that
// requests with the same key share the mesh, media, frequency and lattice, so
// that only the incident wave differs; the assembled matrices (and, for the
// mixed-precision solver, the complex float factors) of the first are kept and
// the next ones only form and solve their right-hand side.  A resident operator
// serves a request only when the request's in-plane wave vector (kx, ky), which
// the lattice sums take as their Bloch phase, has the operator's exact bits:
// where a quadrature point is level with a source (z == 0) the lattice sums
// stop on a rounding residue and so follow the last bits of (kx, ky), and a
// wave vector equal only up to rounding (sin(pi - theta) against sin(theta))
// would move the result.  Such a request is a miss and gets an operator of its
// own.  The rest of the promise is checked (element count, frequency,
// lattice); a mismatch fails the request.
//
// Every kernel sums in a fixed order, so a request's results do not depend on
// what runs beside it, nor on whether its operator was resident.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>
#include "problem.h"

namespace {

// deep copy of a bqd2_model and its host arrays
struct ModelCopy {
  bqd2_model m;
  std::vector<double> rn, cr, cw, ren, wen, Pi, dPi, k0, k2, rho0;
  std::vector<int> ed, eni, md, sd, dom_ne, dom_sid;
  void copy(const bqd2_model &s) {
    m = s;
    const size_t ne1 = (size_t)s.Ne + 1, mn1 = (size_t)s.MN + 1;
    size_t nsid = mn1;
    for (size_t d = 0; d < mn1; d++) nsid += (size_t)s.dom_ne[d];
    rn.assign(s.rn, s.rn + ((size_t)s.Nn + 1) * 3);
    cr.assign(s.cr, s.cr + ne1 * 12);
    cw.assign(s.cw, s.cw + ne1 * 9);
    ren.assign(s.ren, s.ren + ne1 * 12);
    wen.assign(s.wen, s.wen + ne1 * 12);
    Pi.assign(s.Pi, s.Pi + ne1 * 8);
    dPi.assign(s.dPi, s.dPi + ne1 * 8);
    k0.assign(s.k0, s.k0 + mn1);
    k2.assign(s.k2, s.k2 + 2 * mn1);
    rho0.assign(s.rho0, s.rho0 + mn1);
    ed.assign(s.ed, s.ed + ne1 * 4);
    eni.assign(s.eni, s.eni + ne1 * 4);
    md.assign(s.md, s.md + ne1);
    sd.assign(s.sd, s.sd + ne1);
    dom_ne.assign(s.dom_ne, s.dom_ne + mn1);
    dom_sid.assign(s.dom_sid, s.dom_sid + nsid);
    m.rn = rn.data(); m.cr = cr.data(); m.cw = cw.data(); m.ren = ren.data(); m.wen = wen.data();
    m.Pi = Pi.data(); m.dPi = dPi.data(); m.k0 = k0.data(); m.k2 = k2.data(); m.rho0 = rho0.data();
    m.ed = ed.data(); m.eni = eni.data(); m.md = md.data(); m.sd = sd.data();
    m.dom_ne = dom_ne.data(); m.dom_sid = dom_sid.data();
  }
};

// a resident operator: the caller's key and the bits of (kx, ky)
struct OpKey {
  unsigned long long key;
  uint64_t kx, ky;
  bool operator<(const OpKey &o) const {
    if (key != o.key) return key < o.key;
    if (kx != o.kx) return kx < o.kx;
    return ky < o.ky;
  }
};

struct Req {
  long long id;
  ModelCopy model;
  unsigned long long key;  // 0: no reuse
  OpKey okey;
  int solver, npts, force_type;
  double rc[3];
  std::vector<double> pts;
  double *field_out;  // the caller's n x 8 buffer, written at completion
  size_t need;        // footprint when the operator is not resident
  double t_submit, t_start;
  bool done;
  bqd2_result res;
};

struct Op {
  bqd2_problem *p;
  size_t bytes;  // accounted device bytes
  bool busy;
  double last_used;
};

double now_s(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

struct bqd2_server {
  size_t budget = 0, used = 0;
  int slots = 0, running = 0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Req *> pending;
  std::map<long long, Req *> reqs;
  std::map<OpKey, Op> cache;
  std::map<OpKey, int> building;  // operators whose first request is assembling them
  std::vector<std::thread> workers;
  bool stop = false;
  long long next_id = 1;
  long long n_hits = 0, n_misses = 0, n_evictions = 0, n_failed = 0;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  void work();
  bool pick(Req **r, Op **op);
  void evict_idle(size_t need);
};

static size_t sq(size_t n) { return n * n; }

// Device bytes of one request's problem (an upper estimate): mesh, per-domain
// matrices and pair tables, the system, the solver's factors and workspaces,
// and the field and force workspaces, which grow by half again when enlarged.
extern "C" size_t bqd2_footprint(const bqd2_model *m, int solver, int npts, int force_type) {
  if (!m || m->Ne <= 0) return 0;
  const size_t ne1 = (size_t)m->Ne + 1;
  size_t b = ne1 * (45 * 8 + 16 * 8 + 12 * 4) + ((size_t)m->Nn + 1) * 24 + 64 * 1024;
  size_t nmax = 0, ws = 0;
  for (int d = 0; d <= m->MN; d++) {
    const size_t ne = (size_t)m->dom_ne[d], N = 4 * ne;
    nmax = std::max(nmax, N);
    b += 2 * sq(N) * 16;                                   // G, H
    if (!(d == 0 && m->ps == 1)) b += 4 * sq(N) * 16;      // dG, dH
    b += N * 3 * 8 + N * ne * (24 + 1 + 4 + 4) + 4 * PLIST_PAD + 128;  // dF, Fpair, rule, pkey, plist
    b += (ne + 1) * (2 * 64 + 2 * 192) + (ne + 1) * 4 + N * 4 + ne * 4;  // ..., sid, rowmap, fperm
  }
  const size_t na = 2 * (size_t)m->NN;
  b += sq(na) * 16 + na * (16 + 16 + 4) + 8192 + (size_t)m->NN * 48 + 6 * nmax * 16 + 16;
  if (solver == 2) {
    b += sq(na) * 8 + bqd2_cgetrf_wsize((int)na) + na * (48 + 16);
    for (int d = 0; d <= m->MN; d++) {  // float factors of each domain's G (dpv systems)
      const size_t N = 4 * (size_t)m->dom_ne[d];
      if (N) b += sq(N) * 8 + bqd2_cgetrf_wsize((int)N) + N * (48 + 16 + 4);
    }
  }
  if (solver == 1) b += sq(nmax) * 16;                      // dpv LU buffers
  // GMRES: basis of restart + 1 vectors for the system, 3 right-hand sides for dpv
  ws = std::max(ws, 101 * na * 16 + 8 * na * 16);
  ws = std::max(ws, 3 * 101 * nmax * 16 + 8 * 3 * nmax * 16);
  // the far field's amplitudes (2 sides x FAR_NB bands per order), one set per
  // rule type in use, and the band table
  const size_t far = m->ps == 1 ? (size_t)far_grid(m->qd, nullptr, nullptr, nullptr) * 2 * FAR_NB * 64 : 0;
  int far_sets = 0;
  if (m->ps == 1) b += FAR_NB * sizeof(FarBand);
  if (npts > 0) {
    size_t nchunk = 1;
    for (int d = 0; d <= m->MN; d++) {
      const size_t items = (size_t)m->dom_ne[d] * 4, per = (items + 4095) / 4096, chunk = 64 * std::max<size_t>(per, 1);
      nchunk = std::max(nchunk, (items + chunk - 1) / chunk);
    }
    ws = std::max(ws, (size_t)npts * (48 + 64 + 12 + nchunk * 18 * 8) + 7 * 256);
    far_sets++;  // rule type 0
  }
  if (force_type >= 0) {
    const size_t nbd = ne1 * 9, nch = ((size_t)m->Ne * 9 + 1023) / 1024 + 2;
    ws = std::max(ws, ne1 * 48 + nbd * (64 + 4 + 16 + 4) + nbd * nch * 9 * 8 + 8 * 256);
    if (force_type != 0 || npts <= 0) far_sets++;  // the boundary field's rule type
  }
  b += far * far_sets;
  return b + ws + ws / 2 + 2 * CNT_N * 8;
}

void bqd2_server::evict_idle(size_t need) {
  while (used + need > budget) {
    auto victim = cache.end();
    for (auto it = cache.begin(); it != cache.end(); ++it)
      if (!it->second.busy && (victim == cache.end() || it->second.last_used < victim->second.last_used)) victim = it;
    if (victim == cache.end()) return;
    bqd2_destroy(victim->second.p);
    used -= victim->second.bytes;
    cache.erase(victim);
    n_evictions++;
  }
}

// under mu: the next runnable request (resident idle operator, or the first
// miss when it fits), or false
bool bqd2_server::pick(Req **r, Op **op) {
  bool miss_seen = false;
  for (auto it = pending.begin(); it != pending.end();) {
    Req *q = *it;
    if (q->key && building.count(q->okey)) { ++it; continue; }  // wait for the operator being assembled
    auto c = q->key ? cache.find(q->okey) : cache.end();
    if (c != cache.end()) {
      if (c->second.busy) { ++it; continue; }
      c->second.busy = true;
      *r = q;
      *op = &c->second;
      pending.erase(it);
      return true;
    }
    if (miss_seen) { ++it; continue; }
    miss_seen = true;
    if (used + q->need > budget) evict_idle(q->need);
    if (used + q->need <= budget) {
      used += q->need;
      if (q->key) building[q->okey] = 1;
      *r = q;
      *op = nullptr;
      pending.erase(it);
      return true;
    }
    if (running == 0) {
      // cannot fit even with every idle operator evicted and nothing running
      q->res.status = BQD2_E_NOMEM;
      snprintf(q->res.error, sizeof(q->res.error), "footprint of %zu bytes exceeds the server's budget of %zu bytes",
               q->need, budget);
      q->done = true;
      n_failed++;
      it = pending.erase(it);
      cv_done.notify_all();
      miss_seen = false;
      continue;
    }
    ++it;
  }
  return false;
}

static int run_request(Req *q, bqd2_problem *p, bool fresh) {
  int rc;
  const bqd2_model &m = q->model.m;
  if ((rc = bqd2_set_solver(p, q->solver))) return rc;
  if (fresh) {
    for (int d = 0; d <= m.MN; d++)
      if ((rc = bqd2_assemble(p, d, nullptr))) return rc;
  } else {
    // the key's promise: same mesh, frequency and lattice; (kx, ky) has the
    // operator's bits by construction of the cache key; the incidence is set anew
    BQD2_CHECK(p->m.Ne == m.Ne && p->m.MN == m.MN && p->pw_k0 == m.pw_k0 &&
                   !memcmp(p->qd8 + 4, m.qd + 4, 4 * sizeof(double)) &&
                   !memcmp(p->qd8 + 1, m.qd + 1, 2 * sizeof(double)),
               BQD2_E_ARG, "bqd2_server: request %lld does not match the resident operator of its key", q->id);
    if ((rc = bqd2_set_incidence_R(p, m.pw_p, m.pw_tv, m.pw_R))) return rc;
  }
  if ((rc = bqd2_solve_surface(p, nullptr))) return rc;
  if ((rc = bqd2_solve_velocity(p))) return rc;
  int its[17];
  double res[17];
  int lu = 0;
  bqd2_solver_info(p, its, res, &lu);
  q->res.krylov_steps = its[0];
  q->res.relres = res[0];
  if (q->npts > 0 && q->field_out &&
      (rc = bqd2_field(p, q->npts, q->pts.data(), 0, 1, q->field_out, nullptr, nullptr)))
    return rc;
  if (q->force_type >= 0 && (rc = bqd2_force(p, q->rc, q->force_type, q->res.F, q->res.N))) return rc;
  return BQD2_OK;
}

void bqd2_server::work() {
  std::unique_lock<std::mutex> lk(mu);
  for (;;) {
    Req *q = nullptr;
    Op *op = nullptr;
    cv_work.wait(lk, [&] { return (stop && pending.empty()) || pick(&q, &op); });
    if (!q) return;
    running++;
    q->t_start = now_s(t0);
    lk.unlock();
    bqd2_problem *p = op ? op->p : nullptr;
    int rc = BQD2_OK;
    if (op) {
      bqd2_rebind(p);
    } else {
      rc = bqd2_create(&q->model.m, &p);
    }
    if (!rc) rc = run_request(q, p, op == nullptr);
    const size_t actual = p ? bqd2_bytes(p) : 0;
    if (rc) {
      const char *msg = bqd2_last_error();
      strncpy(q->res.error, msg, sizeof(q->res.error) - 1);
    }
    lk.lock();
    const double t1 = now_s(t0);
    running--;
    q->res.status = rc;
    q->res.reused = op ? 1 : 0;
    q->res.t_queue = q->t_start - q->t_submit;
    q->res.t_run = t1 - q->t_start;
    q->res.t_done = t1;
    q->res.device_bytes = actual;
    if (op) {
      n_hits++;
      op->busy = false;
      op->last_used = t1;
      if (actual > op->bytes) { used += actual - op->bytes; op->bytes = actual; }
    } else {
      n_misses++;
      if (q->key) building.erase(q->okey);
      if (rc || q->key == 0 || cache.count(q->okey)) {
        if (p) bqd2_destroy(p);
        used -= q->need;
      } else {
        const size_t bytes = std::max(q->need, actual);
        used += bytes - q->need;
        cache[q->okey] = Op{p, bytes, false, t1};
      }
    }
    if (rc) n_failed++;
    q->done = true;
    cv_done.notify_all();
    cv_work.notify_all();
  }
}

extern "C" int bqd2_server_create(size_t hbm_bytes, int slots, bqd2_server **out) {
  BQD2_CHECK(out && slots >= 1 && slots <= 64 && hbm_bytes > 0, BQD2_E_ARG,
             "bqd2_server_create: need 1..64 slots and a positive memory budget");
  bqd2_server *s = new bqd2_server();
  s->budget = hbm_bytes;
  s->slots = slots;
  for (int i = 0; i < slots; i++) s->workers.emplace_back([s] { s->work(); });
  *out = s;
  return BQD2_OK;
}

extern "C" int bqd2_server_submit(bqd2_server *s, const bqd2_request *rq, double *field_out, long long *ticket) {
  BQD2_CHECK(s && rq && rq->model && ticket, BQD2_E_ARG, "bqd2_server_submit: null argument");
  BQD2_CHECK(rq->solver >= 0 && rq->solver <= 2 && rq->npts >= 0 && (rq->npts == 0 || rq->pts) &&
                 rq->force_type >= -1 && rq->force_type <= 1,
             BQD2_E_ARG, "bqd2_server_submit: bad request");
  Req *q = new Req();
  q->model.copy(*rq->model);
  q->key = rq->op_key;
  q->okey.key = rq->op_key;
  memcpy(&q->okey.kx, &rq->model->qd[1], sizeof(uint64_t));
  memcpy(&q->okey.ky, &rq->model->qd[2], sizeof(uint64_t));
  q->solver = rq->solver;
  q->npts = rq->npts;
  q->force_type = rq->force_type;
  memcpy(q->rc, rq->rc, sizeof(q->rc));
  if (rq->npts) q->pts.assign(rq->pts, rq->pts + 3 * (size_t)rq->npts);
  q->field_out = field_out;
  q->need = bqd2_footprint(rq->model, rq->solver, rq->npts, rq->force_type);
  q->done = false;
  memset(&q->res, 0, sizeof(q->res));
  {
    std::lock_guard<std::mutex> g(s->mu);
    BQD2_CHECK(!s->stop, BQD2_E_STATE, "bqd2_server_submit: server is shutting down");
    q->id = s->next_id++;
    q->t_submit = now_s(s->t0);
    s->reqs[q->id] = q;
    s->pending.push_back(q);
    *ticket = q->id;
  }
  s->cv_work.notify_one();
  return BQD2_OK;
}

extern "C" int bqd2_server_wait(bqd2_server *s, long long ticket, bqd2_result *res) {
  BQD2_CHECK(s && res, BQD2_E_ARG, "bqd2_server_wait: null argument");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  BQD2_CHECK(it != s->reqs.end(), BQD2_E_ARG, "bqd2_server_wait: unknown ticket %lld", ticket);
  Req *q = it->second;
  s->cv_done.wait(lk, [&] { return q->done; });
  *res = q->res;
  s->reqs.erase(it);
  delete q;
  if (res->status) bqd2_set_error("request %lld: %s", ticket, res->error);
  return res->status;
}

extern "C" int bqd2_server_stats(bqd2_server *s, long long *counts, size_t *bytes) {
  BQD2_CHECK(s, BQD2_E_ARG, "bqd2_server_stats: null server");
  std::lock_guard<std::mutex> g(s->mu);
  if (counts) {
    counts[0] = s->n_hits;
    counts[1] = s->n_misses;
    counts[2] = s->n_evictions;
    counts[3] = s->n_failed;
    counts[4] = (long long)s->cache.size();
    counts[5] = (long long)s->pending.size();
  }
  if (bytes) {
    bytes[0] = s->used;
    bytes[1] = s->budget;
  }
  return BQD2_OK;
}

extern "C" int bqd2_server_destroy(bqd2_server *s) {
  if (!s) return BQD2_OK;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (auto &t : s->workers) t.join();
  for (auto &kv : s->cache) bqd2_destroy(kv.second.p);
  for (auto &kv : s->reqs) delete kv.second;
  delete s;
  return BQD2_OK;
}
