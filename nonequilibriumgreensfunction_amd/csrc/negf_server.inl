// Serving / batching of whole transport solves on one GPU (negf_server_* in
// include/negf_capi.h).  Included at the end of negf_engine.hip: it drives the
// same fused per-energy solve (run_batch) and accumulation (accumulate) as
// negf_flush, for many wires at once.
//
// Scheduling.  Requests wait in arrival order until their resident footprint
// (Htotd and the coupling matrices, uploaded through a per-request negf_ctx) fits
// the HBM budget next to the admitted ones.  Up to four worker lanes, each with its
// own HIP stream and a fixed working pool, repeatedly take the next chunk of
// energies from the oldest admitted request that still has unassigned energies; the
// pool is carved for that request's layout, so one lane can run a block-2304 chunk
// while another runs a block-256 chunk of a later request.  A request finishes
// when all its energies are accumulated; its device state is released and the
// next waiting requests are admitted.
//
// Determinism.  Per-energy results are identical in any batch (fixed-order
// reductions, per-matrix kernels), and each request accumulates them strictly in
// energy order (out-of-order chunks are held until their predecessors arrive),
// so a request's outputs do not depend on what else is in flight.
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>

namespace {

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Device bytes one energy of this layout needs in a lane pool.
size_t lane_bytes_per_energy(const Layout& y) {
  return align256(sizeof(zd) * (size_t)y.Es) + align256(sizeof(zd) * 2 * (size_t)y.Cs) +
         align256(sizeof(int) * (size_t)y.Is) + align256(sizeof(int) * (size_t)y.chain_ints()) +
         align256(sizeof(double) * (size_t)y.site_per_e) + 3 * 256 + 2 * align256(sizeof(double) * 8) +
         align256(sizeof(int) * 8);
}

// Point the lane's buffers at `pool` for nb energies of layout y (nb * per-energy
// bytes must fit; checked by the caller).
void bind_lane(Lane& ln, const Layout& y, int nb, char* pool) {
  char* p = pool;
  auto take = [&](size_t bytes) {
    char* q = p;
    p += align256(bytes);
    return q;
  };
  ln.E = (zd*)take(sizeof(zd) * (size_t)y.Es * nb);
  ln.C = (zd*)take(sizeof(zd) * (size_t)y.Cs * 2 * nb);
  ln.Ie = (int*)take(sizeof(int) * (size_t)y.Is * nb);
  ln.Ic = (int*)take(sizeof(int) * (size_t)y.chain_ints() * nb);
  ln.d_site = (double*)take(sizeof(double) * (size_t)y.site_per_e * nb);
  ln.d_energy = (double*)take(sizeof(double) * nb);
  ln.d_scal = (double*)take(sizeof(double) * 8 * nb);
  ln.d_mc = (double*)take(sizeof(double) * 2 * nb);
  ln.d_ref = (double*)take(sizeof(double) * 8 * nb);
  ln.d_active = (int*)take(sizeof(int) * 8 * nb);
  ln.cap = nb;
}

// Device footprint of an admitted request beyond the lane pools.
size_t resident_bytes(int n, int L, int nsites, int precision, bool compact) {
  const size_t nn = (size_t)n * n;
  // X, betaL, betaR, and tmpP (2 n x n, allocated only for a dense HopX; counted
  // always, since admission runs before HopX is inspected); the mirror sectors of the
  // end slice and the couplings (6 n/2 x n/2, likewise counted always)
  size_t b = 5 * sizeof(zd) * nn + 6 * sizeof(zd) * (nn / 4);
  if (compact)
    b += 3 * sizeof(zd) * nn + sizeof(double) * (size_t)L * nsites;  // H0, two end slices, V
  else
    b += sizeof(zd) * nn * (size_t)L;  // Htotd
  if (precision != NEGF_PREC_FP64) b += sizeof(cx<float>) * nn;
  return b;
}

struct SrvRequest {
  long long id = 0;
  negf_request req;
  negf_ctx* ctx = nullptr;
  Layout lay;
  size_t footprint = 0;
  int next_energy = 0;   // next energy to hand to a lane
  int next_accum = 0;    // next energy to accumulate
  std::map<int, EnergyResult> pending;
  int status = 0;
  bool admitting = false, admitted = false, finished = false;
  double t_submit = 0.0, t_done = 0.0;
};

}  // namespace

struct negf_server {
  negf_server_config cfg;
  std::mutex mu;
  std::condition_variable cv;    // work or completion
  std::deque<SrvRequest*> waiting;
  std::vector<SrvRequest*> active;
  std::map<long long, SrvRequest*> all;
  long long next_id = 1;
  size_t budget = 0, used = 0, pool_bytes = 0;
  int nl = 4;  // worker lanes
  Lane lanes[kMaxLanes];
  char* pools[kMaxLanes] = {};
  std::thread workers[kMaxLanes];
  bool stop = false;
};

namespace {

// Admit waiting requests in order while their footprints fit (called with the
// lock held; the uploads run with the lock released by the admitting worker).
SrvRequest* next_admission(negf_server* s) {
  if (s->waiting.empty()) return nullptr;
  SrvRequest* r = s->waiting.front();
  if (r->admitting) return nullptr;
  // `used` includes admissions still uploading; every footprint fits the budget on
  // its own (checked at submission), so an idle server always admits
  if (s->used + r->footprint > s->budget) return nullptr;
  s->waiting.pop_front();
  r->admitting = true;
  s->used += r->footprint;
  return r;
}

int admit(negf_server* s, SrvRequest* r) {
  negf_config cfg = {};
  cfg.npx = r->req.npx; cfg.npy = r->req.npy; cfg.npz = r->req.npz; cfg.norb = r->req.norb;
  cfg.eta = s->cfg.eta;
  cfg.sr_tol = s->cfg.sr_tol;
  cfg.sr_max_iter = s->cfg.sr_max_iter;
  cfg.sr_extra_eta = s->cfg.sr_extra_eta;
  cfg.precision = s->cfg.precision;
  cfg.device = s->cfg.device;
  cfg.flags = s->cfg.flags;
  cfg.lanes = 1;  // the request context only carries state: the server's lanes run its energies
  cfg.gamma_trunc = s->cfg.gamma_trunc;
  int rc = negf_init(&r->ctx, &cfg);
  if (rc) return rc;
  r->ctx->want_currents = wants_currents(&r->req.acc);
  r->ctx->want_pairs = wants_pairs(&r->req.acc);
  if (!r->req.Htotd)
    return negf_set_hamiltonian_compact(r->ctx, r->req.H0, r->req.HopX, r->req.gM, r->req.onsite);
  return negf_set_hamiltonian(r->ctx, r->req.Htotd, r->req.HopX, r->req.gM);
}

// Finish a request (lock held): release its device state, wake waiters.
void finish(negf_server* s, SrvRequest* r) {
  r->finished = true;
  r->t_done = now_ms();
  if (r->ctx) {
    negf_finalize(r->ctx);
    r->ctx = nullptr;
  }
  s->used -= r->footprint;
  s->active.erase(std::find(s->active.begin(), s->active.end(), r));
  s->cv.notify_all();
}

void worker(negf_server* s, int li) {
  hipSetDevice(s->cfg.device);
  Lane& ln = s->lanes[li];
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    // admissions first: they make more work available
    if (SrvRequest* r = next_admission(s)) {
      lk.unlock();
      const int rc = admit(s, r);
      lk.lock();
      r->admitting = false;
      if (rc) {
        r->status = rc;
        s->active.push_back(r);
        finish(s, r);
      } else {
        r->admitted = true;
        s->active.push_back(r);
      }
      s->cv.notify_all();
      continue;
    }
    // the oldest admitted request with unassigned energies
    SrvRequest* r = nullptr;
    for (SrvRequest* q : s->active)
      if (q->admitted && q->next_energy < q->req.n_energies) {
        r = q;
        break;
      }
    if (!r) {
      if (s->stop && s->waiting.empty() && s->active.empty()) return;
      s->cv.wait(lk);
      continue;
    }
    const size_t per_e = lane_bytes_per_energy(r->lay);
    const int fit = (int)std::min<size_t>(s->pool_bytes / per_e, 1 << 20);
    const int i0 = r->next_energy;
    const int nb = std::min(fit, r->req.n_energies - i0);
    r->next_energy += nb;
    lk.unlock();
    bind_lane(ln, r->lay, nb, s->pools[li]);
    std::vector<EnergyResult> res;
    const int rc = run_batch(r->ctx, ln, nb, r->req.energies + i0, res);
    lk.lock();
    if (rc) {
      if (r->status >= 0) r->status = rc;
      res.assign(nb, EnergyResult());
      for (EnergyResult& e : res) e.status = rc;
    }
    for (int b = 0; b < nb; ++b) r->pending.emplace(i0 + b, std::move(res[b]));
    // accumulate in energy order
    for (auto it = r->pending.find(r->next_accum); it != r->pending.end();
         it = r->pending.find(r->next_accum)) {
      const int e = r->next_accum;
      if (it->second.site.size() == (size_t)r->lay.site_per_e)
        accumulate(r->ctx, it->second, r->req.energies[e], e, r->req.deltaE, r->req.mu, r->req.muT,
                   &r->req.acc);
      if (it->second.status && r->status == 0) r->status = it->second.status;
      r->pending.erase(it);
      ++r->next_accum;
    }
    if (r->next_accum == r->req.n_energies) finish(s, r);
  }
}

}  // namespace

extern "C" int negf_server_create(negf_server** out, const negf_server_config* cfg) {
  if (!out || !cfg) {
    negf_set_error("null argument");
    return NEGF_ERR_ARG;
  }
  *out = nullptr;
  if (cfg->lanes < 0 || cfg->lanes > kMaxLanes) {
    negf_set_error("server lanes must be 1..%d (0 = %d)", kMaxLanes, kMaxLanes);
    return NEGF_ERR_ARG;
  }
  if (cfg->precision != NEGF_PREC_FP64 && cfg->precision != NEGF_PREC_MIXED &&
      cfg->precision != NEGF_PREC_REFINED && cfg->precision != NEGF_PREC_LU32) {
    negf_set_error("unknown precision form %d", cfg->precision);
    return NEGF_ERR_ARG;
  }
  if (!(cfg->gamma_trunc >= 0.0 && cfg->gamma_trunc < 1.0)) {
    negf_set_error("gamma_trunc must be in [0, 1)");
    return NEGF_ERR_ARG;
  }
  if (cfg->precision != NEGF_PREC_FP64 && (cfg->flags & NEGF_FLAG_RGF_THREE_PASS)) {
    negf_set_error("the mixed-precision forms run on the two-pass RGF only");
    return NEGF_ERR_ARG;
  }
  // A request's contact cache would live outside the footprint admission counts,
  // and a request is solved once: the cache has nothing to serve.
  if (cfg->flags & NEGF_FLAG_CACHE_CONTACTS) {
    negf_set_error("NEGF_FLAG_CACHE_CONTACTS is not supported by the server");
    return NEGF_ERR_ARG;
  }
  int rc = require_device(cfg->device);
  if (rc) return rc;
  negf_server* s = new negf_server();
  s->cfg = *cfg;
  size_t freeb = 0, totalb = 0;
  if (hipMemGetInfo(&freeb, &totalb) != hipSuccess) {
    delete s;
    negf_set_error("hipMemGetInfo failed");
    return NEGF_ERR_HIP;
  }
  double budget = cfg->hbm_budget_gb > 0 ? cfg->hbm_budget_gb * 1e9 : 0.85 * (double)freeb;
  budget = std::min(budget, 0.92 * (double)freeb);
  s->nl = cfg->lanes == 0 ? kMaxLanes : cfg->lanes;
  double pool = cfg->lane_pool_gb > 0 ? cfg->lane_pool_gb * 1e9 : 0.5 * budget / s->nl;
  pool = std::min(pool, 0.9 * budget / s->nl);
  s->pool_bytes = (size_t)pool;
  s->budget = (size_t)(budget - s->nl * pool);
  for (int li = 0; li < s->nl; ++li) {
    if (hipMalloc(&s->pools[li], s->pool_bytes) != hipSuccess ||
        hipStreamCreateWithFlags(&s->lanes[li].st, hipStreamNonBlocking) != hipSuccess) {
      for (int k = 0; k <= li; ++k) {
        hipFree(s->pools[k]);
        if (s->lanes[k].st) hipStreamDestroy(s->lanes[k].st);
      }
      delete s;
      negf_set_error("server: cannot allocate a %.1f GB lane pool", pool / 1e9);
      return NEGF_ERR_MEMORY;
    }
  }
  for (int li = 0; li < s->nl; ++li) s->workers[li] = std::thread(worker, s, li);
  *out = s;
  return 0;
}

extern "C" int negf_server_submit(negf_server* s, const negf_request* req, long long* id) {
  if (!s || !req || !id || !req->HopX || (!req->Htotd && (!req->H0 || !req->onsite)) ||
      req->n_energies < 0 || (req->n_energies > 0 && !req->energies)) {
    negf_set_error("null argument");
    return NEGF_ERR_ARG;
  }
  if (req->npx < 1 || req->npy < 1 || req->npz < 1 ||
      !(req->norb == 1 || req->norb == 2 || req->norb == 4)) {
    negf_set_error("bad request dimensions");
    return NEGF_ERR_ARG;
  }
  const int n = req->npy * req->npz * req->norb;
  if (n > 3072) {
    negf_set_error("block size n=%d exceeds 3072", n);
    return NEGF_ERR_ARG;
  }
  SrvRequest* r = new SrvRequest();
  r->req = *req;
  r->lay.init(n, req->npx, req->npy * req->npz, !(s->cfg.flags & NEGF_FLAG_RGF_THREE_PASS),
              layout_form(s->cfg.precision));
  r->footprint = resident_bytes(n, req->npx, req->npy * req->npz, s->cfg.precision, !req->Htotd);
  if (lane_bytes_per_energy(r->lay) > s->pool_bytes || r->footprint > s->budget) {
    negf_set_error("request (n=%d, Npx=%d) needs %.1f GB per energy and %.1f GB resident; lane pool "
                   "%.1f GB, resident budget %.1f GB",
                   n, req->npx, lane_bytes_per_energy(r->lay) / 1e9, r->footprint / 1e9,
                   s->pool_bytes / 1e9, s->budget / 1e9);
    delete r;
    return NEGF_ERR_MEMORY;
  }
  std::lock_guard<std::mutex> lk(s->mu);
  r->id = s->next_id++;
  r->t_submit = now_ms();
  s->all[r->id] = r;
  if (req->n_energies == 0) {
    r->finished = true;
    r->t_done = r->t_submit;
  } else {
    s->waiting.push_back(r);
  }
  *id = r->id;
  s->cv.notify_all();
  return 0;
}

extern "C" int negf_server_poll(negf_server* s, long long id, int* done) {
  if (!s || !done) return NEGF_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->all.find(id);
  if (it == s->all.end()) {
    negf_set_error("unknown request %lld", id);
    return NEGF_ERR_ARG;
  }
  *done = it->second->finished ? 1 : 0;
  return 0;
}

extern "C" int negf_server_wait(negf_server* s, long long id, int* status, double* latency_ms) {
  if (!s) return NEGF_ERR_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->all.find(id);
  if (it == s->all.end()) {
    negf_set_error("unknown request %lld", id);
    return NEGF_ERR_ARG;
  }
  SrvRequest* r = it->second;
  s->cv.wait(lk, [&] { return r->finished; });
  if (status) *status = r->status;
  if (latency_ms) *latency_ms = r->t_done - r->t_submit;
  s->all.erase(it);
  delete r;
  return 0;
}

extern "C" int negf_server_destroy(negf_server* s) {
  if (!s) return 0;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    s->stop = true;
    s->cv.notify_all();
  }
  for (std::thread& t : s->workers)
    if (t.joinable()) t.join();
  for (int li = 0; li < s->nl; ++li) {
    hipStreamSynchronize(s->lanes[li].st);
    s->lanes[li].drop_graphs();
    hipFree(s->pools[li]);
    hipStreamDestroy(s->lanes[li].st);
  }
  for (auto& kv : s->all) delete kv.second;
  delete s;
  return 0;
}
