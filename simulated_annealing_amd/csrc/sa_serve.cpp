// sa_serve.cpp -- continuous-batching request loop over the batched engine (the
// C ABI sa_serve_* of include/sa_hip.h).
//
// Requests (an anneal or a parallel-tempering run of one disorder realisation)
// are submitted with an arrival time on the server clock.  The loop admits
// arrived requests in arrival order while the device-memory budget and the slot
// limit allow, so requests that arrive while others run join the very next chunk
// of sweeps; all admitted requests share every launch of the batch.  A request's
// lattice, couplings and random stream stay on the device from admission to
// completion.  After each chunk the finished requests are retired (digest of the
// final spins, final energies, memory released) and their latency recorded.
// Admission work (seeding, coupling draws, uploads) is done while the chunk runs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <vector>

#include "sa_hip.h"
#include "sa_internal.hpp"

namespace {

using clock_type = std::chrono::steady_clock;

uint64_t mix64(uint64_t z) {  // splitmix64 finaliser
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

struct req_state {
  sa_request r{};
  sa_request_result res{};
  int inst = -1;
  int64_t bytes = 0;
  bool done = false;
};

int64_t request_bytes(const sa_request &r) {
  const int nsys = r.mode == SA_REQ_PT ? r.nT : 1;
  return sa_detail::batch_instance_bytes(r.kind, r.L, nsys, r.nsweeps, r.mode == SA_REQ_PT ? 0 : 1);
}

double request_attempts(const sa_request &r) {
  const double N = (double)r.L * r.L * (r.kind == SA_KIND_3D ? r.L : 1);
  return N * 64.0 * (r.mode == SA_REQ_PT ? r.nT : 1) * r.nsweeps;
}

int check_request(const sa_request &r) {
  if (r.kind < SA_KIND_2P || r.kind > SA_KIND_3D) return sa_fail(SA_ERR_ARG, "request: unknown kind %d", r.kind);
  if (r.L < 4 || (r.L & 1)) return sa_fail(SA_ERR_UNSUPPORTED, "request: L=%d (the device sweep needs even L >= 4)", r.L);
  if (r.mode == SA_REQ_ANNEAL) {
    if (r.nsweeps < 1) return sa_fail(SA_ERR_ARG, "request: anneal needs nsweeps >= 1");
  } else if (r.mode == SA_REQ_PT) {
    if (r.nT < 2 || r.every < 1 || r.nsweeps < r.every || r.nsweeps % r.every)
      return sa_fail(SA_ERR_ARG, "request: PT needs nT >= 2, every >= 1 and nsweeps a positive multiple of every");
  } else {
    return sa_fail(SA_ERR_ARG, "request: unknown mode %d", r.mode);
  }
  if (!(r.T0 > 0) || !(r.T1 > 0)) return sa_fail(SA_ERR_ARG, "request: temperatures must be positive");
  return SA_OK;
}

}  // namespace

struct sa_serve {
  sa_serve_config cfg{};
  sa_batch *b = nullptr;
  clock_type::time_point t0;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<int64_t> pending;  // submitted, not admitted, by arrival time
  std::deque<req_state> reqs;  // by id (a deque: submissions keep references valid)
  std::vector<int64_t> active;  // on the device
  int64_t finished = 0;
  int64_t bytes = 0, peak_bytes = 0;
  int peak_active = 0;
  double busy_s = 0;  // wall time with requests on the device

  double now() const { return std::chrono::duration<double>(clock_type::now() - t0).count(); }
};

namespace {

// a request's state; the reference stays valid while other threads submit
req_state &req(sa_serve *s, int64_t id) {
  std::lock_guard<std::mutex> g(s->mu);
  return s->reqs[id];
}

int admit(sa_serve *s, int64_t id) {
  req_state &q = req(s, id);
  const sa_request &r = q.r;
  std::vector<double> T;
  int rc = sa_request_temperatures(&r, nullptr, 0);
  if (rc < 0) return rc;
  T.resize(rc);
  sa_request_temperatures(&r, T.data(), (int64_t)T.size());
  int inst = -1;
  if (r.mode == SA_REQ_PT)
    rc = sa_batch_add_pt(s->b, r.kind, r.L, r.nT, T.data(), r.h, r.seed, r.every, r.nsweeps / r.every, &inst);
  else
    rc = sa_batch_add(s->b, r.kind, r.L, 1, r.seed, r.nsweeps, T.data(), 1, r.h, &inst);
  if (rc) return rc;
  q.inst = inst;
  return SA_OK;
}

void finish_request(sa_serve *s, int64_t id, int status) {
  std::lock_guard<std::mutex> g(s->mu);
  req_state &q = s->reqs[id];
  q.res.status = status;
  q.res.finished = s->now();
  q.done = true;
  s->finished += 1;
}

// final results of a finished request, then its device memory is released
int retire(sa_serve *s, int64_t id) {
  req_state &q = req(s, id);
  sa_request_result &res = q.res;
  sa_ladder *l = sa_batch_ladder(s->b, q.inst);
  if (!l) return SA_ERR_STATE;
  const int nsys = sa_ladder_nsys(l);
  std::vector<int32_t> E((size_t)nsys * 64), M((size_t)nsys * 64);
  int rc = sa_ladder_measure(l, E.data(), M.data());
  if (rc) return rc;
  std::memcpy(res.E, E.data(), sizeof res.E);
  std::memcpy(res.M, M.data(), sizeof res.M);
  if (s->cfg.digests) {
    const int64_t N = sa_ladder_sites(l);
    std::vector<uint64_t> w((size_t)N);
    uint64_t d = 0;
    for (int j = 0; j < nsys; ++j) {
      rc = sa_ladder_get_spins(l, j, w.data());
      if (rc) return rc;
      d ^= sa_digest(w.data(), N, (uint64_t)j * (uint64_t)N);
    }
    res.spins_digest = d;
  }
  if (q.r.mode == SA_REQ_PT) {
    const int n = sa_batch_pt_result(s->b, q.inst, 0, nullptr, nullptr, nullptr, nullptr, nullptr);
    if (n < 0) return n;
    std::vector<int32_t> pairs(std::max(n, 1));
    std::vector<double> us(std::max(n, 1));
    std::vector<uint64_t> masks(std::max(n, 1));
    rc = sa_batch_pt_result(s->b, q.inst, n, pairs.data(), us.data(), masks.data(), res.E_min, nullptr);
    if (rc < 0) return rc;
    uint64_t h = 0;
    for (int k = 0; k < n; ++k) {
      uint64_t ub;
      std::memcpy(&ub, &us[k], 8);
      h = mix64(h ^ (uint64_t)(uint32_t)pairs[k]);
      h = mix64(h ^ ub);
      h = mix64(h ^ masks[k]);
    }
    res.log_digest = h;
    res.steps = n;
  }
  rc = sa_batch_remove(s->b, q.inst);
  if (rc) return rc;
  s->bytes -= q.bytes;
  q.inst = -1;
  return SA_OK;
}

// admit arrived requests in arrival order while memory and slots allow
void admit_arrived(sa_serve *s) {
  for (;;) {
    int64_t id;
    {
      std::lock_guard<std::mutex> g(s->mu);
      if (s->pending.empty()) return;
      id = s->pending.front();
      req_state &q = s->reqs[id];
      if (q.r.arrival > s->now()) return;
      if (q.bytes <= s->cfg.hbm_budget &&
          ((int)s->active.size() >= s->cfg.max_active || s->bytes + q.bytes > s->cfg.hbm_budget))
        return;  // wait for memory / a slot (arrival order is kept)
      s->pending.pop_front();
    }
    req_state &q = req(s, id);
    if (q.bytes > s->cfg.hbm_budget) {
      sa_fail(SA_ERR_NOMEM, "request %lld needs %lld bytes, over the budget", (long long)id, (long long)q.bytes);
      finish_request(s, id, SA_ERR_NOMEM);
      continue;
    }
    {
      std::lock_guard<std::mutex> g(s->mu);
      q.res.admitted = s->now();
    }
    const int rc = admit(s, id);
    if (rc) {
      finish_request(s, id, rc);
      continue;
    }
    s->active.push_back(id);
    s->bytes += q.bytes;
    s->peak_bytes = std::max(s->peak_bytes, s->bytes);
    s->peak_active = std::max(s->peak_active, (int)s->active.size());
  }
}

}  // namespace

extern "C" {

uint64_t sa_digest(const uint64_t *w, int64_t n, uint64_t first_index) {
  uint64_t d = 0;
  for (int64_t i = 0; i < n; ++i) d ^= mix64(w[i] + (first_index + (uint64_t)i) * 0x9E3779B97F4A7C15ULL);
  return d;
}

int sa_request_temperatures(const sa_request *r, double *T, int64_t capacity) {
  if (!r) return sa_fail(SA_ERR_ARG, "null request");
  int rc = check_request(*r);
  if (rc) return rc;
  const int n = r->mode == SA_REQ_PT ? r->nT : r->nsweeps;
  if (T) {
    if (capacity < n) return sa_fail(SA_ERR_ARG, "temperature buffer too small");
    for (int j = 0; j < n; ++j) {
      if (r->mode == SA_REQ_PT)  // geometric ladder T0 .. T1
        T[j] = r->T0 * std::pow(r->T1 / r->T0, (double)j / (double)(n - 1));
      else  // linear anneal T0 -> T1 over the sweeps
        T[j] = n == 1 ? r->T0 : r->T0 + ((r->T1 - r->T0) * (double)j) / (double)(n - 1);
    }
  }
  return n;
}

int64_t sa_request_bytes(const sa_request *r) {
  if (!r) return sa_fail(SA_ERR_ARG, "null request");
  const int rc = check_request(*r);
  return rc ? rc : request_bytes(*r);
}

int sa_serve_create(const sa_serve_config *cfg, sa_serve **out) {
  if (!out) return sa_fail(SA_ERR_ARG, "null output handle");
  sa_serve *s = new sa_serve();
  if (cfg) s->cfg = *cfg;
  if (s->cfg.hbm_budget <= 0) {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
      delete s;
      return sa_fail(SA_ERR_HIP, "hipMemGetInfo failed");
    }
    s->cfg.hbm_budget = (int64_t)(free_b / 2);
  }
  if (s->cfg.max_active <= 0) s->cfg.max_active = 512;
  if (s->cfg.chunk <= 0) s->cfg.chunk = 10;
  int rc = sa_batch_create(&s->b);
  if (rc) {
    delete s;
    return rc;
  }
  sa_batch_set_tuning(s->b, SA_TUNE_BATCH_CHUNK, s->cfg.chunk);
  s->t0 = clock_type::now();
  *out = s;
  return SA_OK;
}

int sa_serve_destroy(sa_serve *s) {
  if (!s) return SA_OK;
  sa_batch_destroy(s->b);
  delete s;
  return SA_OK;
}

double sa_serve_clock(const sa_serve *s) { return s ? s->now() : 0.0; }

int sa_serve_submit(sa_serve *s, const sa_request *r, int64_t *id) {
  if (!s || !r) return sa_fail(SA_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> g(s->mu);
  req_state q;
  q.r = *r;
  if (q.r.arrival < 0) q.r.arrival = s->now();
  const int64_t k = (int64_t)s->reqs.size();
  q.res.id = k;
  q.res.arrival = q.r.arrival;
  q.res.attempts = request_attempts(q.r);
  const int rc = check_request(q.r);
  if (rc) {
    q.res.status = rc;
    q.res.admitted = q.res.finished = q.r.arrival;
    q.done = true;
    s->reqs.push_back(q);
    s->finished += 1;
  } else {
    q.bytes = request_bytes(q.r);
    q.res.bytes = q.bytes;
    s->reqs.push_back(q);
    // keep the pending queue in arrival order (submissions usually come in order)
    auto it = s->pending.end();
    while (it != s->pending.begin() && s->reqs[*(it - 1)].r.arrival > q.r.arrival) --it;
    s->pending.insert(it, k);
  }
  if (id) *id = k;
  s->cv.notify_all();
  return rc ? rc : SA_OK;
}

int sa_serve_run(sa_serve *s, int64_t n_finished) {
  if (!s) return sa_fail(SA_ERR_ARG, "null server");
  for (;;) {
    {
      std::unique_lock<std::mutex> g(s->mu);
      const int64_t target = n_finished > 0 ? n_finished : (int64_t)s->reqs.size();
      if (s->finished >= target) return SA_OK;
      if (s->active.empty()) {
        // idle: wait for the next arrival (or a submission)
        if (s->pending.empty()) {
          if (n_finished <= 0) return SA_OK;
          s->cv.wait_for(g, std::chrono::milliseconds(5));
          continue;
        }
        const double wait = s->reqs[s->pending.front()].r.arrival - s->now();
        if (wait > 0) {
          s->cv.wait_for(g, std::chrono::duration<double>(std::min(wait, 0.005)));
          continue;
        }
      }
    }
    admit_arrived(s);
    if (s->active.empty()) continue;
    const double c0 = s->now();
    int k = 0;
    int rc = sa_detail::batch_launch_chunk(s->b, s->cfg.chunk, &k);
    if (rc) return rc;
    admit_arrived(s);  // host work of new arrivals overlaps the chunk
    if (k > 0) {
      rc = sa_detail::batch_finish_chunk(s->b, k);
      if (rc) return rc;
    }
    rc = sa_detail::batch_wait(s->b);
    if (rc) return rc;
    s->busy_s += s->now() - c0;
    // retire finished requests
    std::vector<int64_t> still;
    for (int64_t id : s->active) {
      if (sa_detail::batch_done(s->b, req(s, id).inst)) {
        rc = retire(s, id);  // the result is complete once read back
        finish_request(s, id, rc);
      } else {
        still.push_back(id);
      }
    }
    s->active.swap(still);
  }
}

int64_t sa_serve_results(sa_serve *s, int64_t first, int64_t capacity, sa_request_result *out) {
  if (!s) return sa_fail(SA_ERR_ARG, "null server");
  std::lock_guard<std::mutex> g(s->mu);
  int64_t n = 0;
  for (int64_t id = std::max<int64_t>(first, 0); id < (int64_t)s->reqs.size() && n < capacity; ++id, ++n)
    if (out) {
      out[n] = s->reqs[id].res;
      if (!s->reqs[id].done) out[n].status = SA_ERR_STATE;  // not finished yet
    }
  return n;
}

int sa_serve_get_stats(sa_serve *s, sa_serve_stats *st) {
  if (!s || !st) return sa_fail(SA_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> g(s->mu);
  std::memset(st, 0, sizeof *st);
  st->submitted = (int64_t)s->reqs.size();
  st->finished = s->finished;
  st->active = (int)s->active.size();
  st->pending = (int64_t)s->pending.size();
  st->peak_active = s->peak_active;
  st->bytes = s->bytes;
  st->peak_bytes = s->peak_bytes;
  st->hbm_budget = s->cfg.hbm_budget;
  st->busy_s = s->busy_s;
  st->clock_s = s->now();
  return SA_OK;
}

}  // extern "C"
