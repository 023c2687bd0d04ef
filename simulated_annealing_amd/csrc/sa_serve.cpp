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
// Admission work (seeding, coupling draws, uploads) runs on a pool of preparation
// threads; the loop thread only attaches prepared instances between chunks, and
// reads finished requests back while the next chunk runs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
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
  if (r.L < 3) return sa_fail(SA_ERR_UNSUPPORTED, "request: L=%d (the device sweep needs L >= 3)", r.L);
  if (r.mode == SA_REQ_ANNEAL) {
    if (r.nsweeps < 1) return sa_fail(SA_ERR_ARG, "request: anneal needs nsweeps >= 1");
  } else if (r.mode == SA_REQ_PT) {
    if (r.nT < 2 || r.every < 1 || r.nsweeps < r.every || r.nsweeps % r.every)
      return sa_fail(SA_ERR_ARG, "request: PT needs nT >= 2, every >= 1 and nsweeps a positive multiple of every");
  } else {
    return sa_fail(SA_ERR_ARG, "request: unknown mode %d", r.mode);
  }
  if (!(r.T0 > 0) || !(r.T1 > 0)) return sa_fail(SA_ERR_ARG, "request: temperatures must be positive");
  return sa_detail::batch_geometry(r.kind, r.L, r.mode == SA_REQ_PT ? r.nT : 1);
}

}  // namespace

struct sa_serve {
  sa_serve_config cfg{};
  sa_batch *b = nullptr;
  int device = 0;
  // preparation pool: jobs = admitted request ids, ready = prepared instances
  std::vector<std::thread> workers;
  std::deque<int64_t> jobs;
  struct ready_item {
    int64_t id;
    int rc;
    sa_detail::prepared_instance *p;
  };
  std::vector<ready_item> ready;
  std::condition_variable cv_jobs;
  bool stopping = false;
  int preparing = 0;  // admitted, not attached yet
  clock_type::time_point t0;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<int64_t> pending;  // submitted, not admitted, by arrival time
  std::deque<req_state> reqs;  // by id (a deque: submissions keep references valid)
  std::vector<int64_t> active;  // on the device
  int64_t finished = 0;
  int64_t bytes = 0, peak_bytes = 0;
  int peak_active = 0;
  double busy_s = 0;    // wall time from chunk launch to chunk completion
  double admit_s = 0;   // host time admitting requests
  double retire_s = 0;  // host time reading results back
  int64_t chunks = 0;

  double now() const { return std::chrono::duration<double>(clock_type::now() - t0).count(); }
};

namespace {

// a request's state; the reference stays valid while other threads submit
req_state &req(sa_serve *s, int64_t id) {
  std::lock_guard<std::mutex> g(s->mu);
  return s->reqs[id];
}

// the device state of a request, on a preparation thread
int prepare_request(const sa_request &r, sa_detail::prepared_instance **out) {
  int rc = sa_request_temperatures(&r, nullptr, 0);
  if (rc < 0) return rc;
  std::vector<double> T(rc);
  sa_request_temperatures(&r, T.data(), (int64_t)T.size());
  if (r.mode == SA_REQ_PT)
    return sa_detail::batch_prepare(r.kind, r.L, r.nT, r.seed, r.nsweeps, T.data(), 0, r.h, r.every, out);
  return sa_detail::batch_prepare(r.kind, r.L, 1, r.seed, r.nsweeps, T.data(), 1, r.h, 0, out);
}

void worker_main(sa_serve *s) {
  (void)hipSetDevice(s->device);
  for (;;) {
    int64_t id;
    sa_request r;
    {
      std::unique_lock<std::mutex> g(s->mu);
      s->cv_jobs.wait(g, [&] { return s->stopping || !s->jobs.empty(); });
      if (s->jobs.empty()) return;  // stopping
      id = s->jobs.front();
      s->jobs.pop_front();
      r = s->reqs[id].r;
    }
    sa_detail::prepared_instance *p = nullptr;
    const int rc = prepare_request(r, &p);
    {
      std::lock_guard<std::mutex> g(s->mu);
      s->ready.push_back({id, rc, p});
    }
    s->cv.notify_all();
  }
}

void finish_locked(sa_serve *s, int64_t id, int status) {
  req_state &q = s->reqs[id];
  q.res.status = status;
  q.res.finished = s->now();
  q.done = true;
  s->finished += 1;
}

void finish_request(sa_serve *s, int64_t id, int status) {
  std::lock_guard<std::mutex> g(s->mu);
  finish_locked(s, id, status);
}

// final results of a finished request.  Runs while the next chunk is on the GPU:
// the instance is finished, so the chunk skips it, and everything here uses the
// instance's own stream (measure, spin copies).
int read_results(sa_serve *s, int64_t id) {
  req_state &q = req(s, id);
  sa_request_result &res = q.res;
  sa_ladder *l = sa_detail::batch_instance_ladder(s->b, q.inst);
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
  return SA_OK;
}

// release a retired request's device memory (waits for the batch stream)
int release(sa_serve *s, int64_t id) {
  req_state &q = req(s, id);
  const int rc = sa_batch_remove(s->b, q.inst);
  s->bytes -= q.bytes;
  q.inst = -1;
  return rc;
}

// admit arrived requests in arrival order while memory and slots allow: reserve
// their memory and hand them to the preparation threads
void dispatch_arrived(sa_serve *s) {
  bool any = false;
  {
    std::lock_guard<std::mutex> g(s->mu);
    while (!s->pending.empty()) {
      const int64_t id = s->pending.front();
      req_state &q = s->reqs[id];
      if (q.r.arrival > s->now()) break;
      if (q.bytes > s->cfg.hbm_budget) {  // can never fit
        s->pending.pop_front();
        sa_fail(SA_ERR_NOMEM, "request %lld needs %lld bytes, over the budget", (long long)id, (long long)q.bytes);
        q.res.admitted = s->now();
        finish_locked(s, id, SA_ERR_NOMEM);
        continue;
      }
      if ((int)s->active.size() + s->preparing >= s->cfg.max_active || s->bytes + q.bytes > s->cfg.hbm_budget)
        break;  // wait for memory / a slot (arrival order is kept)
      s->pending.pop_front();
      q.res.admitted = s->now();
      s->bytes += q.bytes;
      s->peak_bytes = std::max(s->peak_bytes, s->bytes);
      s->preparing += 1;
      s->jobs.push_back(id);
      any = true;
    }
  }
  if (any) s->cv_jobs.notify_all();
}

// prepared instances join the batch (loop thread, between chunks)
void attach_ready(sa_serve *s) {
  std::vector<sa_serve::ready_item> got;
  {
    std::lock_guard<std::mutex> g(s->mu);
    got.swap(s->ready);
  }
  for (auto &it : got) {
    req_state &q = req(s, it.id);
    int rc = it.rc;
    int inst = -1;
    if (!rc) rc = sa_detail::batch_attach(s->b, it.p, &inst);
    std::lock_guard<std::mutex> g(s->mu);
    s->preparing -= 1;
    if (rc) {
      s->bytes -= q.bytes;
      finish_locked(s, it.id, rc);
      continue;
    }
    q.inst = inst;
    s->active.push_back(it.id);
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
  (void)hipGetDevice(&s->device);
  int nw = s->cfg.prep_threads;
  if (nw <= 0) nw = (int)std::min<unsigned>(8, std::max(1u, std::thread::hardware_concurrency() / 2));
  for (int i = 0; i < nw; ++i) s->workers.emplace_back(worker_main, s);
  s->t0 = clock_type::now();
  *out = s;
  return SA_OK;
}

int sa_serve_destroy(sa_serve *s) {
  if (!s) return SA_OK;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stopping = true;
    s->jobs.clear();
  }
  s->cv_jobs.notify_all();
  for (auto &t : s->workers) t.join();
  for (auto &it : s->ready) sa_detail::batch_discard(it.p);
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
  std::vector<int64_t> done_dev;  // finished on the device, results not read yet
  std::vector<int64_t> to_free;   // results read, memory not released yet
  for (;;) {
    if (!to_free.empty()) {  // the batch stream is idle here: releasing costs no overlap
      for (int64_t id : to_free) {
        const int rc = release(s, id);
        if (rc) return rc;
      }
      to_free.clear();
    }
    double h0 = s->now();
    dispatch_arrived(s);
    attach_ready(s);
    s->admit_s += s->now() - h0;
    {
      std::unique_lock<std::mutex> g(s->mu);
      const int64_t target = n_finished > 0 ? n_finished : (int64_t)s->reqs.size();
      if (s->finished >= target && done_dev.empty()) return SA_OK;
      if (s->active.empty() && done_dev.empty()) {
        // idle: wait for a prepared instance, the next arrival or a submission
        if (s->pending.empty() && s->preparing == 0 && s->ready.empty() && n_finished <= 0) return SA_OK;
        double wait = 0.005;
        if (!s->pending.empty()) wait = std::min(wait, s->reqs[s->pending.front()].r.arrival - s->now());
        if (s->ready.empty() && wait > 0) s->cv.wait_for(g, std::chrono::duration<double>(wait));
        continue;
      }
    }
    int k = 0;
    const double c0 = s->now();
    if (!s->active.empty()) {
      const int rc = sa_detail::batch_launch_chunk(s->b, s->cfg.chunk, &k);
      if (rc) return rc;
    }
    // host work that overlaps the chunk: results of the requests that finished in
    // the previous chunk, then admission of new arrivals
    h0 = s->now();
    for (int64_t id : done_dev) {
      const int rc = read_results(s, id);
      finish_request(s, id, rc);
      to_free.push_back(id);
    }
    done_dev.clear();
    s->retire_s += s->now() - h0;
    dispatch_arrived(s);
    if (k > 0) {
      int rc = sa_detail::batch_finish_chunk(s->b, k);
      if (rc) return rc;
      rc = sa_detail::batch_wait(s->b);
      if (rc) return rc;
      s->busy_s += s->now() - c0;
      s->chunks += 1;
    }
    std::vector<int64_t> still;
    for (int64_t id : s->active) {
      if (sa_detail::batch_done(s->b, req(s, id).inst)) done_dev.push_back(id);
      else still.push_back(id);
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
  st->admit_s = s->admit_s;
  st->retire_s = s->retire_s;
  st->chunks = s->chunks;
  st->clock_s = s->now();
  return SA_OK;
}

}  // extern "C"
