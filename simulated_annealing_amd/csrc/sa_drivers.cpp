// sa_drivers.cpp -- C ABI over the C++ drivers of include/sa: parallel tempering
// and annealing run on hip_sys systems, for callers that cannot instantiate the
// templates (Python via ctypes, C, other FFIs).
#include <algorithm>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "sa/annealing.hpp"
#include "sa/hip_sys.hpp"
#include "sa/parallel_tempering.hpp"
#include "sa_hip.h"

struct sa_pt {
  std::shared_ptr<sa::ladder> lad;
  std::vector<std::unique_ptr<sa::hip_sys>> owned;
  std::vector<gpu_sys *> sys;
  std::unique_ptr<sa::parallel_tempering<gpu_sys>> pt;
  std::vector<sa::swap_record> log;
  std::string fname = "sa_pt_%.dat";
  int64_t done = 0;
};

namespace {

// every system of a main()-style run: set_seed, init_J, set_seed, init_rand
// (src/main_gpu.cu:262-276).  The slots of one ladder share couplings and stream,
// so doing it once for the ladder is the same as doing it for each system.
void seed_like_main(sa::hip_sys &s, uint64_t seed) {
  s.set_seed((long long)seed);
  s.init_J();
  s.set_seed((long long)seed);
  s.init_rand();
}

}  // namespace

extern "C" {

int sa_pt_create(int kind, int L, int nT, const double *T, double h, uint64_t seed, int sweeps, int rng_mode,
                 sa_pt **out) {
  if (!out || !T || nT < 2 || sweeps < 1) return SA_ERR_ARG;
  *out = nullptr;
  std::unique_ptr<sa_pt> p(new sa_pt());
  try {
    p->lad = std::make_shared<sa::ladder>(kind, L, nT);
  } catch (const std::exception &) {
    return SA_ERR_NOMEM;
  }
  for (int j = 0; j < nT; ++j) {
    p->owned.emplace_back(new sa::hip_sys(p->lad, j, T[0], h));
    p->sys.push_back(p->owned.back().get());
  }
  p->owned[0]->set_rng(rng_mode);
  seed_like_main(*p->owned[0], seed);
  for (auto &s : p->owned) s->set_rng(rng_mode);
  std::vector<double> Tv(T, T + nT);
  p->pt.reset(new sa::parallel_tempering<gpu_sys>(p->sys, Tv, sweeps, p->fname));
  p->pt->set_swap_log(&p->log);
  *out = p.release();
  return SA_OK;
}

int sa_pt_step(sa_pt *p, int nsteps) {
  if (!p || nsteps < 0) return SA_ERR_ARG;
  for (int k = 0; k < nsteps; ++k) p->pt->step(0);
  p->done += nsteps;
  return SA_OK;
}

int sa_pt_sync(sa_pt *p) { return p ? sa_ladder_sync(p->lad->get()) : SA_ERR_ARG; }
int64_t sa_pt_steps_done(const sa_pt *p) { return p ? p->done : 0; }
sa_ladder *sa_pt_ladder(sa_pt *p) { return p ? p->lad->get() : nullptr; }

int sa_pt_result(sa_pt *p, int64_t capacity, int32_t *pairs, double *us, uint64_t *masks, float *E_last,
                 uint64_t *spins) {
  if (!p) return SA_ERR_ARG;
  const int64_t n = std::min<int64_t>(capacity, (int64_t)p->log.size());
  for (int64_t k = 0; k < n; ++k) {
    if (pairs) pairs[k] = p->log[k].pair;
    if (us) us[k] = p->log[k].u;
    if (masks) masks[k] = p->log[k].mask;
  }
  const auto &E = p->pt->get_E();  // measures systems not measured at the last step
  const auto &sys = p->pt->systems();  // temperature order, inserted systems included
  if (E_last)
    for (size_t j = 0; j < sys.size(); ++j)
      for (int k = 0; k < 64; ++k) E_last[j * 64 + k] = E[j].size() == 64 ? E[j][k] : 0.f;
  if (spins) {
    const int64_t N = p->lad->N();
    for (size_t j = 0; j < sys.size(); ++j) {
      sa::hip_sys *h = dynamic_cast<sa::hip_sys *>(sys[j]);
      if (!h) return SA_ERR_ARG;
      const std::vector<uint64_t> v = h->spins();
      std::memcpy(spins + j * N, v.data(), (size_t)N * 8);
    }
  }
  return SA_OK;
}

int sa_pt_update_temperatures(sa_pt *p) {
  if (!p) return SA_ERR_ARG;
  p->pt->update_T();
  return (int)p->pt->get_T().size();
}

int sa_pt_reset(sa_pt *p) {
  if (!p) return SA_ERR_ARG;
  p->pt->reset();
  return SA_OK;
}

int sa_pt_temperatures(sa_pt *p, int capacity, double *T) {
  if (!p || capacity < 0 || (capacity && !T)) return SA_ERR_ARG;
  const auto &v = p->pt->get_T();
  for (int j = 0; j < capacity && j < (int)v.size(); ++j) T[j] = v[j];
  return (int)v.size();
}

int sa_pt_acceptance(sa_pt *p, int capacity, double *rates) {
  if (!p || capacity < 0 || (capacity && !rates)) return SA_ERR_ARG;
  const std::vector<double> v = p->pt->acceptance();
  for (int j = 0; j < capacity && j < (int)v.size(); ++j) rates[j] = v[j];
  return (int)v.size();
}

int sa_pt_E_min(sa_pt *p, float *E_min) {
  if (!p || !E_min) return SA_ERR_ARG;
  const std::vector<float> v = p->pt->get_E_min();
  for (int k = 0; k < 64; ++k) E_min[k] = k < (int)v.size() ? v[k] : 0.f;
  return (int)v.size();
}

int sa_pt_set_output(sa_pt *p, const char *pattern) {
  if (!p || !pattern) return SA_ERR_ARG;
  p->pt->close_file();
  p->fname = pattern;
  p->pt->set_filename(p->fname);
  return SA_OK;
}

int sa_pt_step_stride(sa_pt *p, int nsteps, int stride) {
  if (!p || nsteps < 0 || stride < 0) return SA_ERR_ARG;
  for (int k = 0; k < nsteps; ++k) p->pt->step(stride);
  p->done += nsteps;
  return SA_OK;
}

int sa_pt_print_min(sa_pt *p) {
  if (!p) return SA_ERR_ARG;
  p->pt->print(sa::PRINT_MIN);
  return SA_OK;
}

int64_t sa_pt_print_swap(sa_pt *p, char *buf, int64_t capacity) {
  if (!p) return SA_ERR_ARG;
  std::ostringstream os;
  p->pt->print_swap(os);
  const std::string t = os.str();
  if (buf && capacity > 0) {
    const int64_t n = std::min<int64_t>(capacity - 1, (int64_t)t.size());
    std::memcpy(buf, t.data(), n);
    buf[n] = '\0';
  }
  return (int64_t)t.size();
}

int sa_pt_close_output(sa_pt *p) {
  if (!p) return SA_ERR_ARG;
  p->pt->close_file();
  return SA_OK;
}

int sa_pt_save(sa_pt *p, const char *path) {
  if (!p || !path) return SA_ERR_ARG;
  std::ofstream f(path, std::ios::binary);
  if (!f) return SA_ERR_ARG;
  p->pt->save(f);
  return f ? SA_OK : SA_ERR_STATE;
}

int sa_pt_load(sa_pt *p, const char *path) {
  if (!p || !path) return SA_ERR_ARG;
  std::ifstream f(path, std::ios::binary);
  if (!f) return SA_ERR_ARG;
  p->pt->load(f);
  if (!f) return SA_ERR_STATE;
  p->log.clear();
  p->done = 0;
  return SA_OK;
}

int sa_pt_destroy(sa_pt *p) {
  delete p;
  return SA_OK;
}

int sa_pt_run(int kind, int L, int nT, const double *T, double h, uint64_t seed, int sweeps, int nsteps,
              int rng_mode, int32_t *pairs, double *us, uint64_t *masks, float *E_last, uint64_t *spins) {
  sa_pt *p = nullptr;
  int rc = sa_pt_create(kind, L, nT, T, h, seed, sweeps, rng_mode, &p);
  if (rc != SA_OK) return rc;
  rc = sa_pt_step(p, nsteps);
  if (rc == SA_OK) rc = sa_pt_result(p, nsteps, pairs, us, masks, E_last, spins);
  sa_pt_destroy(p);
  return rc;
}

int sa_anneal_run(int kind, int L, const double *Ts, int nsweeps, double h, uint64_t seed, int stride,
                  int rng_mode, float *E_log, uint64_t *spins_out) {
  if (!Ts || nsweeps < 1) return SA_ERR_ARG;
  std::unique_ptr<sa::hip_sys> s;
  try {
    s.reset(new sa::hip_sys(kind, L, Ts[0], h));
  } catch (const std::exception &) {
    return SA_ERR_NOMEM;
  }
  s->set_rng(rng_mode);
  seed_like_main(*s, seed);
  sa::annealing<sa::hip_sys> a(Ts[0], 1, 1.0, s.get());
  std::vector<double> sched(Ts, Ts + nsweeps);
  auto log = a.run_schedule(sched, stride);
  if (E_log)
    for (size_t r = 0; r < log.size(); ++r) std::memcpy(E_log + r * 64, log[r].data(), 64 * sizeof(float));
  if (spins_out) {
    std::vector<uint64_t> v = s->spins();
    std::memcpy(spins_out, v.data(), v.size() * 8);
  }
  return (int)log.size();
}

int64_t sa_anneal_geometric(int kind, int L, double T0, double alfa, int sweeps, int nsteps, int stride, double h,
                            uint64_t seed, int rng_mode, int precision, char *buf, int64_t capacity) {
  if (sweeps < 1 || nsteps < 0 || stride < 0 || alfa < 0 || alfa > 1 || precision < 0 || capacity < 0 ||
      (capacity && !buf))
    return SA_ERR_ARG;
  std::unique_ptr<sa::hip_sys> s;
  try {
    s.reset(new sa::hip_sys(kind, L, T0, h));
  } catch (const std::exception &) {
    return SA_ERR_NOMEM;
  }
  s->set_rng(rng_mode);
  seed_like_main(*s, seed);
  std::ostringstream os;
  if (precision) os.precision(precision);
  {
    sa::annealing<sa::hip_sys> a(T0, sweeps, alfa, s.get(), os);
    for (int k = 0; k < nsteps; ++k) a.step(stride, sa::PRINT_E);
  }
  const std::string t = os.str();
  if (capacity) {
    const int64_t n = std::min<int64_t>((int64_t)t.size(), capacity - 1);
    std::memcpy(buf, t.data(), (size_t)n);
    buf[n] = '\0';
  }
  return (int64_t)t.size();
}

}  // extern "C"
