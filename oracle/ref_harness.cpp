// ref_harness.cpp -- extern "C" access to the reference's own CPU classes and
// drivers, compiled from the reference sources where they lie (see
// oracle/build_ref.sh).  TEST INFRASTRUCTURE ONLY: used to pin the oracle and to
// write the golden fixtures under tests/golden/.  Nothing of the reference is
// copied here; this file only calls it.
//
// Standard headers first, then the reference headers with class internals made
// visible so the harness can read spins, couplings and tables (GCC lays class
// members out in declaration order whatever their access, so objects compiled
// with and without the define agree).
#include <cstdint>
#include <cstring>
#include <fstream>
#include <iostream>
#include <memory>
#include <random>
#include <regex>
#include <sstream>
#include <string>
#include <vector>
#include <algorithm>
#include <chrono>
#include <omp.h>

#define private public
#define protected public
#include "Logging.hpp"
#include "annealing.hpp"
#include "cpu_2d.hpp"
#include "cpu_2dp.hpp"
#include "cpu_3d.hpp"
#include "cpu_sys.hpp"
#include "parallel_tempering.hpp"
#undef private
#undef protected

namespace {
cpu_sys *make_sys(int kind, int L, double T, double h) {
  if (kind == 0) return new cpu_2dp(L, T, h);
  if (kind == 1) return new cpu_2d(L, T, h);
  return new cpu_3d(L, T, h);
}
}  // namespace

extern "C" {

void *ref_create(int kind, int L, double T, double h) {
  Logger::verbosity = 0;
  return make_sys(kind, L, T, h);
}
void ref_destroy(void *p) { delete static_cast<cpu_sys *>(p); }
void ref_set_seed(void *p, long seed) { static_cast<cpu_sys *>(p)->set_seed(seed); }
void ref_init_J(void *p) { static_cast<cpu_sys *>(p)->init_J(); }
void ref_init_rand(void *p) { static_cast<cpu_sys *>(p)->init_rand(); }
void ref_init_order(void *p) { static_cast<cpu_sys *>(p)->init_order(); }
void ref_set_T(void *p, double T) { static_cast<cpu_sys *>(p)->set_T(T); }
void ref_sweep(void *p, int n) {
  for (int i = 0; i < n; ++i) static_cast<cpu_sys *>(p)->sweep();
}
void ref_measure(void *p, float *out) {
  std::vector<float> r = static_cast<cpu_sys *>(p)->measure();
  std::memcpy(out, r.data(), r.size() * sizeof(float));
}
long ref_get_N(void *p) { return static_cast<cpu_sys *>(p)->get_N(); }

// state access: spins, duplicated couplings (2N per axis), Boltzmann table, engine draw
void ref_get_state(void *p, int kind, uint64_t *s, uint64_t *Jx, uint64_t *Jy, uint64_t *Jz,
                   double *boltz) {
  if (kind == 2) {
    cpu_3d *q = static_cast<cpu_3d *>(p);
    long N = q->N;
    if (s) std::memcpy(s, q->s, N * 8);
    if (Jx) std::memcpy(Jx, q->Jx, 2 * N * 8);
    if (Jy) std::memcpy(Jy, q->Jy, 2 * N * 8);
    if (Jz) std::memcpy(Jz, q->Jz, 2 * N * 8);
    if (boltz) std::memcpy(boltz, q->boltz, sizeof(q->boltz));
  } else {
    cpu_2dp *q = static_cast<cpu_2dp *>(p);
    long N = q->N;
    if (s) std::memcpy(s, q->s, N * 8);
    if (Jx) std::memcpy(Jx, q->Jx, 2 * N * 8);
    if (Jy) std::memcpy(Jy, q->Jy, 2 * N * 8);
    if (boltz) std::memcpy(boltz, q->boltz, sizeof(q->boltz));
  }
}
void ref_set_spins(void *p, int kind, const uint64_t *s) {
  if (kind == 2) {
    cpu_3d *q = static_cast<cpu_3d *>(p);
    std::memcpy(q->s, s, q->N * 8);
  } else {
    cpu_2dp *q = static_cast<cpu_2dp *>(p);
    std::memcpy(q->s, s, q->N * 8);
  }
}
// next raw word of the system's mt19937_64, without disturbing it
uint64_t ref_peek_word(void *p, int kind) {
  if (kind == 2) {
    std::mt19937_64 g = static_cast<cpu_3d *>(p)->gen;
    return g();
  }
  std::mt19937_64 g = static_cast<cpu_2dp *>(p)->gen;
  return g();
}
// the system's J file (cpu_3d::save_J / cpu_2dp::save_J, src/sys_file.cpp)
void ref_save_J(void *p, const char *path) { static_cast<cpu_sys *>(p)->save_J(path); }
void ref_load_J(void *p, const char *path) { static_cast<cpu_sys *>(p)->load_J(path); }
void ref_swap(void *a, void *b, uint64_t mask) {
  std::unique_ptr<spin_t[]> m(new spin_t[1]);
  m[0] = mask;
  static_cast<cpu_sys *>(a)->swap(static_cast<cpu_sys *>(b), std::move(m));
}

// std::mt19937_64 / std::mt19937 / distributions, for the engine pins
void ref_mt64(uint64_t seed, uint64_t *out, long n) {
  std::mt19937_64 g(seed);
  for (long i = 0; i < n; ++i) out[i] = g();
}
void ref_mt32(uint32_t seed, uint32_t *out, long n) {
  std::mt19937 g(seed);
  for (long i = 0; i < n; ++i) out[i] = g();
}
void ref_uniform_real64(uint64_t seed, double *out, long n) {
  std::mt19937_64 g(seed);
  std::uniform_real_distribution<double> d(0., 1.);
  for (long i = 0; i < n; ++i) out[i] = d(g);
}
void ref_pt_draws(int nT, int32_t *pairs, double *us, long n) {
  // the draw pattern of parallel_tempering::step (src/parallel_tempering.hpp:112-114)
  std::mt19937 gen(123);
  std::uniform_int_distribution<int> uni_int(0, nT - 2);
  std::uniform_real_distribution<double> uni_double(0., 1.);
  for (long k = 0; k < n; ++k) {
    pairs[k] = uni_int(gen);
    us[k] = uni_double(gen);
  }
}

// Full parallel-tempering run as main_cpu builds it (src/main_cpu.cpp:267-296),
// step(0) repeated nsteps times.  Records pair, shared uniform, mask per step, the
// acceptance fractions (num_switch), the last energies and the final spins.
int ref_pt_run(int kind, int L, int nT, const double *Tin, double h, long seed, int sweeps,
               int nsteps, int32_t *pairs, double *us, uint64_t *masks, double *accept,
               float *E_last, uint64_t *spins) {
  Logger::verbosity = 0;
  std::vector<cpu_sys *> sys;
  for (int i = 0; i < nT; ++i) {
    cpu_sys *q = make_sys(kind, L, Tin[0], h);
    q->set_seed(seed);
    q->init_J();
    q->set_seed(seed);
    q->init_rand();
    sys.push_back(q);
  }
  std::vector<double> T(Tin, Tin + nT);
  std::string fname = "/tmp/sa_ref_harness_%.dat";
  parallel_tempering<cpu_sys> pt(sys, T, sweeps, fname);
  for (int k = 0; k < nsteps; ++k) {
    std::mt19937 g = pt.gen;
    int i = pt.uni_int(g);
    double u = pt.uni_double(g);
    pt.step(0);
    std::vector<accumulator<double>> keep = pt.num_switch;
    std::unique_ptr<spin_t[]> m = pt.creat_mask(i, u);
    pt.num_switch = keep;
    if (pairs) pairs[k] = i;
    if (us) us[k] = u;
    if (masks) masks[k] = m[0];
  }
  if (accept)
    for (int i = 0; i + 1 < nT; ++i) accept[i] = pt.num_switch[i].get();
  if (E_last)
    for (int j = 0; j < nT; ++j) std::memcpy(E_last + 64 * j, pt.E[j].data(), 64 * sizeof(float));
  if (spins)
    for (int j = 0; j < nT; ++j) {
      long N = sys[j]->get_N();
      if (kind == 2) std::memcpy(spins + j * N, static_cast<cpu_3d *>(sys[j])->s, N * 8);
      else std::memcpy(spins + j * N, static_cast<cpu_2dp *>(sys[j])->s, N * 8);
    }
  for (cpu_sys *q : sys) delete q;
  return 0;
}

// Checkpoint of a parallel-tempering run as main_cpu writes it (.tempsys,
// parallel_tempering::save, src/parallel_tempering.hpp:280-315): nsteps1 x step(0),
// save to `path`, then nsteps2 more steps recording pair, shared uniform and mask
// per step and the final spins of every system.
int ref_pt_checkpoint(int kind, int L, int nT, const double *Tin, double h, long seed, int sweeps, int nsteps1,
                      int nsteps2, const char *fname, const char *path, int32_t *pairs, double *us,
                      uint64_t *masks, uint64_t *spins) {
  Logger::verbosity = 0;
  std::vector<cpu_sys *> sys;
  for (int i = 0; i < nT; ++i) {
    cpu_sys *q = make_sys(kind, L, Tin[0], h);
    q->set_seed(seed);
    q->init_J();
    q->set_seed(seed);
    q->init_rand();
    sys.push_back(q);
  }
  std::vector<double> T(Tin, Tin + nT);
  std::string fn = fname;
  parallel_tempering<cpu_sys> pt(sys, T, sweeps, fn);
  for (int k = 0; k < nsteps1; ++k) pt.step(0);
  {
    std::ofstream f(path, std::ios::binary);
    pt.save(f);
    if (!f) return -1;
  }
  for (int k = 0; k < nsteps2; ++k) {
    std::mt19937 g = pt.gen;
    int i = pt.uni_int(g);
    double u = pt.uni_double(g);
    pt.step(0);
    std::vector<accumulator<double>> keep = pt.num_switch;
    std::unique_ptr<spin_t[]> m = pt.creat_mask(i, u);
    pt.num_switch = keep;
    if (pairs) pairs[k] = i;
    if (us) us[k] = u;
    if (masks) masks[k] = m[0];
  }
  if (spins)
    for (int j = 0; j < nT; ++j) {
      long N = sys[j]->get_N();
      if (kind == 2) std::memcpy(spins + j * N, static_cast<cpu_3d *>(sys[j])->s, N * 8);
      else std::memcpy(spins + j * N, static_cast<cpu_2dp *>(sys[j])->s, N * 8);
    }
  for (cpu_sys *q : sys) delete q;
  return 0;
}

// The reference PT driver's output files: nsteps x step(stride) with energy files
// named by `pattern` (% -> T<temperature>), then print(PRINT_MIN) and print_swap
// into swapbuf.  The files are complete when this returns.
int ref_pt_files(int kind, int L, int nT, const double *Tin, double h, long seed, int sweeps, int nsteps, int stride,
                 const char *pattern, char *swapbuf, long buflen) {
  Logger::verbosity = 0;
  std::vector<cpu_sys *> sys;
  for (int i = 0; i < nT; ++i) {
    cpu_sys *q = make_sys(kind, L, Tin[0], h);
    q->set_seed(seed);
    q->init_J();
    q->set_seed(seed);
    q->init_rand();
    sys.push_back(q);
  }
  std::vector<double> T(Tin, Tin + nT);
  std::string fn = pattern;
  std::string swaps;
  {
    parallel_tempering<cpu_sys> pt(sys, T, sweeps, fn);
    for (int k = 0; k < nsteps; ++k) pt.step(stride);
    pt.print(PRINT_MIN);
    std::ostringstream os;
    pt.print_swap(os);
    swaps = os.str();
    pt.close_file();
  }
  for (cpu_sys *q : sys) delete q;
  if ((long)swaps.size() + 1 > buflen) return -1;
  std::memcpy(swapbuf, swaps.c_str(), swaps.size() + 1);
  return (int)swaps.size();
}

// The reference's ground-state loop (main_cpu -r, src/main_cpu.cpp:344-434) on one
// parallel_tempering object: `rounds` times nsteps x step(0), then get_E_min,
// update_T (a temperature inserted between neighbours whose exchange rate is below
// 0.3, src/parallel_tempering.hpp:384-413) and reset (E_min cleared, new spins
// drawn, src/parallel_tempering.hpp:265-277); then nsteps x step(0) once more.
// Per round: the exchange rates before update_T (acc[r*cap + i], nT-1 of them), the
// temperature count and temperatures after it (nT_out[r], T_out[r*cap + j]) and
// E_min (E_min_out[r*64]).  At the end: every system's spins (spins_out[cap*N]) and
// last energies (E_out[cap*64]).  Returns the final temperature count, -1 if over
// cap.
int ref_pt_ground(int kind, int L, int nT, const double *Tin, double h, long seed, int sweeps, int nsteps, int rounds,
                  int cap, int *nT_out, double *acc, float *E_min_out, double *T_out, uint64_t *spins_out,
                  float *E_out) {
  Logger::verbosity = 0;
  std::vector<cpu_sys *> sys;
  for (int i = 0; i < nT; ++i) {
    cpu_sys *q = make_sys(kind, L, Tin[0], h);
    q->set_seed(seed);
    q->init_J();
    q->set_seed(seed);
    q->init_rand();
    sys.push_back(q);
  }
  std::vector<double> T(Tin, Tin + nT);
  std::string fn = "ground_%.dat";
  int n = 0;
  {
    parallel_tempering<cpu_sys> pt(sys, T, sweeps, fn);
    for (int r = 0; r < rounds; ++r) {
      for (int k = 0; k < nsteps; ++k) pt.step(0);
      std::vector<float> em = pt.get_E_min();
      for (int k = 0; k < 64; ++k) E_min_out[r * 64 + k] = k < (int)em.size() ? em[k] : 0.f;
      for (size_t i = 0; i < pt.num_switch.size() && (int)i < cap; ++i) acc[r * cap + i] = pt.num_switch[i].get();
      pt.update_T();
      nT_out[r] = (int)pt.T.size();
      for (size_t j = 0; j < pt.T.size() && (int)j < cap; ++j) T_out[r * cap + j] = pt.T[j];
      pt.reset();
    }
    for (int k = 0; k < nsteps; ++k) pt.step(0);
    n = (int)pt.T.size();
    if (n <= cap) {
      for (int j = 0; j < n; ++j) {
        const long N = pt.sys[j]->get_N();
        if (kind == 2) std::memcpy(spins_out + j * N, static_cast<cpu_3d *>(pt.sys[j])->s, N * 8);
        else std::memcpy(spins_out + j * N, static_cast<cpu_2dp *>(pt.sys[j])->s, N * 8);
        for (int k = 0; k < 64; ++k) E_out[j * 64 + k] = k < (int)pt.E[j].size() ? pt.E[j][k] : 0.f;
      }
    }
    for (cpu_sys *q : pt.sys)
      if (std::find(sys.begin(), sys.end(), q) == sys.end()) delete q;  // the clones update_T made
  }
  for (cpu_sys *q : sys) delete q;
  return n <= cap ? n : -1;
}

// annealing<cpu_sys> with geometric cooling, exactly the reference driver:
// annealing(T0, sweeps, alfa, sys, out); nsteps x step(stride, PRINT_E).  Writes the
// driver's text output (one line of 64 energies per measurement) into buf.
int ref_anneal_geometric(int kind, int L, double T0, double alfa, int sweeps, int nsteps,
                         int stride, double h, long seed, char *buf, long buflen) {
  Logger::verbosity = 0;
  cpu_sys *q = make_sys(kind, L, T0, h);
  q->set_seed(seed);
  q->init_J();
  q->set_seed(seed);
  q->init_rand();
  std::ostringstream os;
  os.precision(9);
  {
    annealing<cpu_sys> a(T0, sweeps, alfa, q, os);
    for (int k = 0; k < nsteps; ++k) a.step(stride, PRINT_E);
  }
  delete q;
  std::string s = os.str();
  if ((long)s.size() + 1 > buflen) return -1;
  std::memcpy(buf, s.c_str(), s.size() + 1);
  return (int)s.size();
}

// Timing of the reference's own code, for the CPU baseline: seconds spent in
// `nsweeps` cpu_*::sweep() calls of one system (setup excluded), the measurement
// the reference's -t mode makes (src/main_cpu.cpp:183-195).
double ref_time_sweeps(int kind, int L, double T, long seed, int nsweeps) {
  Logger::verbosity = 0;
  cpu_sys *q = make_sys(kind, L, T, 0.0);
  q->set_seed(seed);
  q->init_J();
  q->set_seed(seed);
  q->init_rand();
  auto t0 = std::chrono::steady_clock::now();
  for (int i = 0; i < nsweeps; ++i) q->sweep();
  auto t1 = std::chrono::steady_clock::now();
  delete q;
  return std::chrono::duration<double>(t1 - t0).count();
}

// Seconds spent in `nsteps` x parallel_tempering::step(0) (OpenMP over the
// temperatures, src/parallel_tempering.hpp:86) with `nthreads` threads (0 = the
// OpenMP default); setup excluded.
double ref_time_pt(int kind, int L, int nT, const double *Tin, long seed, int sweeps, int nsteps, int nthreads) {
  Logger::verbosity = 0;
  if (nthreads > 0) omp_set_num_threads(nthreads);
  std::vector<cpu_sys *> sys;
  for (int i = 0; i < nT; ++i) {
    cpu_sys *q = make_sys(kind, L, Tin[0], 0.0);
    q->set_seed(seed);
    q->init_J();
    q->set_seed(seed);
    q->init_rand();
    sys.push_back(q);
  }
  std::vector<double> T(Tin, Tin + nT);
  std::string fname = "/tmp/sa_ref_harness_%.dat";
  double dt;
  {
    parallel_tempering<cpu_sys> pt(sys, T, sweeps, fname);
    auto t0 = std::chrono::steady_clock::now();
    for (int k = 0; k < nsteps; ++k) pt.step(0);
    auto t1 = std::chrono::steady_clock::now();
    dt = std::chrono::duration<double>(t1 - t0).count();
  }
  for (cpu_sys *q : sys) delete q;
  return dt;
}

int ref_omp_max_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
