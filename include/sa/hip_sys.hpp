// hip_sys.hpp -- MI355X systems behind the reference's gpu_sys interface.
//
//   hip_3d          3D periodic           replaces gpu_3d (src/gpu_3d.cu), semantics of cpu_3d
//   hip_2d          2D open x, periodic y replaces gpu_2d (src/gpu_2d.cu), semantics of cpu_2dp
//   hip_2d_periodic 2D fully periodic     semantics of cpu_2d (CPU-only in the reference)
//
// Each system is one slot of a device ladder (sa_hip.h).  A system built with
// (L, T, h) owns a one-slot ladder.  Systems built with (ladder, index) are
// slots of a shared ladder: they share couplings and the random stream, exactly
// what the reference's main programs produce by seeding every temperature alike
// (src/main_gpu.cu:262-276); sweep() on such a slot advances the whole ladder,
// because a shared stream can only advance for all slots together, and the
// drivers sweep all temperatures in lock step anyway.
//
// Results are bit-exact with the reference CPU classes: the couplings and initial
// spins come from mt19937_64 as in cpu_3d::init_J / init_rand, and the sweep
// consumes the same raw words in the same order (SA_RNG_REPLAY) or Philox words
// (SA_RNG_COUNTER).  Errors are reported on stderr and execution continues, as in
// the reference (LOG(LOG_ERROR), src/setup.cuh:16-20); constructors throw.
#pragma once

#include <cstdint>
#include <cstring>
#include <fstream>
#include <iostream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "sa/gpu_sys.hpp"
#include "sa_hip.h"

namespace sa {

inline void log_error(const std::string &what, int rc) {
  std::cerr << "Error:   " << what << ": " << sa_status_string(rc) << " (" << sa_last_error() << ")\n";
}

// Owning handle of a device ladder.
class ladder {
 public:
  ladder(int kind, int L, int nsys) : kind_(kind), L_(L) {
    const int rc = sa_ladder_create(kind, L, nsys, &l_);
    if (rc != SA_OK) throw std::runtime_error(std::string("sa_ladder_create: ") + sa_last_error());
    sa_ladder_seed(l_, 1234);  // the reference classes seed their engine with 1234 (src/cpu_3d.cpp:22,45)
  }
  ~ladder() { release(); }
  ladder(const ladder &) = delete;
  ladder &operator=(const ladder &) = delete;
  void release() {
    if (l_) sa_ladder_destroy(l_);
    l_ = nullptr;
  }
  sa_ladder *get() const { return l_; }
  int kind() const { return kind_; }
  int L() const { return L_; }
  int nsys() const { return l_ ? sa_ladder_nsys(l_) : 0; }
  long N() const { return l_ ? (long)sa_ladder_sites(l_) : 0; }
  // measure every slot: E/M per slot and lane (lane order)
  int measure_all(std::vector<int32_t> &E, std::vector<int32_t> &M) {
    E.resize((size_t)nsys() * 64);
    M.resize((size_t)nsys() * 64);
    return sa_ladder_measure(l_, E.data(), M.data());
  }
  // Set when a system record loaded into one slot wrote the state all slots share
  // (couplings, engine, counter stream): a later slot's record must agree with it.
  // parallel_tempering::load clears it at the start of a checkpoint.
  bool shared_loaded = false;
  uint64_t shared_rng = 0;  // the random source that record set (sa_rng_mode)

 private:
  sa_ladder *l_ = nullptr;
  int kind_, L_;
};

class hip_sys : public gpu_sys {
 public:
  // a system with its own one-slot ladder (the reference's gpu_3d(L, T, h))
  hip_sys(int kind, int L, double T, double h) : lad_(std::make_shared<ladder>(kind, L, 1)), idx_(0), T_(T), h_(h) {
    apply_T();
  }
  // slot `index` of a shared ladder
  hip_sys(std::shared_ptr<ladder> lad, int index, double T, double h) : lad_(std::move(lad)), idx_(index), T_(T), h_(h) {
    if (!lad_ || index < 0 || index >= lad_->nsys()) throw std::invalid_argument("hip_sys: bad ladder slot");
    apply_T();
  }

  // ---- gpu_sys ------------------------------------------------------------------
  void sweep() override { check("sweep", sa_ladder_sweep(l(), 1)); }
  void set_T(double T) override {
    T_ = T;
    apply_T();
  }
  void set_h(double h) override {
    h_ = h;
    apply_T();
  }
  void set_seed(long long seed) override {
    seed_ = (uint64_t)seed;
    check("set_seed", sa_ladder_seed(l(), seed_));
  }
  // float(E + h*M) per lane, index k = lane k (src/cpu_3d.cpp:360-371)
  std::vector<float> measure() override {
    std::vector<int32_t> E, M;
    std::vector<float> r(64, 0.f);
    if (!check("measure", lad_->measure_all(E, M))) return r;
    for (int k = 0; k < 64; ++k) r[k] = (float)((double)E[idx_ * 64 + k] + h_ * (double)M[idx_ * 64 + k]);
    return r;
  }
  void gpu_free() override {
    if (lad_ && lad_.use_count() == 1) lad_->release();
  }
  // reference J file: int d; long N; Jx[2N]; Jy[2N]; [Jz[2N]] (src/sys_file.cpp:10-110).
  // As the reference, an error is logged and the couplings stay as they were
  // (src/cpu_3d.cpp:379-399).
  void load_J(std::string fname) override {
    if (sa_ladder_load_J(l(), fname.c_str()) != SA_OK)
      std::cerr << "Error:   failed to load J from \"" << fname << "\" (" << sa_last_error() << ")\n";
  }
  void save_J(std::string fname) override {
    if (sa_ladder_save_J(l(), fname.c_str()) != SA_OK)
      std::cerr << "Error:   failed to save J to \"" << fname << "\" (" << sa_last_error() << ")\n";
  }
  void init_J() override { check("init_J", sa_ladder_init_couplings(l())); }
  void init_rand() override { check("init_rand", sa_ladder_init_spins(l())); }
  // one PBM image per lane (src/cpu_3d.cpp:405-425)
  void save_sys(std::string prefix) override {
    std::vector<uint64_t> s = spins();
    const int L = lad_->L();
    const long rows = dim() == 3 ? (long)L * L : L;
    for (int i = 0; i < 64; ++i) {
      std::ofstream file(prefix + std::to_string(i) + ".pbm");
      file << "P1\n" << L << " " << rows << "\n";
      for (long j = 0; j < rows; ++j) {
        for (int k = 0; k < L; ++k) file << (((s[j * L + k] >> i) & 1ULL) ? "1 " : "0 ");
        file << "\n";
      }
    }
  }
  long get_N() override { return lad_->N(); }
  void swap(gpu_sys *other, std::unique_ptr<spin_t[]> mask) override {
    hip_sys *o = dynamic_cast<hip_sys *>(other);
    // the reference casts to its own class (cpu_3d / cpu_2dp, src/cpu_3d.cpp:429-444,
    // src/cpu_2dp.cpp:426-442): any two 2D systems swap (cpu_2d is a cpu_2dp)
    if (!o || o->get_N() != get_N() || (o->lad_->kind() == SA_KIND_3D) != (lad_->kind() == SA_KIND_3D)) {
      std::cerr << "Error:   conversion error\n";
      return;
    }
    if (o->lad_ == lad_) {
      check("swap", sa_ladder_swap(l(), idx_, o->idx_, (uint64_t)mask[0]));
      return;
    }
    std::vector<uint64_t> a = spins(), b = o->spins();  // systems on different ladders
    const uint64_t m = mask[0];
    for (size_t i = 0; i < a.size(); ++i) {
      const uint64_t d = (a[i] ^ b[i]) & m;
      a[i] ^= d;
      b[i] ^= d;
    }
    set_spins(a);
    o->set_spins(b);
  }
  // Record: int L; double T, h; s[N]; Jx[2N]; Jy[2N]; [Jz[2N]]; the engine
  // (312 words + position); uniform_real (0,1) and uniform_int (0,N-1) parameters
  // -- the bytes of cpu_3d::save / cpu_2dp::save (src/cpu_3d.cpp:449-461,
  // src/cpu_2dp.cpp:444-455).  With the counter RNG a trailer follows ("SAHIPX01",
  // rng mode, counter seed, sweep index): that state has no place in the reference's
  // record.  With the replay RNG the record is the reference's byte for byte.
  std::ostream &save(std::ostream &os) override {
    int L = lad_->L();
    os.write(reinterpret_cast<const char *>(&L), sizeof L);
    os.write(reinterpret_cast<const char *>(&T_), sizeof T_);
    os.write(reinterpret_cast<const char *>(&h_), sizeof h_);
    std::vector<uint64_t> s = spins();
    os.write(reinterpret_cast<const char *>(s.data()), s.size() * 8);
    for (auto &a : duplicated_couplings()) os.write(reinterpret_cast<const char *>(a.data()), a.size() * 8);
    uint64_t eng[313];
    sa_ladder_get_engine(l(), eng);
    os.write(reinterpret_cast<const char *>(eng), sizeof eng);
    const double ud[2] = {0.0, 1.0};
    const int ui[2] = {0, (int)(get_N() - 1)};
    os.write(reinterpret_cast<const char *>(ud), sizeof ud);
    os.write(reinterpret_cast<const char *>(ui), sizeof ui);
    if (rng_mode_ == SA_RNG_REPLAY) return os;
    uint64_t trailer[4] = {0, (uint64_t)rng_mode_, 0, 0};
    std::memcpy(&trailer[0], "SAHIPX01", 8);
    sa_ladder_get_sweep_index(l(), &trailer[3]);
    sa_ladder_get_seed(l(), &trailer[2]);  // the ladder's key (slots share it)
    os.write(reinterpret_cast<const char *>(trailer), sizeof trailer);
    return os;
  }
  std::istream &load(std::istream &is) override {
    int L = 0;
    double T = 0, h = 0;
    is.read(reinterpret_cast<char *>(&L), sizeof L);
    is.read(reinterpret_cast<char *>(&T), sizeof T);
    is.read(reinterpret_cast<char *>(&h), sizeof h);
    if (!is || L != lad_->L()) {
      std::cerr << "Error:   load: lattice size " << L << " does not match " << lad_->L() << "\n";
      is.setstate(std::ios::failbit);
      return is;
    }
    const long N = get_N();
    std::vector<uint64_t> s(N);
    is.read(reinterpret_cast<char *>(s.data()), N * 8);
    std::vector<uint64_t> dup[3], j[3];
    for (int a = 0; a < dim(); ++a) {
      dup[a].resize(2 * N);
      is.read(reinterpret_cast<char *>(dup[a].data()), 16 * N);
      j[a].resize(N);
      for (long i = 0; i < N; ++i) j[a][i] = dup[a][2 * i];
    }
    uint64_t eng[313];
    is.read(reinterpret_cast<char *>(eng), sizeof eng);
    double ud[2];
    int ui[2];
    is.read(reinterpret_cast<char *>(ud), sizeof ud);
    is.read(reinterpret_cast<char *>(ui), sizeof ui);
    if (!is) {
      std::cerr << "Error:   load: truncated system record\n";
      return is;
    }
    // optional trailer (counter RNG state); a reference record has none
    uint64_t trailer[4] = {0, (uint64_t)SA_RNG_REPLAY, 0, 0};
    const std::streampos mark = is.tellg();
    char magic[8];
    is.read(magic, 8);
    if (is && std::memcmp(magic, "SAHIPX01", 8) == 0) {
      is.read(reinterpret_cast<char *>(&trailer[1]), 3 * sizeof(uint64_t));
      if (!is) {
        std::cerr << "Error:   load: truncated system record\n";
        return is;
      }
    } else {
      is.clear();
      is.seekg(mark);
    }
    // Slots of one ladder share couplings and the random stream.  A record that
    // disagrees with what an earlier slot's record put there (a system inserted by
    // update_T restarts its engine at seed 1234, src/cpu_3d.cpp:47-80, and the
    // reference restores every system's engine, src/parallel_tempering.hpp:362-372)
    // moves this system onto a ladder of its own instead of overwriting the others'.
    if (lad_->nsys() > 1 && lad_->shared_loaded && !matches_shared(j, eng, trailer)) {
      auto own = std::make_shared<ladder>(lad_->kind(), lad_->L(), 1);
      lad_ = std::move(own);
      idx_ = 0;
    }
    T_ = T;
    h_ = h;
    apply_T();
    set_spins(s);
    check("load", sa_ladder_set_couplings(l(), j[0].data(), j[1].data(), dim() == 3 ? j[2].data() : nullptr));
    set_rng((int)trailer[1]);
    if (trailer[1] != SA_RNG_REPLAY) {  // the counter stream's key and position
      seed_ = trailer[2];
      check("load", sa_ladder_seed(l(), seed_));
      check("load", sa_ladder_set_sweep_index(l(), trailer[3]));
    }
    check("load", sa_ladder_set_engine(l(), eng));
    lad_->shared_loaded = true;
    lad_->shared_rng = trailer[1];
    return is;
  }
  double get_T() override { return T_; }
  // An independent copy on its own one-slot ladder, as the reference's clone()
  // makes it (cpu_3d::clone / cpu_2dp::clone, src/cpu_3d.cpp:494-496,
  // src/cpu_2dp.cpp:481-483, through the copy constructors src/cpu_3d.cpp:47-80,
  // src/cpu_2dp.cpp:53-86): spins, couplings, T and h are copied, the engine is not
  // -- the copy's stream restarts from seed 1234 (here the counter stream too, at
  // sweep 0).  cpu_2d has no clone of its own, so the copy of a 2D periodic system
  // is a cpu_2dp: open in x.  The copy follows that.
  gpu_sys *clone() const override {
    const int kind = lad_->kind() == SA_KIND_2D ? SA_KIND_2P : lad_->kind();
    hip_sys *c = new hip_sys(kind, lad_->L(), T_, h_);  // a new ladder is seeded with 1234
    c->set_rng(rng_mode_);
    c->set_spins(spins());
    const long N = lad_->N();
    std::vector<uint64_t> j[3];
    for (auto &a : j) a.resize(N);
    check("clone", sa_ladder_get_couplings(l(), j[0].data(), j[1].data(), j[2].data()));
    check("clone", sa_ladder_set_couplings(c->l(), j[0].data(), j[1].data(), j[2].data()));
    return c;
  }

  // ---- beyond gpu_sys -------------------------------------------------------------
  void set_rng(int mode) {
    rng_mode_ = mode;
    check("set_rng", sa_ladder_set_rng(l(), mode));
  }
  std::vector<uint64_t> spins() const {
    std::vector<uint64_t> s(lad_->N());
    check("get_spins", sa_ladder_get_spins(l(), idx_, s.data()));
    return s;
  }
  void set_spins(const std::vector<uint64_t> &s) { check("set_spins", sa_ladder_set_spins(l(), idx_, s.data())); }
  std::shared_ptr<ladder> get_ladder() const { return lad_; }
  int index() const { return idx_; }
  double get_h() const { return h_; }
  int dim() const { return lad_->kind() == SA_KIND_3D ? 3 : 2; }

 protected:
  sa_ladder *l() const { return lad_->get(); }
  static bool check(const char *what, int rc) {
    if (rc == SA_OK) return true;
    log_error(what, rc);
    return false;
  }
  void apply_T() { check("set_T", sa_ladder_set_temperature(l(), idx_, T_, h_, nullptr)); }
  // does a loaded record's shared state (deduplicated couplings j, engine, counter
  // trailer) equal the ladder's?
  bool matches_shared(const std::vector<uint64_t> (&j)[3], const uint64_t (&eng)[313],
                      const uint64_t (&trailer)[4]) const {
    uint64_t cur[313];
    if (sa_ladder_get_engine(l(), cur) != SA_OK || std::memcmp(cur, eng, sizeof cur) != 0) return false;
    if (lad_->shared_rng != trailer[1]) return false;
    if (trailer[1] != SA_RNG_REPLAY) {
      uint64_t seed = 0, t = 0;
      sa_ladder_get_seed(l(), &seed);
      sa_ladder_get_sweep_index(l(), &t);
      if (seed != trailer[2] || t != trailer[3]) return false;
    }
    const long N = lad_->N();
    std::vector<uint64_t> c[3];
    for (auto &a : c) a.resize(N);
    if (sa_ladder_get_couplings(l(), c[0].data(), c[1].data(), c[2].data()) != SA_OK) return false;
    for (int a = 0; a < dim(); ++a)
      if (c[a] != j[a]) return false;
    return true;
  }
  // couplings in the reference's duplicated layout: [2i] = own bond, [2i+1] = the
  // +axis neighbour's bond (src/cpu_3d.cpp:290-299)
  std::vector<std::vector<uint64_t>> duplicated_couplings() const {
    const long N = lad_->N();
    const int L = lad_->L();
    std::vector<uint64_t> j[3];
    for (auto &a : j) a.resize(N);
    check("get_couplings", sa_ladder_get_couplings(l(), j[0].data(), j[1].data(), j[2].data()));
    std::vector<std::vector<uint64_t>> out(dim(), std::vector<uint64_t>(2 * N));
    const long L2 = (long)L * L;
    for (long i = 0; i < N; ++i) {
      const long x = i % L, y = (i / L) % L;
      const long east = i - x + (x + 1) % L;
      const long south = i - y * L + ((y + 1) % L) * L;
      out[0][2 * i] = j[0][i];
      out[0][2 * i + 1] = j[0][east];
      out[1][2 * i] = j[1][i];
      out[1][2 * i + 1] = j[1][south];
      if (dim() == 3) {
        const long down = (i + L2) % N;
        out[2][2 * i] = j[2][i];
        out[2][2 * i + 1] = j[2][down];
      }
    }
    return out;
  }

  std::shared_ptr<ladder> lad_;
  int idx_;
  double T_, h_;
  int rng_mode_ = SA_RNG_REPLAY;
  uint64_t seed_ = 1234;
};

}  // namespace sa

class hip_3d : public sa::hip_sys {
 public:
  hip_3d(int L, double T, double h) : hip_sys(SA_KIND_3D, L, T, h) {}
  hip_3d(std::shared_ptr<sa::ladder> lad, int index, double T, double h) : hip_sys(std::move(lad), index, T, h) {}
};
class hip_2d : public sa::hip_sys {
 public:
  hip_2d(int L, double T, double h) : hip_sys(SA_KIND_2P, L, T, h) {}
  hip_2d(std::shared_ptr<sa::ladder> lad, int index, double T, double h) : hip_sys(std::move(lad), index, T, h) {}
};
class hip_2d_periodic : public sa::hip_sys {
 public:
  hip_2d_periodic(int L, double T, double h) : hip_sys(SA_KIND_2D, L, T, h) {}
  hip_2d_periodic(std::shared_ptr<sa::ladder> lad, int index, double T, double h)
      : hip_sys(std::move(lad), index, T, h) {}
};

#ifdef SA_REFERENCE_NAMES
// drop-in spellings for code written against the reference's CUDA classes
typedef hip_3d gpu_3d;
typedef hip_2d gpu_2d;
#endif
