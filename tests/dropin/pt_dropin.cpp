// The drop-in of the integration notes, compiled as a program: the reference's
// main_gpu system set-up (src/main_gpu.cu:262-276) -- one gpu_3d / gpu_2d per
// temperature, set_seed, init_J, set_seed, init_rand -- run by
// parallel_tempering<gpu_sys> (src/parallel_tempering.hpp:54-115), the systems
// released with gpu_free() and deleted through the base pointer.  The spellings
// gpu_3d / gpu_2d come from SA_REFERENCE_NAMES.  A J-file round trip (save_J, load_J)
// runs before the tempering.
//
//   pt_dropin <kind 0|2> <L> <seed> <sweeps> <nsteps> <spins_out> T1 T2 ...
//
// prints one line per exchange attempt ("step <pair> <uniform bits> <mask>") and one
// line per system with the driver's energies of the last step ("E <index> <64 float
// bit patterns>"), and writes the systems' spin words to spins_out.
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#define SA_REFERENCE_NAMES
#include "sa/annealing.hpp"
#include "sa/hip_sys.hpp"
#include "sa/parallel_tempering.hpp"

// a user class written against the reference's gpu_sys, without get_T / clone:
// it must still compile
class user_sys : public gpu_sys {
 public:
  void sweep() override {}
  void set_T(double) override {}
  void set_h(double) override {}
  void set_seed(long long) override {}
  std::vector<float> measure() override { return std::vector<float>(64, 0.f); }
  void gpu_free() override {}
  void load_J(std::string) override {}
  void save_J(std::string) override {}
  void init_J() override {}
  void init_rand() override {}
  void save_sys(std::string) override {}
  long get_N() override { return 0; }
  void swap(gpu_sys *, std::unique_ptr<spin_t[]>) override {}
  std::ostream &save(std::ostream &s) override { return s; }
  std::istream &load(std::istream &s) override { return s; }
};

int main(int argc, char **argv) {
  if (argc < 9) {
    std::fprintf(stderr, "usage: pt_dropin kind L seed sweeps nsteps spins_out T1 T2 ...\n");
    return 2;
  }
  const int kind = std::atoi(argv[1]), L = std::atoi(argv[2]);
  const long long seed = std::atoll(argv[3]);
  const long sweeps = std::atol(argv[4]);
  const int nsteps = std::atoi(argv[5]);
  const std::string spins_out = argv[6];
  std::vector<double> Ts;
  for (int i = 7; i < argc; ++i) Ts.push_back(std::atof(argv[i]));
  const double h = 0.0;
  gpu_sys *probe = new user_sys();  // the reference-style class links and runs
  delete probe;

  std::vector<gpu_sys *> sys;
  for (size_t i = 0; i < Ts.size(); ++i) {
    if (kind == SA_KIND_3D) sys.push_back(new gpu_3d(L, Ts[i], h));
    else sys.push_back(new gpu_2d(L, Ts[i], h));
    sys[i]->set_seed(seed);
    sys[i]->init_J();
    sys[i]->set_seed(seed);
    sys[i]->init_rand();
  }
  // J file round trip on the first system: the couplings must come back unchanged
  const std::string jf = spins_out + ".J";
  sys[0]->save_J(jf);
  sys[0]->init_J();  // draws new couplings ...
  sys[0]->load_J(jf);  // ... replaced by the saved ones
  sys[0]->set_seed(seed);
  sys[0]->init_rand();
  std::string fname = spins_out + "_%.dat";
  sa::parallel_tempering<gpu_sys> tempering(sys, Ts, sweeps, fname);
  std::vector<sa::swap_record> log;
  tempering.set_swap_log(&log);
  for (int k = 0; k < nsteps; ++k) tempering.step(0);
  for (const auto &r : log) {
    uint64_t ub;
    std::memcpy(&ub, &r.u, 8);
    std::printf("step %d %016" PRIx64 " %016" PRIx64 "\n", r.pair, ub, r.mask);
  }
  std::ofstream out(spins_out, std::ios::binary);
  // the driver's energies of the last step (measured before its exchange, as the
  // reference's parallel_tempering keeps them)
  const std::vector<std::vector<float>> &E = tempering.get_E();
  for (size_t i = 0; i < sys.size(); ++i) {
    const std::vector<float> &e = E[i];
    std::printf("E %zu", i);
    for (float v : e) {
      uint32_t b;
      std::memcpy(&b, &v, 4);
      std::printf(" %08x", b);
    }
    std::printf("\n");
    std::vector<uint64_t> s = static_cast<sa::hip_sys *>(tempering.systems()[i])->spins();
    out.write(reinterpret_cast<const char *>(s.data()), s.size() * 8);
  }
  for (gpu_sys *s : sys) {
    s->gpu_free();
    delete s;  // through the base pointer
  }
  return 0;
}
