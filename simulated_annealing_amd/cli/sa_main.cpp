// sa_main -- command-line driver with the reference's flags (annealing_gpu, src/main_gpu.cu:56-209),
// running the MI355X systems of include/sa/hip_sys.hpp through the annealing and
// parallel-tempering drivers of include/sa.
//
//   sa_main -d 3 -L 64 -T 3.0 -a 0.9 -S 100 -s 100 [-o out.txt]        geometric anneal (PRINT_E)
//   sa_main -d 3 -L 32 -P 0.5 -P 0.8 -P 1.2 -S 10 -s 500 [--swap]       parallel tempering
//   sa_main -d 3 -L 128 -s 200 -t                                       seconds per spin flip
//
// Flags (reference meaning and defaults): -s/--steps cooling or tempering steps (100),
// -a/--a cooling factor (0.9), -S/--sweeps sweeps per step (100), -L/--size (64),
// -T/--T start temperature (2.26) | -P/--parallel_tempering temperature (repeatable),
// -m/--h field (0), -o/--output (stdout), --outputT tempering file pattern (out_%.dat),
// --seed (1234), -j/--load_J, -J/--save_J, -d/--dim 2p | 3 (2p), -t/--timing,
// --swap (print exchange rates), -i/--image, -v/--verbosity (accepted, unused).
// Additions: -d 2 selects the periodic 2D system (the reference's cpu_2d), and
// --rng replay|counter picks the random source (replay: the reference's mt19937_64
// stream, the default; counter: Philox computed on the device).
// Not provided: -r/--repeat and its --tmp/--home/--maxiter/--reload file juggling,
// --Pb/--P_file temperature blocks; they are refused with a message.
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "sa/annealing.hpp"
#include "sa/hip_sys.hpp"
#include "sa/parallel_tempering.hpp"
#include "sa_hip.h"

namespace {

struct options {
  int steps = 100, sweeps = 100, L = 64;
  double a = 0.9, T = 2.26, h = 0.0;
  bool T_set = false, timing = false, swap = false, image = false;
  std::vector<double> P;
  std::string out, outT = "out_%.dat", load_J, save_J, dim = "2p", rng = "replay";
  long seed = 1234;
};

[[noreturn]] void usage(const std::string &why) {
  std::cerr << "sa_main: " << why << "\n"
            << "usage: sa_main [-d 2p|2|3] [-L size] (-T T0 [-a alfa] | -P T -P T ...) [-S sweeps] [-s steps]\n"
            << "               [-m h] [-o file] [--outputT pattern] [--seed n] [-j Jfile] [-J Jfile]\n"
            << "               [--swap] [-i] [-t] [--rng replay|counter]\n";
  std::exit(2);
}

options parse(int argc, char **argv) {
  options o;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto val = [&]() -> std::string {
      if (i + 1 >= argc) usage("missing value after " + a);
      return argv[++i];
    };
    if (a == "-s" || a == "--steps") o.steps = std::atoi(val().c_str());
    else if (a == "-a" || a == "--a") o.a = std::atof(val().c_str());
    else if (a == "-S" || a == "--sweeps") o.sweeps = std::atoi(val().c_str());
    else if (a == "-L" || a == "--size") o.L = std::atoi(val().c_str());
    else if (a == "-T" || a == "--T") o.T = std::atof(val().c_str()), o.T_set = true;
    else if (a == "-P" || a == "--parallel_tempering") o.P.push_back(std::atof(val().c_str()));
    else if (a == "-m" || a == "--h") o.h = std::atof(val().c_str());
    else if (a == "-o" || a == "--output") o.out = val();
    else if (a == "--outputT") o.outT = val();
    else if (a == "--seed") o.seed = std::atol(val().c_str());
    else if (a == "-j" || a == "--load_J") o.load_J = val();
    else if (a == "-J" || a == "--save_J") o.save_J = val();
    else if (a == "-d" || a == "--dim") o.dim = val();
    else if (a == "-v" || a == "--verbosity") (void)val();
    else if (a == "--rng") o.rng = val();
    else if (a == "-t" || a == "--timing") o.timing = true;
    else if (a == "--swap") o.swap = true;
    else if (a == "-i" || a == "--image") o.image = true;
    else if (a == "-r" || a == "--repeat" || a == "--Pb" || a == "--P_file" || a == "--reload" || a == "--tmp" ||
             a == "--home" || a == "--maxiter")
      usage(a + " is not provided by this driver (the ground-state loop lives in the C ABI: sa_pt_*)");
    else
      usage("unknown argument " + a);
  }
  if (o.dim != "2p" && o.dim != "2" && o.dim != "3") usage("-d must be 2p, 2 or 3");
  if (o.rng != "replay" && o.rng != "counter") usage("--rng must be replay or counter");
  if (!o.timing && o.T_set == !o.P.empty()) usage("exactly one of -T and -P is required");
  if (!o.P.empty() && o.P.size() < 2) usage("parallel tempering needs at least two -P temperatures");
  return o;
}

int kind_of(const options &o) { return o.dim == "3" ? SA_KIND_3D : (o.dim == "2" ? SA_KIND_2D : SA_KIND_2P); }

sa::hip_sys *make_system(const options &o, double T) {
  const int k = kind_of(o);
  sa::hip_sys *s = k == SA_KIND_3D ? static_cast<sa::hip_sys *>(new hip_3d(o.L, T, o.h))
                   : k == SA_KIND_2D ? static_cast<sa::hip_sys *>(new hip_2d_periodic(o.L, T, o.h))
                                     : static_cast<sa::hip_sys *>(new hip_2d(o.L, T, o.h));
  s->set_rng(o.rng == "counter" ? SA_RNG_COUNTER : SA_RNG_REPLAY);
  return s;
}

// -t: seconds per replica spin flip over `steps` single sweeps, each timed with HIP
// events on the system's stream (src/main_gpu.cu:152-209).  One word holds 64
// replicas, so a sweep is N*64 flip attempts (the reference's CUDA classes keep two
// lattices and divide by N*64*2).
int timing(const options &o) {
  sa_ladder *l = nullptr;
  if (sa_ladder_create(kind_of(o), o.L, 1, &l) != SA_OK) {
    std::cerr << "sa_main: " << sa_last_error() << "\n";
    return 1;
  }
  sa_ladder_set_rng(l, SA_RNG_COUNTER);  // the device path (the reference times its GPU classes)
  sa_ladder_seed(l, (uint64_t)o.seed);
  sa_ladder_init_couplings(l);
  sa_ladder_init_spins(l);
  sa_ladder_set_temperature(l, 0, o.T, o.h, nullptr);
  sa_ladder_sweep(l, 2);  // first launches (code load) outside the timing
  sa_ladder_sync(l);
  hipStream_t st = static_cast<hipStream_t>(sa_ladder_stream(l));
  hipEvent_t a, b;
  (void)hipEventCreate(&a);
  (void)hipEventCreate(&b);
  const double flips = (double)sa_ladder_sites(l) * 64;
  double t = 0, t2 = 0, tmin = NAN, tmax = NAN;
  for (int i = 0; i < o.steps; ++i) {
    (void)hipEventRecord(a, st);
    sa_ladder_sweep(l, 1);
    (void)hipEventRecord(b, st);
    (void)hipEventSynchronize(b);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, a, b);
    t += ms;
    t2 += (double)ms * ms;
    tmin = std::fmin(tmin, ms);
    tmax = std::fmax(tmax, ms);
  }
  hipDeviceProp_t prop;
  int dev = 0;
  (void)hipGetDevice(&dev);
  (void)hipGetDeviceProperties(&prop, dev);
  const double mean = t / (o.steps * 1000.0 * flips);
  const double var = t2 / (o.steps * 1e6 * flips * flips) - mean * mean;
  std::cout << "libsa_hip " << sa_version() << "\n" << prop.name << "\n"
            << o.dim << "\t" << o.L << "\t" << o.steps << "\n"
            << mean << "\t" << std::sqrt(std::fmax(var, 0.0)) << "\t" << tmin / (1000.0 * flips) << "\t"
            << tmax / (1000.0 * flips) << "\n";
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  sa_ladder_destroy(l);
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  const options o = parse(argc, argv);
  if (o.timing) return timing(o);
  std::string info = "#";
  for (int i = 0; i < argc; ++i) info += std::string(" ") + argv[i];
  info += "\n# Version : libsa_hip " + std::to_string(sa_version()) + "\n";

  std::ostream *fp = &std::cout;
  std::ofstream fout;
  if (!o.out.empty()) {
    fout.open(o.out);
    if (!fout.is_open()) {
      std::cerr << "sa_main: cannot open output file " << o.out << "\n";
      return 1;
    }
    fp = &fout;
  }
  // systems as main_gpu sets them up (src/main_gpu.cu:250-276)
  const int num = o.P.empty() ? 1 : (int)o.P.size();
  std::vector<gpu_sys *> sys;
  for (int i = 0; i < num; ++i) {
    sys.push_back(make_system(o, o.T));
    if (!o.load_J.empty()) {
      sys[i]->load_J(o.load_J);
    } else {
      sys[i]->set_seed(o.seed);
      sys[i]->init_J();
      if (!o.save_J.empty()) sys[i]->save_J(o.save_J);
    }
    sys[i]->set_seed(o.seed);
    sys[i]->init_rand();
  }
  if (o.T_set) {
    sa::annealing<gpu_sys> annealer(o.T, o.sweeps, o.a, sys[0], *fp);
    (*fp) << info;
    for (int i = 0; i < o.steps; ++i) annealer.step(1, sa::PRINT_E);
  } else {
    if (o.swap) (*fp) << info;
    std::string outT = o.outT;
    sa::parallel_tempering<gpu_sys> tempering(sys, o.P, o.sweeps, outT);
    tempering.set_info(info, sa::PRINT_E);
    for (int j = 0; j < o.steps; ++j) tempering.step(1);
    if (o.swap) tempering.print_swap(*fp);
    tempering.print(sa::PRINT_E);
  }
  if (o.image)
    for (gpu_sys *s : sys) s->save_sys("gpu_sys_");
  for (gpu_sys *s : sys) {
    s->gpu_free();
    delete s;
  }
  return 0;
}
