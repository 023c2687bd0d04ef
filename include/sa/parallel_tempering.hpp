// parallel_tempering.hpp -- replica-exchange driver over any system type.
//
// Same interface and results as the reference's parallel_tempering<system>
// (src/parallel_tempering.hpp:17-413): temperatures sorted ascending, `steps`
// sweeps per step() with measurements when (i % stride == 0) and on the last
// sweep, then one attempted exchange between a random adjacent pair drawn from
// mt19937(123), accepted lane by lane with one shared uniform
// (creat_mask, src/parallel_tempering.hpp:141-162).  The engines and the
// uniform/int mappings are this project's restatement of libstdc++'s
// (sa/rng.hpp), so the pair sequence does not depend on the host C++ library.
//
// MI355X batching: systems that are slots of one device ladder (hip_sys.hpp) are
// swept by one ladder-wide launch per sweep and measured by one launch, instead
// of one launch per temperature; results are the same as per-system sweeps
// because the slots are independent apart from the exchange step.
#pragma once

#include <algorithm>
#include <fstream>
#include <iostream>
#include <map>
#include <memory>
#include <regex>
#include <string>
#include <type_traits>
#include <vector>

#include "sa/annealing.hpp"
#include "sa/exchange.hpp"
#include "sa/hip_sys.hpp"
#include "sa/rng.hpp"

namespace sa {

// acceptance-rate accumulator (src/accumulator.hpp); reset() is a no-op there
// (it constructs a temporary) and is kept so for identical print_swap output.
struct accumulator {
  double data = 0;
  long long N = 0;
  void operator()(double x, long long n = 1) {
    data += x;
    N += n;
  }
  double get() const { return data != 0 ? data / (double)N : 0.; }
  void reset() {}
};

struct swap_record {
  int pair;
  double u;
  uint64_t mask;
};

template <class system>
class parallel_tempering {
 public:
  parallel_tempering(std::vector<system *> &sys, const std::vector<double> &T, long steps, std::string &filename)
      : sys_(sys), T_(T), steps_(steps), fname_(filename), uni_a_(0), uni_b_((int)T.size() - 2) {
    gen_.seed(123);
    num_switch_.assign(T_.size() - 1, accumulator{});
    if (T_.size() != sys_.size())
      std::cerr << "Error:   size difference: sys " << sys_.size() << " T " << T_.size() << "\n";
    std::sort(T_.begin(), T_.end());
    for (size_t i = 0; i < T_.size(); ++i) {
      sys_[i]->set_T(T_[i]);
      E_.push_back(std::vector<float>());
    }
  }

  // src/parallel_tempering.hpp:81-115
  void step(int stride) {
    const bool grouped = stride == 0 && all_on_ladders();
    int i;
    double u;
    if (grouped) {
      // Whole step at once: one launch sequence per ladder.  With stride 0 the step
      // uses only the coldest system's energies (E_min) and the exchanging pair's;
      // the pair is drawn first (the draws do not depend on the energies), those
      // systems are measured now and the others when E is next read -- their spins
      // do not change before that, so the values are the reference's.
      // The reference folds E[0] into E_min after every sweep of the step, and E[0]
      // changes only at the last one: the sweeps before it fold the previous step's
      // E[0] in again.  That is a no-op except after reset() (E_min zeroed, E[0]
      // still the energies from before), so it is done once here.
      if (steps_ > 1) min_E_update();
      sweep_all(steps_);
      i = uniform_int(gen_, uni_a_, uni_b_);
      u = canonical32(gen_);
      measure_some({0, i, i + 1});
      min_E_update();
    } else {
      for (long s = 0; s < steps_; ++s) {
        sweep_all(1);
        if (((stride != 0) && ((s % stride) == 0)) || s == (steps_ - 1)) measure_all();
        if (stride != 0 && (s % stride) == 0) print(PRINT_E);
        else if (stride == 0) min_E_update();
      }
      i = uniform_int(gen_, uni_a_, uni_b_);
      u = canonical32(gen_);
    }
    std::unique_ptr<spin_t[]> mask = creat_mask(i, u);
    if (log_) log_->push_back(swap_record{i, u, (uint64_t)mask[0]});
    sys_[i]->swap(sys_[i + 1], std::move(mask));
  }

  void print(print_select select) {
    refresh();
    open_file(select, false);
    if (select == PRINT_E) {
      for (size_t i = 0; i < T_.size(); ++i) {
        for (float v : E_[i]) (*files_[i]) << v << "\t";
        (*files_[i]) << "\n";
      }
    } else {
      for (float v : E_min_) (*file_E_min_) << v << "\t";
      (*file_E_min_) << std::endl;
    }
  }
  void print_swap(std::ostream &f) {
    for (size_t i = 0; i + 1 < T_.size(); ++i) f << T_[i] << " -> " << T_[i + 1] << "\t: " << num_switch_[i].get() << "\n";
  }
  void set_info(const std::string &info, print_select select) {
    open_file(select, false);
    if (select == PRINT_E) {
      for (auto &f : files_) (*f) << info;
    } else {
      (*file_E_min_) << info;
    }
  }
  // src/parallel_tempering.hpp:265-277: every system draws new spins from its own
  // engine.  The reference's engines all sit at one stream position, so each
  // system gets the same N words and every engine advances by N; slots of one
  // device ladder share one engine, so the ladder draws once for all its slots.
  void reset() {
    // the reference keeps the energies measured at the last step (all systems):
    // bring the ones not measured then up to date before the spins change
    refresh();
    for (float &e : E_min_) e = 0.f;
    std::vector<ladder *> done;
    for (system *s : sys_) {
      hip_sys *h = as_hip(s);
      if (h) {
        ladder *l = h->get_ladder().get();
        if (std::find(done.begin(), done.end(), l) != done.end()) continue;
        done.push_back(l);
      }
      s->init_rand();
    }
    for (auto &a : num_switch_) a.reset();
  }
  std::vector<float> get_E_min() const { return E_min_; }
  const std::vector<std::vector<float>> &get_E() {
    refresh();
    return E_;
  }
  const std::vector<double> &get_T() const { return T_; }
  // the systems in temperature order (update_T inserts clones)
  const std::vector<system *> &systems() const { return sys_; }
  std::vector<double> acceptance() const {
    std::vector<double> r;
    for (auto &a : num_switch_) r.push_back(a.get());
    return r;
  }
  // records (pair, shared uniform, mask) of every exchange attempt when set
  void set_swap_log(std::vector<swap_record> *log) { log_ = log; }
  // file-name pattern of the energy files (% -> T<temperature> / E_min); the
  // reference fixes it at construction
  void set_filename(const std::string &f) { fname_ = f; }
  // flush and close the energy files (parallel_tempering::close_file,
  // src/parallel_tempering.hpp:250-262)
  void close_file() {
    files_.clear();
    if (file_E_min_) file_E_min_->close();
    file_E_min_.reset();
    print_E_ = print_min_ = false;
  }

  // Insert a temperature between neighbours whose exchange rate is below 0.3,
  // cloning the hotter neighbour (src/parallel_tempering.hpp:381-413).
  void update_T() {
    refresh();
    stale_.clear();
    for (size_t i = 0; i < num_switch_.size(); i++) {
      if (num_switch_[i].get() < 0.3) {
        i++;
        system *c = static_cast<system *>(sys_[i]->clone());
        if (!c) {  // a system class without clone() (gpu_sys's default body)
          std::cerr << "Error:   update_T: system " << i << " cannot be cloned; no temperature inserted\n";
          return;
        }
        num_switch_.insert(num_switch_.begin() + i, accumulator{});
        T_.insert(T_.begin() + i, (T_[i] + T_[i - 1]) / 2.0);
        E_.insert(E_.begin() + i, std::vector<float>());
        sys_.insert(sys_.begin() + i, c);
        owned_.emplace_back(c);
        sys_[i]->set_T(T_[i]);
        if (print_E_) files_.insert(files_.begin() + i, open_T_file(T_[i], std::ios_base::out));
        uni_b_ += 1;
      }
    }
  }

  // field order of parallel_tempering::save (src/parallel_tempering.hpp:280-315)
  std::ostream &save(std::ostream &os) {
    refresh();
    auto w = [&](const void *p, size_t n) { os.write(reinterpret_cast<const char *>(p), n); };
    w(&steps_, sizeof steps_);
    w(&print_E_, sizeof print_E_);
    w(&print_min_, sizeof print_min_);
    w(&gen_, sizeof gen_);
    const int ui[2] = {uni_a_, uni_b_};
    w(ui, sizeof ui);
    const double ud[2] = {0.0, 1.0};
    w(ud, sizeof ud);
    size_t n = fname_.size();
    w(&n, sizeof n);
    w(fname_.data(), n);
    n = T_.size();
    w(&n, sizeof n);
    w(T_.data(), n * sizeof(double));
    n = E_min_.size();
    w(&n, sizeof n);
    w(E_min_.data(), n * sizeof(float));
    n = num_switch_.size();
    w(&n, sizeof n);
    w(num_switch_.data(), n * sizeof(accumulator));
    n = E_.size();
    w(&n, sizeof n);
    for (size_t i = 0; i < n; ++i) {
      size_t m = E_[i].size();
      w(&m, sizeof m);
      w(E_[i].data(), m * sizeof(float));
      sys_[i]->save(os);
    }
    return os;
  }
  std::istream &load(std::istream &is) {
    auto r = [&](void *p, size_t n) { is.read(reinterpret_cast<char *>(p), n); };
    files_.clear();
    file_E_min_.reset();
    r(&steps_, sizeof steps_);
    r(&print_E_, sizeof print_E_);
    r(&print_min_, sizeof print_min_);
    r(&gen_, sizeof gen_);
    int ui[2];
    r(ui, sizeof ui);
    uni_a_ = ui[0];
    uni_b_ = ui[1];
    double ud[2];
    r(ud, sizeof ud);
    size_t n = 0;
    r(&n, sizeof n);
    fname_.assign(n, '\0');
    r(&fname_[0], n);
    r(&n, sizeof n);
    T_.assign(n, 0.0);
    r(T_.data(), n * sizeof(double));
    const bool pe = print_E_, pm = print_min_;
    print_E_ = print_min_ = false;
    if (pe) open_file(PRINT_E, true);
    if (pm) open_file(PRINT_MIN, true);
    r(&n, sizeof n);
    E_min_.assign(n, 0.f);
    r(E_min_.data(), n * sizeof(float));
    r(&n, sizeof n);
    num_switch_.assign(n, accumulator{});
    r(num_switch_.data(), n * sizeof(accumulator));
    r(&n, sizeof n);
    if (n != sys_.size()) {  // the reference reads one record per system it has
      std::cerr << "Error:   load: the checkpoint holds " << n << " systems, the driver " << sys_.size() << "\n";
      is.setstate(std::ios::failbit);
      return is;
    }
    // a new checkpoint: every shared ladder takes its state from the first of its
    // slots' records again (hip_sys::load)
    for (system *s : sys_)
      if (hip_sys *h = as_hip(s)) h->get_ladder()->shared_loaded = false;
    E_.assign(n, std::vector<float>());
    for (size_t i = 0; i < n && i < sys_.size(); ++i) {
      size_t m = 0;
      r(&m, sizeof m);
      E_[i].assign(m, 0.f);
      r(E_[i].data(), m * sizeof(float));
      sys_[i]->load(is);
    }
    stale_.assign(sys_.size(), 0);  // the loaded energies are current
    return is;
  }

 private:
  std::unique_ptr<spin_t[]> creat_mask(int i, double rand) {
    const size_t words = E_[i].size() / 64;
    std::unique_ptr<spin_t[]> mask(new spin_t[words > 0 ? words : 1]);
    for (size_t j = 0; j < (words > 0 ? words : 1); ++j) mask[j] = 0ULL;
    int a = 0;
    const double dB = (1. / T_[i] - 1. / T_[i + 1]);
    for (size_t j = 0; j < E_[i].size(); j += 64) {
      const int n = (int)std::min<size_t>(64, E_[i].size() - j);
      mask[j / 64] = exchange_word(E_[i].data() + j, E_[i + 1].data() + j, dB, rand, n, a);
    }
    num_switch_[i](a / (double)64.);
    return mask;
  }

  void min_E_update() {
    if (E_min_.empty()) E_min_ = E_[0];
    for (size_t j = 0; j < E_min_.size() && j < E_[0].size(); ++j) E_min_[j] = std::max(E_[0][j], E_min_[j]);
  }

  static hip_sys *as_hip(system *s) {
    if constexpr (std::is_polymorphic<system>::value) return dynamic_cast<hip_sys *>(s);
    else return nullptr;
  }
  bool all_on_ladders() const {
    for (system *s : sys_)
      if (!as_hip(s)) return false;
    return true;
  }
  // One sweep (or n) of every system: each device ladder is swept once for all of
  // its slots; other systems one by one.
  void sweep_all(long n) {
    std::vector<ladder *> done;
    for (system *s : sys_) {
      hip_sys *h = as_hip(s);
      if (h) {
        ladder *l = h->get_ladder().get();
        if (std::find(done.begin(), done.end(), l) != done.end()) continue;
        done.push_back(l);
        const int rc = sa_ladder_sweep(l->get(), (int)n);
        if (rc != SA_OK) log_error("sweep", rc);
      } else {
        for (long k = 0; k < n; ++k) s->sweep();
      }
    }
  }
  // energies of the listed systems now; the rest become stale until refresh()
  void measure_some(std::vector<int> want) {
    std::sort(want.begin(), want.end());
    want.erase(std::unique(want.begin(), want.end()), want.end());
    stale_.assign(sys_.size(), 1);
    measure_list(want);
  }
  void refresh() {
    std::vector<int> want;
    for (size_t j = 0; j < stale_.size() && j < sys_.size(); ++j)
      if (stale_[j]) want.push_back((int)j);
    if (!want.empty()) measure_list(want);
  }
  void measure_list(const std::vector<int> &want) {
    std::map<ladder *, std::vector<int>> by;  // ladder -> driver indices
    for (int j : want) {
      hip_sys *h = as_hip(sys_[j]);
      if (!h || h->get_ladder()->nsys() == 1) {
        E_[j] = sys_[j]->measure();
        if ((size_t)j < stale_.size()) stale_[j] = 0;
      } else {
        by[h->get_ladder().get()].push_back(j);
      }
    }
    for (auto &kv : by) {
      std::vector<int> slots;
      for (int j : kv.second) slots.push_back(as_hip(sys_[j])->index());
      std::vector<int32_t> E(slots.size() * 64), M(slots.size() * 64);
      const int rc = sa_ladder_measure_systems(kv.first->get(), slots.data(), (int)slots.size(), E.data(), M.data());
      if (rc != SA_OK) log_error("measure", rc);
      for (size_t q = 0; q < kv.second.size(); ++q) {
        const int j = kv.second[q];
        const double h = as_hip(sys_[j])->get_h();
        std::vector<float> r(64);
        for (int k = 0; k < 64; ++k) r[k] = (float)((double)E[q * 64 + k] + h * (double)M[q * 64 + k]);
        E_[j] = r;
        if ((size_t)j < stale_.size()) stale_[j] = 0;
      }
    }
  }
  void measure_all() {
    stale_.assign(sys_.size(), 0);
    std::map<ladder *, std::pair<std::vector<int32_t>, std::vector<int32_t>>> got;
    for (size_t j = 0; j < sys_.size(); ++j) {
      hip_sys *h = as_hip(sys_[j]);
      if (!h || h->get_ladder()->nsys() == 1) {
        E_[j] = sys_[j]->measure();
        continue;
      }
      ladder *l = h->get_ladder().get();
      auto it = got.find(l);
      if (it == got.end()) {
        it = got.emplace(l, std::make_pair(std::vector<int32_t>(), std::vector<int32_t>())).first;
        const int rc = l->measure_all(it->second.first, it->second.second);
        if (rc != SA_OK) log_error("measure", rc);
      }
      std::vector<float> r(64);
      const int idx = h->index();
      for (int k = 0; k < 64; ++k)
        r[k] = (float)((double)it->second.first[idx * 64 + k] + h->get_h() * (double)it->second.second[idx * 64 + k]);
      E_[j] = r;
    }
  }

  std::unique_ptr<std::ofstream> open_T_file(double T, std::ios_base::openmode mode) {
    std::string f = std::regex_replace(fname_, std::regex("%", std::regex_constants::grep), "T" + std::to_string(T),
                                       std::regex_constants::match_default);
    if (f == fname_) std::cerr << "Error:   file name hasn't changed, missing % to replace\n";
    std::unique_ptr<std::ofstream> p(new std::ofstream(f, mode));
    if (!p->is_open()) std::cerr << "Error:   cannot open energy file " << f << "\n";
    return p;
  }
  void open_file(print_select select, bool append) {
    const std::ios_base::openmode mode = append ? std::ios_base::app : std::ios_base::out;
    if (select == PRINT_E && !print_E_) {
      for (double T : T_) files_.push_back(open_T_file(T, mode));
      print_E_ = true;
    } else if (select == PRINT_MIN && !print_min_) {
      std::string f = std::regex_replace(fname_, std::regex("%", std::regex_constants::grep), "E_min",
                                         std::regex_constants::match_default);
      file_E_min_.reset(new std::ofstream(f, mode));
      if (!file_E_min_->is_open()) std::cerr << "Error:   cannot open E_min file " << f << "\n";
      print_min_ = true;
    }
  }

  std::vector<system *> sys_;
  std::vector<double> T_;
  std::vector<std::vector<float>> E_;
  std::vector<float> E_min_;
  std::vector<std::unique_ptr<std::ofstream>> files_;
  std::unique_ptr<std::ofstream> file_E_min_;
  long steps_;
  std::vector<accumulator> num_switch_;
  mt19937 gen_;
  std::string fname_;
  int uni_a_, uni_b_;
  bool print_E_ = false, print_min_ = false;
  std::vector<swap_record> *log_ = nullptr;
  std::vector<char> stale_;  // E entries not measured at the last step (see step())
  std::vector<std::unique_ptr<gpu_sys>> owned_;
};

}  // namespace sa
