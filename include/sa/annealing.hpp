// annealing.hpp -- simulated annealing driver over any system type.
//
// Same interface and behaviour as the reference's annealing<system>
// (src/annealing.hpp:10-139): `steps` sweeps per step(), a measurement when
// (i % stride == 0), the per-lane maximum of -H kept in min_E, geometric cooling
// T *= alfa after every step.  run_schedule() adds what the reference lacks: an
// arbitrary per-sweep temperature schedule (e.g. a linear anneal), implemented
// with the same set_T / sweep / measure calls.
#pragma once

#include <algorithm>
#include <iostream>
#include <stdexcept>
#include <vector>

namespace sa {

enum print_select { PRINT_E, PRINT_MIN };

template <class system>
class annealing {
 public:
  annealing(double T_s, int steps, double alfa, system *S, std::ostream &file = std::cout)
      : T_(T_s), alfa_(alfa), steps_(steps), file_(&file), sys_(S) {
    if (alfa_ > 1 || alfa_ < 0) throw std::invalid_argument("alfa out of range");
    sys_->set_T(T_);
  }

  // src/annealing.hpp:67-100
  void step(int stride, print_select select = PRINT_E) {
    for (int i = 0; i < steps_; ++i) {
      sys_->sweep();
      if (stride != 0 && i % stride == 0) record(select);
    }
    T_ = T_ * alfa_;
    sys_->set_T(T_);
  }

  // set_T(Ts[k]) before sweep k; measure after sweep k when k % stride == 0.
  // Returns the measurements made.
  std::vector<std::vector<float>> run_schedule(const std::vector<double> &Ts, int stride,
                                               print_select select = PRINT_E, bool print_rows = false) {
    std::vector<std::vector<float>> log;
    for (size_t k = 0; k < Ts.size(); ++k) {
      T_ = Ts[k];
      sys_->set_T(T_);
      sys_->sweep();
      if (stride != 0 && k % (size_t)stride == 0) {
        record(select, print_rows);
        log.push_back(result_);
      }
    }
    return log;
  }

  void print(print_select select) {
    const std::vector<float> &data = (select == PRINT_MIN) ? min_E_ : result_;
    for (float v : data) (*file_) << v << "\t";
    (*file_) << std::endl;
  }
  double get_T() const { return T_; }
  void set_T(double T_s) {  // resets min_E as the reference does
    T_ = T_s;
    min_E_.clear();
  }
  const std::vector<float> &get_min_E() const { return min_E_; }
  const std::vector<float> &get_result() const { return result_; }

 private:
  void record(print_select select, bool print_rows = true) {
    result_ = sys_->measure();
    if (min_E_.empty()) min_E_ = result_;
    if (min_E_.size() == result_.size())
      for (size_t j = 0; j < min_E_.size(); ++j) min_E_[j] = std::max(result_[j], min_E_[j]);
    if (select == PRINT_E && print_rows) print(select);
  }

  double T_, alfa_;
  int steps_;
  std::ostream *file_;
  system *sys_;
  std::vector<float> min_E_, result_;
};

}  // namespace sa
