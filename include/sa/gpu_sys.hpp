// gpu_sys.hpp -- the reference's system interface, kept as is so that the
// annealing and parallel-tempering drivers (and user code written against the
// reference) run unchanged on the MI355X systems of hip_sys.hpp.
// Reference: src/gpu_sys.cuh:12-29 (gpu_sys) and src/cpu_sys.hpp:12-31 (cpu_sys,
// which adds get_T / clone / a virtual destructor; they are added here too, the
// reference's missing virtual destructor being the one change of substance).
// get_T and clone have default bodies, so a class written against the reference's
// gpu_sys (which has neither) still compiles: get_T() then reports NaN and clone()
// nullptr, which parallel_tempering::update_T reports as an error.
#pragma once

#include <iostream>
#include <limits>
#include <memory>
#include <string>
#include <vector>

typedef unsigned long long spin_t;

class gpu_sys {
 public:
  virtual void sweep() = 0;
  virtual void set_T(double T_) = 0;
  virtual void set_h(double h_) = 0;
  virtual void set_seed(long long seed_) = 0;
  virtual std::vector<float> measure() = 0;
  virtual void gpu_free() = 0;
  virtual void load_J(std::string fname) = 0;
  virtual void save_J(std::string fname) = 0;
  virtual void init_J() = 0;
  virtual void init_rand() = 0;
  virtual void save_sys(std::string prefix) = 0;
  virtual long get_N() = 0;
  virtual void swap(gpu_sys *sys, std::unique_ptr<spin_t[]> mask) = 0;
  virtual std::ostream &save(std::ostream &stream) = 0;
  virtual std::istream &load(std::istream &stream) = 0;
  virtual double get_T() { return std::numeric_limits<double>::quiet_NaN(); }
  virtual gpu_sys *clone() const { return nullptr; }
  virtual ~gpu_sys() = default;
};
