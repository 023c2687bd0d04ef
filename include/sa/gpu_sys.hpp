// gpu_sys.hpp -- the reference's system interface, kept as is so that the
// annealing and parallel-tempering drivers (and user code written against the
// reference) run unchanged on the MI355X systems of hip_sys.hpp.
// Reference: src/gpu_sys.cuh:12-29 (gpu_sys) and src/cpu_sys.hpp:12-31 (cpu_sys,
// which adds get_T / clone / a virtual destructor; they are added here too, the
// reference's missing virtual destructor being the one change of substance).
#pragma once

#include <iostream>
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
  virtual double get_T() = 0;
  virtual gpu_sys *clone() const = 0;
  virtual ~gpu_sys() = default;
};
