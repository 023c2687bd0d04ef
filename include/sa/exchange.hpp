// exchange.hpp -- the lane-wise replica-exchange decision shared by the
// parallel-tempering driver and the batched engine.
#pragma once

#include <cmath>
#include <cstdint>

namespace sa {

// Lanes k < n of one 64-lane word: lane k of systems i and i+1 exchange when
// dE*dB > 0 or u < exp(dB*dE), with dE = -(E_i[k] - E_{i+1}[k]) in float and
// dB = 1/T_i - 1/T_{i+1} (creat_mask, src/parallel_tempering.hpp:141-162).
// `accepted` is increased by the number of exchanging lanes.
inline uint64_t exchange_word(const float *Ei, const float *Ej, double dB, double u, int n, int &accepted) {
  uint64_t m = 0;
  for (int k = 0; k < n; ++k) {
    const float dE = -(Ei[k] - Ej[k]);
    if (((dE * dB) > 0.) || u < std::exp(dB * dE)) {
      m |= 1ULL << k;
      ++accepted;
    }
  }
  return m;
}

}  // namespace sa
