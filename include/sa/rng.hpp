// rng.hpp -- random-number pieces of the sweep, host and device.
//
//  * Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
//    as 1, 2, 3", SC'11): the counter-based generator of the counter RNG mode.
//    One call per (site pair, sweep); the pairing is fixed by the site
//    coordinates, never by launch geometry, so a lattice's random stream is the
//    same whatever kernel, tile or batch processes it.
//  * mt19937_64 / mt19937 (Matsumoto & Nishimura) and the libstdc++ 11 mapping of
//    their output to doubles and bounded ints (generate_canonical, Lemire's
//    nearly divisionless method).  The reference draws every random number from
//    these (src/cpu_3d.cpp:22,130,261,286; src/parallel_tempering.hpp:59,112-114);
//    restating them here keeps results independent of the host's C++ library.
//  * integer acceptance thresholds: the reference accepts a flip when
//    u < boltz with u = double(g)/2^64 (g the raw 64-bit draw).  u is monotone in
//    g, so u < b  <=>  g < t(b) for one integer t(b) in [0, 2^64]; the kernels
//    compare raw words against t(b) and never touch floating point.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>

#ifndef SA_HD
#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define SA_HD __host__ __device__ __forceinline__
#else
#define SA_HD inline
#endif
#endif

namespace sa {

// ---------------------------------------------------------------- Philox4x32-10
struct u32x4 {
  uint32_t x, y, z, w;
};

SA_HD uint32_t xor3_u32(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);  // one gfx950 instruction
#else
  return a ^ b ^ c;
#endif
}

SA_HD u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
    u32x4 n;
    n.x = xor3_u32((uint32_t)(p1 >> 32), c.y, k0);
    n.y = (uint32_t)p1;
    n.z = xor3_u32((uint32_t)(p0 >> 32), c.w, k1);
    n.w = (uint32_t)p0;
    c = n;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

// Counter word of a site pair.  3D pairs sites (x,y,z) and (x,y,z+1) for even z
// (same checkerboard colour, consecutive steps of one column walk); 2D pairs x and
// x+2 inside each group of four columns.  `base` is the linear index of the even
// member of the pair; the two 64-bit halves of one Philox block serve the pair.
SA_HD u32x4 counter_block(uint32_t base, uint64_t seed, uint64_t sweep) {
  return philox4x32_10(u32x4{base, (uint32_t)sweep, (uint32_t)(sweep >> 32), 0u}, (uint32_t)seed,
                       (uint32_t)(seed >> 32));
}
SA_HD uint64_t counter_half(const u32x4 &r, int sel) {
  return sel ? (((uint64_t)r.w << 32) | r.z) : (((uint64_t)r.y << 32) | r.x);
}

// ---------------------------------------------------------------- Mersenne twisters
// Same object layout as libstdc++'s engines (state words of the engine's
// result_type, then the position), so a saved engine reads back byte for byte.
struct mt19937_64 {
  uint64_t x[312];
  uint64_t p;
  void seed(uint64_t s) {
    x[0] = s;
    for (int i = 1; i < 312; ++i) x[i] = 6364136223846793005ULL * (x[i - 1] ^ (x[i - 1] >> 62)) + (uint64_t)i;
    p = 312;
  }
  void twist() {
    for (int i = 0; i < 312; ++i) {
      const uint64_t y = (x[i] & 0xFFFFFFFF80000000ULL) | (x[(i + 1) % 312] & 0x7FFFFFFFULL);
      x[i] = x[(i + 156) % 312] ^ (y >> 1) ^ ((y & 1) ? 0xB5026F5AA96619E9ULL : 0ULL);
    }
    p = 0;
  }
  uint64_t operator()() {
    if (p >= 312) twist();
    uint64_t y = x[p++];
    y ^= (y >> 29) & 0x5555555555555555ULL;
    y ^= (y << 17) & 0x71D67FFFEDA60000ULL;
    y ^= (y << 37) & 0xFFF7EEE000000000ULL;
    return y ^ (y >> 43);
  }
  void fill(uint64_t *out, int64_t n) {
    for (int64_t i = 0; i < n; ++i) out[i] = (*this)();
  }
};

struct mt19937 {
  uint64_t x[624];  // uint_fast32_t on LP64, as in libstdc++
  uint64_t p;
  void seed(uint32_t s) {
    x[0] = s;
    for (int i = 1; i < 624; ++i) {
      const uint32_t prev = (uint32_t)x[i - 1];
      x[i] = (uint32_t)(1812433253u * (prev ^ (prev >> 30)) + (uint32_t)i);
    }
    p = 624;
  }
  uint32_t operator()() {
    if (p >= 624) {
      for (int i = 0; i < 624; ++i) {
        const uint32_t y = ((uint32_t)x[i] & 0x80000000u) | ((uint32_t)x[(i + 1) % 624] & 0x7FFFFFFFu);
        x[i] = (uint32_t)x[(i + 397) % 624] ^ (y >> 1) ^ ((y & 1) ? 0x9908B0DFu : 0u);
      }
      p = 0;
    }
    uint32_t y = (uint32_t)x[p++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9D2C5680u;
    y ^= (y << 15) & 0xEFC60000u;
    return y ^ (y >> 18);
  }
};

// generate_canonical<double, 53> of one 64-bit draw: double(g) / 2^64, clamped
// below one.
inline double canonical64(uint64_t g) {
  const double r = (double)g / 18446744073709551616.0;
  return r >= 1.0 ? std::nextafter(1.0, 0.0) : r;
}
// ... of two 32-bit draws: (g1 + g2*2^32) / 2^64 accumulated in double.
inline double canonical32(mt19937 &g) {
  double sum = (double)g();
  sum += (double)g() * 4294967296.0;
  const double r = sum / 18446744073709551616.0;
  return r >= 1.0 ? std::nextafter(1.0, 0.0) : r;
}
// uniform_int_distribution<int>(a, b) over a 32-bit engine (Lemire 2019).
inline int uniform_int(mt19937 &g, int a, int b) {
  const uint32_t range = (uint32_t)((int64_t)b - (int64_t)a) + 1u;
  uint64_t prod = (uint64_t)g() * range;
  uint32_t low = (uint32_t)prod;
  if (low < range) {
    const uint32_t thr = (uint32_t)(0u - range) % range;
    while (low < thr) {
      prod = (uint64_t)g() * range;
      low = (uint32_t)prod;
    }
  }
  return (int)((uint32_t)(prod >> 32) + (uint32_t)a);
}

// ---------------------------------------------------------------- thresholds
// Smallest g in [0, 2^64] with canonical64(g) >= b; a flip is accepted iff
// g < threshold.  Returned as (value, always): always=true encodes 2^64 (b = 1,
// every draw accepts); value = 0 encodes "never" (b <= 0 or NaN).
struct threshold {
  uint64_t value;
  bool always;
};
inline threshold accept_threshold(double b) {
  if (!(b > 0.0)) return {0, false};
  if (b > std::nextafter(1.0, 0.0)) return {0, true};
  uint64_t lo = 0, hi = ~0ULL;  // canonical64(hi) = 1 - 2^-53 >= b
  while (lo < hi) {
    const uint64_t mid = lo + ((hi - lo) >> 1);
    if (canonical64(mid) >= b) hi = mid; else lo = mid + 1;
  }
  return {lo, false};
}

// Boltzmann tables exactly as the reference computes them: 3D
// (src/cpu_3d.cpp:32-43) 14 entries [c] / [7+c]; 2D (src/cpu_2dp.cpp:33-47) 20
// entries [c] / [5+c] bulk and [10+c] / [15+c] open-x border.
inline void boltz_3d(double T, double h, double *b) {
  for (int i = 0; i < 7; ++i) {
    const int j = 2 * i - 6;
    b[i] = std::fmin(1., std::exp(-2 * (1 / T) * (j - h)));
    b[7 + i] = std::fmin(1., std::exp(-2 * (1 / T) * (j + h)));
  }
}
inline void boltz_2d(double T, double h, double *b) {
  for (int i = 0; i < 5; ++i) {
    const int j = 2 * i - 4;
    b[i] = std::fmin(1., std::exp(-2 * (1 / T) * (j - h)));
    b[5 + i] = std::fmin(1., std::exp(-2 * (1 / T) * (j + h)));
    b[10 + i] = std::fmin(1., (i == 0) ? 0. : std::exp(-2 * (1 / T) * (j - 1 - h)));
    b[15 + i] = std::fmin(1., (i == 0) ? 0. : std::exp(-2 * (1 / T) * (j - 1 + h)));
  }
}

}  // namespace sa
