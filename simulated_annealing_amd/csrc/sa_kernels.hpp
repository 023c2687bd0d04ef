// sa_kernels.hpp -- gfx950 kernels of the multispin +-J Metropolis sweep.
//
// Data layout in HBM (one ladder = `nsys` systems of one geometry that share the
// couplings and the random stream, i.e. the temperatures of a parallel-tempering
// run or a single annealed system):
//   s   : nsys x N uint64 words, system-major, site i = x + L*y (+ L*L*z); bit k of
//         a word is lane k (64 independent disorder samples).
//   jx, jy, jz : N words each, deduplicated: jx[i] couples i to its -x neighbour,
//         jy[i] to -y, jz[i] to -z (the reference keeps each bond twice,
//         src/gpu_3d.cuh:16, src/cpu_3d.cpp:281-301).
//   tab : one sys_table per system: integer acceptance thresholds.
//
// Update order.  The reference CPU sweep (src/cpu_3d.cpp:94-105) visits, for
// offset 0 then 1, every site with (x+y)%2 == offset in z-major order and updates in
// place.  In 2D (and in-plane in 3D) neighbours of a site have the other colour
// (even L), so a half-sweep is data-parallel.  Along z, however, neighbours share
// the colour: site z reads the NEW value of z-1 and the OLD value of z+1; z = 0
// reads the old plane L-1 and z = L-1 the new plane 0.  The 3D kernel therefore
// gives each thread one (x,y) column of the current colour and walks z in order,
// carrying the new z-1 word and the new plane-0 word in registers: this is the
// reference's update order exactly, not a 3D checkerboard.
//
// Arithmetic is integer only: sat = J ^ s_i ^ s_j (bit set = satisfied bond), a
// bit-sliced full-adder tree gives the 3-bit satisfied count c per lane, the
// site's random word g is compared with the per-level thresholds, and a mux tree
// over c picks the flip mask.  One random word per site update is shared by the 64
// lanes, as in the reference (src/cpu_3d.cpp:130).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sa/rng.hpp"

namespace sa {

enum { KIND_2P = 0, KIND_2D = 1, KIND_3D = 2 };

// Acceptance thresholds of one system.  Table 0/1: bulk sites with spin bit 0/1;
// table 2/3: open-x border sites (2p only).  Level c accepts iff g < thr[t][c]
// or bit c of always[t] is set.
struct sys_table {
  uint64_t thr[4][8];
  uint32_t always[4];
  uint32_t pad[4];
};

struct sweep_args {
  uint64_t *s;
  const uint64_t *jx, *jy, *jz;
  const sys_table *tab;
  const uint64_t *words;  // replay: raw words of this sweep in visit order (N)
  uint64_t seed;          // counter mode key
  uint64_t sweep;         // counter mode: sweeps done since seeding
  int64_t N;
  int L;
  int nsys;
};

// ---------------------------------------------------------------- bit-sliced logic
struct count3 {
  uint64_t b0, b1, b2;
};

__device__ __forceinline__ uint64_t maj3(uint64_t a, uint64_t b, uint64_t c) {
  return (a & b) | (c & (a | b));
}

// number of set bits among six words, per lane (0..6)
__device__ __forceinline__ count3 count6(uint64_t a, uint64_t b, uint64_t c, uint64_t d, uint64_t e,
                                         uint64_t f) {
  const uint64_t s1 = a ^ b ^ c, k1 = maj3(a, b, c);
  const uint64_t s2 = d ^ e ^ f, k2 = maj3(d, e, f);
  const uint64_t k3 = s1 & s2;
  return count3{s1 ^ s2, k1 ^ k2 ^ k3, maj3(k1, k2, k3)};
}

// ... among four words (0..4)
__device__ __forceinline__ count3 count4(uint64_t a, uint64_t b, uint64_t c, uint64_t d) {
  const uint64_t s1 = a ^ b ^ c, k1 = maj3(a, b, c);
  const uint64_t k2 = s1 & d;
  return count3{s1 ^ d, k1 ^ k2, k1 & k2};
}

__device__ __forceinline__ uint64_t sel(uint64_t m, uint64_t a, uint64_t b) {  // m ? a : b, per bit
  return (a & m) | (b & ~m);
}

__device__ __forceinline__ uint64_t lvl(bool acc) { return acc ? ~0ULL : 0ULL; }

// per-level accept masks of one table for draw g (levels 0..7)
__device__ __forceinline__ void level_masks(const sys_table &t, int tb, uint64_t g, uint64_t a[8], int nlev) {
#pragma unroll
  for (int c = 0; c < 8; ++c)
    a[c] = (c < nlev) ? lvl((g < t.thr[tb][c]) || ((t.always[tb] >> c) & 1u)) : 0ULL;
}

// lane-wise a[c] with c given bit-sliced
__device__ __forceinline__ uint64_t mux8(const count3 &c, const uint64_t a[8]) {
  const uint64_t m0 = sel(c.b0, a[1], a[0]), m1 = sel(c.b0, a[3], a[2]);
  const uint64_t m2 = sel(c.b0, a[5], a[4]), m3 = sel(c.b0, a[7], a[6]);
  const uint64_t n0 = sel(c.b1, m1, m0), n1 = sel(c.b1, m3, m2);
  return sel(c.b2, n1, n0);
}

// Flip mask of a site: general tables (any T, any h).
__device__ __forceinline__ uint64_t flip_general(const sys_table &t, int tb0, uint64_t g, const count3 &c,
                                                 uint64_t si, int nlev) {
  uint64_t a[8];
  level_masks(t, tb0, g, a, nlev);
  const uint64_t f0 = mux8(c, a);
  level_masks(t, tb0 + 1, g, a, nlev);
  const uint64_t f1 = mux8(c, a);
  return sel(si, f1, f0);
}

// Flip mask when both spin tables are equal and every level c <= nb/2 always
// accepts (h = 0, T > 0): only the levels above nb/2 need a comparison.
//   3D (levels 4,5,6 thresholded): flip = ~c2 | (c1 ? a6 : (c0 ? a5 : a4))
//   2D (levels 3,4):               flip = c2 ? a4 : ((c1 & c0) ? a3 : ~0)
__device__ __forceinline__ uint64_t flip_fast3(const sys_table &t, uint64_t g, const count3 &c) {
  const uint64_t a4 = lvl(g < t.thr[0][4]), a5 = lvl(g < t.thr[0][5]), a6 = lvl(g < t.thr[0][6]);
  return ~c.b2 | sel(c.b1, a6, sel(c.b0, a5, a4));
}
__device__ __forceinline__ uint64_t flip_fast2(const sys_table &t, uint64_t g, const count3 &c) {
  const uint64_t a3 = lvl(g < t.thr[0][3]), a4 = lvl(g < t.thr[0][4]);
  return sel(c.b2, a4, ~(c.b1 & c.b0) | a3);
}

// ---------------------------------------------------------------- 3D half-sweep
// Thread = one column (x, y) of colour `color` and MB systems; grid.y = system
// block (so thresholds are wave-uniform and live in scalar registers).
//
// Software pipeline: nothing a step loads depends on the step before it (the
// in-plane neighbours have the other colour and do not change during the half;
// the column's own next word is still old), so the loads of step z+1 are issued
// before the arithmetic and the store of step z and overlap its latency.
template <int MB>
struct plane_loads {
  uint64_t jw, je, jn, js, jd;  // couplings: -x, +x, -y, +y, +z bonds
  uint64_t w[MB], e[MB], n[MB], s[MB], d[MB];
};

template <bool COUNTER, bool FAST, int MB>
__global__ void __launch_bounds__(256) k_sweep3d_half(sweep_args p, int color) {
  const int L = p.L;
  const int half = L >> 1;
  const int cols = L * half;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= cols) return;
  const int y = k / half;
  const int x = 2 * (k - y * half) + ((y + color) & 1);
  const int xm = x == 0 ? L - 1 : x - 1, xp = x == L - 1 ? 0 : x + 1;
  const int ym = y == 0 ? L - 1 : y - 1, yp = y == L - 1 ? 0 : y + 1;
  const int64_t L2 = (int64_t)L * L;
  const int c0 = x + L * y;               // in-plane offsets
  const int cw = xm + L * y, ce = xp + L * y, cn = x + L * ym, cs = x + L * yp;
  const int sys0 = blockIdx.y * MB;
  const int64_t N = p.N;
  const int64_t rank0 = (int64_t)color * (N >> 1) + k;  // visit rank of (x,y,0)
  const int64_t rstep = L2 >> 1;

  uint64_t *sp[MB];
  uint64_t up[MB], own[MB], first[MB];
#pragma unroll
  for (int m = 0; m < MB; ++m) {
    sp[m] = p.s + (int64_t)(sys0 + m) * N;
    up[m] = sp[m][c0 + (L - 1) * L2];     // old plane L-1
    own[m] = sp[m][c0];
    first[m] = 0;
  }
  // loads of plane z (the +z neighbour of plane L-1 is the new plane 0: not loaded)
  auto load = [&](int z, plane_loads<MB> &q) {
    const int64_t P = z * L2;
    const int64_t Pd = (z == L - 1) ? 0 : P + L2;
    q.jw = p.jx[P + c0];
    q.je = p.jx[P + ce];
    q.jn = p.jy[P + c0];
    q.js = p.jy[P + cs];
    q.jd = p.jz[Pd + c0];
#pragma unroll
    for (int m = 0; m < MB; ++m) {
      const uint64_t *s = sp[m];
      q.w[m] = s[P + cw];
      q.e[m] = s[P + ce];
      q.n[m] = s[P + cn];
      q.s[m] = s[P + cs];
      q.d[m] = (z == L - 1) ? 0 : s[Pd + c0];
    }
  };
  uint64_t jz_up = p.jz[c0];
  u32x4 blk{0, 0, 0, 0};
  plane_loads<MB> cur, nxt;
  load(0, cur);
  for (int z = 0; z < L; ++z) {
    if (z + 1 < L) load(z + 1, nxt);
    uint64_t g;
    if (COUNTER) {
      if ((z & 1) == 0) blk = counter_block((uint32_t)(c0 + z * L2), p.seed, p.sweep);
      g = counter_half(blk, z & 1);
    } else {
      g = p.words[rank0 + z * rstep];
    }
    const int64_t P = z * L2;
#pragma unroll
    for (int m = 0; m < MB; ++m) {
      const uint64_t si = own[m];
      const uint64_t down = (z == L - 1) ? first[m] : cur.d[m];
      const uint64_t bw = cur.jw ^ si ^ cur.w[m];
      const uint64_t be = cur.je ^ si ^ cur.e[m];
      const uint64_t bn = cur.jn ^ si ^ cur.n[m];
      const uint64_t bs = cur.js ^ si ^ cur.s[m];
      const uint64_t bu = jz_up ^ si ^ up[m];
      const uint64_t bd = cur.jd ^ si ^ down;
      const count3 c = count6(bw, be, bn, bs, bu, bd);
      const sys_table &t = p.tab[sys0 + m];
      const uint64_t f = FAST ? flip_fast3(t, g, c) : flip_general(t, 0, g, c, si, 7);
      const uint64_t sn = si ^ f;
      sp[m][P + c0] = sn;
      up[m] = sn;
      if (z == 0) first[m] = sn;
      own[m] = down;
    }
    jz_up = cur.jd;
    cur = nxt;
  }
}

// ---------------------------------------------------------------- 2D half-sweep
template <bool COUNTER, bool FAST, bool OPEN, int MB>
__global__ void __launch_bounds__(256) k_sweep2d_half(sweep_args p, int color) {
  const int L = p.L;
  const int half = L >> 1;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= L * half) return;
  const int y = k / half;
  const int x = 2 * (k - y * half) + ((y + color) & 1);
  const int xm = x == 0 ? L - 1 : x - 1, xp = x == L - 1 ? 0 : x + 1;
  const int ym = y == 0 ? L - 1 : y - 1, yp = y == L - 1 ? 0 : y + 1;
  const int c0 = x + L * y;
  const int cw = xm + L * y, ce = xp + L * y, cn = x + L * ym, cs = x + L * yp;
  const int sys0 = blockIdx.y * MB;
  const uint64_t jw = p.jx[c0], je = p.jx[ce], jn = p.jy[c0], js = p.jy[cs];
  uint64_t g;
  if (COUNTER) {
    const u32x4 r = counter_block((uint32_t)(((x & ~3) | (x & 1)) + L * y), p.seed, p.sweep);
    g = counter_half(r, (x >> 1) & 1);
  } else {
    g = p.words[(int64_t)color * (p.N >> 1) + k];
  }
  const bool border = OPEN && (x == 0 || x == L - 1);
#pragma unroll
  for (int m = 0; m < MB; ++m) {
    uint64_t *s = p.s + (int64_t)(sys0 + m) * p.N;
    const uint64_t si = s[c0];
    uint64_t bw = jw ^ si ^ s[cw];
    uint64_t be = je ^ si ^ s[ce];
    if (OPEN) {  // missing open-x bond counts as satisfied (src/cpu_2dp.cpp:338-339)
      if (x == 0) bw = ~0ULL;
      if (x == L - 1) be = ~0ULL;
    }
    const uint64_t bn = jn ^ si ^ s[cn];
    const uint64_t bs = js ^ si ^ s[cs];
    const count3 c = count4(bw, be, bn, bs);
    const sys_table &t = p.tab[sys0 + m];
    uint64_t f;
    if (FAST && !OPEN) f = flip_fast2(t, g, c);
    else f = flip_general(t, border ? 2 : 0, g, c, si, 5);
    s[c0] = si ^ f;
  }
}

// ---------------------------------------------------------------- measure
// Per system and lane: number of satisfied -x,-y[,-z] bonds and of set spins.
// A wave stages 64 site words (and their bond words) in LDS; lane k then walks
// the 64 staged words and counts bit k.  Counts are reduced in LDS per block and
// added to cnt[sys][0|1][lane] with one atomic per lane per block.
template <int KIND>
__global__ void __launch_bounds__(256) k_measure(const uint64_t *__restrict__ s_all, const uint64_t *__restrict__ jx,
                                                 const uint64_t *__restrict__ jy, const uint64_t *__restrict__ jz,
                                                 int L, int64_t N, int *cnt) {
  constexpr int NB = (KIND == KIND_3D) ? 3 : 2;
  __shared__ uint64_t stage[4][NB + 1][64];
  __shared__ int red[2][64];
  const int sys = blockIdx.y;
  const uint64_t *s = s_all + (int64_t)sys * N;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (threadIdx.x < 128) red[threadIdx.x >> 6][threadIdx.x & 63] = 0;
  int sat = 0, up = 0;
  const int64_t L2 = (int64_t)L * L;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x + wv * 64; base < N; base += stride) {
    const int64_t i = base + lane;
    uint64_t w = 0, bw = 0, bn = 0, bu = 0;
    if (i < N) {
      const int x = (int)(i % L);
      const int64_t y = (i / L) % L;
      const int64_t z = i / L2;
      w = s[i];
      const int64_t iw = i - x + (x == 0 ? L - 1 : x - 1);
      const int64_t in = i + ((y == 0 ? L - 1 : y - 1) - y) * L;
      bw = jx[i] ^ w ^ s[iw];
      if (KIND == KIND_2P && x == 0) bw = 0;  // no open-x west bond
      bn = jy[i] ^ w ^ s[in];
      if (KIND == KIND_3D) bu = jz[i] ^ w ^ s[i + ((z == 0 ? L - 1 : z - 1) - z) * L2];
    }
    stage[wv][0][lane] = w;
    stage[wv][1][lane] = bw;
    stage[wv][2][lane] = bn;
    if (KIND == KIND_3D) stage[wv][3][lane] = bu;
    __builtin_amdgcn_wave_barrier();
    const int n = (int)((N - base) < 64 ? (N - base) : 64);
    for (int r = 0; r < n; ++r) {
      up += (int)((stage[wv][0][r] >> lane) & 1ULL);
      int b = (int)((stage[wv][1][r] >> lane) & 1ULL) + (int)((stage[wv][2][r] >> lane) & 1ULL);
      if (KIND == KIND_3D) b += (int)((stage[wv][3][r] >> lane) & 1ULL);
      sat += b;
    }
    __builtin_amdgcn_wave_barrier();
  }
  __syncthreads();
  atomicAdd(&red[0][lane], sat);
  atomicAdd(&red[1][lane], up);
  __syncthreads();
  if (threadIdx.x < 128) atomicAdd(&cnt[(sys * 2 + (threadIdx.x >> 6)) * 64 + (threadIdx.x & 63)],
                                   red[threadIdx.x >> 6][threadIdx.x & 63]);
}

__global__ void k_fill_zero(int *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// ---------------------------------------------------------------- masked swap
// a ^= (a^b) & m, b ^= (a^b) & m (src/device.cuh:18-23)
__global__ void __launch_bounds__(256) k_swap(uint64_t *a, uint64_t *b, uint64_t mask, int64_t N) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += stride) {
    const uint64_t va = a[i], vb = b[i];
    const uint64_t d = (va ^ vb) & mask;
    a[i] = va ^ d;
    b[i] = vb ^ d;
  }
}

}  // namespace sa
