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
// gfx950's v_bitop3_b32 evaluates any 3-input bitwise function in one VALU
// instruction; the 8-bit truth table of f is f(0xF0, 0xCC, 0xAA) for operands
// (a, b, c).  64-bit words are handled as two 32-bit halves.
constexpr int B3_XOR3 = 0x96;  // a ^ b ^ c
constexpr int B3_MAJ3 = 0xE8;  // majority(a, b, c)
constexpr int B3_MUX = 0xCA;   // a ? b : c, per bit
constexpr int B3_FLIP = 0x65;  // c ^ (~a | b): spin c flipped where ~a | b

template <int IMM>
__device__ __forceinline__ uint32_t b3(uint32_t a, uint32_t b, uint32_t c) {
  return __builtin_amdgcn_bitop3_b32(a, b, c, IMM);
}
template <int IMM>
__device__ __forceinline__ uint64_t b3(uint64_t a, uint64_t b, uint64_t c) {
  const uint32_t lo = __builtin_amdgcn_bitop3_b32((uint32_t)a, (uint32_t)b, (uint32_t)c, IMM);
  const uint32_t hi = __builtin_amdgcn_bitop3_b32((uint32_t)(a >> 32), (uint32_t)(b >> 32), (uint32_t)(c >> 32), IMM);
  return ((uint64_t)hi << 32) | lo;
}

struct count3 {
  uint64_t b0, b1, b2;
};

__device__ __forceinline__ uint64_t xor3(uint64_t a, uint64_t b, uint64_t c) { return b3<B3_XOR3>(a, b, c); }
__device__ __forceinline__ uint64_t maj3(uint64_t a, uint64_t b, uint64_t c) { return b3<B3_MAJ3>(a, b, c); }

// number of set bits among six words, per lane (0..6): two full adders and a
// third on the carries
__device__ __forceinline__ count3 count6(uint64_t a, uint64_t b, uint64_t c, uint64_t d, uint64_t e,
                                         uint64_t f) {
  const uint64_t s1 = xor3(a, b, c), k1 = maj3(a, b, c);
  const uint64_t s2 = xor3(d, e, f), k2 = maj3(d, e, f);
  const uint64_t k3 = s1 & s2;
  return count3{s1 ^ s2, xor3(k1, k2, k3), maj3(k1, k2, k3)};
}

// ... among four words (0..4)
__device__ __forceinline__ count3 count4(uint64_t a, uint64_t b, uint64_t c, uint64_t d) {
  const uint64_t s1 = xor3(a, b, c), k1 = maj3(a, b, c);
  const uint64_t k2 = s1 & d;
  return count3{s1 ^ d, k1 ^ k2, k1 & k2};
}

__device__ __forceinline__ uint64_t sel(uint64_t m, uint64_t a, uint64_t b) {  // m ? a : b, per bit
  return b3<B3_MUX>(m, a, b);
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

// New spin word of a site: general tables (any T, any h).
__device__ __forceinline__ uint64_t update_general(const sys_table &t, int tb0, uint64_t g, const count3 &c,
                                                   uint64_t si, int nlev) {
  uint64_t a[8];
  level_masks(t, tb0, g, a, nlev);
  const uint64_t f0 = mux8(c, a);
  level_masks(t, tb0 + 1, g, a, nlev);
  const uint64_t f1 = mux8(c, a);
  return si ^ sel(si, f1, f0);
}

// New spin word when both spin tables are equal and every level c <= nb/2 always
// accepts (h = 0, T > 0): only the levels above nb/2 need a comparison.  The
// accept decisions are per site, so they are built once as 32-bit masks and
// applied to both halves of the word.
//   3D (levels 4,5,6 thresholded): flip = ~c2 | (c1 ? a6 : (c0 ? a5 : a4))
//   2D (levels 3,4):               flip = c2 ? a4 : ((c1 & c0) ? a3 : ~0)
__device__ __forceinline__ uint64_t update_fast3t(uint64_t t4, uint64_t t5, uint64_t t6, uint64_t g, const count3 &c,
                                                  uint64_t si) {
  const uint32_t a4 = (g < t4) ? ~0u : 0u, a5 = (g < t5) ? ~0u : 0u;
  const uint32_t a6 = (g < t6) ? ~0u : 0u;
  const uint32_t vlo = b3<B3_MUX>((uint32_t)c.b1, a6, b3<B3_MUX>((uint32_t)c.b0, a5, a4));
  const uint32_t vhi = b3<B3_MUX>((uint32_t)(c.b1 >> 32), a6, b3<B3_MUX>((uint32_t)(c.b0 >> 32), a5, a4));
  const uint32_t lo = b3<B3_FLIP>((uint32_t)c.b2, vlo, (uint32_t)si);
  const uint32_t hi = b3<B3_FLIP>((uint32_t)(c.b2 >> 32), vhi, (uint32_t)(si >> 32));
  return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint64_t update_fast3(const sys_table &t, uint64_t g, const count3 &c, uint64_t si) {
  return update_fast3t(t.thr[0][4], t.thr[0][5], t.thr[0][6], g, c, si);
}
__device__ __forceinline__ uint64_t update_fast2(const sys_table &t, uint64_t g, const count3 &c, uint64_t si) {
  const uint64_t a3 = lvl(g < t.thr[0][3]), a4 = lvl(g < t.thr[0][4]);
  return si ^ sel(c.b2, a4, ~(c.b1 & c.b0) | a3);
}

// ---------------------------------------------------------------- buffer access
// Raw buffer loads/stores: a 128-bit resource (base, size) in scalar registers, a
// per-thread 32-bit byte offset that stays fixed for the whole z-walk, and the
// plane offset as a wave-uniform scalar -- no per-load 64-bit address arithmetic.
typedef unsigned int sa_v2u32 __attribute__((ext_vector_type(2)));
typedef unsigned int sa_v4u32 __attribute__((ext_vector_type(4)));
typedef __amdgpu_buffer_rsrc_t buf_t;

// A buffer access whose per-thread offset is at or beyond the resource size is
// discarded by the hardware range check: loads return 0 and stores write nothing.
// Lanes and planes with nothing to load or store use this offset, so every load
// and store instruction of the z-walk is issued unconditionally and the compiler
// can count the outstanding loads exactly (no drain of the prefetch).
constexpr uint32_t kOutOfRange = 0x80000000u;

__device__ __forceinline__ buf_t make_buf(const void *p, uint32_t bytes) {
  // dword 3 = 0x00020000: raw (untyped) dword access on gfx9-family GPUs
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), 0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ uint64_t bld(buf_t b, uint32_t voff, uint32_t soff) {
  const sa_v2u32 v = __builtin_amdgcn_raw_buffer_load_b64(b, (int)voff, (int)soff, 0);
  return ((uint64_t)v.y << 32) | v.x;
}
// Cache policy of a load / store (gfx950: bit 1 = "nt", non-temporal).  The fused
// sweep reads each old spin pair once and writes each new pair once per sweep, and
// a ladder's spins are far larger than the last-level cache, so that traffic is
// marked non-temporal; couplings (read by every system block) and the halo rows
// (read by two bands) keep the default policy.  Measured on the 3D L=128, 32-system
// sweep: 309 -> 283 us per launch.
constexpr int kCpolNT = 2;
#ifndef SA_FUSED_PAIR_CPOL
#define SA_FUSED_PAIR_CPOL kCpolNT
#endif
template <int CPOL = 0>
__device__ __forceinline__ void bld2(buf_t b, uint32_t voff, uint32_t soff, uint64_t &a, uint64_t &c) {
  const sa_v4u32 v = __builtin_amdgcn_raw_buffer_load_b128(b, (int)voff, (int)soff, CPOL);
  a = ((uint64_t)v.y << 32) | v.x;
  c = ((uint64_t)v.w << 32) | v.z;
}
template <int CPOL = 0>
__device__ __forceinline__ void bst2(buf_t b, uint32_t voff, uint32_t soff, uint64_t a, uint64_t c) {
  sa_v4u32 v;
  v.x = (uint32_t)a;
  v.y = (uint32_t)(a >> 32);
  v.z = (uint32_t)c;
  v.w = (uint32_t)(c >> 32);
  __builtin_amdgcn_raw_buffer_store_b128(v, b, (int)voff, (int)soff, CPOL);
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

// column k (0 <= k < L*L/2) of colour `color`, systems sys0 .. sys0+MB-1
// DEEP: plane loads issued two planes ahead instead of one (more registers; pays
// where few long column walks set the pace, as in the batched engine)
template <bool COUNTER, bool FAST, int MB, bool DEEP = false>
__device__ __forceinline__ void half3d_column(const sweep_args &p, int color, int k, int sys0) {
  const int L = p.L;
  const int half = L >> 1;
  const int y = k / half;
  const int x = 2 * (k - y * half) + ((y + color) & 1);
  const int xm = x == 0 ? L - 1 : x - 1, xp = x == L - 1 ? 0 : x + 1;
  const int ym = y == 0 ? L - 1 : y - 1, yp = y == L - 1 ? 0 : y + 1;
  const int64_t L2 = (int64_t)L * L;
  const int c0 = x + L * y;               // in-plane offsets
  const int cw = xm + L * y, ce = xp + L * y, cn = x + L * ym, cs = x + L * yp;
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
  // DEEP: plane z's loads are issued two planes ahead, three buffers rotating
  // through the z loop unrolled by 3; otherwise one plane ahead.
  auto plane = [&](int z, const plane_loads<MB> &cur, plane_loads<MB> &far, int ahead) {
    if (z + ahead < L) load(z + ahead, far);
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
      const count3 c = count6(xor3(cur.jw, si, cur.w[m]), xor3(cur.je, si, cur.e[m]), xor3(cur.jn, si, cur.n[m]),
                              xor3(cur.js, si, cur.s[m]), xor3(jz_up, si, up[m]), xor3(cur.jd, si, down));
      const sys_table &t = p.tab[sys0 + m];
      const uint64_t sn = FAST ? update_fast3(t, g, c, si) : update_general(t, 0, g, c, si, 7);
      sp[m][P + c0] = sn;
      up[m] = sn;
      if (z == 0) first[m] = sn;
      own[m] = down;
    }
    jz_up = cur.jd;
  };
  if constexpr (DEEP) {
    plane_loads<MB> b0, b1, b2;
    load(0, b0);
    if (1 < L) load(1, b1);
    for (int z = 0; z < L; z += 3) {
      plane(z, b0, b2, 2);
      if (z + 1 < L) plane(z + 1, b1, b0, 2);
      if (z + 2 < L) plane(z + 2, b2, b1, 2);
    }
  } else {
    plane_loads<MB> cur, nxt;
    load(0, cur);
    for (int z = 0; z < L; ++z) {
      plane(z, cur, nxt, 1);
      cur = nxt;
    }
  }
}

template <bool COUNTER, bool FAST, int MB>
__global__ void __launch_bounds__(256) k_sweep3d_half(sweep_args p, int color) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= p.L * (p.L >> 1)) return;
  half3d_column<COUNTER, FAST, MB>(p, color, k, blockIdx.y * MB);
}

// ---------------------------------------------------------------- 3D fused sweep
// One launch per sweep.  A workgroup owns a band of TH full rows (y0 .. y0+TH-1)
// of MB systems and walks z once; at every plane it performs the offset-0 half
// (colour 0) and then the offset-1 half (colour 1):
//   * colour-0 site at plane z needs old colour-1 in-plane neighbours, the new
//     colour-0 word at z-1 and the old one at z+1 -- exactly what the reference's
//     first half-sweep sees there, since that half never touches colour 1 and walks
//     z upwards (src/cpu_3d.cpp:94-105);
//   * colour-1 site at plane z needs the NEW colour-0 in-plane neighbours of the
//     same plane, the new colour-1 word at z-1 and the old one at z+1.
// The colour-0 updates of the rows just outside the band (y0-1, y0+TH) are
// recomputed by the workgroup (identical to the neighbour's), so no workgroup
// waits for another.  Spins are read from one array and written to another
// (ping-pong per sweep), so every "old" value a workgroup reads is untouched by
// any other workgroup during the sweep.
//
// Threads: (TH+2) rows x L/2 site pairs (2k, 2k+1); row 0 and row TH+1 are the
// halo rows.  LDS holds, per plane buffer, row and system, the colour-0 and the
// colour-1 word of every pair: c0[k], c1[k].  In a row of parity P the colour-0
// site of pair k is x0 = 2k+P and the colour-1 site x1 = 2k+1-P, so
//   colour-0 x-neighbours: c1[k], c1[k + (P ? 1 : -1)]   y-neighbours: c1[k] of rows r-1, r+1
//   colour-1 x-neighbours: c0[k], c0[k + (P ? -1 : 1)]   y-neighbours: c0[k] of rows r-1, r+1
// One barrier per plane: iteration z updates colour 1 of plane z and colour 0 of
// plane z+1 (which reads only old colour-1 words of plane z+1 and the column's own
// colour-0 chain) and stores the old words of plane z+2; three LDS plane buffers
// rotate.  A thread's row parity is fixed, so the z-walk is instantiated per
// parity when a row spans whole waves (L/2 a multiple of 64), where the choice is
// wave-uniform; smaller lattices keep the parity as a run-time value.
template <int MB>
struct fused_loads {
  uint64_t na[MB], nb[MB];  // own pair at plane z+1 (old)
  uint64_t ring2[MB];       // halo rows: colour-1 word one row further out (old)
  uint64_t jxa, jxb, jxn;   // jx of 2k, 2k+1, 2k+2
  uint64_t jya, jyb;        // jy of the pair (bond to y-1)
  uint64_t jsa, jsb;        // jy of the pair in row y+1 (bond to y+1)
  uint64_t jda, jdb;        // jz of the pair at plane z+1 (bond to z+1)
  uint64_t g0, g1;          // replay words of the colour-0 / colour-1 site
};

template <bool COUNTER, bool FAST, int MB, int TH>
struct fused_ctx {
  const sweep_args &p;
  uint64_t *lds;
  int L, PR, r, rn, rs, k, kn, kp, y, row, rowS, yr2, sys0;
  uint32_t t_rk, t_rkp, t_rkn, t_rnk, t_rsk;  // per-thread LDS byte offsets (see at())
  uint64_t t4[MB], t5[MB], t6[MB];  // FAST: the thresholded levels of each system's table
  bool halo;
  int64_t L2, N, rank0;
  int par;  // row parity (the template argument P is -1 when it is only known at run time)
  buf_t sin[MB], sout[MB], bjx, bjy, bjz, bw;
  uint32_t vpair, vpairS, vnext, vring, vrank, vrank1, vstore;  // per-thread byte offsets
  uint32_t plane_bytes;                          // L*L*8

  // LDS word of (plane buffer, system m, colour c) at a thread position.  The
  // element index (((buf*(TH+2) + row)*MB + m)*2 + c)*PR + col splits into a
  // wave-uniform slab offset (buf, m, c) and a per-thread offset (row, col) that is
  // fixed for the walk (t_rk .. t_rsk), so an access costs one add.
  __device__ __forceinline__ uint32_t slab(int buf, int m, int c) const {
    return (uint32_t)((buf * (TH + 2) * MB * 2 + m * 2 + c) * PR) * 8u;
  }
  __device__ __forceinline__ uint64_t &at(uint32_t slab_off, uint32_t toff) const {
    return *reinterpret_cast<uint64_t *>(reinterpret_cast<char *>(lds) + slab_off + toff);
  }
  // WROW: a row is one whole wave (L/2 == 64); jx of pair k+1 then comes from the
  // next lane at use time instead of a load
  template <bool WROW>
  __device__ __forceinline__ void load(int z, fused_loads<MB> &q) const {
    const bool top = z == L - 1;
    const uint32_t P = (uint32_t)z * plane_bytes;
    const uint32_t Pd = top ? 0u : P + plane_bytes;
    const uint32_t vn = top ? kOutOfRange : vpair;
#pragma unroll
    for (int m = 0; m < MB; ++m) {
      bld2<SA_FUSED_PAIR_CPOL>(sin[m], vn, Pd, q.na[m], q.nb[m]);
      q.ring2[m] = bld(sin[m], vring, P);
    }
    bld2(bjx, vpair, P, q.jxa, q.jxb);
    if (!WROW) q.jxn = bld(bjx, vnext, P);
    bld2(bjy, vpair, P, q.jya, q.jyb);
    bld2(bjy, vpairS, P, q.jsa, q.jsb);
    bld2(bjz, vpair, Pd, q.jda, q.jdb);
    if (!COUNTER) {
      const uint32_t half_plane = (uint32_t)z * (plane_bytes >> 1);
      q.g0 = bld(bw, vrank, half_plane);
      q.g1 = bld(bw, vrank1, half_plane + (uint32_t)(N >> 1) * 8u);
    }
  }

  // What the colour-1 update of a plane needs, captured right after the colour-0
  // update of the same plane.
  struct bstate {
    uint64_t jw, je, jn, js, jd, g;
    uint64_t si[MB], down[MB], n0[MB];
  };

  // in-plane neighbour words of the thread's site of one colour, per system
  struct nbrs {
    uint64_t w[MB], e[MB], n[MB], s[MB];
  };
  // Both phases' LDS reads are issued together at the top of an iteration (they
  // read different plane buffers than the iteration writes), so a wave waits for
  // LDS once per plane.  Rows rn/rs are r-1/r+1 clamped to the band: the halo
  // rows take ring2 instead, or do not use the values (phase B).
  template <int PT>
  __device__ __forceinline__ void read_a(int q, const fused_loads<MB> &ld, nbrs &v) const {
    const int P = PT >= 0 ? PT : par;
    const int buf = q % 3;
#pragma unroll
    for (int m = 0; m < MB; ++m) {
      const uint32_t sl = slab(buf, m, 1);
      v.w[m] = at(sl, P ? t_rk : t_rkp);
      v.e[m] = at(sl, P ? t_rkn : t_rk);
      const uint64_t nn = at(sl, t_rnk), ss = at(sl, t_rsk);
      v.n[m] = (r == 0) ? ld.ring2[m] : nn;
      v.s[m] = (r == TH + 1) ? ld.ring2[m] : ss;
    }
  }
  template <int PT>
  __device__ __forceinline__ void read_b(int q, nbrs &v) const {
    const int P = PT >= 0 ? PT : par;
    const int buf = q % 3;
#pragma unroll
    for (int m = 0; m < MB; ++m) {
      const uint32_t sl = slab(buf, m, 0);
      v.w[m] = at(sl, P ? t_rkp : t_rk);
      v.e[m] = at(sl, P ? t_rk : t_rkn);
      v.n[m] = at(sl, t_rnk);
      v.s[m] = at(sl, t_rsk);
    }
  }

  // colour-0 update of plane q (all rows, halo included).  q_loads holds the loads
  // of plane q, pair[] the old pair at plane q, v the colour-1 neighbours.  Writes
  // the new colour-0 words into LDS buffer q%3 and fills bs for the colour-1 update
  // of plane q.
  template <int PT>
  __device__ __forceinline__ void phase_a(int q, const fused_loads<MB> &ld, const nbrs &v, const uint64_t *pa,
                                          const uint64_t *pb, uint64_t g0, uint64_t g1, uint64_t *up0,
                                          uint64_t *first0, const uint64_t *first1, uint64_t &ju0, bstate &bs) const {
    const int P = PT >= 0 ? PT : par;
    const int buf = q % 3;
    const bool last = q == L - 1;
    // jx of 2k+2: the next pair's jxa (lane k+1 of the same row) when a row is a wave
    const uint64_t jxn = PT >= 0 ? __shfl(ld.jxa, (int)((threadIdx.x + 1) & 63)) : ld.jxn;
    const uint64_t j0w = P ? ld.jxb : ld.jxa, j0e = P ? jxn : ld.jxb;
    const uint64_t j0n = P ? ld.jyb : ld.jya, j0s = P ? ld.jsb : ld.jsa;
    const uint64_t j0d = P ? ld.jdb : ld.jda;
#pragma unroll
    for (int m = 0; m < MB; ++m) {
      const uint64_t si = P ? pb[m] : pa[m];
      const uint64_t w = v.w[m], e = v.e[m], nn = v.n[m], ss = v.s[m];
      const uint64_t down = last ? first0[m] : (P ? ld.nb[m] : ld.na[m]);
      const count3 c = count6(xor3(j0w, si, w), xor3(j0e, si, e), xor3(j0n, si, nn), xor3(j0s, si, ss),
                              xor3(ju0, si, up0[m]), xor3(j0d, si, down));
      const uint64_t n0 =
          FAST ? update_fast3t(t4[m], t5[m], t6[m], g0, c, si) : update_general(p.tab[sys0 + m], 0, g0, c, si, 7);
      at(slab(buf, m, 0), t_rk) = n0;
      up0[m] = n0;
      if (q == 0) first0[m] = n0;
      bs.si[m] = P ? pa[m] : pb[m];
      bs.down[m] = last ? first1[m] : (P ? ld.na[m] : ld.nb[m]);
      bs.n0[m] = n0;
    }
    ju0 = j0d;
    bs.jw = P ? ld.jxa : ld.jxb;
    bs.je = P ? ld.jxb : jxn;
    bs.jn = P ? ld.jya : ld.jyb;
    bs.js = P ? ld.jsa : ld.jsb;
    bs.jd = P ? ld.jda : ld.jdb;
    bs.g = g1;
  }

  // colour-1 update of plane q (band rows): reads the new colour-0 words of plane q
  // from LDS buffer q%3; oa/ob = the pair's new words (x = 2k, 2k+1).
  template <int PT>
  __device__ __forceinline__ void phase_b(int q, const nbrs &v, const bstate &bs, uint64_t *up1, uint64_t *first1,
                                          uint64_t &ju1, uint64_t *oa, uint64_t *ob) const {
    const int P = PT >= 0 ? PT : par;
#pragma unroll
    for (int m = 0; m < MB; ++m) {
      const uint64_t si = bs.si[m];
      const uint64_t w = v.w[m], e = v.e[m], nn = v.n[m], ss = v.s[m];
      const count3 c = count6(xor3(bs.jw, si, w), xor3(bs.je, si, e), xor3(bs.jn, si, nn), xor3(bs.js, si, ss),
                              xor3(ju1, si, up1[m]), xor3(bs.jd, si, bs.down[m]));
      const uint64_t n1 = FAST ? update_fast3t(t4[m], t5[m], t6[m], bs.g, c, si)
                               : update_general(p.tab[sys0 + m], 0, bs.g, c, si, 7);
      up1[m] = n1;
      if (q == 0) first1[m] = n1;
      oa[m] = P ? n1 : bs.n0[m];
      ob[m] = P ? bs.n0[m] : n1;
    }
    ju1 = bs.jd;
  }

  // old colour-1 word of plane q into LDS buffer q%3.  The old colour-0 word is not
  // stored: nothing reads it from LDS (phase A of plane q writes the new one into the
  // same slot before any phase B reads the slot).
  template <int PT>
  __device__ __forceinline__ void put_old(int q, const uint64_t *pa, const uint64_t *pb) const {
    const int P = PT >= 0 ? PT : par;
    const int buf = q % 3;
#pragma unroll
    for (int m = 0; m < MB; ++m) at(slab(buf, m, 1), t_rk) = P ? pa[m] : pb[m];
  }

  // random words of the colour-0 / colour-1 site at plane q
  __device__ __forceinline__ void words(int q, const fused_loads<MB> &ld, int x0, int x1, u32x4 &blk0, u32x4 &blk1,
                                        uint64_t &g0, uint64_t &g1) const {
    if (COUNTER) {
      if ((q & 1) == 0) {
        blk0 = counter_block((uint32_t)(x0 + row + q * L2), p.seed, p.sweep);
        if (!halo) blk1 = counter_block((uint32_t)(x1 + row + q * L2), p.seed, p.sweep);
      }
      g0 = counter_half(blk0, q & 1);
      g1 = counter_half(blk1, q & 1);
    } else {
      g0 = ld.g0;
      g1 = ld.g1;
    }
  }

  // One pass over z.  Iteration z: colour-1 update of plane z, colour-0 update of
  // plane z+1, old words of plane z+2 into LDS, one barrier.  Phase A(z+1) may run
  // before phase B(z) finishes elsewhere: it reads only old colour-1 words of plane
  // z+1 and the column's own colour-0 chain.  Three LDS plane buffers: z (read by
  // phase B), z+1 (phase A), z+2 (being filled).
  template <int PT>
  __device__ __forceinline__ void walk() const {
    const int P = PT >= 0 ? PT : par;
    const int x0 = 2 * k + P, x1 = 2 * k + 1 - P;
    uint64_t pa[MB], pb[MB], up0[MB], up1[MB], first0[MB], first1[MB];
    const uint32_t top = (uint32_t)(L - 1) * plane_bytes;
#pragma unroll
    for (int m = 0; m < MB; ++m) {
      bld2(sin[m], vpair, 0u, pa[m], pb[m]);
      uint64_t ta, tb;
      bld2(sin[m], vpair, top, ta, tb);  // old plane L-1
      up0[m] = P ? tb : ta;
      up1[m] = P ? ta : tb;
      first0[m] = first1[m] = 0;
    }
    uint64_t ja, jb;
    bld2(bjz, vpair, 0u, ja, jb);  // -z bonds at plane 0
    uint64_t ju0 = P ? jb : ja, ju1 = P ? ja : jb;
    u32x4 blk0{0, 0, 0, 0}, blk1{0, 0, 0, 0};
    fused_loads<MB> A, B;  // loads of planes z+1 / z+2, used alternately
    bstate bs;
    // prologue: old planes 0 and 1 into LDS, colour-0 update of plane 0
    load<PT >= 0>(0, A);
    load<PT >= 0>(1, B);
    put_old<PT>(0, pa, pb);
    put_old<PT>(1, A.na, A.nb);
    __syncthreads();
    {
      uint64_t g0, g1;
      nbrs va;
      read_a<PT>(0, A, va);
      words(0, A, x0, x1, blk0, blk1, g0, g1);
      phase_a<PT>(0, A, va, pa, pb, g0, g1, up0, first0, first1, ju0, bs);
    }
#pragma unroll
    for (int m = 0; m < MB; ++m) {
      pa[m] = A.na[m];  // old pair at plane 1
      pb[m] = A.nb[m];
    }
    __syncthreads();
    // iteration z: near = loads of plane z+1, far = loads of plane z+2 (prefetched)
    auto iteration = [&](int z, fused_loads<MB> &near, fused_loads<MB> &far) {
      load<PT >= 0>(z + 2 < L ? z + 2 : L - 1, far);  // the last two loads are not used
      nbrs vb;
      read_b<PT>(z, vb);
      {
        uint64_t oa[MB], ob[MB];
        if (!halo) {
          phase_b<PT>(z, vb, bs, up1, first1, ju1, oa, ob);
        } else {
#pragma unroll
          for (int m = 0; m < MB; ++m) oa[m] = ob[m] = 0;
        }
#pragma unroll
        for (int m = 0; m < MB; ++m) bst2<SA_FUSED_PAIR_CPOL>(sout[m], vstore, (uint32_t)z * plane_bytes, oa[m], ob[m]);
      }
      if (z + 1 < L) {
        uint64_t g0, g1;
        nbrs va;
        read_a<PT>(z + 1, near, va);
        words(z + 1, near, x0, x1, blk0, blk1, g0, g1);
        phase_a<PT>(z + 1, near, va, pa, pb, g0, g1, up0, first0, first1, ju0, bs);
        if (z + 2 < L) put_old<PT>(z + 2, near.na, near.nb);
#pragma unroll
        for (int m = 0; m < MB; ++m) {
          pa[m] = near.na[m];  // old pair at plane z+2
          pb[m] = near.nb[m];
        }
      }
      __syncthreads();
    };
    for (int z = 0; z < L; z += 2) {  // L is even
      iteration(z, B, A);
      iteration(z + 1, A, B);
    }
  }
};

// One workgroup's band `tile` of system block `sb` (systems sb*MB .. sb*MB+MB-1):
// the whole z-walk.  blockDim.x must be (TH+2)*L/2 and the dynamic LDS
// 3*(TH+2)*MB*2*(L/2)*8 bytes.
// WALK: 1 = rows are whole waves (L/2 a multiple of 64), 0 = they are not, 2 = decide
// at run time (both walks compiled in).
template <bool COUNTER, bool FAST, int MB, int TH, int WALK = 2>
__device__ __forceinline__ void fused_band(const sweep_args &p, uint64_t *__restrict__ s_out, int sb, int tile,
                                           uint64_t *lds) {
  const int L = p.L, PR = L >> 1;
  fused_ctx<COUNTER, FAST, MB, TH> c{p, lds};
  c.L = L;
  c.PR = PR;
  c.r = threadIdx.x / PR;
  c.k = threadIdx.x - c.r * PR;
  c.kn = c.k == PR - 1 ? 0 : c.k + 1;
  c.kp = c.k == 0 ? PR - 1 : c.k - 1;
  c.halo = (c.r == 0) || (c.r == TH + 1);
  c.rn = c.r == 0 ? 0 : c.r - 1;
  c.rs = c.r == TH + 1 ? TH + 1 : c.r + 1;
  {
    const uint32_t rowst = (uint32_t)(MB * 2 * PR);  // elements per LDS row (all systems, both colours)
    c.t_rk = ((uint32_t)c.r * rowst + (uint32_t)c.k) * 8u;
    c.t_rkp = ((uint32_t)c.r * rowst + (uint32_t)c.kp) * 8u;
    c.t_rkn = ((uint32_t)c.r * rowst + (uint32_t)c.kn) * 8u;
    c.t_rnk = ((uint32_t)c.rn * rowst + (uint32_t)c.k) * 8u;
    c.t_rsk = ((uint32_t)c.rs * rowst + (uint32_t)c.k) * 8u;
  }
  c.y = (tile * TH - 1 + c.r + L) % L;
  const int ym = c.y == 0 ? L - 1 : c.y - 1, yp = c.y == L - 1 ? 0 : c.y + 1;
  c.row = c.y * L;
  c.rowS = yp * L;
  c.yr2 = (c.r == 0) ? ym : yp;  // halo rows: the row beyond the halo
  c.sys0 = sb * MB;
#pragma unroll
  for (int m = 0; m < MB; ++m) {  // loaded once for the whole walk
    const sys_table &t = p.tab[c.sys0 + m];
    c.t4[m] = t.thr[0][4];
    c.t5[m] = t.thr[0][5];
    c.t6[m] = t.thr[0][6];
  }
  c.L2 = (int64_t)L * L;
  c.N = p.N;
  c.rank0 = (int64_t)c.y * PR + c.k;  // visit rank inside a plane half
  const uint32_t bytes = (uint32_t)(c.N * 8);
#pragma unroll
  for (int m = 0; m < MB; ++m) {
    c.sin[m] = make_buf(p.s + (int64_t)(c.sys0 + m) * c.N, bytes);
    c.sout[m] = make_buf(s_out + (int64_t)(c.sys0 + m) * c.N, bytes);
  }
  c.bjx = make_buf(p.jx, bytes);
  c.bjy = make_buf(p.jy, bytes);
  c.bjz = make_buf(p.jz, bytes);
  c.bw = make_buf(p.words, COUNTER ? 0u : bytes);
  c.plane_bytes = (uint32_t)(c.L2 * 8);
  c.vpair = (uint32_t)(c.row + 2 * c.k) * 8u;
  c.vpairS = (uint32_t)(c.rowS + 2 * c.k) * 8u;
  c.vnext = (uint32_t)(c.row + 2 * c.kn) * 8u;
  c.vring = c.halo ? (uint32_t)(c.yr2 * L + 2 * c.k + (c.y & 1)) * 8u : kOutOfRange;
  c.vrank = (uint32_t)c.rank0 * 8u;
  c.vrank1 = c.halo ? kOutOfRange : c.vrank;  // halo rows do not update colour 1
  c.vstore = c.halo ? kOutOfRange : c.vpair;  // nor store anything
  c.par = c.y & 1;
  // The walk contains barriers, so a parity-specialised walk may only be chosen
  // per wave: when a row is a whole number of waves (L/2 a multiple of 64) every
  // wave has one parity; otherwise the parity stays a run-time value.
  if (WALK == 1 || (WALK == 2 && (PR & 63) == 0)) {
    if (c.par) c.template walk<1>(); else c.template walk<0>();
  } else {
    c.template walk<-1>();
  }
}

// Compiled once per walk (WALK 0 or 1, see fused_band): one kernel holding both
// walks allocates registers for both and runs slower.
template <bool COUNTER, bool FAST, int MB, int TH, int WALK>
__global__ void __launch_bounds__(640) k_sweep3d_fused(sweep_args p, uint64_t *__restrict__ s_out, int xcd_map) {
  extern __shared__ uint64_t lds[];
  const int ntile = p.L / TH;
  // workgroup -> (system block, band).  Workgroups are dispatched round-robin over
  // the 8 XCDs (L2 is per XCD).  xcd_map = 0: b = sb*ntile + band, so with ntile a
  // multiple of 8 every system block of a band runs on one XCD and shares its
  // couplings in L2.  xcd_map = 1: XCD b%8 gets ntile/8 adjacent bands of every
  // system block, so the halo rows between those bands are L2 hits too.
  // xcd_map = 2: XCD b%8 gets every band of nsb/8 whole system blocks, so every halo
  // row a band reads is read by its neighbour band on the same XCD (the couplings
  // are then read once per XCD).
  const int b = blockIdx.x;
  int sb, tile;
  if (xcd_map == 2) {
    const int per = (p.nsys / MB) >> 3, xcd = b & 7, q = b >> 3;
    sb = xcd * per + q / ntile;
    tile = q % ntile;
  } else if (xcd_map) {
    const int per = ntile >> 3, xcd = b & 7, q = b >> 3;
    tile = xcd * per + q % per;
    sb = q / per;
  } else {
    sb = b / ntile;
    tile = b - sb * ntile;
  }
  fused_band<COUNTER, FAST, MB, TH, WALK>(p, s_out, sb, tile, lds);
}

// ---------------------------------------------------------------- 3D half-sweep, z in parallel
// For a lattice too small to fill the GPU with column walks (one system offers L*L/2
// columns per half-sweep), each column is cut into G segments of S = L/G planes, one
// thread per segment, the G threads of a column in one wave: a wave holds C = 64/G
// consecutive columns, lane = segment*C + column, so C neighbouring lanes read
// neighbouring columns of one plane.
// In a half-sweep the in-plane neighbours have the other colour and stay fixed, and
// the old +z word is fixed, so the new word of plane z is a per-lane function of the
// new word of plane z-1 only: one of the four maps of a bit, kept as the pair
// (f0, f1) = (result for a 0 below, result for a 1 below) -- two evaluations of the
// update with the -z neighbour all zeros and all ones, sharing the site's random
// word.  Plane 0 reads the old plane L-1, so its map is constant; plane L-1 reads the
// new plane 0 as its +z neighbour, which the segment-0 lane has computed first and
// passes on.  Maps compose as (F o G)(b) = G(b) ? F1 : F0 per bit; a scan over the
// G lanes gives every segment the new word below it, and each segment then applies
// its maps in order and stores.  The result is the reference's in-order z walk bit
// for bit (src/cpu_3d.cpp:94-105), with the same random word per site.  All loads of
// a wave complete before its scan, so no segment's store meets a read of another.
struct zmap {
  uint64_t f0, f1;
};
__device__ __forceinline__ zmap zcompose(const zmap &F, const zmap &G) {  // F o G
  return zmap{sel(G.f0, F.f1, F.f0), sel(G.f1, F.f1, F.f0)};
}
__device__ __forceinline__ uint64_t shfl_up64(uint64_t v, int d, int width) {
  const uint32_t lo = __shfl_up((uint32_t)v, d, width), hi = __shfl_up((uint32_t)(v >> 32), d, width);
  return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint64_t shfl64(uint64_t v, int src) {
  const uint32_t lo = __shfl((uint32_t)v, src), hi = __shfl((uint32_t)(v >> 32), src);
  return ((uint64_t)hi << 32) | lo;
}

// segments per column for an even L: a power of two G <= 64 with L/G <= kZsegMax
// planes per segment and at least two planes; 0 when there is none.  Two planes per
// segment while a wave still holds two or more columns (G <= 32), else the fewest
// segments of at most 4 planes, else of at most 8.  Measured per sweep of one system:
// L=64 13.0 us with 2-plane segments, 13.9 with 4; L=32 7.9 / 9.6; L=128 63 with 4
// (32 segments, two columns per wave) against 94 with 2 (one column per wave).
constexpr int kZsegMax = 8;
__host__ __device__ __forceinline__ int zscan_groups(int L) {
  const int h = L / 2;
  if (L >= 4 && L % 2 == 0 && h <= 32 && (h & (h - 1)) == 0) return h;
  for (int G = 2; G <= 64; G *= 2)
    if (L % G == 0 && L / G <= 4 && L / G >= 2) return G;
  for (int G = 2; G <= 64; G *= 2)
    if (L % G == 0 && L / G <= kZsegMax && L / G >= 2) return G;
  return 0;
}

template <bool COUNTER, bool FAST, int S>
__device__ __forceinline__ void zscan_segment(const sweep_args &p, int color, int k, int g, int G, int sys, bool live) {
  const int L = p.L, half = L >> 1;
  const int y = k / half;
  const int x = 2 * (k - y * half) + ((y + color) & 1);
  const int xm = x == 0 ? L - 1 : x - 1, xp = x == L - 1 ? 0 : x + 1;
  const int ym = y == 0 ? L - 1 : y - 1, yp = y == L - 1 ? 0 : y + 1;
  const int64_t L2 = (int64_t)L * L, N = p.N;
  const int c0 = x + L * y, cw = xm + L * y, ce = xp + L * y, cn = x + L * ym, cs = x + L * yp;
  uint64_t *s = p.s + (int64_t)sys * N;
  const sys_table &t = p.tab[sys];
  const int z0 = g * S;
  const int lane = threadIdx.x & 63, C = 64 / G, base = lane % C;  // base: the column's segment-0 lane
  auto eval = [&](uint64_t g64, const uint64_t *q, uint64_t si, uint64_t bd, uint64_t bu) {
    const count3 cc = count6(q[0], q[1], q[2], q[3], bd, bu);
    return FAST ? update_fast3(t, g64, cc, si) : update_general(t, 0, g64, cc, si, 7);
  };
  // every load of the segment first (a dead lane reads column 0: valid addresses),
  // then the arithmetic: the planes' memory round trips overlap
  uint64_t si[S], nb[S][4], jb[S][4], ju[S], jd[S], dn[S], g64[S], q[S][4];
#pragma unroll
  for (int j = 0; j < S; ++j) {
    const int z = z0 + j;
    const int64_t P = z * L2;
    si[j] = s[P + c0];
    nb[j][0] = s[P + cw];
    nb[j][1] = s[P + ce];
    nb[j][2] = s[P + cn];
    nb[j][3] = s[P + cs];
    jb[j][0] = p.jx[P + c0];
    jb[j][1] = p.jx[P + ce];
    jb[j][2] = p.jy[P + c0];
    jb[j][3] = p.jy[P + cs];
    ju[j] = p.jz[P + c0];
    jd[j] = p.jz[(z == L - 1 ? 0 : P + L2) + c0];
    dn[j] = s[(z == L - 1 ? 0 : P + L2) + c0];  // plane L-1: replaced by the new plane 0 below
    if (!COUNTER) g64[j] = p.words[(int64_t)color * (N >> 1) + k + z * (L2 >> 1)];
  }
  const uint64_t up0 = s[(L - 1) * L2 + c0];  // old plane L-1: plane 0's -z neighbour
  if (COUNTER) {
    u32x4 blk{0, 0, 0, 0};
#pragma unroll
    for (int j = 0; j < S; ++j) {
      const int z = z0 + j;
      if (j == 0 || (z & 1) == 0) blk = counter_block((uint32_t)(c0 + (z & ~1) * L2), p.seed, p.sweep);
      g64[j] = counter_half(blk, z & 1);
    }
  }
  zmap F[S];
#pragma unroll
  for (int j = 0; j < S; ++j) {
    const int z = z0 + j;
#pragma unroll
    for (int r = 0; r < 4; ++r) q[j][r] = xor3(jb[j][r], si[j], nb[j][r]);
    const uint64_t bu0 = z == 0 ? xor3(ju[j], si[j], up0) : (ju[j] ^ si[j]);  // -z bond word, 0 below
    const uint64_t bd = xor3(jd[j], si[j], dn[j]);
    F[j].f0 = eval(g64[j], q[j], si[j], bd, bu0);
    F[j].f1 = z == 0 ? F[j].f0 : eval(g64[j], q[j], si[j], bd, ~bu0);  // plane 0: a constant map
  }
  // the new plane 0 (segment 0's constant map) to every lane of the column
  const uint64_t new0 = shfl64(F[0].f0, base);
  if (z0 + S == L) {  // plane L-1: its +z neighbour is the new plane 0
    constexpr int j = S - 1;
    const uint64_t bd = xor3(jd[j], si[j], new0), bu0 = ju[j] ^ si[j];
    F[j] = zmap{eval(g64[j], q[j], si[j], bd, bu0), eval(g64[j], q[j], si[j], bd, ~bu0)};
  }
  // the segment's map: F[S-1] o ... o F[0]
  zmap M = F[0];
#pragma unroll
  for (int j = 1; j < S; ++j) M = zcompose(F[j], M);
  // inclusive scan over the column's G lanes (C apart): segment g ends with
  // G_g o ... o G_0, a constant map (G_0 is one), i.e. the new word of its last plane
  for (int d = 1; d < G; d *= 2) {
    const zmap T{shfl_up64(M.f0, d * C, 64), shfl_up64(M.f1, d * C, 64)};
    if (g >= d) M = zcompose(M, T);
  }
  uint64_t v = shfl_up64(M.f0, C, 64);  // the new word below the segment's first plane
  if (!live) return;
#pragma unroll
  for (int j = 0; j < S; ++j) {
    v = (g == 0 && j == 0) ? F[0].f0 : sel(v, F[j].f1, F[j].f0);
    s[(int64_t)(z0 + j) * L2 + c0] = v;
  }
}

// one launch per colour; thread = (column, segment), grid.y = system
template <bool COUNTER, bool FAST, int S>
__global__ void __launch_bounds__(256) k_sweep3d_zscan(sweep_args p, int color, int G) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int C = 64 / G, lane = i & 63;
  const int k = (i >> 6) * C + lane % C, g = lane / C;
  const bool live = k < p.L * (p.L >> 1);
  zscan_segment<COUNTER, FAST, S>(p, color, live ? k : 0, g, G, blockIdx.y, live);
}

// ---------------------------------------------------------------- 2D half-sweep
// site k (0 <= k < L*L/2) of colour `color`, systems sys0 .. sys0+MB-1
template <bool COUNTER, bool FAST, bool OPEN, int MB>
__device__ __forceinline__ void half2d_site(const sweep_args &p, int color, int k, int sys0) {
  const int L = p.L;
  const int half = L >> 1;
  const int y = k / half;
  const int x = 2 * (k - y * half) + ((y + color) & 1);
  const int xm = x == 0 ? L - 1 : x - 1, xp = x == L - 1 ? 0 : x + 1;
  const int ym = y == 0 ? L - 1 : y - 1, yp = y == L - 1 ? 0 : y + 1;
  const int c0 = x + L * y;
  const int cw = xm + L * y, ce = xp + L * y, cn = x + L * ym, cs = x + L * yp;
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
    uint64_t bw = xor3(jw, si, s[cw]);
    uint64_t be = xor3(je, si, s[ce]);
    if (OPEN) {  // missing open-x bond counts as satisfied (src/cpu_2dp.cpp:338-339)
      if (x == 0) bw = ~0ULL;
      if (x == L - 1) be = ~0ULL;
    }
    const uint64_t bn = xor3(jn, si, s[cn]);
    const uint64_t bs = xor3(js, si, s[cs]);
    const count3 c = count4(bw, be, bn, bs);
    const sys_table &t = p.tab[sys0 + m];
    s[c0] = (FAST && !OPEN) ? update_fast2(t, g, c, si) : update_general(t, border ? 2 : 0, g, c, si, 5);
  }
}

template <bool COUNTER, bool FAST, bool OPEN, int MB>
__global__ void __launch_bounds__(256) k_sweep2d_half(sweep_args p, int color) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= p.L * (p.L >> 1)) return;
  half2d_site<COUNTER, FAST, OPEN, MB>(p, color, k, blockIdx.y * MB);
}

// ---------------------------------------------------------------- odd L
// With odd L the wrap neighbours in x and in y have the same colour: in a half-sweep
// the site (L-1, y) reads the new (0, y), the site (x, L-1) the new (x, 0), and
// (L-1, L-1) the new (0, L-1) and (L-1, 0) -- each is visited after its partner in
// the reference's row-major order (src/cpu_3d.cpp:94-105, src/cpu_2dp.cpp:106-114),
// which sees the partner's new value and the other way round the old one.  These
// groups (a site of the (L-1) x (L-1) corner block [0, L-1)^2 with its partners in
// the last column / last row / corner) are closed: every other neighbour of a group
// site has the other colour and does not change in the half-sweep, and along z a
// site's same-colour neighbours are its own column.  A group's sites are updated
// in visit order with the new words of its own sites passed on in registers (the
// partners' +x / +y neighbours); every other neighbour is an old word.  The corner
// block holds (L-1)^2/2 sites of each colour.  Open x (2p): the x-wrap has no bond,
// so the x partner is merely handled in the same group.

// visit rank of colour-c site (x, y) in plane z (the replay word it consumes): all
// colour-0 sites come first; a row holds (L-1)/2 sites plus one when it starts at x=0
__device__ __forceinline__ int64_t odd_rank(int L, int nz, int c, int x, int y, int z) {
  const int64_t L2 = (int64_t)L * L;
  const int64_t P0 = (L2 + 1) / 2, Pc = c ? L2 / 2 : P0;  // colour-0 / colour-c sites per plane
  const int64_t rows_before = (int64_t)y * ((L - 1) / 2) + (c == 0 ? (y + 1) / 2 : y / 2);
  return (c ? nz * P0 : 0) + z * Pc + rows_before + (x - ((y + c) & 1)) / 2;
}

// groups of one colour: (L-1)^2/2 corner-block sites, one per thread
__host__ __device__ __forceinline__ int odd_groups(int L) { return (L - 1) * ((L - 1) >> 1); }

// Group k's members: 0 = (x, y) in the corner block, 1 = (L-1, y) when x = 0, 2 = (x, L-1)
// when y = 0, 3 = (L-1, L-1) when x = y = 0.  Forwarded new words: member 1's +x
// neighbour is member 0, member 2's +y neighbour is member 0, member 3's are members 2
// (+x) and 1 (+y).
struct odd_member {
  int c0, cw, ce, cn, cs;  // in-plane offsets of the site and its -x, +x, -y, +y neighbours
};
__device__ __forceinline__ odd_member odd_offsets(int L, int x, int y) {
  const int xm = x == 0 ? L - 1 : x - 1, xp = x == L - 1 ? 0 : x + 1;
  const int ym = y == 0 ? L - 1 : y - 1, yp = y == L - 1 ? 0 : y + 1;
  return odd_member{x + L * y, xm + L * y, xp + L * y, x + L * ym, x + L * yp};
}
struct odd_loads {
  uint64_t jw, je, jn, js, jd;  // couplings: -x, +x, -y, +y, +z bonds
  uint64_t w, e, n, s, d;       // old neighbour words (e / s not loaded where forwarded)
};

// 3D: one lane per member.  Lane i of a system's odd_width(L) lanes (systems start
// at a wave boundary):
//   i < 4h (h = (L-1)/2): lane pairs (2q, 2q+1) = (corner-block site, its partner)
//     colour 0: q = 0: (0,0) & (0,L-1);  q = 1: (L-1,0) & (L-1,L-1);
//               q = 2..h: (2(q-1),0) & (2(q-1),L-1);  q = h+1..2h-1: (0,y) & (L-1,y), y = 2(q-h)
//     colour 1: q < h: (1+2q,0) & (1+2q,L-1);  q = h..2h-1: (0,y) & (L-1,y), y = 1+2(q-h)
//   then the other corner-block sites (1 <= x, y <= L-2), two rows at a time;
//   lanes from the colour's site count on are idle.
// Member 0 sites walk z in step tau = z, members 1 and 2 in tau = z+1, member 3 in
// tau = z+2: a partner updates plane z one step after the site it reads the new
// word of, which then sits in that lane's `up` register and is passed by a wave
// shuffle.  The other same-colour wrap neighbours (a member 0 site reads its
// partners, members 1 and 2 read (L-1,L-1)) are updated one step or more after the
// reader has used their old word, and all other neighbours have the other colour,
// so every lane prefetches its plane loads two steps ahead like the even walk.
__host__ __device__ __forceinline__ int odd_width(int L) { return ((L * L + 1) / 2 + 63) & ~63; }

template <bool COUNTER, bool FAST>
__device__ __forceinline__ void odd3d_lane(const sweep_args &p, int color, int i, int sys) {
  const int L = p.L, h = (L - 1) >> 1;
  const int64_t L2 = (int64_t)L * L;
  const int64_t Pc = color ? L2 / 2 : (L2 + 1) / 2;  // sites of this colour per plane
  const int lane = i & 63;
  const bool live = i < Pc;
  int x = 0, y = 0, skew = 0, srcE = lane, srcS = lane;  // lanes of the forwarded +x / +y words
  if (i < 4 * h) {
    const int q = i >> 1, r = i & 1;
    int xr = -1, yr = -1;  // row-0 pair at x = xr, or x = 0 pair at row yr
    if (color == 0) {
      if (q == 1) {  // (L-1,0) reads the new (0,0); (L-1,L-1) reads (0,L-1) and (L-1,0)
        x = L - 1;
        y = r ? L - 1 : 0;
        skew = 1 + r;
        srcE = lane - 2;
        if (r) srcS = lane - 1;
      } else if (q <= h) {
        xr = q == 0 ? 0 : 2 * (q - 1);
      } else {
        yr = 2 * (q - h);
      }
    } else {
      if (q < h) xr = 1 + 2 * q;
      else yr = 1 + 2 * (q - h);
    }
    if (xr >= 0) {  // (xr, 0) and (xr, L-1), which reads the new (xr, 0) as its +y neighbour
      x = xr;
      y = r ? L - 1 : 0;
      skew = r;
      if (r) srcS = lane - 1;
    } else if (yr >= 0) {  // (0, yr) and (L-1, yr), which reads the new (0, yr) as its +x neighbour
      x = r ? L - 1 : 0;
      y = yr;
      skew = r;
      if (r) srcE = lane - 1;
    }
  } else {
    // rows 1 .. L-2 in pairs; a row whose x = 0 site has this colour holds h-1
    // sites (x = 2, 4, ..), the other h (x = 1, 3, ..); row 1 is of the second kind
    // for colour 0 and of the first for colour 1
    const int sidx = i - 4 * h, r2 = sidx / (2 * h - 1), rem = sidx - r2 * (2 * h - 1);
    const int y0 = 1 + 2 * r2;
    const int n0 = color ? h - 1 : h;  // sites in row y0
    if (rem < n0) {
      y = y0;
      x = (color ? 2 : 1) + 2 * rem;
    } else {
      y = y0 + 1;
      x = (color ? 1 : 2) + 2 * (rem - n0);
    }
  }
  const odd_member o = odd_offsets(L, x, y);
  const bool fwdE = srcE != lane, fwdS = srcS != lane;
  uint64_t *s = p.s + (int64_t)sys * p.N;
  const sys_table &t = p.tab[sys];
  const int64_t rk = live ? odd_rank(L, L, color, x, y, 0) : 0;
  uint64_t own = 0, up = 0, first = 0, jzu = 0;
  if (live) {
    own = s[o.c0];
    up = s[o.c0 + (L - 1) * L2];  // old plane L-1
    jzu = p.jz[o.c0];
  }
  auto load = [&](int z, odd_loads &q) {
    if (!live || z < 0 || z >= L) return;
    const int64_t P = z * L2, Pd = z == L - 1 ? 0 : P + L2;
    q.jw = p.jx[P + o.c0];
    q.je = p.jx[P + o.ce];
    q.jn = p.jy[P + o.c0];
    q.js = p.jy[P + o.cs];
    q.jd = p.jz[Pd + o.c0];
    q.w = s[P + o.cw];
    q.n = s[P + o.cn];
    q.e = fwdE ? 0 : s[P + o.ce];
    q.s = fwdS ? 0 : s[P + o.cs];
    q.d = z == L - 1 ? 0 : s[Pd + o.c0];
  };
  u32x4 blk{0, 0, 0, 0};
  // step tau: the forwarded words are the source lanes' `up` (their new word of the
  // plane this lane updates now); plane tau-skew uses the buffer filled at tau-2
  auto step = [&](int tau, const odd_loads &cur, odd_loads &far) {
    const uint64_t fe = __shfl(up, srcE), fs = __shfl(up, srcS);
    load(tau + 2 - skew, far);
    const int z = tau - skew;
    if (!live || z < 0 || z >= L) return;
    const uint64_t si = own;
    const uint64_t down = z == L - 1 ? first : cur.d;
    const count3 c = count6(xor3(cur.jw, si, cur.w), xor3(cur.je, si, fwdE ? fe : cur.e), xor3(cur.jn, si, cur.n),
                            xor3(cur.js, si, fwdS ? fs : cur.s), xor3(jzu, si, up), xor3(cur.jd, si, down));
    uint64_t g;
    if (COUNTER) {
      if ((z & 1) == 0) blk = counter_block((uint32_t)(o.c0 + L2 * z), p.seed, p.sweep);
      g = counter_half(blk, z & 1);
    } else {
      g = p.words[rk + z * Pc];
    }
    const uint64_t sn = FAST ? update_fast3(t, g, c, si) : update_general(t, 0, g, c, si, 7);
    s[z * L2 + o.c0] = sn;
    up = sn;
    if (z == 0) first = sn;
    own = down;
    jzu = cur.jd;
  };
  odd_loads b0{}, b1{}, b2{};
  load(0 - skew, b0);
  load(1 - skew, b1);
  const int steps = L + 2;
  for (int tau = 0; tau < steps; tau += 3) {
    step(tau, b0, b2);
    if (tau + 1 < steps) step(tau + 1, b1, b0);
    if (tau + 2 < steps) step(tau + 2, b2, b1);
  }
}

template <bool COUNTER, bool FAST>
__global__ void __launch_bounds__(256) k_sweep3d_odd(sweep_args p, int color) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < odd_width(p.L)) odd3d_lane<COUNTER, FAST>(p, color, i, blockIdx.y);  // whole waves
}

// 2D: all members' words are loaded first, then updated in visit order
template <bool COUNTER, bool FAST, bool OPEN>
__device__ __forceinline__ void odd2d_group(const sweep_args &p, int color, int k, int sys) {
  const int L = p.L, h = (L - 1) >> 1;
  const int y = k / h, x = 2 * (k - y * h) + ((y + color) & 1);
  uint64_t *s = p.s + (int64_t)sys * p.N;
  const sys_table &t = p.tab[sys];
  const bool pres[4] = {true, x == 0, y == 0, x == 0 && y == 0};
  const int mx[4] = {x, L - 1, x, L - 1}, my[4] = {y, y, L - 1, L - 1};
  odd_member o[4];
  odd_loads b[4];
  uint64_t own[4], nw[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    nw[j] = 0;
    if (!pres[j]) continue;
    o[j] = odd_offsets(L, mx[j], my[j]);
    odd_loads &q = b[j];
    q.jw = p.jx[o[j].c0];
    q.je = p.jx[o[j].ce];
    q.jn = p.jy[o[j].c0];
    q.js = p.jy[o[j].cs];
    q.w = s[o[j].cw];
    q.n = s[o[j].cn];
    q.e = (j == 1 || j == 3) ? 0 : s[o[j].ce];
    q.s = (j >= 2) ? 0 : s[o[j].cs];
    own[j] = s[o[j].c0];
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    if (!pres[j]) continue;
    const odd_loads &q = b[j];
    const int xj = mx[j], yj = my[j];
    const uint64_t si = own[j];
    const uint64_t e = j == 1 ? nw[0] : (j == 3 ? nw[2] : q.e);
    const uint64_t sv = j == 2 ? nw[0] : (j == 3 ? nw[1] : q.s);
    uint64_t bw = xor3(q.jw, si, q.w), be = xor3(q.je, si, e);
    if (OPEN) {  // missing open-x bond counts as satisfied (src/cpu_2dp.cpp:338-339)
      if (xj == 0) bw = ~0ULL;
      if (xj == L - 1) be = ~0ULL;
    }
    const count3 c = count4(bw, be, xor3(q.jn, si, q.n), xor3(q.js, si, sv));
    uint64_t g;
    if (COUNTER) {
      const u32x4 r = counter_block((uint32_t)(((xj & ~3) | (xj & 1)) + L * yj), p.seed, p.sweep);
      g = counter_half(r, (xj >> 1) & 1);
    } else {
      g = p.words[odd_rank(L, 1, color, xj, yj, 0)];
    }
    const bool border = OPEN && (xj == 0 || xj == L - 1);
    const uint64_t sn = (FAST && !OPEN) ? update_fast2(t, g, c, si) : update_general(t, border ? 2 : 0, g, c, si, 5);
    s[o[j].c0] = sn;
    nw[j] = sn;
  }
}

template <bool COUNTER, bool FAST, bool OPEN>
__global__ void __launch_bounds__(256) k_sweep2d_odd(sweep_args p, int color) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < odd_groups(p.L)) odd2d_group<COUNTER, FAST, OPEN>(p, color, k, blockIdx.y);
}

// ---------------------------------------------------------------- measure
// Per system and lane: number of satisfied -x,-y[,-z] bonds and of set spins.
// Each lane keeps bit-sliced counters (bit k of plane b = bit b of lane k's count):
// a full adder folds the bond words of one site into a 2-bit sum that is rippled
// into six planes; the spin word ripples into five.  Every 16 site groups the
// planes are turned into ordinary per-lane counts by a 64x64 bit transpose across
// the wave (six butterfly stages of lane shuffles) and a popcount, so lane k ends
// up holding the counts of bit k.  Blocks reduce in LDS and add one atomic per
// lane and counter.

// Wave-wide 64x64 bit transpose: lane i holds row i on entry, column i on exit.
// Stage d exchanges bit d of the lane index with bit d of the bit index.
__device__ __forceinline__ uint64_t wave_transpose64(uint64_t x, int lane) {
  constexpr uint64_t M32 = 0x00000000FFFFFFFFULL, M16 = 0x0000FFFF0000FFFFULL, M8 = 0x00FF00FF00FF00FFULL,
                     M4 = 0x0F0F0F0F0F0F0F0FULL, M2 = 0x3333333333333333ULL, M1 = 0x5555555555555555ULL;
#define SA_TSTAGE(D, M)                                                            \
  {                                                                                \
    const uint64_t o = __shfl_xor(x, D);                                           \
    x = (lane & D) ? ((x & ~M) | ((o & ~M) >> D)) : ((x & M) | ((o & M) << D));    \
  }
  SA_TSTAGE(32, M32)
  SA_TSTAGE(16, M16)
  SA_TSTAGE(8, M8)
  SA_TSTAGE(4, M4)
  SA_TSTAGE(2, M2)
  SA_TSTAGE(1, M1)
#undef SA_TSTAGE
  return x;
}

// The measurement's per-lane counters are kept in kMeasureCopies copies
// (cnt[copy][listed system][128], copy = block index mod kMeasureCopies), summed on
// the host: blocks then contend for an address 8 times less.
constexpr int kMeasureCopies = 8;

// up to four listed systems, passed by value
struct sys4 {
  int v[4];
};

template <int KIND, int SB>
__global__ void __launch_bounds__(256) k_measure(const uint64_t *__restrict__ s_all, const uint64_t *__restrict__ jx,
                                                 const uint64_t *__restrict__ jy, const uint64_t *__restrict__ jz,
                                                 int L, int64_t N, int nsys, const int *__restrict__ sys_list,
                                                 sys4 sel, int *cnt) {
  // SB: systems per pass over the sites (couplings loaded once per pass); with few
  // systems one per block row gives more waves in flight
  constexpr int FLUSH = 16;  // groups between flushes: E <= 3*16 < 64, M <= 16 < 32
  __shared__ int red[SB][2][64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int sys0 = blockIdx.y * SB;
  const int ns = (nsys - sys0) < SB ? (nsys - sys0) : SB;
  for (int q = threadIdx.x; q < SB * 128; q += blockDim.x) (&red[0][0][0])[q] = 0;
  __syncthreads();
  const int64_t L2 = (int64_t)L * L;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  uint64_t e[SB][6], mc[SB][5];
  int sat[SB], up[SB];
#pragma unroll
  for (int m = 0; m < SB; ++m) {
    sat[m] = up[m] = 0;
#pragma unroll
    for (int b = 0; b < 6; ++b) e[m][b] = 0;
#pragma unroll
    for (int b = 0; b < 5; ++b) mc[m][b] = 0;
  }
  int pending = 0;
  auto flush = [&]() {
#pragma unroll
    for (int m = 0; m < SB; ++m) {
#pragma unroll
      for (int b = 0; b < 6; ++b) {
        sat[m] += __popcll(wave_transpose64(e[m][b], lane)) << b;
        e[m][b] = 0;
      }
#pragma unroll
      for (int b = 0; b < 5; ++b) {
        up[m] += __popcll(wave_transpose64(mc[m][b], lane)) << b;
        mc[m][b] = 0;
      }
    }
  };
  // U groups per pass: all their loads are issued before any of them is counted, so
  // a wave waits for memory once per pass (with one system per block row; with four
  // the loads of one group already fill the registers)
  constexpr int U = SB == 1 ? 4 : 1;
  uint64_t Jw[U], Jn[U], Ju[U], w[U][SB], sw[U][SB], sn[U][SB], su[U][SB];
  bool west[U];
  for (int64_t base0 = (int64_t)blockIdx.x * blockDim.x + wv * 64; base0 < N; base0 += U * stride) {
#pragma unroll
    for (int g = 0; g < U; ++g) {
      const int64_t base = base0 + g * stride;
      // coordinates of the group's first site once per wave (scalar), then per lane
      // by stepping along the row: no per-site division
      const uint32_t b32 = (uint32_t)base, uL = (uint32_t)L, uL2 = (uint32_t)L2;
      const uint32_t zb = b32 / uL2, rem = b32 - zb * uL2, yb = rem / uL, xb = rem - yb * uL;
      uint32_t x = xb + (uint32_t)lane, y = yb, z = zb;
      while (x >= uL) {  // at most 64/L steps
        x -= uL;
        if (++y == uL) {
          y = 0;
          ++z;
        }
      }
      const uint32_t i = b32 + (uint32_t)lane;
      const bool in = base < N && (int64_t)i < N;
      Jw[g] = Jn[g] = Ju[g] = 0;
      west[g] = true;
#pragma unroll
      for (int m = 0; m < SB; ++m) w[g][m] = sw[g][m] = sn[g][m] = su[g][m] = 0;
      if (in) {
        const uint32_t iw = i - x + (x == 0 ? uL - 1 : x - 1);
        const uint32_t inn = (y == 0) ? i + uL2 - uL : i - uL;
        const uint32_t iu = (KIND == KIND_3D) ? ((z == 0) ? i + (uint32_t)(N - L2) : i - uL2) : 0;
        Jw[g] = jx[i];
        Jn[g] = jy[i];
        west[g] = !(KIND == KIND_2P && x == 0);  // no open-x west bond
        if (KIND == KIND_3D) Ju[g] = jz[i];
#pragma unroll
        for (int m = 0; m < SB; ++m) {
          if (m >= ns) break;
          const int q = sys0 + m;
          const int sys = sys_list ? sys_list[q] : (sel.v[0] >= 0 ? sel.v[q & 3] : q);  // sel: at most 4 systems
          const uint64_t *s = s_all + (int64_t)sys * N;
          w[g][m] = s[i];
          sw[g][m] = s[iw];
          sn[g][m] = s[inn];
          if (KIND == KIND_3D) su[g][m] = s[iu];
        }
      }
    }
#pragma unroll
    for (int g = 0; g < U; ++g) {
      if (base0 + g * stride >= N) break;  // wave-uniform
#pragma unroll
      for (int m = 0; m < SB; ++m) {
        if (m >= ns) break;
        const uint64_t ww = w[g][m];
        const uint64_t bw = west[g] ? (Jw[g] ^ ww ^ sw[g][m]) : 0;
        const uint64_t bn = Jn[g] ^ ww ^ sn[g][m];
        const uint64_t bu = (KIND == KIND_3D) ? (Ju[g] ^ ww ^ su[g][m]) : 0;
        // satisfied bonds of this site: sum bit sb (weight 1) and carry cb (weight 2);
        // lanes outside the lattice hold zero words and count nothing
        const uint64_t sb = bw ^ bn ^ bu, cb = maj3(bw, bn, bu);
        uint64_t t = e[m][0] & sb;
        e[m][0] ^= sb;
        uint64_t u = maj3(e[m][1], cb, t);
        e[m][1] ^= cb ^ t;
#pragma unroll
        for (int b = 2; b < 6; ++b) {
          const uint64_t v = e[m][b] & u;
          e[m][b] ^= u;
          u = v;
        }
        t = ww;
#pragma unroll
        for (int b = 0; b < 5; ++b) {
          const uint64_t v = mc[m][b] & t;
          mc[m][b] ^= t;
          t = v;
        }
      }
      if (++pending == FLUSH) {
        flush();
        pending = 0;
      }
    }
  }
  if (pending) flush();
#pragma unroll
  for (int m = 0; m < SB; ++m) {
    atomicAdd(&red[m][0][lane], sat[m]);
    atomicAdd(&red[m][1][lane], up[m]);
  }
  __syncthreads();
  for (int q = threadIdx.x; q < ns * 128; q += blockDim.x)
    atomicAdd(&cnt[((int64_t)(blockIdx.x % kMeasureCopies) * nsys + sys0) * 128 + q], (&red[0][0][0])[q]);
}

static __global__ void k_fill_zero(int *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// ---------------------------------------------------------------- masked swap
// a ^= (a^b) & m, b ^= (a^b) & m (src/device.cuh:18-23)
static __global__ void __launch_bounds__(256) k_swap(uint64_t *a, uint64_t *b, uint64_t mask, int64_t N) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += stride) {
    const uint64_t va = a[i], vb = b[i];
    const uint64_t d = (va ^ vb) & mask;
    a[i] = va ^ d;
    b[i] = vb ^ d;
  }
}

}  // namespace sa
