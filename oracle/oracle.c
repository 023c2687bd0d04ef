/*
 * oracle.c -- plain C restatement of the reference's CPU Metropolis path.
 * TEST INFRASTRUCTURE ONLY (see oracle.h for the rules and the pins).
 *
 * Written from the published algorithms and from reading the reference; every
 * function cites the reference lines it follows.  Deliberately simple: scalar
 * loops, one site at a time, in the reference's visit order.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------- */
/* mt19937_64: Matsumoto & Nishimura, "Mersenne Twister 64-bit", parameters
 * w=64 n=312 m=156 r=31 a=0xB5026F5AA96619E9 u=29 d=0x5555555555555555 s=17
 * b=0x71D67FFFEDA60000 t=37 c=0xFFF7EEE000000000 l=43 f=6364136223846793005. */
void or_mt64_seed(or_mt64 *g, uint64_t seed) {
  g->mt[0] = seed;
  for (int i = 1; i < 312; ++i)
    g->mt[i] = 6364136223846793005ULL * (g->mt[i - 1] ^ (g->mt[i - 1] >> 62)) + (uint64_t)i;
  g->idx = 312;
}

static void mt64_twist(or_mt64 *g) {
  const uint64_t upper = 0xFFFFFFFF80000000ULL, lower = 0x7FFFFFFFULL;
  for (int i = 0; i < 312; ++i) {
    uint64_t y = (g->mt[i] & upper) | (g->mt[(i + 1) % 312] & lower);
    uint64_t v = g->mt[(i + 156) % 312] ^ (y >> 1);
    if (y & 1) v ^= 0xB5026F5AA96619E9ULL;
    g->mt[i] = v;
  }
  g->idx = 0;
}

uint64_t or_mt64_next(or_mt64 *g) {
  if (g->idx >= 312) mt64_twist(g);
  uint64_t y = g->mt[g->idx++];
  y ^= (y >> 29) & 0x5555555555555555ULL;
  y ^= (y << 17) & 0x71D67FFFEDA60000ULL;
  y ^= (y << 37) & 0xFFF7EEE000000000ULL;
  y ^= y >> 43;
  return y;
}

void or_mt64_fill(or_mt64 *g, uint64_t *out, int64_t n) {
  for (int64_t i = 0; i < n; ++i) out[i] = or_mt64_next(g);
}

/* mt19937: n=624 m=397 a=0x9908B0DF u=11 s=7 b=0x9D2C5680 t=15 c=0xEFC60000 l=18
 * f=1812433253. */
void or_mt32_seed(or_mt32 *g, uint32_t seed) {
  g->mt[0] = seed;
  for (int i = 1; i < 624; ++i)
    g->mt[i] = 1812433253U * (g->mt[i - 1] ^ (g->mt[i - 1] >> 30)) + (uint32_t)i;
  g->idx = 624;
}

uint32_t or_mt32_next(or_mt32 *g) {
  if (g->idx >= 624) {
    for (int i = 0; i < 624; ++i) {
      uint32_t y = (g->mt[i] & 0x80000000U) | (g->mt[(i + 1) % 624] & 0x7FFFFFFFU);
      uint32_t v = g->mt[(i + 397) % 624] ^ (y >> 1);
      if (y & 1) v ^= 0x9908B0DFU;
      g->mt[i] = v;
    }
    g->idx = 0;
  }
  uint32_t y = g->mt[g->idx++];
  y ^= y >> 11;
  y ^= (y << 7) & 0x9D2C5680U;
  y ^= (y << 15) & 0xEFC60000U;
  y ^= y >> 18;
  return y;
}

/* generate_canonical<double,53>: ceil(53 / log2(range)) draws, summed in double
 * with weights 1, r, r^2..., divided by r^k, clamped below 1
 * (libstdc++ bits/random.tcc:3348-3380). */
double or_canonical64(uint64_t g) {
  double sum = (double)g;                       /* one 64-bit draw */
  double ret = sum / 18446744073709551616.0;    /* / 2^64 */
  if (ret >= 1.0) ret = nextafter(1.0, 0.0);
  return ret;
}

double or_canonical32(uint32_t g1, uint32_t g2) {
  double sum = 0.0, tmp = 1.0;
  sum += (double)g1 * tmp;
  tmp *= 4294967296.0;
  sum += (double)g2 * tmp;
  tmp *= 4294967296.0;
  double ret = sum / tmp;
  if (ret >= 1.0) ret = nextafter(1.0, 0.0);
  return ret;
}

double or_mt32_uniform(or_mt32 *g) {
  uint32_t a = or_mt32_next(g);
  uint32_t b = or_mt32_next(g);
  return or_canonical32(a, b);
}

/* Lemire, "Fast Random Integer Generation in an Interval" (TOMACS 2019), as
 * libstdc++ 11 applies it to a 32-bit engine (uniform_int_dist.h:246-316). */
int32_t or_mt32_uniform_int(or_mt32 *g, int32_t a, int32_t b) {
  uint32_t range = (uint32_t)((uint64_t)(uint32_t)b - (uint64_t)(uint32_t)a) + 1U;
  uint64_t product = (uint64_t)or_mt32_next(g) * (uint64_t)range;
  uint32_t low = (uint32_t)product;
  if (low < range) {
    uint32_t threshold = (uint32_t)(-range) % range;
    while (low < threshold) {
      product = (uint64_t)or_mt32_next(g) * (uint64_t)range;
      low = (uint32_t)product;
    }
  }
  return (int32_t)((uint32_t)(product >> 32) + (uint32_t)a);
}

/* ------------------------------------------------------------------------- */
/* Philox4x32-10: round (c0,c1,c2,c3) -> (hi(M1*c2)^c1^k0, lo(M1*c2),
 * hi(M0*c0)^c3^k1, lo(M0*c0)), key += (W0, W1) between rounds. */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53U * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57U * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    uint32_t n1 = (uint32_t)p1;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    uint32_t n3 = (uint32_t)p0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9U;
    k1 += 0xBB67AE85U;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

uint64_t or_counter_word(int dim, int L, int x, int y, int z, uint64_t seed, uint64_t t) {
  uint32_t base;
  int sel;
  if (dim == 3) {
    base = (uint32_t)x + (uint32_t)L * (uint32_t)y + (uint32_t)L * (uint32_t)L * (uint32_t)(z & ~1);
    sel = z & 1;
  } else {
    base = (uint32_t)((x & ~3) | (x & 1)) + (uint32_t)L * (uint32_t)y;
    sel = (x >> 1) & 1;
  }
  uint32_t ctr[4] = {base, (uint32_t)t, (uint32_t)(t >> 32), 0U};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t r[4];
  or_philox4x32_10(ctr, key, r);
  return sel ? (((uint64_t)r[3] << 32) | r[2]) : (((uint64_t)r[1] << 32) | r[0]);
}

/* ------------------------------------------------------------------------- */
/* Boltzmann tables.  The reference loops over size_t i and computes
 * int j = (i-3)*2 (3D) / (i-2)*2 (2D); the unsigned wrap-around converts to the
 * negative j one expects, so j = 2i-6 / 2i-4.  The product -2*(1/T)*(j-h) is
 * evaluated left to right in double as written (src/cpu_3d.cpp:36-39). */
static double minf1(double a) { return a < 1.0 ? a : 1.0; }

void or_boltz_3d(double T, double h, double out[14]) {
  for (int i = 0; i < 7; ++i) {
    int j = 2 * i - 6;
    out[i] = minf1(exp(-2 * (1 / T) * (j - h)));
    out[7 + i] = minf1(exp(-2 * (1 / T) * (j + h)));
  }
}

void or_boltz_2d(double T, double h, double out[20]) {
  for (int i = 0; i < 5; ++i) {
    int j = 2 * i - 4;
    out[i] = minf1(exp(-2 * (1 / T) * (j - h)));
    out[5 + i] = minf1(exp(-2 * (1 / T) * (j + h)));
    out[10 + i] = minf1((i == 0) ? 0. : exp(-2 * (1 / T) * (j - 1 - h)));
    out[15 + i] = minf1((i == 0) ? 0. : exp(-2 * (1 / T) * (j - 1 + h)));
  }
}

/* ------------------------------------------------------------------------- */
/* per-lane count of set bits among n words -> n+1 "exactly c" masks p[0..n] */
static void exact_count_masks(const uint64_t *b, int n, uint64_t *p) {
  for (int c = 0; c <= n; ++c) p[c] = 0;
  for (int lane = 0; lane < 64; ++lane) {
    int c = 0;
    for (int k = 0; k < n; ++k) c += (int)((b[k] >> lane) & 1U);
    p[c] |= (uint64_t)1 << lane;
  }
}

/* One site update.  sat[k] = J ^ s_i ^ s_neighbour (bit set = satisfied bond);
 * lane flips iff u < boltz[c] (spin bit 0) or u < boltz[z+1+c] (spin bit 1), c =
 * number of satisfied bonds of the lane (src/cpu_3d.cpp:107-161: the p0..p6 masks
 * are "exactly c of the six bits set"; d_c selects the table by the spin bit). */
static uint64_t site_flip(const uint64_t *sat, int nb, uint64_t si, double u, const double *tab) {
  uint64_t p[7];
  exact_count_masks(sat, nb, p);
  uint64_t flip = 0;
  for (int c = 0; c <= nb; ++c) {
    uint64_t d0 = (u < tab[c]) ? ~(uint64_t)0 : 0;
    uint64_t d1 = (u < tab[nb + 1 + c]) ? ~(uint64_t)0 : 0;
    uint64_t d = (d0 & ~si) | (d1 & si);
    flip |= p[c] & d;
  }
  return flip;
}

static void update_3d(int L, uint64_t *s, const uint64_t *jx, const uint64_t *jy,
                      const uint64_t *jz, const double *boltz, int x, int y, int z, uint64_t w) {
  const int64_t L2 = (int64_t)L * L;
  int64_t i = x + (int64_t)L * y + L2 * z;
  int xm = (x + L - 1) % L, xp = (x + 1) % L, ym = (y + L - 1) % L, yp = (y + 1) % L;
  int zm = (z + L - 1) % L, zp = (z + 1) % L;
  int64_t iw = xm + (int64_t)L * y + L2 * z, ie = xp + (int64_t)L * y + L2 * z;
  int64_t in = x + (int64_t)L * ym + L2 * z, is = x + (int64_t)L * yp + L2 * z;
  int64_t iu = x + (int64_t)L * y + L2 * zm, id = x + (int64_t)L * y + L2 * zp;
  uint64_t si = s[i];
  uint64_t sat[6];
  sat[0] = jx[i] ^ si ^ s[iw];
  sat[1] = jx[ie] ^ si ^ s[ie];
  sat[2] = jy[i] ^ si ^ s[in];
  sat[3] = jy[is] ^ si ^ s[is];
  sat[4] = jz[i] ^ si ^ s[iu];
  sat[5] = jz[id] ^ si ^ s[id];
  s[i] = si ^ site_flip(sat, 6, si, or_canonical64(w), boltz);
}

static void update_2d(int kind, int L, uint64_t *s, const uint64_t *jx, const uint64_t *jy,
                      const double *boltz, int x, int y, uint64_t w) {
  int64_t i = x + (int64_t)L * y;
  int xm = (x + L - 1) % L, xp = (x + 1) % L, ym = (y + L - 1) % L, yp = (y + 1) % L;
  int64_t iw = xm + (int64_t)L * y, ie = xp + (int64_t)L * y;
  int64_t in = x + (int64_t)L * ym, is = x + (int64_t)L * yp;
  uint64_t si = s[i];
  uint64_t sat[4];
  const double *tab = boltz;
  sat[0] = jx[i] ^ si ^ s[iw];
  sat[1] = jx[ie] ^ si ^ s[ie];
  sat[2] = jy[i] ^ si ^ s[in];
  sat[3] = jy[is] ^ si ^ s[is];
  if (kind == OR_2P) {
    /* open x boundary: the missing bond is counted as satisfied and the border
     * table is used (src/cpu_2dp.cpp:338-362) */
    if (x == 0) sat[0] = ~(uint64_t)0;
    if (x == L - 1) sat[1] = ~(uint64_t)0;
    if (x == 0 || x == L - 1) tab = boltz + 10;
  }
  s[i] = si ^ site_flip(sat, 4, si, or_canonical64(w), tab);
}

void or_visit_order(int kind, int L, int64_t *site) {
  int64_t k = 0;
  int nz = (kind == OR_3D) ? L : 1;
  for (int offset = 0; offset < 2; ++offset)
    for (int z = 0; z < nz; ++z)
      for (int y = 0; y < L; ++y)
        for (int x = (y + offset) % 2; x < L; x += 2)
          site[k++] = x + (int64_t)L * y + (int64_t)L * L * z;
}

void or_sweep(int kind, int L, uint64_t *s, const uint64_t *jx, const uint64_t *jy,
              const uint64_t *jz, const double *boltz, const uint64_t *words) {
  int64_t k = 0;
  if (kind == OR_3D) {
    /* src/cpu_3d.cpp:94-105 */
    for (int offset = 0; offset < 2; ++offset)
      for (int z = 0; z < L; ++z)
        for (int y = 0; y < L; ++y)
          for (int x = (y + offset) % 2; x < L; x += 2)
            update_3d(L, s, jx, jy, jz, boltz, x, y, z, words[k++]);
  } else {
    /* src/cpu_2dp.cpp:106-114 */
    for (int offset = 0; offset < 2; ++offset)
      for (int y = 0; y < L; ++y)
        for (int x = (y + offset) % 2; x < L; x += 2)
          update_2d(kind, L, s, jx, jy, boltz, x, y, words[k++]);
  }
}

void or_counter_words(int kind, int L, uint64_t seed, uint64_t t, uint64_t *out) {
  int64_t k = 0;
  int nz = (kind == OR_3D) ? L : 1;
  int dim = (kind == OR_3D) ? 3 : 2;
  for (int offset = 0; offset < 2; ++offset)
    for (int z = 0; z < nz; ++z)
      for (int y = 0; y < L; ++y)
        for (int x = (y + offset) % 2; x < L; x += 2)
          out[k++] = or_counter_word(dim, L, x, y, z, seed, t);
}

/* ------------------------------------------------------------------------- */
void or_measure(int kind, int L, const uint64_t *s, const uint64_t *jx, const uint64_t *jy,
                const uint64_t *jz, int32_t E[64], int32_t M[64]) {
  int nz = (kind == OR_3D) ? L : 1;
  for (int k = 0; k < 64; ++k) E[k] = M[k] = 0;
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < L; ++y)
      for (int x = 0; x < L; ++x) {
        int64_t i = x + (int64_t)L * y + (int64_t)L * L * z;
        int xm = (x + L - 1) % L, ym = (y + L - 1) % L, zm = (z + L - 1) % L;
        uint64_t si = s[i];
        uint64_t bw = jx[i] ^ si ^ s[xm + (int64_t)L * y + (int64_t)L * L * z];
        uint64_t bn = jy[i] ^ si ^ s[x + (int64_t)L * ym + (int64_t)L * L * z];
        uint64_t bu = (kind == OR_3D) ? (jz[i] ^ si ^ s[x + (int64_t)L * y + (int64_t)L * L * zm]) : 0;
        for (int k = 0; k < 64; ++k) {
          int w = (int)((bw >> k) & 1U), n = (int)((bn >> k) & 1U), u = (int)((bu >> k) & 1U);
          if (!(kind == OR_2P && x == 0)) E[k] += 2 * w - 1;
          E[k] += 2 * n - 1;
          if (kind == OR_3D) E[k] += 2 * u - 1;
          M[k] += 2 * (int)((si >> k) & 1U) - 1;
        }
      }
}

void or_measure_ref_order(int kind, int L, const uint64_t *s, const uint64_t *jx,
                          const uint64_t *jy, const uint64_t *jz, double h, float out[64]) {
  int32_t E[64], M[64];
  or_measure(kind, L, s, jx, jy, jz, E, M);
  for (int k = 0; k < 64; ++k) out[k] = (float)((double)E[k] + h * (double)M[k]);
}

void or_init_J(int kind, int L, or_mt64 *g, uint64_t *jx, uint64_t *jy, uint64_t *jz) {
  int64_t N = (int64_t)L * L * ((kind == OR_3D) ? L : 1);
  for (int64_t i = 0; i < N; ++i) {
    jx[i] = or_mt64_next(g);
    jy[i] = or_mt64_next(g);
    if (kind == OR_3D) jz[i] = or_mt64_next(g);
  }
}

void or_swap(int64_t N, uint64_t *a, uint64_t *b, uint64_t mask) {
  for (int64_t i = 0; i < N; ++i) {
    uint64_t d = (a[i] ^ b[i]) & mask;
    a[i] ^= d;
    b[i] ^= d;
  }
}

/* ------------------------------------------------------------------------- */
uint64_t or_pt_mask(const float *Ei, const float *Ei1, double Ti, double Ti1, double u) {
  uint64_t mask = 0;
  double dB = (1. / Ti - 1. / Ti1);
  for (int j = 0; j < 64; ++j) {
    float dE = -(Ei[j] - Ei1[j]);
    if (((dE * dB) > 0.) || u < exp(dB * dE)) mask |= (uint64_t)1 << j;
  }
  return mask;
}

static int cmp_double(const void *a, const void *b) {
  double x = *(const double *)a, y = *(const double *)b;
  return (x > y) - (x < y);
}

typedef struct {
  int64_t N;
  uint64_t *s;
  double boltz[20];
  or_mt64 gen;
  uint64_t t;       /* sweeps done since seeding (counter mode) */
} or_member;

static void member_boltz(int kind, double T, double h, double *b) {
  if (kind == OR_3D) or_boltz_3d(T, h, b); else or_boltz_2d(T, h, b);
}

static void member_sweep(int kind, int L, or_member *m, const uint64_t *jx, const uint64_t *jy,
                         const uint64_t *jz, int rng_mode, uint64_t seed, uint64_t *words) {
  if (rng_mode == 0) or_mt64_fill(&m->gen, words, m->N);
  else or_counter_words(kind, L, seed, m->t, words);
  or_sweep(kind, L, m->s, jx, jy, jz, m->boltz, words);
  m->t++;
}

int or_pt_run(int kind, int L, int nT, const double *Tin, double h, uint64_t seed,
              int sweeps, int nsteps, int rng_mode, int32_t *pairs, double *us,
              uint64_t *masks, float *E_last, uint64_t *spins) {
  if (nT < 2) return -1;
  int64_t N = (int64_t)L * L * ((kind == OR_3D) ? L : 1);
  uint64_t *jx = malloc(N * 8), *jy = malloc(N * 8), *jz = malloc(N * 8);
  or_member *m = calloc((size_t)nT, sizeof(or_member));
  float *E = calloc((size_t)nT * 64, sizeof(float));
  double *T = malloc(nT * sizeof(double));
  memcpy(T, Tin, nT * sizeof(double));
  /* every system: set_seed, init_J, set_seed, init_rand (src/main_cpu.cpp:284-295);
   * all of them end up with the same couplings, spins and stream */
  for (int j = 0; j < nT; ++j) {
    m[j].N = N;
    m[j].s = malloc(N * 8);
    or_mt64_seed(&m[j].gen, seed);
    or_init_J(kind, L, &m[j].gen, jx, jy, jz);
    or_mt64_seed(&m[j].gen, seed);
    or_mt64_fill(&m[j].gen, m[j].s, N);
    m[j].t = 0;
  }
  /* constructor: sort T, sys[i]->set_T(T[i]) (src/parallel_tempering.hpp:54-79) */
  qsort(T, nT, sizeof(double), cmp_double);
  for (int j = 0; j < nT; ++j) member_boltz(kind, T[j], h, m[j].boltz);
  or_mt32 pg;
  or_mt32_seed(&pg, 123);
  for (int step = 0; step < nsteps; ++step) {
    /* step(0): sweeps, measure on the last one (src/parallel_tempering.hpp:81-115) */
    for (int i = 0; i < sweeps; ++i) {
      /* the members are independent until the exchange: one OpenMP thread per
       * member, as the reference's step does (src/parallel_tempering.hpp:86) */
#pragma omp parallel
      {
        uint64_t *w = malloc(N * 8);
#pragma omp for schedule(dynamic)
        for (int j = 0; j < nT; ++j) {
          member_sweep(kind, L, &m[j], jx, jy, jz, rng_mode, seed, w);
          if (i == sweeps - 1) or_measure_ref_order(kind, L, m[j].s, jx, jy, jz, h, &E[j * 64]);
        }
        free(w);
      }
    }
    int i = or_mt32_uniform_int(&pg, 0, nT - 2);
    double u = or_mt32_uniform(&pg);
    uint64_t mask = or_pt_mask(&E[i * 64], &E[(i + 1) * 64], T[i], T[i + 1], u);
    or_swap(N, m[i].s, m[i + 1].s, mask);
    if (pairs) pairs[step] = i;
    if (us) us[step] = u;
    if (masks) masks[step] = mask;
  }
  if (E_last) memcpy(E_last, E, (size_t)nT * 64 * sizeof(float));
  if (spins)
    for (int j = 0; j < nT; ++j) memcpy(spins + (int64_t)j * N, m[j].s, N * 8);
  for (int j = 0; j < nT; ++j) free(m[j].s);
  free(m); free(E); free(T); free(jx); free(jy); free(jz);
  return 0;
}

int or_anneal_run(int kind, int L, const double *Ts, int nsweeps, double h, uint64_t seed,
                  int stride, int rng_mode, float *E_log, uint64_t *spins_out) {
  int64_t N = (int64_t)L * L * ((kind == OR_3D) ? L : 1);
  uint64_t *jx = malloc(N * 8), *jy = malloc(N * 8), *jz = malloc(N * 8), *words = malloc(N * 8);
  or_member m;
  m.N = N;
  m.s = malloc(N * 8);
  or_mt64_seed(&m.gen, seed);
  or_init_J(kind, L, &m.gen, jx, jy, jz);
  or_mt64_seed(&m.gen, seed);
  or_mt64_fill(&m.gen, m.s, N);
  m.t = 0;
  int rec = 0;
  for (int k = 0; k < nsweeps; ++k) {
    member_boltz(kind, Ts[k], h, m.boltz);
    member_sweep(kind, L, &m, jx, jy, jz, rng_mode, seed, words);
    if (stride > 0 && k % stride == 0) {
      if (E_log) or_measure_ref_order(kind, L, m.s, jx, jy, jz, h, &E_log[rec * 64]);
      rec++;
    }
  }
  if (spins_out) memcpy(spins_out, m.s, N * 8);
  free(m.s); free(jx); free(jy); free(jz); free(words);
  return rec;
}
