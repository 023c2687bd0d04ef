/*
 * oracle.h -- CPU restatement of the reference's multispin +-J Metropolis path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or the timed CPU baseline).  The product path (simulated_annealing_amd) never
 * links or calls it.
 *
 * What each function restates (reference file:line, relative to the reference
 * repository root) and what pins it is listed beside each declaration.  Pins:
 *   REF   = compared against the compiled reference (oracle/_ref, built from the
 *           reference sources) and against golden fixtures made from it
 *           (the fixtures under tests/golden), see tests/test_oracle_vs_ref.py;
 *   KAT   = published known-answer values;
 *   MATH  = closed forms / invariants / brute force in tests/test_oracle_math.py.
 * No function of this file is "parity unpinned".
 *
 * Lattice conventions (same as the reference CPU classes):
 *   site index i = x + L*y (2D) or x + L*y + L*L*z (3D), one uint64 word per site,
 *   bit k of the word = lane k (64 independent disorder samples).
 *   Couplings are kept deduplicated: jx[i] is the bond between i and its -x
 *   neighbour, jy[i] towards -y, jz[i] towards -z.  The reference stores the same
 *   bonds twice (Jx[2i] = jx[i], Jx[2i+1] = jx[east(i)], src/cpu_3d.cpp:281-301).
 */
#ifndef SA_ORACLE_H
#define SA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- random engines (libstdc++ 11 semantics) -------------------------------- */

/* std::mt19937_64 (Matsumoto & Nishimura 2004).  Used by cpu_3d/cpu_2dp
 * (src/cpu_3d.hpp:35, src/cpu_2dp.hpp:35).  Pin: KAT (10000th output of the
 * default seed 5489 = 9981545732273789042, C++ standard [rand.predef]) + REF. */
typedef struct { uint64_t mt[312]; int idx; } or_mt64;
void     or_mt64_seed(or_mt64 *g, uint64_t seed);
uint64_t or_mt64_next(or_mt64 *g);
void     or_mt64_fill(or_mt64 *g, uint64_t *out, int64_t n);

/* std::mt19937 (32-bit).  Used by parallel_tempering (src/parallel_tempering.hpp:28,59).
 * Pin: KAT (10000th output of seed 5489 = 4123659995) + REF. */
typedef struct { uint32_t mt[624]; int idx; } or_mt32;
void     or_mt32_seed(or_mt32 *g, uint32_t seed);
uint32_t or_mt32_next(or_mt32 *g);

/* uniform_real_distribution<double>(0,1) over a 64-bit engine: one draw g,
 * u = double(g) / 2^64, clamped to nextafter(1,0)  (libstdc++ generate_canonical,
 * bits/random.tcc:3348-3380).  Drawn once per site update, src/cpu_3d.cpp:130.
 * Pin: MATH (exhaustive threshold neighbourhoods) + REF (sweeps). */
double or_canonical64(uint64_t g);
/* same over a 32-bit engine: two draws, u = (g1 + g2*2^32) / 2^64 in double. */
double or_canonical32(uint32_t g1, uint32_t g2);
double or_mt32_uniform(or_mt32 *g);
/* uniform_int_distribution<int>(a,b) over mt19937: Lemire's nearly divisionless
 * method as in libstdc++ 11 (bits/uniform_int_dist.h:246-316).  Pin: REF (PT pair
 * sequence) + MATH (range/unbiasedness on small ranges). */
int32_t or_mt32_uniform_int(or_mt32 *g, int32_t a, int32_t b);

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).  Pin: KAT (Random123
 * kat_vectors) + cross-check against rocRAND's host implementation. */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* The counter-mode random word of one site update (the product's counter RNG,
 * defined in include/sa/rng.hpp):
 *   3D: base = x + L*y + L*L*(z & ~1), sel = z & 1
 *   2D: base = ((x & ~3) | (x & 1)) + L*y, sel = (x >> 1) & 1
 *   r = philox(ctr = {base, lo32(t), hi32(t), 0}, key = {lo32(seed), hi32(seed)})
 *   word = sel ? (r3 << 32 | r2) : (r1 << 32 | r0)
 * t = number of sweeps done since the stream was seeded. */
uint64_t or_counter_word(int dim, int L, int x, int y, int z, uint64_t seed, uint64_t t);

/* ---- Boltzmann tables ------------------------------------------------------- */
/* 3D, 14 entries: [c] spin bit 0, [7+c] spin bit 1, c = satisfied bonds
 * (src/cpu_3d.cpp:32-43, set_T :218-251). */
void or_boltz_3d(double T, double h, double out[14]);
/* 2D, 20 entries: [c], [5+c] bulk; [10+c], [15+c] open-x border
 * (src/cpu_2dp.cpp:33-47, set_T :122-164).  The periodic cpu_2d uses [0..9]. */
void or_boltz_2d(double T, double h, double out[20]);

/* ---- systems ---------------------------------------------------------------- */
enum { OR_2P = 0, OR_2D = 1, OR_3D = 2 };   /* -d 2p / -d 2 / -d 3 */

/* One in-place sweep in the reference loop order (offset -> [z] -> y -> x step 2).
 * words[k] is the raw 64-bit random word consumed by the k-th site update of the
 * sweep (N words); the acceptance test is or_canonical64(words[k]) < boltz[...].
 *   OR_3D: src/cpu_3d.cpp:94-161      OR_2D: src/cpu_2d.cpp:5-41 (loop src/cpu_2dp.cpp:106-114)
 *   OR_2P: src/cpu_2dp.cpp:106-114,336-383 */
void or_sweep(int kind, int L, uint64_t *s, const uint64_t *jx, const uint64_t *jy,
              const uint64_t *jz, const double *boltz, const uint64_t *words);
/* visit order of the sweep: site[k] = site index of the k-th update. */
void or_visit_order(int kind, int L, int64_t *site);
/* counter-mode words of sweep t in visit order (N words). */
void or_counter_words(int kind, int L, uint64_t seed, uint64_t t, uint64_t *out);

/* Per-lane energy and magnetisation in LANE order (lane k = bit k):
 *   E[k] = sum over the -x,-y[,-z] bond of every site of (+1 satisfied, -1 not),
 *          open-x: the fictitious x=0 west bond is excluded (src/cpu_2dp.cpp:194-214)
 *   M[k] = sum over sites of (+1 bit set, -1 clear)
 * The reference returns float(E + h*M) at index k: its loop writes index 63-j from
 * bit_to_double(s, j), which reads bit 63-j (src/tools_inline.hpp:27-31), so the two
 * reversals cancel (src/cpu_3d.cpp:310-371). */
void or_measure(int kind, int L, const uint64_t *s, const uint64_t *jx, const uint64_t *jy,
                const uint64_t *jz, int32_t E[64], int32_t M[64]);
/* the reference's measure() vector: out[k] = (float)((double)E[k] + h*(double)M[k]) */
void or_measure_ref_order(int kind, int L, const uint64_t *s, const uint64_t *jx,
                          const uint64_t *jy, const uint64_t *jz, double h, float out[64]);

/* couplings: init_J draws d*N raw words (src/cpu_3d.cpp:281-301, src/cpu_2dp.cpp:243-260) */
void or_init_J(int kind, int L, or_mt64 *g, uint64_t *jx, uint64_t *jy, uint64_t *jz);
/* masked lane exchange a ^= (a^b)&m, b ^= (a^b)&m (src/tools_inline.hpp:4-9) */
void or_swap(int64_t N, uint64_t *a, uint64_t *b, uint64_t mask);

/* ---- drivers ---------------------------------------------------------------- */
/* parallel_tempering::creat_mask (src/parallel_tempering.hpp:141-162): bit j set iff
 * dE*dB > 0 || u < exp(dB*dE) with dE = -(Ei[j] - Ei1[j]) in float and
 * dB = 1/Ti - 1/Ti1.  Ei/Ei1 are measure() vectors (reference index order). */
uint64_t or_pt_mask(const float *Ei, const float *Ei1, double Ti, double Ti1, double u);

/* A complete parallel-tempering run as main_cpu builds it (src/main_cpu.cpp:267-296,
 * src/parallel_tempering.hpp:54-115): nT systems of one kind, all seeded with `seed`
 * (set_seed, init_J, set_seed, init_rand), temperatures sorted ascending, then
 * `nsteps` calls of step(0) with `sweeps` sweeps each.  rng_mode 0 = the reference's
 * mt19937_64 stream, 1 = counter words (or_counter_words, one stream shared by all
 * temperatures since they share the seed).
 * Outputs (any may be NULL): pairs[nsteps], us[nsteps], masks[nsteps],
 * E_last[nT*64] (measure() vectors after the last step, reference order),
 * spins[nT*N] final spins.  Returns 0 on success. */
int or_pt_run(int kind, int L, int nT, const double *T, double h, uint64_t seed,
              int sweeps, int nsteps, int rng_mode, int32_t *pairs, double *us,
              uint64_t *masks, float *E_last, uint64_t *spins);

/* annealing over a per-sweep temperature schedule: one system seeded like main_cpu,
 * for k in 0..nsweeps-1: set_T(Ts[k]); sweep(); if (k % stride == 0) record
 * measure() (reference order) into E_log[(k/stride)*64 ...].  rng_mode as above.
 * This is annealing<>::step with a per-sweep set_T (src/annealing.hpp:67-100).
 * Returns the number of records. */
int or_anneal_run(int kind, int L, const double *Ts, int nsweeps, double h, uint64_t seed,
                  int stride, int rng_mode, float *E_log, uint64_t *spins_out);

#ifdef __cplusplus
}
#endif
#endif
