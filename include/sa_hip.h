/*
 * sa_hip.h -- C ABI of the MI355X (gfx950) multispin +-J Metropolis engine.
 *
 * A "ladder" is a set of `nsys` systems (replica lattices) of one geometry that
 * live in HBM and share the couplings J and the sweep random stream -- exactly
 * the temperatures of the reference's parallel tempering, which main_cpu/main_gpu
 * seed identically (src/main_cpu.cpp:267-296), or a single annealed system
 * (nsys = 1).  Every system holds one uint64 word per site; bit k of a word is
 * lane k, one of 64 independent disorder samples (spin_t, src/gpu_sys.cuh:10).
 *
 * Results are bit-exact against the reference CPU classes (cpu_2dp, cpu_2d,
 * cpu_3d): same seeds and random words give the same spin words, energies and
 * parallel-tempering swap decisions.  Each entry point names the reference
 * interface it replaces (file:line relative to the reference repository).
 *
 * Conventions: every function returns SA_OK (0) or a negative sa_status; a
 * failing call leaves the ladder unchanged and sa_last_error() describes the
 * cause.  Host buffers are caller-allocated.  Calls on different ladders may come
 * from different threads (each ladder owns a HIP stream and no device-global
 * mutable state exists); calls on one ladder must not overlap.
 */
#ifndef SA_HIP_H
#define SA_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sa_ladder sa_ladder;

enum sa_status {
  SA_OK = 0,
  SA_ERR_ARG = -1,          /* invalid argument */
  SA_ERR_UNSUPPORTED = -2,  /* geometry not supported (L < 3) */
  SA_ERR_HIP = -3,          /* HIP runtime error */
  SA_ERR_NOMEM = -4,        /* device or pinned allocation failed */
  SA_ERR_STATE = -5         /* call not valid in the current state */
};

/* lattice kinds, the reference's -d flag (src/main_cpu.cpp:110-112) */
enum sa_kind {
  SA_KIND_2P = 0, /* 2D, open x / periodic y: cpu_2dp, gpu_2d (src/cpu_2dp.cpp:336-383) */
  SA_KIND_2D = 1, /* 2D fully periodic: cpu_2d (src/cpu_2d.cpp:5-41) */
  SA_KIND_3D = 2  /* 3D periodic: cpu_3d, gpu_3d (src/cpu_3d.cpp:107-161) */
};

/* source of the per-site random words of a sweep */
enum sa_rng_mode {
  SA_RNG_REPLAY = 0,  /* the reference's own stream: raw mt19937_64 words drawn on the
                         host in the reference's visit order (src/cpu_3d.cpp:130) */
  SA_RNG_COUNTER = 1  /* Philox4x32-10 keyed by (seed, sweep, site pair), computed
                         inline by the sweep kernels */
};

const char *sa_last_error(void);
const char *sa_status_string(int status);
int sa_version(void);

/* Create a ladder of nsys systems (device memory for spins, couplings, tables).
 * Replaces the gpu_2d/gpu_3d constructors (src/gpu_3d.cu:20-77, src/gpu_2d.cu:20-84).
 * L >= 3 (SA_ERR_UNSUPPORTED otherwise); odd L sweeps each wrap group (a site of the
 * corner block [0, L-1)^2 with its same-colour partners in the last column / row) in
 * the reference's visit order. */
int sa_ladder_create(int kind, int L, int nsys, sa_ladder **out);
/* Release everything (gpu_free, src/gpu_3d.cu:79-98).  NULL is accepted. */
int sa_ladder_destroy(sa_ladder *l);
/* sites per system: L*L or L*L*L (get_N, src/gpu_3d.cu:301-303) */
int64_t sa_ladder_sites(const sa_ladder *l);
int sa_ladder_nsys(const sa_ladder *l);

/* Random streams. */
int sa_ladder_set_rng(sa_ladder *l, int mode);
/* Re-seed: the host mt19937_64 restarts from `seed` and the counter stream restarts
 * at sweep 0 (set_seed, src/gpu_sys.cuh:18, CPU src/cpu_3d.cpp:170). */
int sa_ladder_seed(sa_ladder *l, uint64_t seed);
/* the seed of the last sa_ladder_seed (the counter stream's key) */
int sa_ladder_get_seed(const sa_ladder *l, uint64_t *seed);
/* sweeps done since the last seed (the counter-mode sweep index) */
int sa_ladder_get_sweep_index(const sa_ladder *l, uint64_t *t);
int sa_ladder_set_sweep_index(sa_ladder *l, uint64_t t);
/* the host engine state: 312 words + position, libstdc++ mt19937_64 layout */
int sa_ladder_get_engine(const sa_ladder *l, uint64_t state[313]);
int sa_ladder_set_engine(sa_ladder *l, const uint64_t state[313]);

/* Couplings drawn from the host engine exactly as cpu_3d::init_J / cpu_2dp::init_J
 * (d*N raw words, src/cpu_3d.cpp:281-301, src/cpu_2dp.cpp:243-260); replaces
 * gpu_3d::init_J (src/gpu_3d.cu:251-271). */
int sa_ladder_init_couplings(sa_ladder *l);
/* Spins of every system = the next N raw words (cpu_3d::init_rand,
 * src/cpu_3d.cpp:257-263); replaces gpu_3d::init_rand (src/gpu_3d.cu:273-279). */
int sa_ladder_init_spins(sa_ladder *l);
/* Deduplicated couplings (N words per axis; jz ignored in 2D).  load_J / save_J
 * (src/gpu_3d.cu:212-249) move the same bonds in the duplicated file layout. */
int sa_ladder_set_couplings(sa_ladder *l, const uint64_t *jx, const uint64_t *jy, const uint64_t *jz);
int sa_ladder_get_couplings(const sa_ladder *l, uint64_t *jx, uint64_t *jy, uint64_t *jz);
/* The reference's J file (src/sys_file.cpp:10-110: int d; long N; Jx[2N]; Jy[2N];
 * [Jz[2N]], every bond stored twice, [2i] toward -axis and [2i+1] toward +axis):
 * load_J / save_J of gpu_3d, gpu_2d, cpu_3d, cpu_2dp (src/gpu_3d.cu:212-249,
 * src/cpu_3d.cpp:379-399, src/cpu_2dp.cpp:268-288).  Loading checks the dimension
 * and N (the reference's errors -2, -3) and that both copies of every bond agree; on
 * any error the couplings stay unchanged and SA_ERR_ARG is returned. */
int sa_ladder_load_J(sa_ladder *l, const char *path);
int sa_ladder_save_J(const sa_ladder *l, const char *path);
/* spins of system `sys` (N words) */
int sa_ladder_set_spins(sa_ladder *l, int sys, const uint64_t *s);
int sa_ladder_get_spins(const sa_ladder *l, int sys, uint64_t *s);

/* Temperature and field of one system; computes the reference's Boltzmann table in
 * double (set_T / set_h, src/gpu_3d.cu:117-166, src/cpu_3d.cpp:218-251) and turns
 * it into integer acceptance thresholds.  `boltz_out` (may be NULL) receives the
 * table: 14 doubles (3D) or 20 (2D). */
int sa_ladder_set_temperature(sa_ladder *l, int sys, double T, double h, double *boltz_out);

/* `nsweeps` sweeps of every system (sweep(), src/gpu_3d.cu:100-114 and
 * src/cpu_3d.cpp:94-105): per sweep, the offset-0 then the offset-1 sites in the
 * reference CPU order.  Asynchronous on the ladder's stream in counter mode; in
 * replay mode the host draws the words and the call returns when they are
 * consumed. */
int sa_ladder_sweep(sa_ladder *l, int nsweeps);

/* An anneal on the ladder: before sweep t, system j gets temperature T[t*nsys+j]
 * and field h, then one sweep (the loop of annealing::run with set_T per sweep,
 * src/annealing.hpp:67-100, without its printing).  Asynchronous in counter mode. */
int sa_ladder_anneal(sa_ladder *l, const double *T, int nsweeps, double h);

/* Per system and lane k (bit k): E[sys*64+k] = sum over the -x,-y[,-z] bond of
 * every site of +1 (satisfied) / -1, M[sys*64+k] = sum of +1 (set) / -1 (clear).
 * The reference's measure() returns float(E + h*M) at index k
 * (src/cpu_3d.cpp:310-371; GPU src/gpu_3d.cu:168-203).  Synchronous. */
int sa_ladder_measure(sa_ladder *l, int32_t *E, int32_t *M);
/* The same for the n systems listed in sys[] (E[j*64+k] for the j-th listed one):
 * the replica-exchange step needs only the exchanging pair and the coldest system
 * (src/parallel_tempering.hpp:91-114). */
int sa_ladder_measure_systems(sa_ladder *l, const int *sys, int n, int32_t *E, int32_t *M);

/* sa_ladder_measure without any allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device or host memory (code under HIP graph
 * capture, serving loops and harnesses that forbid hipMalloc).  The caller passes
 * device arrays: s (nsys*N words, system-major), jx/jy/jz (N words each; jz unused
 * in 2D) and receives E[sys*64+k], M[sys*64+k] (nsys*64 int32 each, device memory)
 * with the values sa_ladder_measure returns.  Asynchronous on `stream` (a
 * hipStream_t; NULL = the null stream).  The partial counters live in one fixed
 * device buffer per GPU, sized for at most SA_NOALLOC_MAX_SYSTEMS systems (256 KiB,
 * held until the process ends); larger nsys returns SA_ERR_ARG.  Every call on a GPU
 * uses that GPU's buffer, so calls on one GPU that overlap in time race. */
#define SA_NOALLOC_MAX_SYSTEMS 64
int sa_measure_noalloc(int kind, int L, int nsys, const uint64_t *s, const uint64_t *jx, const uint64_t *jy,
                       const uint64_t *jz, int32_t *E, int32_t *M, void *stream);

/* Exchange the lanes selected by `mask` between systems a and b
 * (swap, src/gpu_3d.cu:310-328, src/cpu_3d.cpp:429-444). */
int sa_ladder_swap(sa_ladder *l, int a, int b, uint64_t mask);

/* Kernel tuning (no effect on results; 0 = automatic everywhere).
 *   SA_TUNE_SYS_PER_THREAD  systems one thread updates per site (1, 2 or 4, must
 *                           divide nsys): more share coupling loads and random words
 *   SA_TUNE_SWEEP3D_KERNEL  1 = two half-sweep launches, 2 = fused single pass,
 *                           3 = two z-parallel half-sweep launches (even L; column
 *                           walks split into segments joined by a scan)
 *   SA_TUNE_ROWS_PER_TILE   rows of the band a fused workgroup owns (4, 8, 16)
 *   SA_TUNE_XCD_MAP         fused sweep: 1 = adjacent bands of every system on one
 *                           XCD, 2 = whole systems on one XCD, 0 = neither (-1 automatic) */
enum sa_tune_key {
  SA_TUNE_SYS_PER_THREAD = 1,
  SA_TUNE_SWEEP3D_KERNEL = 2,
  SA_TUNE_ROWS_PER_TILE = 3,
  SA_TUNE_XCD_MAP = 4
};
int sa_ladder_set_tuning(sa_ladder *l, int key, int value);
int sa_ladder_get_tuning(const sa_ladder *l, int key, int *value);

/* wait for the ladder's stream */
int sa_ladder_sync(sa_ladder *l);
/* the HIP stream (hipStream_t) the ladder launches on, for event timing */
void *sa_ladder_stream(sa_ladder *l);

/* ---- drivers (C++ templates of include/sa, exposed for non-C++ callers) ------- */

/* Parallel tempering as main_gpu runs it (src/main_gpu.cu:262-416,
 * src/parallel_tempering.hpp:54-115): nT systems seeded with `seed`
 * (set_seed, init_J, set_seed, init_rand), temperatures sorted ascending,
 * `nsteps` x step(0) with `sweeps` sweeps per step, swap decisions on the host from
 * mt19937(123).  One ladder holds all temperatures.  Outputs may be NULL:
 * pairs[nsteps], us[nsteps], masks[nsteps], E_last[nT*64] (measure() vectors after
 * the last step), spins[nT*N]. */
int sa_pt_run(int kind, int L, int nT, const double *T, double h, uint64_t seed, int sweeps, int nsteps,
              int rng_mode, int32_t *pairs, double *us, uint64_t *masks, float *E_last, uint64_t *spins);

/* Annealing over a per-sweep temperature schedule (annealing<gpu_sys>,
 * src/annealing.hpp:67-100, with set_T before every sweep): records measure()
 * after sweep k when k % stride == 0.  Returns the number of records (>= 0) or an
 * sa_status. */
int sa_anneal_run(int kind, int L, const double *Ts, int nsweeps, double h, uint64_t seed, int stride,
                  int rng_mode, float *E_log, uint64_t *spins_out);
/* The annealing driver with geometric cooling as main runs it
 * (annealing<system>(T0, sweeps, alfa, sys, out); nsteps x step(stride, PRINT_E),
 * src/annealing.hpp:42-100, src/main_gpu.cu:279-300): each step is `sweeps` sweeps
 * at the current T, a line of 64 energies every `stride` sweeps, then T *= alfa.
 * The system is seeded like main.  Returns the text's length and copies it
 * (NUL-terminated, truncated to capacity) when buf is not NULL; `precision` is the
 * stream precision (0: the stream default, as main's output file). */
int64_t sa_anneal_geometric(int kind, int L, double T0, double alfa, int sweeps, int nsteps, int stride, double h,
                            uint64_t seed, int rng_mode, int precision, char *buf, int64_t capacity);

/* A parallel-tempering run kept alive between calls (what sa_pt_run does in one
 * call), for callers that time or interleave steps: create seeds the systems,
 * sa_pt_step runs nsteps x step(0), sa_pt_result copies the swap log of the
 * steps run so far (pairs/us/masks, each capacity entries) and the current
 * energies and spins. */
typedef struct sa_pt sa_pt;
int sa_pt_create(int kind, int L, int nT, const double *T, double h, uint64_t seed, int sweeps, int rng_mode,
                 sa_pt **out);
int sa_pt_step(sa_pt *pt, int nsteps);
int sa_pt_sync(sa_pt *pt);
int64_t sa_pt_steps_done(const sa_pt *pt);
int sa_pt_result(sa_pt *pt, int64_t capacity, int32_t *pairs, double *us, uint64_t *masks, float *E_last,
                 uint64_t *spins);
sa_ladder *sa_pt_ladder(sa_pt *pt);
/* The driver's text output (parallel_tempering::print / print_swap,
 * src/parallel_tempering.hpp:117-139,164-170,203-262): energy files named by
 * `pattern` with % replaced by T<temperature> (E_min for the minimum file).
 * sa_pt_step_stride runs step(stride): stride > 0 measures every system and
 * appends a line of 64 energies to each file every `stride` sweeps (stride 0 is
 * sa_pt_step).  sa_pt_print_swap returns the acceptance report's length and copies
 * it (NUL-terminated, truncated to capacity) when buf is not NULL. */
int sa_pt_set_output(sa_pt *pt, const char *pattern);
int sa_pt_step_stride(sa_pt *pt, int nsteps, int stride);
int sa_pt_print_min(sa_pt *pt);
int64_t sa_pt_print_swap(sa_pt *pt, char *buf, int64_t capacity);
int sa_pt_close_output(sa_pt *pt);
/* Checkpoint (.tempsys) of the run: parallel_tempering::save / load
 * (src/parallel_tempering.hpp:280-377; main_cpu's -s / --reload files).  With the
 * replay RNG the file is the reference's byte for byte and a reference checkpoint
 * loads and continues identically; with the counter RNG each system record carries
 * a trailer with the counter state.  Loading clears the swap log. */
int sa_pt_save(sa_pt *pt, const char *path);
int sa_pt_load(sa_pt *pt, const char *path);
/* The pieces of main_cpu's ground-state loop (-r, src/main_cpu.cpp:344-434):
 * sa_pt_update_temperatures = parallel_tempering::update_T
 * (src/parallel_tempering.hpp:384-413): a temperature (T_i + T_i+1)/2 is inserted
 * between neighbours whose exchange rate is below 0.3, its system a copy of the
 * hotter neighbour on a ladder of its own; returns the new temperature count.
 * sa_pt_reset = parallel_tempering::reset (src/parallel_tempering.hpp:265-277):
 * E_min cleared, new spins drawn (init_rand), rates kept (the reference's
 * accumulator::reset does nothing).  sa_pt_temperatures / sa_pt_acceptance copy
 * up to `capacity` temperatures / exchange rates (num_switch, get()) and return
 * the count; sa_pt_E_min = get_E_min (64 values, 0 before the first step) and
 * returns the count.  sa_pt_result's E_last and spins cover every current system
 * in temperature order. */
int sa_pt_update_temperatures(sa_pt *pt);
int sa_pt_reset(sa_pt *pt);
int sa_pt_temperatures(sa_pt *pt, int capacity, double *T);
int sa_pt_acceptance(sa_pt *pt, int capacity, double *rates);
int sa_pt_E_min(sa_pt *pt, float *E_min);
int sa_pt_destroy(sa_pt *pt);

/* ---- batched instances (no counterpart in the reference) --------------------
 * Many independent instances of uneven size and kind, each a ladder seeded like
 * main_gpu (set_seed, init_J, set_seed, init_rand; src/main_gpu.cu:262-276) with
 * its own temperatures, swept together: one launch per half-sweep covers every
 * unfinished instance, chunks of sweeps are replayed as captured hipGraphs, and
 * block size / chunk length are tuned online.  Counter RNG only.  An instance's
 * spins after k sweeps are bit for bit those of its ladder swept k times alone
 * (sa_ladder_set_temperature + sa_ladder_sweep per sweep). */
typedef struct sa_batch sa_batch;
int sa_batch_create(sa_batch **out);
int sa_batch_destroy(sa_batch *b);
/* Add an instance of nsys systems.  per_sweep = 0: T[nsys] fixed temperatures;
 * per_sweep = 1: T[nsweeps*nsys], row t = the temperatures of sweep t (an anneal
 * schedule, annealing::set_T before each sweep, src/annealing.hpp:67-100).
 * *id receives the instance id. */
int sa_batch_add(sa_batch *b, int kind, int L, int nsys, uint64_t seed, int nsweeps, const double *T, int per_sweep,
                 double h, int *id);
/* Add a parallel-tempering instance: nT systems at the temperatures T (sorted
 * ascending, as parallel_tempering does), field h, `nsteps` steps of `every` sweeps
 * each followed by one exchange attempt between a random adjacent pair drawn from
 * the instance's own mt19937(123) (parallel_tempering::step(0),
 * src/parallel_tempering.hpp:81-115; same pair sequence, masks and spins as
 * sa_pt_run with the counter RNG). */
int sa_batch_add_pt(sa_batch *b, int kind, int L, int nT, const double *T, double h, uint64_t seed, int every,
                    int nsteps, int *id);
/* Run up to `sweeps` sweeps of every unfinished instance (sweeps <= 0: until all
 * are done); chunks stop at every exchange point.  Returns the number of
 * unfinished instances (>= 0) or an sa_status. */
int sa_batch_run(sa_batch *b, int sweeps);
int sa_batch_sweeps_done(sa_batch *b, int id, int *t);
int sa_batch_get_spins(sa_batch *b, int id, int sys, uint64_t *s);
int sa_batch_measure(sa_batch *b, int id, int32_t *E, int32_t *M);
/* A parallel-tempering instance's exchange log so far: pairs/us/masks (capacity
 * entries each, any may be NULL), E_min[64] (per lane the largest measure() value of
 * the coldest system over the steps, parallel_tempering::min_E_update) and
 * acceptance[nT-1] (mean accepted lane fraction per pair).  Returns the number of
 * steps logged (>= 0) or an sa_status. */
int sa_batch_pt_result(sa_batch *b, int id, int64_t capacity, int32_t *pairs, double *us, uint64_t *masks,
                       float *E_min, double *acceptance);
/* device bytes held by the batch's live instances */
int64_t sa_batch_bytes(const sa_batch *b);
/* the instance's ladder (owned by the batch; valid until sa_batch_remove) */
sa_ladder *sa_batch_ladder(sa_batch *b, int id);
/* free the instance's device memory; its id is not reused */
int sa_batch_remove(sa_batch *b, int id);
/* the stream the batch launches on (hipStream_t) */
void *sa_batch_stream(sa_batch *b);
/* tuning (no effect on results): threads per block 64/128/256 of the half-sweep
 * kernel, sweeps per captured graph, and whether 3D instances with symmetric tables
 * and L divisible by 4 take the fused single-pass sweep (1 off, 2 on); 0 = picked
 * online by timing the first chunks.  The fused form is switched off in this build:
 * SA_TUNE_BATCH_FUSED = 2 returns SA_ERR_ARG. */
enum sa_batch_tune_key { SA_TUNE_BATCH_BLOCK = 101, SA_TUNE_BATCH_CHUNK = 102, SA_TUNE_BATCH_FUSED = 103 };
int sa_batch_set_tuning(sa_batch *b, int key, int value);
int sa_batch_get_tuning(sa_batch *b, int key, int *value);

/* ---- serving loop (no counterpart in the reference) --------------------------
 * Continuous batching of anneal / parallel-tempering requests over one sa_batch:
 * requests arrive on the server clock, are admitted in arrival order as device
 * memory (hbm_budget) and slots (max_active) free up, join the next chunk of
 * sweeps, and are retired with a digest of their final spins and their energies.
 * Each request's results are those of running it alone (sa_ladder_anneal on a
 * ladder seeded like main, or sa_pt_run with the counter RNG). */
enum sa_request_mode { SA_REQ_ANNEAL = 0, SA_REQ_PT = 1 };
typedef struct sa_request {
  int kind, L;
  int mode;      /* SA_REQ_ANNEAL or SA_REQ_PT */
  int nT;        /* PT: temperatures (>= 2); ignored for an anneal */
  int nsweeps;   /* total sweeps; PT: a multiple of `every` */
  int every;     /* PT: sweeps between exchange attempts */
  double T0, T1; /* anneal: linear T0 -> T1 over the sweeps; PT: geometric ladder T0 .. T1 */
  double h;
  uint64_t seed; /* set_seed, init_J, set_seed, init_rand (src/main_gpu.cu:262-276) */
  double arrival; /* seconds on the server clock; < 0: the time of submission */
} sa_request;
typedef struct sa_request_result {
  int64_t id;
  int status; /* SA_OK, an sa_status for a refused request, SA_ERR_STATE while unfinished */
  int steps;  /* PT exchange attempts logged */
  double arrival, admitted, finished; /* server clock, seconds */
  int64_t bytes;   /* device memory held while admitted */
  double attempts; /* spin-flip attempts: sites x 64 x systems x sweeps */
  uint64_t spins_digest; /* sa_digest over all systems' spin words (system j from index j*N) */
  uint64_t log_digest;   /* PT: chained splitmix64 of (pair, uniform bits, mask) per step */
  int32_t E[64], M[64];  /* final integer energy / magnetisation of system 0 per lane */
  float E_min[64];       /* PT: largest coldest-system measure() value over the steps */
} sa_request_result;
typedef struct sa_serve_config {
  int64_t hbm_budget; /* device bytes for admitted requests (0: half the free memory) */
  int max_active;     /* requests on the device at once (0: 512) */
  int chunk;          /* sweeps between admission points (0: 10) */
  int digests;        /* 1: read final spins back and digest them */
  int prep_threads;   /* host threads preparing admitted requests (0: automatic) */
} sa_serve_config;
typedef struct sa_serve_stats {
  int64_t submitted, finished, pending;
  int active, peak_active;
  int64_t bytes, peak_bytes, hbm_budget;
  double busy_s, clock_s; /* time from chunk launch to completion; server clock */
  double admit_s, retire_s; /* host time admitting requests / reading results back */
  int64_t chunks;
} sa_serve_stats;
typedef struct sa_serve sa_serve;
int sa_serve_create(const sa_serve_config *cfg, sa_serve **out);
int sa_serve_destroy(sa_serve *s);
/* seconds since the server was created (the clock arrivals are given on) */
double sa_serve_clock(const sa_serve *s);
/* thread-safe; *id = request id (results are indexed by it) */
int sa_serve_submit(sa_serve *s, const sa_request *r, int64_t *id);
/* run the loop until n_finished requests have finished in total (<= 0: until every
 * submitted request has) */
int sa_serve_run(sa_serve *s, int64_t n_finished);
/* results of requests first .. first+capacity-1; returns how many were copied */
int64_t sa_serve_results(sa_serve *s, int64_t first, int64_t capacity, sa_request_result *out);
int sa_serve_get_stats(sa_serve *s, sa_serve_stats *st);
/* the temperatures a request runs at (nsweeps for an anneal, nT for PT); returns the
 * count, T may be NULL */
int sa_request_temperatures(const sa_request *r, double *T, int64_t capacity);
/* device bytes a request holds while admitted (what admission charges against
 * hbm_budget), or an sa_status for an invalid request */
int64_t sa_request_bytes(const sa_request *r);
/* XOR over i of splitmix64(w[i] + (first_index + i) * 0x9E3779B97F4A7C15) */
uint64_t sa_digest(const uint64_t *w, int64_t n, uint64_t first_index);

#ifdef __cplusplus
}
#endif
#endif
