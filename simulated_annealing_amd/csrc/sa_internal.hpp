// sa_internal.hpp -- definitions shared by the translation units of libsa_hip.so
// (not part of the public interface).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "sa/rng.hpp"
#include "sa_hip.h"
#include "sa_kernels.hpp"

// error reporting: records the message for sa_last_error() and returns `status`
int sa_fail(int status, const char *fmt, ...);

#define SA_HIP_TRY(expr)                                                                     \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess) return sa_fail(SA_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

struct sa_batch;

struct sa_ladder {
  int kind = 0, L = 0, nsys = 0, dim = 0;
  int64_t N = 0;
  int device = 0;
  hipStream_t stream = nullptr;
  uint64_t *s = nullptr, *jx = nullptr, *jy = nullptr, *jz = nullptr;
  uint64_t *s_alt = nullptr;  // second spin array of the fused 3D sweep (ping-pong)
  sa::sys_table *tab = nullptr;
  int *cnt = nullptr;
  std::vector<sa::sys_table> htab;
  std::vector<char> fast;  // per system: symmetric tables, low levels always accept
  int rng_mode = SA_RNG_REPLAY;
  uint64_t seed = 0, t = 0;
  sa::mt19937_64 eng{};
  uint64_t *words_d = nullptr, *words_h = nullptr;
  int64_t words_cap = 0;  // sweeps that fit in the staging buffers
  std::vector<int> cnt_h;  // host copy of the measurement counters
  int *cnt_pin = nullptr;  // pinned landing buffer of the counters (allocated at the first measurement)
  hipEvent_t meas_ev = nullptr;  // recorded after the counters' copy; polled
  int *sys_list = nullptr;  // device copy of a measured subset (more than 4 systems)
  int sys_per_thread = 0;  // SA_TUNE_SYS_PER_THREAD, 0 = automatic
  int sweep3d_kernel = 0;  // SA_TUNE_SWEEP3D_KERNEL, 0 = automatic
  int rows_per_tile = 0;   // SA_TUNE_ROWS_PER_TILE, 0 = automatic
  int xcd_map = -1;        // SA_TUNE_XCD_MAP, -1 = automatic
};

namespace sa_detail {
// acceptance thresholds of one temperature/field for a ladder of `kind`; fast =
// the table qualifies for the symmetric low-levels-always kernels
void make_table(int kind, double T, double h, sa::sys_table &tab, bool &fast, double *boltz_out);
// Measurement of n systems of s (nsys x N words; listed by dlist in device memory,
// by sel for up to four, else systems 0..n-1): clears cnt (kMeasureCopies x n x 128
// ints) with a fill kernel, then per copy, listed system and lane the satisfied
// -x,-y[,-z] bonds (cnt[..][j][0][lane]) and set spins (cnt[..][j][1][lane]).
// Asynchronous on st; the caller sums the copies.
void launch_measure(int kind, int L, int64_t N, const uint64_t *s, const uint64_t *jx, const uint64_t *jy,
                    const uint64_t *jz, int n, const int *dlist, sa::sys4 sel, int *cnt, hipStream_t st);
// device bytes sa_ladder_create allocates for a ladder (the fused 3D sweep's second
// spin array, allocated on first use, not included)
int64_t ladder_bytes(int kind, int L, int nsys);
// SA_OK when an instance of nsys systems of (kind, L) fits the batch kernels: the
// batch indexes an instance's work items (half_items(kind, L) * nsys) in 32 bits;
// otherwise an sa_status (recorded for sa_last_error), checked before any device work
int batch_geometry(int kind, int L, int nsys);
// device bytes of a batch instance: its ladder plus its temperature tables
int64_t batch_instance_bytes(int kind, int L, int nsys, int nsweeps, int per_sweep);
// One chunk of a batch in two halves, so a caller can do host work while it runs:
// launch (asynchronous; *k = sweeps launched, 0 when nothing is left), then
// finish (host sweep indices, exchange steps; waits for the chunk when an
// exchange is due).  batch_wait waits for everything launched.
int batch_launch_chunk(sa_batch *b, int max_sweeps, int *k);
int batch_finish_chunk(sa_batch *b, int k);
int batch_wait(sa_batch *b);
bool batch_done(const sa_batch *b, int id);
// An instance prepared off the batch (ladder seeded like main, tables uploaded),
// from any thread; attaching it (loop thread) makes it part of the next chunk.
// pt_every > 0: a parallel-tempering instance of nsys temperatures T exchanging
// every pt_every sweeps (nsweeps a multiple of it).
struct prepared_instance;
int batch_prepare(int kind, int L, int nsys, uint64_t seed, int nsweeps, const double *T, int per_sweep, double h,
                  int pt_every, prepared_instance **out);
int batch_attach(sa_batch *b, prepared_instance *p, int *id);
void batch_discard(prepared_instance *p);
// an instance's ladder without waiting for the batch (only for an instance that
// has finished: running chunks skip it and never touch its memory)
sa_ladder *batch_instance_ladder(sa_batch *b, int id);
}  // namespace sa_detail
