// sa_ladder.hip -- the C ABI of include/sa_hip.h: device-resident ladders and the
// launches of the sweep, measure and swap kernels (sa_kernels.hpp).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "sa_internal.hpp"

namespace {

thread_local std::string g_last_error;

constexpr int64_t kReplayWordsCap = int64_t(1) << 24;  // 128 MiB of staged random words

}  // namespace

int sa_fail(int status, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return status;
}

namespace {

int bonds_per_site(int kind) { return kind == SA_KIND_3D ? 3 : 2; }

void free_all(sa_ladder *l) {
  if (l->s) (void)hipFree(l->s);
  if (l->jx) (void)hipFree(l->jx);  // jy, jz live in the same allocation
  if (l->tab) (void)hipFree(l->tab);
  if (l->cnt) (void)hipFree(l->cnt);
  if (l->s_alt) (void)hipFree(l->s_alt);
  if (l->words_d) (void)hipFree(l->words_d);
  if (l->words_h) (void)hipHostFree(l->words_h);
  if (l->sys_list) (void)hipFree(l->sys_list);
  if (l->cnt_pin) (void)hipHostFree(l->cnt_pin);
  if (l->meas_ev) (void)hipEventDestroy(l->meas_ev);
  if (l->stream) (void)hipStreamDestroy(l->stream);
}

int check_sys(const sa_ladder *l, int sys) {
  if (!l) return sa_fail(SA_ERR_ARG, "null ladder");
  if (sys < 0 || sys >= l->nsys) return sa_fail(SA_ERR_ARG, "system index %d out of range [0,%d)", sys, l->nsys);
  return SA_OK;
}

int ensure_words(sa_ladder *l, int64_t sweeps) {
  if (l->words_cap >= sweeps) return SA_OK;
  if (l->words_d) (void)hipFree(l->words_d);
  if (l->words_h) (void)hipHostFree(l->words_h);
  l->words_d = nullptr;
  l->words_h = nullptr;
  l->words_cap = 0;
  if (hipMalloc(&l->words_d, sweeps * l->N * 8) != hipSuccess) return sa_fail(SA_ERR_NOMEM, "replay buffer");
  if (hipHostMalloc(&l->words_h, sweeps * l->N * 8, hipHostMallocDefault) != hipSuccess)
    return sa_fail(SA_ERR_NOMEM, "pinned replay buffer");
  l->words_cap = sweeps;
  return SA_OK;
}

template <bool COUNTER, bool FAST, int MB>
void launch3d(const sa::sweep_args &a, int color, hipStream_t st) {
  const int cols = a.L * (a.L / 2);
  dim3 grid((cols + 255) / 256, a.nsys / MB);
  hipLaunchKernelGGL((sa::k_sweep3d_half<COUNTER, FAST, MB>), grid, dim3(256), 0, st, a, color);
}

template <bool COUNTER, bool FAST, bool OPEN, int MB>
void launch2d(const sa::sweep_args &a, int color, hipStream_t st) {
  const int n = a.L * (a.L / 2);
  dim3 grid((n + 255) / 256, a.nsys / MB);
  hipLaunchKernelGGL((sa::k_sweep2d_half<COUNTER, FAST, OPEN, MB>), grid, dim3(256), 0, st, a, color);
}

int pick_mb(const sa_ladder *l) {
  if (l->sys_per_thread > 0) return l->sys_per_thread;
  return (l->nsys % 4 == 0) ? 4 : (l->nsys % 2 == 0) ? 2 : 1;
}

template <bool COUNTER, bool FAST, int S>
void launch_zscan_s(const sa::sweep_args &a, int color, int G, hipStream_t st) {
  const int64_t items = (int64_t)a.L * (a.L / 2) * G;
  const dim3 grid((unsigned)((items + 255) / 256), (unsigned)a.nsys);
  hipLaunchKernelGGL((sa::k_sweep3d_zscan<COUNTER, FAST, S>), grid, dim3(256), 0, st, a, color, G);
}

// z-parallel half-sweep (3D, even L with zscan_groups(L) > 0)
template <bool COUNTER, bool FAST>
void launch_zscan(const sa_ladder *l, const sa::sweep_args &a, int color) {
  const int G = sa::zscan_groups(l->L), S = l->L / G;
  switch (S) {
    case 2: launch_zscan_s<COUNTER, FAST, 2>(a, color, G, l->stream); break;
    case 3: launch_zscan_s<COUNTER, FAST, 3>(a, color, G, l->stream); break;
    case 4: launch_zscan_s<COUNTER, FAST, 4>(a, color, G, l->stream); break;
    case 5: launch_zscan_s<COUNTER, FAST, 5>(a, color, G, l->stream); break;
    case 6: launch_zscan_s<COUNTER, FAST, 6>(a, color, G, l->stream); break;
    case 7: launch_zscan_s<COUNTER, FAST, 7>(a, color, G, l->stream); break;
    default: launch_zscan_s<COUNTER, FAST, 8>(a, color, G, l->stream); break;
  }
}

bool zscan_ok(const sa_ladder *l) {
  return l->kind == SA_KIND_3D && !(l->L & 1) && sa::zscan_groups(l->L) > 0 && l->N < ((int64_t)1 << 31);
}

// the z-parallel half-sweeps: asked for, or automatically for ladders of at most 2^21
// site words in all, where they beat the fused walk (us per sweep, fused vs z-parallel:
// one L=64 system 46.7 / 23.1, four 47.1 / 41.6; one L=128 system 107 / 70; two L=128
// systems 108 / 134)
bool use_zscan(const sa_ladder *l) {
  if (!zscan_ok(l)) return false;
  if (l->sweep3d_kernel == 3) return true;
  return l->sweep3d_kernel == 0 && l->N * l->nsys <= ((int64_t)1 << 21);
}

template <bool COUNTER, bool FAST>
void launch_half(const sa_ladder *l, const sa::sweep_args &a, int color) {
  if (use_zscan(l)) {
    launch_zscan<COUNTER, FAST>(l, a, color);
    return;
  }
  if (l->L & 1) {  // odd L: 3D one lane per member, 2D one thread per wrap group; one system per grid row
    const int items = l->kind == SA_KIND_3D ? sa::odd_width(l->L) : sa::odd_groups(l->L);
    const dim3 grid((unsigned)((items + 255) / 256), (unsigned)l->nsys);
    if (l->kind == SA_KIND_3D)
      hipLaunchKernelGGL((sa::k_sweep3d_odd<COUNTER, FAST>), grid, dim3(256), 0, l->stream, a, color);
    else if (l->kind == SA_KIND_2D)
      hipLaunchKernelGGL((sa::k_sweep2d_odd<COUNTER, FAST, false>), grid, dim3(256), 0, l->stream, a, color);
    else
      hipLaunchKernelGGL((sa::k_sweep2d_odd<COUNTER, false, true>), grid, dim3(256), 0, l->stream, a, color);
    return;
  }
  const int mb = pick_mb(l);
  if (l->kind == SA_KIND_3D) {
    if (mb == 4) launch3d<COUNTER, FAST, 4>(a, color, l->stream);
    else if (mb == 2) launch3d<COUNTER, FAST, 2>(a, color, l->stream);
    else launch3d<COUNTER, FAST, 1>(a, color, l->stream);
  } else if (l->kind == SA_KIND_2D) {
    if (mb == 4) launch2d<COUNTER, FAST, false, 4>(a, color, l->stream);
    else if (mb == 2) launch2d<COUNTER, FAST, false, 2>(a, color, l->stream);
    else launch2d<COUNTER, FAST, false, 1>(a, color, l->stream);
  } else {
    if (mb == 4) launch2d<COUNTER, false, true, 4>(a, color, l->stream);
    else if (mb == 2) launch2d<COUNTER, false, true, 2>(a, color, l->stream);
    else launch2d<COUNTER, false, true, 1>(a, color, l->stream);
  }
}

// ---- fused 3D sweep: configuration (rows per tile TH, systems per workgroup MB)
struct fused_cfg {
  int th, mb, xcd;
};

size_t fused_lds_bytes(int L, int th, int mb) { return (size_t)3 * (th + 2) * mb * 2 * (L / 2) * 8; }  // 3 plane buffers

bool fused_ok(const sa_ladder *l, int th, int mb) {
  if (l->kind != SA_KIND_3D || l->L % th != 0 || th > l->L) return false;
  if (l->nsys % mb != 0) return false;
  // buffer byte offsets (and the out-of-range marker 2^31 above them) are 32-bit
  if ((int64_t)l->N * 8 >= (int64_t)1 << 31) return false;
  if ((th + 2) * (l->L / 2) > 640) return false;  // the kernel's launch bound
  return fused_lds_bytes(l->L, th, mb) <= 64 * 1024;
}

// automatic choice: the largest band that fits (less halo recomputation) among those
// that give at least one workgroup per CU, two systems per workgroup when the ladder
// has enough of them; a ladder too small to fill the GPU takes the smallest band
// (most workgroups).  Measured (us per sweep): one L=64 system 46.7 with 4-row
// bands, 61.0 with 8, 88.3 with 16; one L=128 system 107 / 159; 32 L=128 systems
// 283 with 8-row bands and two systems per workgroup.
fused_cfg pick_fused(const sa_ladder *l) {
  fused_cfg c{0, 1, 0};
  // general tables (h != 0) need more registers: one system per thread unless asked
  bool all_fast = true;
  for (char f : l->fast) all_fast = all_fast && f;
  const int mbs[2] = {l->sys_per_thread > 0 ? l->sys_per_thread : (all_fast ? 2 : 1), 1};
  const int ths[3] = {16, 8, 4};
  bool fill = false;  // some configuration reaches one workgroup per CU
  for (int mb : mbs)
    for (int th : ths)
      if ((l->rows_per_tile <= 0 || th == l->rows_per_tile) && fused_ok(l, th, mb) &&
          (l->L / th) * (l->nsys / mb) >= 256)
        fill = true;
  for (int mb : mbs) {
    if (mb != 1 && mb != 2) continue;
    for (int th : ths) {
      if (l->rows_per_tile > 0 ? th != l->rows_per_tile : (!fill && th != 4)) continue;
      if (!fused_ok(l, th, mb)) continue;
      // two systems per workgroup share couplings and random words, unless that
      // leaves fewer workgroups than CUs
      if (mb == 2 && l->sys_per_thread != 2 && (l->L / th) * (l->nsys / mb) < 256) continue;
      c.th = th;
      c.mb = mb;
      const int ntile = l->L / th;
      const bool can = ntile % 8 == 0, can2 = (l->nsys / mb) % 8 == 0;
      if (l->xcd_map == 2) c.xcd = can2 ? 2 : 0;
      else c.xcd = (l->xcd_map < 0) ? can : (l->xcd_map && can);  // measured faster where it applies
      return c;
    }
  }
  return c;
}

bool use_fused(const sa_ladder *l) {
  if (l->kind != SA_KIND_3D || l->sweep3d_kernel == 1 || use_zscan(l)) return false;
  return pick_fused(l).th > 0;
}

template <bool COUNTER, bool FAST, int MB, int TH>
void launch_fused_t(const sa::sweep_args &a, uint64_t *out, const fused_cfg &c, hipStream_t st) {
  const int G = (a.L / TH) * (a.nsys / MB);
  const int threads = (TH + 2) * (a.L / 2);
  if ((a.L / 2) % 64 == 0)
    hipLaunchKernelGGL((sa::k_sweep3d_fused<COUNTER, FAST, MB, TH, 1>), dim3(G), dim3(threads),
                       fused_lds_bytes(a.L, TH, MB), st, a, out, c.xcd);
  else
    hipLaunchKernelGGL((sa::k_sweep3d_fused<COUNTER, FAST, MB, TH, 0>), dim3(G), dim3(threads),
                       fused_lds_bytes(a.L, TH, MB), st, a, out, c.xcd);
}

template <bool COUNTER, bool FAST>
void launch_fused(const sa::sweep_args &a, uint64_t *out, const fused_cfg &c, hipStream_t st) {
  if (c.mb == 2) {
    if (c.th == 16) launch_fused_t<COUNTER, FAST, 2, 16>(a, out, c, st);
    else if (c.th == 8) launch_fused_t<COUNTER, FAST, 2, 8>(a, out, c, st);
    else launch_fused_t<COUNTER, FAST, 2, 4>(a, out, c, st);
  } else {
    if (c.th == 16) launch_fused_t<COUNTER, FAST, 1, 16>(a, out, c, st);
    else if (c.th == 8) launch_fused_t<COUNTER, FAST, 1, 8>(a, out, c, st);
    else launch_fused_t<COUNTER, FAST, 1, 4>(a, out, c, st);
  }
}

void launch_sweep(sa_ladder *l, const uint64_t *words, uint64_t t) {
  sa::sweep_args a;
  a.s = l->s;
  a.jx = l->jx;
  a.jy = l->jy;
  a.jz = l->jz;
  a.tab = l->tab;
  a.words = words;
  a.seed = l->seed;
  a.sweep = t;
  a.N = l->N;
  a.L = l->L;
  a.nsys = l->nsys;
  bool fast = true;
  for (char f : l->fast) fast = fast && f;
  if (use_fused(l)) {
    const fused_cfg c = pick_fused(l);
    if (l->rng_mode == SA_RNG_COUNTER) {
      if (fast) launch_fused<true, true>(a, l->s_alt, c, l->stream); else launch_fused<true, false>(a, l->s_alt, c, l->stream);
    } else {
      if (fast) launch_fused<false, true>(a, l->s_alt, c, l->stream); else launch_fused<false, false>(a, l->s_alt, c, l->stream);
    }
    std::swap(l->s, l->s_alt);
    return;
  }
  for (int color = 0; color < 2; ++color) {
    if (l->rng_mode == SA_RNG_COUNTER) {
      if (fast) launch_half<true, true>(l, a, color); else launch_half<true, false>(l, a, color);
    } else {
      if (fast) launch_half<false, true>(l, a, color); else launch_half<false, false>(l, a, color);
    }
  }
}

// thresholds of one Boltzmann table into sys_table t (levels 0..nb)
void fill_table(sa::sys_table &t, int tb, const double *b, int nlev) {
  t.always[tb] = 0;
  for (int c = 0; c < 8; ++c) t.thr[tb][c] = 0;
  for (int c = 0; c < nlev; ++c) {
    const sa::threshold th = sa::accept_threshold(b[c]);
    t.thr[tb][c] = th.value;
    if (th.always) t.always[tb] |= 1u << c;
  }
}

}  // namespace

extern "C" {

const char *sa_last_error(void) { return g_last_error.c_str(); }

const char *sa_status_string(int status) {
  switch (status) {
    case SA_OK: return "ok";
    case SA_ERR_ARG: return "invalid argument";
    case SA_ERR_UNSUPPORTED: return "unsupported geometry";
    case SA_ERR_HIP: return "HIP runtime error";
    case SA_ERR_NOMEM: return "out of memory";
    case SA_ERR_STATE: return "invalid state";
    default: return "unknown status";
  }
}

int sa_version(void) { return 1; }

int sa_ladder_create(int kind, int L, int nsys, sa_ladder **out) {
  if (!out) return sa_fail(SA_ERR_ARG, "null output handle");
  *out = nullptr;
  if (kind < SA_KIND_2P || kind > SA_KIND_3D) return sa_fail(SA_ERR_ARG, "unknown kind %d", kind);
  if (nsys < 1) return sa_fail(SA_ERR_ARG, "nsys must be >= 1");
  if (L < 3) return sa_fail(SA_ERR_UNSUPPORTED, "device path needs L >= 3 (got %d)", L);
  if ((kind == SA_KIND_3D && L > 1024) || (kind != SA_KIND_3D && L > 32768))
    return sa_fail(SA_ERR_ARG, "L=%d too large", L);
  sa_ladder *l = new sa_ladder();
  l->kind = kind;
  l->L = L;
  l->nsys = nsys;
  l->dim = kind == SA_KIND_3D ? 3 : 2;
  l->N = (int64_t)L * L * (kind == SA_KIND_3D ? L : 1);
  const int64_t N = l->N;
  int rc = SA_OK;
  auto bad = [&](hipError_t e, const char *what) {
    if (e == hipSuccess) return false;
    rc = (e == hipErrorOutOfMemory) ? sa_fail(SA_ERR_NOMEM, "%s: %s", what, hipGetErrorString(e))
                                    : sa_fail(SA_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
    return true;
  };
  if (bad(hipGetDevice(&l->device), "hipGetDevice") || bad(hipStreamCreateWithFlags(&l->stream, hipStreamNonBlocking), "hipStreamCreate") ||
      bad(hipMalloc(&l->s, (size_t)nsys * N * 8), "spins") || bad(hipMalloc(&l->jx, 3 * N * 8), "couplings") ||
      bad(hipMalloc(&l->tab, sizeof(sa::sys_table) * nsys), "tables") ||
      bad(hipMalloc(&l->cnt, sizeof(int) * 128 * nsys * sa::kMeasureCopies), "counters") ||
      bad(hipMalloc(&l->sys_list, sizeof(int) * nsys), "system list") ||
      bad(hipMemsetAsync(l->s, 0, (size_t)nsys * N * 8, l->stream), "memset") ||
      bad(hipMemsetAsync(l->jx, 0, 3 * N * 8, l->stream), "memset")) {
    free_all(l);
    delete l;
    return rc;
  }
  l->jy = l->jx + N;  // one allocation for the three coupling arrays
  l->jz = l->jx + 2 * N;
  l->htab.assign(nsys, sa::sys_table{});
  l->fast.assign(nsys, 0);
  l->cnt_h.assign((size_t)128 * nsys * sa::kMeasureCopies, 0);
  l->eng.seed(5489u);
  *out = l;
  for (int i = 0; i < nsys; ++i) {
    rc = sa_ladder_set_temperature(l, i, 1.0, 0.0, nullptr);
    if (rc != SA_OK) {
      sa_ladder_destroy(l);
      *out = nullptr;
      return rc;
    }
  }
  return SA_OK;
}

int sa_ladder_destroy(sa_ladder *l) {
  if (!l) return SA_OK;
  if (l->stream) (void)hipStreamSynchronize(l->stream);
  free_all(l);
  delete l;
  return SA_OK;
}

int64_t sa_ladder_sites(const sa_ladder *l) { return l ? l->N : 0; }
int sa_ladder_nsys(const sa_ladder *l) { return l ? l->nsys : 0; }
void *sa_ladder_stream(sa_ladder *l) { return l ? (void *)l->stream : nullptr; }

int sa_ladder_set_rng(sa_ladder *l, int mode) {
  if (!l) return sa_fail(SA_ERR_ARG, "null ladder");
  if (mode != SA_RNG_REPLAY && mode != SA_RNG_COUNTER) return sa_fail(SA_ERR_ARG, "unknown rng mode %d", mode);
  l->rng_mode = mode;
  return SA_OK;
}

int sa_ladder_seed(sa_ladder *l, uint64_t seed) {
  if (!l) return sa_fail(SA_ERR_ARG, "null ladder");
  l->eng.seed(seed);
  l->seed = seed;
  l->t = 0;
  return SA_OK;
}

int sa_ladder_get_seed(const sa_ladder *l, uint64_t *seed) {
  if (!l || !seed) return sa_fail(SA_ERR_ARG, "null argument");
  *seed = l->seed;
  return SA_OK;
}

int sa_ladder_get_sweep_index(const sa_ladder *l, uint64_t *t) {
  if (!l || !t) return sa_fail(SA_ERR_ARG, "null argument");
  *t = l->t;
  return SA_OK;
}

int sa_ladder_set_sweep_index(sa_ladder *l, uint64_t t) {
  if (!l) return sa_fail(SA_ERR_ARG, "null ladder");
  l->t = t;
  return SA_OK;
}

int sa_ladder_get_engine(const sa_ladder *l, uint64_t state[313]) {
  if (!l || !state) return sa_fail(SA_ERR_ARG, "null argument");
  std::memcpy(state, l->eng.x, 312 * 8);
  state[312] = l->eng.p;
  return SA_OK;
}

int sa_ladder_set_engine(sa_ladder *l, const uint64_t state[313]) {
  if (!l || !state) return sa_fail(SA_ERR_ARG, "null argument");
  if (state[312] > 312) return sa_fail(SA_ERR_ARG, "engine position %llu > 312", (unsigned long long)state[312]);
  std::memcpy(l->eng.x, state, 312 * 8);
  l->eng.p = state[312];
  return SA_OK;
}

int sa_ladder_set_couplings(sa_ladder *l, const uint64_t *jx, const uint64_t *jy, const uint64_t *jz) {
  if (!l || !jx || !jy || (l->kind == SA_KIND_3D && !jz)) return sa_fail(SA_ERR_ARG, "null argument");
  const size_t bytes = l->N * 8;
  SA_HIP_TRY(hipMemcpyAsync(l->jx, jx, bytes, hipMemcpyHostToDevice, l->stream));
  SA_HIP_TRY(hipMemcpyAsync(l->jy, jy, bytes, hipMemcpyHostToDevice, l->stream));
  if (l->kind == SA_KIND_3D) SA_HIP_TRY(hipMemcpyAsync(l->jz, jz, bytes, hipMemcpyHostToDevice, l->stream));
  SA_HIP_TRY(hipStreamSynchronize(l->stream));
  return SA_OK;
}

int sa_ladder_get_couplings(const sa_ladder *l, uint64_t *jx, uint64_t *jy, uint64_t *jz) {
  if (!l || !jx || !jy || (l->kind == SA_KIND_3D && !jz)) return sa_fail(SA_ERR_ARG, "null argument");
  const size_t bytes = l->N * 8;
  SA_HIP_TRY(hipMemcpyAsync(jx, l->jx, bytes, hipMemcpyDeviceToHost, l->stream));
  SA_HIP_TRY(hipMemcpyAsync(jy, l->jy, bytes, hipMemcpyDeviceToHost, l->stream));
  if (l->kind == SA_KIND_3D) SA_HIP_TRY(hipMemcpyAsync(jz, l->jz, bytes, hipMemcpyDeviceToHost, l->stream));
  SA_HIP_TRY(hipStreamSynchronize(l->stream));
  return SA_OK;
}

// J file of the reference (src/sys_file.cpp:10-110): int d; long N; then per axis 2N
// words, [2i] = the bond of site i toward -axis, [2i+1] = toward +axis (the +axis
// neighbour's -axis bond, periodic wrap; src/cpu_3d.cpp:290-299).
namespace {
int64_t neighbour_plus(int64_t i, int axis, int L, int64_t N) {
  const int64_t x = i % L, y = (i / L) % L, L2 = (int64_t)L * L;
  if (axis == 0) return i - x + (x + 1) % L;
  if (axis == 1) return i - y * L + ((y + 1) % L) * L;
  return (i + L2) % N;
}
}  // namespace

int sa_ladder_load_J(sa_ladder *l, const char *path) {
  if (!l || !path) return sa_fail(SA_ERR_ARG, "null argument");
  FILE *f = std::fopen(path, "rb");
  if (!f) return sa_fail(SA_ERR_ARG, "load_J: cannot open \"%s\"", path);
  int d = 0;
  long n = 0;
  const int dim = l->kind == SA_KIND_3D ? 3 : 2;
  int rc = SA_OK;
  if (std::fread(&d, sizeof d, 1, f) != 1 || std::fread(&n, sizeof n, 1, f) != 1)
    rc = sa_fail(SA_ERR_ARG, "load_J: \"%s\" is shorter than its header", path);
  else if (d != dim)
    rc = sa_fail(SA_ERR_ARG, "load_J: \"%s\" holds d=%d couplings, the ladder is %dD", path, d, dim);
  else if (n != l->N)
    rc = sa_fail(SA_ERR_ARG, "load_J: \"%s\" holds N=%ld sites, the ladder has %lld", path, n, (long long)l->N);
  std::vector<uint64_t> dup((size_t)2 * l->N), j[3];
  for (int a = 0; a < dim && rc == SA_OK; ++a) {
    if (std::fread(dup.data(), 8, dup.size(), f) != dup.size()) {
      rc = sa_fail(SA_ERR_ARG, "load_J: \"%s\" ends inside the couplings", path);
      break;
    }
    j[a].resize(l->N);
    for (int64_t i = 0; i < l->N; ++i) j[a][i] = dup[2 * i];
    // each bond is stored twice; a file whose copies disagree has no deduplicated form
    for (int64_t i = 0; i < l->N && rc == SA_OK; ++i)
      if (dup[2 * i + 1] != dup[2 * neighbour_plus(i, a, l->L, l->N)])
        rc = sa_fail(SA_ERR_ARG, "load_J: \"%s\": axis %d, site %lld: the two copies of a bond differ", path, a,
                     (long long)i);
  }
  std::fclose(f);
  if (rc != SA_OK) return rc;  // couplings unchanged
  return sa_ladder_set_couplings(l, j[0].data(), j[1].data(), dim == 3 ? j[2].data() : nullptr);
}

int sa_ladder_save_J(const sa_ladder *l, const char *path) {
  if (!l || !path) return sa_fail(SA_ERR_ARG, "null argument");
  const int dim = l->kind == SA_KIND_3D ? 3 : 2;
  std::vector<uint64_t> j[3];
  for (auto &a : j) a.resize(l->N);
  int rc = sa_ladder_get_couplings(l, j[0].data(), j[1].data(), j[2].data());
  if (rc != SA_OK) return rc;
  FILE *f = std::fopen(path, "wb");
  if (!f) return sa_fail(SA_ERR_ARG, "save_J: cannot open \"%s\"", path);
  const int d = dim;
  const long n = (long)l->N;
  bool ok = std::fwrite(&d, sizeof d, 1, f) == 1 && std::fwrite(&n, sizeof n, 1, f) == 1;
  std::vector<uint64_t> dup((size_t)2 * l->N);
  for (int a = 0; a < dim && ok; ++a) {
    for (int64_t i = 0; i < l->N; ++i) {
      dup[2 * i] = j[a][i];
      dup[2 * i + 1] = j[a][neighbour_plus(i, a, l->L, l->N)];
    }
    ok = std::fwrite(dup.data(), 8, dup.size(), f) == dup.size();
  }
  ok = (std::fclose(f) == 0) && ok;
  return ok ? SA_OK : sa_fail(SA_ERR_ARG, "save_J: writing \"%s\" failed", path);
}

int sa_ladder_init_couplings(sa_ladder *l) {
  if (!l) return sa_fail(SA_ERR_ARG, "null ladder");
  const int64_t N = l->N;
  const int d = bonds_per_site(l->kind);
  std::vector<uint64_t> jx(N), jy(N), jz(l->kind == SA_KIND_3D ? N : 1);
  // the reference draws d*N words into buffer[d*i + axis] (src/cpu_3d.cpp:285-300)
  for (int64_t i = 0; i < N; ++i) {
    jx[i] = l->eng();
    jy[i] = l->eng();
    if (d == 3) jz[i] = l->eng();
  }
  return sa_ladder_set_couplings(l, jx.data(), jy.data(), d == 3 ? jz.data() : nullptr);
}

int sa_ladder_init_spins(sa_ladder *l) {
  if (!l) return sa_fail(SA_ERR_ARG, "null ladder");
  std::vector<uint64_t> s(l->N);
  l->eng.fill(s.data(), l->N);
  for (int i = 0; i < l->nsys; ++i)
    SA_HIP_TRY(hipMemcpyAsync(l->s + (int64_t)i * l->N, s.data(), l->N * 8, hipMemcpyHostToDevice, l->stream));
  SA_HIP_TRY(hipStreamSynchronize(l->stream));
  return SA_OK;
}

int sa_ladder_set_spins(sa_ladder *l, int sys, const uint64_t *s) {
  int rc = check_sys(l, sys);
  if (rc) return rc;
  if (!s) return sa_fail(SA_ERR_ARG, "null spins");
  SA_HIP_TRY(hipMemcpyAsync(l->s + (int64_t)sys * l->N, s, l->N * 8, hipMemcpyHostToDevice, l->stream));
  SA_HIP_TRY(hipStreamSynchronize(l->stream));
  return SA_OK;
}

int sa_ladder_get_spins(const sa_ladder *l, int sys, uint64_t *s) {
  int rc = check_sys(l, sys);
  if (rc) return rc;
  if (!s) return sa_fail(SA_ERR_ARG, "null spins");
  SA_HIP_TRY(hipMemcpyAsync(s, l->s + (int64_t)sys * l->N, l->N * 8, hipMemcpyDeviceToHost, l->stream));
  SA_HIP_TRY(hipStreamSynchronize(l->stream));
  return SA_OK;
}

}  // extern "C"

int64_t sa_detail::ladder_bytes(int kind, int L, int nsys) {
  const int64_t N = (int64_t)L * L * (kind == SA_KIND_3D ? L : 1);
  return (int64_t)nsys * N * 8 + 3 * N * 8 + (int64_t)sizeof(sa::sys_table) * nsys +
         4 * 128 * nsys * sa::kMeasureCopies + 4 * nsys;
}

void sa_detail::make_table(int kind, double T, double h, sa::sys_table &t, bool &fast, double *boltz_out) {
  double b[20];
  t = sa::sys_table{};
  const int nb = kind == SA_KIND_3D ? 6 : 4;
  if (kind == SA_KIND_3D) {
    sa::boltz_3d(T, h, b);
    fill_table(t, 0, b, 7);
    fill_table(t, 1, b + 7, 7);
  } else {
    sa::boltz_2d(T, h, b);
    fill_table(t, 0, b, 5);
    fill_table(t, 1, b + 5, 5);
    fill_table(t, 2, b + 10, 5);
    fill_table(t, 3, b + 15, 5);
  }
  if (boltz_out) std::memcpy(boltz_out, b, sizeof(double) * (kind == SA_KIND_3D ? 14 : 20));
  // fast kernels: identical spin tables, levels 0..nb/2 always accepted
  fast = kind != SA_KIND_2P && std::memcmp(t.thr[0], t.thr[1], sizeof t.thr[0]) == 0 && t.always[0] == t.always[1];
  const uint32_t low = (1u << (nb / 2 + 1)) - 1u;
  fast = fast && (t.always[0] & low) == low && (t.always[0] >> (nb / 2 + 1)) == 0;
}

void sa_detail::launch_measure(int kind, int L, int64_t N, const uint64_t *s, const uint64_t *jx, const uint64_t *jy,
                               const uint64_t *jz, int n, const int *dlist, sa::sys4 sel, int *cnt, hipStream_t st) {
  const int all_n = 128 * n * sa::kMeasureCopies;
  hipLaunchKernelGGL(sa::k_fill_zero, dim3((all_n + 255) / 256), dim3(256), 0, st, cnt, all_n);
  // cnt layout [copy][listed system][2][64] (the caller sums the copies); one block
  // row per SB systems (4, or 1 when there are at most 4: more waves in flight for
  // the few-system measurement of an exchange step).  A wave converts its bit-sliced
  // counters (a 64x64 transpose per plane) after 16 site groups or at its end, so
  // each wave gets at least 16 groups where possible (32 with one system per row,
  // four of them loaded at a time): more blocks would only add conversions.
  const int sb = n <= 4 ? 1 : 4;
  const int yb = (n + sb - 1) / sb;
  const int64_t want = std::max<int64_t>(1, (sb == 1 ? 4096 : 2048) / yb);
  const int64_t per = sb == 1 ? 32 : 16;  // site groups per wave (one system per row: 4 in flight at a time)
  const int64_t blocks = std::min<int64_t>((N + 256 * per - 1) / (256 * per), want);
  dim3 grid((unsigned)blocks, yb);
#define SA_MEASURE(KINDV)                                                                                        \
  if (sb == 1)                                                                                                   \
    hipLaunchKernelGGL((sa::k_measure<KINDV, 1>), grid, dim3(256), 0, st, s, jx, jy, jz, L, N, n, dlist, sel, cnt); \
  else                                                                                                           \
    hipLaunchKernelGGL((sa::k_measure<KINDV, 4>), grid, dim3(256), 0, st, s, jx, jy, jz, L, N, n, dlist, sel, cnt);
  if (kind == SA_KIND_3D) {
    SA_MEASURE(sa::KIND_3D)
  } else if (kind == SA_KIND_2D) {
    SA_MEASURE(sa::KIND_2D)
  } else {
    SA_MEASURE(sa::KIND_2P)
  }
#undef SA_MEASURE
}

extern "C" {

int sa_ladder_set_temperature(sa_ladder *l, int sys, double T, double h, double *boltz_out) {
  int rc = check_sys(l, sys);
  if (rc) return rc;
  sa::sys_table t;
  bool fast = false;
  sa_detail::make_table(l->kind, T, h, t, fast, boltz_out);
  l->htab[sys] = t;
  l->fast[sys] = fast;
  // stream-ordered: a pageable-source copy is staged before the call returns, so no
  // wait is needed; every later launch on the stream sees the new table
  SA_HIP_TRY(hipMemcpyAsync(l->tab + sys, &l->htab[sys], sizeof(sa::sys_table), hipMemcpyHostToDevice, l->stream));
  return SA_OK;
}

int sa_ladder_sweep(sa_ladder *l, int nsweeps) {
  if (!l) return sa_fail(SA_ERR_ARG, "null ladder");
  if (nsweeps < 0) return sa_fail(SA_ERR_ARG, "negative sweep count");
  if (nsweeps == 0) return SA_OK;
  if (use_fused(l) && !l->s_alt) {
    const hipError_t e = hipMalloc(&l->s_alt, (size_t)l->nsys * l->N * 8);
    if (e != hipSuccess) {
      l->s_alt = nullptr;
      return sa_fail(SA_ERR_NOMEM, "second spin array: %s", hipGetErrorString(e));
    }
  }
  if (l->rng_mode == SA_RNG_COUNTER) {
    for (int k = 0; k < nsweeps; ++k) launch_sweep(l, nullptr, l->t + k);
    SA_HIP_TRY(hipGetLastError());
    l->t += nsweeps;
    return SA_OK;
  }
  const int64_t per_chunk = std::max<int64_t>(1, std::min<int64_t>(nsweeps, kReplayWordsCap / l->N));
  int rc = ensure_words(l, per_chunk);
  if (rc) return rc;
  for (int64_t done = 0; done < nsweeps;) {
    const int64_t n = std::min<int64_t>(per_chunk, nsweeps - done);
    l->eng.fill(l->words_h, n * l->N);
    SA_HIP_TRY(hipMemcpyAsync(l->words_d, l->words_h, n * l->N * 8, hipMemcpyHostToDevice, l->stream));
    for (int64_t k = 0; k < n; ++k) launch_sweep(l, l->words_d + k * l->N, l->t + done + k);
    SA_HIP_TRY(hipGetLastError());
    SA_HIP_TRY(hipStreamSynchronize(l->stream));
    done += n;
  }
  l->t += nsweeps;
  return SA_OK;
}

namespace {

// energies/magnetisations of `n` systems (list = nullptr: systems 0..n-1); results
// land in cnt_h[(2*j)*64 ...] for the j-th listed system
int measure_list(sa_ladder *l, const int *list, int n) {
  const int cnt_n = 128 * n, all_n = cnt_n * sa::kMeasureCopies;
  // up to 4 listed systems travel as a kernel argument; more through device memory
  const int *dlist = nullptr;
  sa::sys4 sel{{-1, -1, -1, -1}};
  if (list && n <= 4) {
    for (int j = 0; j < n; ++j) sel.v[j] = list[j];
  } else if (list) {
    SA_HIP_TRY(hipMemcpyAsync(l->sys_list, list, sizeof(int) * n, hipMemcpyHostToDevice, l->stream));
    dlist = l->sys_list;
  }
  sa_detail::launch_measure(l->kind, l->L, l->N, l->s, l->jx, l->jy, l->jz, n, dlist, sel, l->cnt, l->stream);
  SA_HIP_TRY(hipGetLastError());
  // The counters land in pinned memory allocated at the first measurement (so
  // creating a ladder, e.g. per served request, needs no pinned allocation) and the
  // host polls an event instead of sleeping in a stream synchronisation: the
  // replica-exchange step waits on this between sweeps.
  if (!l->cnt_pin &&
      hipHostMalloc((void **)&l->cnt_pin, sizeof(int) * 128 * l->nsys * sa::kMeasureCopies, hipHostMallocDefault) !=
          hipSuccess)
    l->cnt_pin = nullptr;
  if (!l->meas_ev && hipEventCreateWithFlags(&l->meas_ev, hipEventDisableTiming) != hipSuccess) l->meas_ev = nullptr;
  if (l->cnt_pin && l->meas_ev) {
    SA_HIP_TRY(hipMemcpyAsync(l->cnt_pin, l->cnt, sizeof(int) * all_n, hipMemcpyDeviceToHost, l->stream));
    SA_HIP_TRY(hipEventRecord(l->meas_ev, l->stream));
    hipError_t e;
    while ((e = hipEventQuery(l->meas_ev)) == hipErrorNotReady) {
    }
    SA_HIP_TRY(e);
    std::memcpy(l->cnt_h.data(), l->cnt_pin, sizeof(int) * all_n);
  } else {
    SA_HIP_TRY(hipMemcpyAsync(l->cnt_h.data(), l->cnt, sizeof(int) * all_n, hipMemcpyDeviceToHost, l->stream));
    SA_HIP_TRY(hipStreamSynchronize(l->stream));
  }
  // sum the copies into the first
  for (int c = 1; c < sa::kMeasureCopies; ++c)
    for (int q = 0; q < cnt_n; ++q) l->cnt_h[q] += l->cnt_h[(size_t)c * cnt_n + q];
  return SA_OK;
}

void counts_to_EM(const sa_ladder *l, int j, int32_t *E, int32_t *M) {
  const int64_t nb = bonds_per_site(l->kind) * l->N - (l->kind == SA_KIND_2P ? l->L : 0);
  for (int k = 0; k < 64; ++k) {
    E[k] = (int32_t)(2 * (int64_t)l->cnt_h[(2 * j) * 64 + k] - nb);
    M[k] = (int32_t)(2 * (int64_t)l->cnt_h[(2 * j + 1) * 64 + k] - l->N);
  }
}

}  // namespace

int sa_ladder_anneal(sa_ladder *l, const double *T, int nsweeps, double h) {
  if (!l || (!T && nsweeps > 0) || nsweeps < 0) return sa_fail(SA_ERR_ARG, "bad argument");
  for (int t = 0; t < nsweeps; ++t) {
    for (int j = 0; j < l->nsys; ++j) {
      const int rc = sa_ladder_set_temperature(l, j, T[(int64_t)t * l->nsys + j], h, nullptr);
      if (rc) return rc;
    }
    const int rc = sa_ladder_sweep(l, 1);
    if (rc) return rc;
  }
  return SA_OK;
}

int sa_ladder_measure(sa_ladder *l, int32_t *E, int32_t *M) {
  if (!l || !E || !M) return sa_fail(SA_ERR_ARG, "null argument");
  const int rc = measure_list(l, nullptr, l->nsys);
  if (rc) return rc;
  for (int i = 0; i < l->nsys; ++i) counts_to_EM(l, i, E + i * 64, M + i * 64);
  return SA_OK;
}

int sa_ladder_measure_systems(sa_ladder *l, const int *sys, int n, int32_t *E, int32_t *M) {
  if (!l || !sys || !E || !M) return sa_fail(SA_ERR_ARG, "null argument");
  if (n < 1 || n > l->nsys) return sa_fail(SA_ERR_ARG, "measure of %d systems out of %d", n, l->nsys);
  for (int j = 0; j < n; ++j)
    if (sys[j] < 0 || sys[j] >= l->nsys) return sa_fail(SA_ERR_ARG, "system index %d out of range", sys[j]);
  const int rc = measure_list(l, sys, n);
  if (rc) return rc;
  for (int j = 0; j < n; ++j) counts_to_EM(l, j, E + j * 64, M + j * 64);
  return SA_OK;
}

int sa_ladder_swap(sa_ladder *l, int a, int b, uint64_t mask) {
  int rc = check_sys(l, a);
  if (rc) return rc;
  rc = check_sys(l, b);
  if (rc) return rc;
  if (a == b || mask == 0) return SA_OK;
  const int64_t blocks = std::min<int64_t>((l->N + 255) / 256, 2048);
  hipLaunchKernelGGL(sa::k_swap, dim3((unsigned)blocks), dim3(256), 0, l->stream, l->s + (int64_t)a * l->N,
                     l->s + (int64_t)b * l->N, mask, l->N);
  SA_HIP_TRY(hipGetLastError());
  return SA_OK;
}

int sa_ladder_set_tuning(sa_ladder *l, int key, int value) {
  if (!l) return sa_fail(SA_ERR_ARG, "null ladder");
  switch (key) {
    case SA_TUNE_SYS_PER_THREAD:
      if (value != 0 && value != 1 && value != 2 && value != 4)
        return sa_fail(SA_ERR_ARG, "systems per thread must be 0 (auto), 1, 2 or 4");
      if (value > 0 && l->nsys % value != 0)
        return sa_fail(SA_ERR_ARG, "%d systems per thread do not divide nsys=%d", value, l->nsys);
      l->sys_per_thread = value;
      return SA_OK;
    case SA_TUNE_SWEEP3D_KERNEL:
      if (value < 0 || value > 3) return sa_fail(SA_ERR_ARG, "3D sweep kernel must be 0 (auto), 1, 2 or 3");
      l->sweep3d_kernel = value;
      return SA_OK;
    case SA_TUNE_ROWS_PER_TILE:
      if (value != 0 && value != 4 && value != 8 && value != 16)
        return sa_fail(SA_ERR_ARG, "rows per tile must be 0 (auto), 4, 8 or 16");
      l->rows_per_tile = value;
      return SA_OK;
    case SA_TUNE_XCD_MAP:
      if (value < -1 || value > 2) return sa_fail(SA_ERR_ARG, "XCD map must be -1 (auto), 0, 1 or 2");
      l->xcd_map = value;
      return SA_OK;
    default:
      return sa_fail(SA_ERR_ARG, "unknown tuning key %d", key);
  }
}

int sa_ladder_get_tuning(const sa_ladder *l, int key, int *value) {
  if (!l || !value) return sa_fail(SA_ERR_ARG, "null argument");
  const bool fused = use_fused(l);
  const fused_cfg c = fused ? pick_fused(l) : fused_cfg{0, 0, 0};
  switch (key) {
    case SA_TUNE_SYS_PER_THREAD: *value = fused ? c.mb : pick_mb(l); return SA_OK;
    case SA_TUNE_SWEEP3D_KERNEL:
      *value = l->kind != SA_KIND_3D ? 0 : (fused ? 2 : use_zscan(l) ? 3 : 1);
      return SA_OK;
    case SA_TUNE_ROWS_PER_TILE: *value = c.th; return SA_OK;
    case SA_TUNE_XCD_MAP: *value = c.xcd; return SA_OK;
    default: return sa_fail(SA_ERR_ARG, "unknown tuning key %d", key);
  }
}

int sa_ladder_sync(sa_ladder *l) {
  if (!l) return sa_fail(SA_ERR_ARG, "null ladder");
  SA_HIP_TRY(hipStreamSynchronize(l->stream));
  return SA_OK;
}

}  // extern "C"
