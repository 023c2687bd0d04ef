// sa_batch.hip -- many independent instances of uneven size swept together (the
// C ABI sa_batch_* of include/sa_hip.h).
//
// Each instance is a ladder (couplings, spins, nsys systems) with its own counter
// RNG seed and either an anneal (temperatures per sweep) or parallel tempering
// (fixed ladder of temperatures, an exchange attempt every `every` sweeps).
//
// Sweeps.  A half-sweep launch covers every unfinished instance.  The work items
// (3D: (column, system), 2D: (site, system)) of all instances are cut into blocks
// of one instance each; the block map and the per-instance descriptors live in
// device memory and a fixed grid walks the block map (grid-stride), so the launch
// sequence of a chunk of sweeps does not depend on which instances are present: it
// is captured once per chunk length as a hipGraph and replayed while instances
// come and go.  Every block runs the per-column / per-site code of the
// single-instance kernels, so an instance's spins are bit for bit those of its
// ladder swept alone.  The descriptors hold each instance's sweep index at the
// start of the chunk; every launch of the chunk gets its sweep's offset as a
// constant argument, and one one-block kernel at the end of the chunk advances the
// indices (no index update, and no drain of the whole grid for one, per sweep).
//
// Exchanges.  Chunks end on every parallel-tempering instance's exchange points.
// There the host draws the pair and the uniform of each due instance (its own
// mt19937(123), as parallel_tempering::step), one launch measures the coldest
// system and the pair of all due instances, the host decides the lane masks with
// the reference's rule, and one launch applies all masked swaps.
//
// Tuning.  Threads per block and sweeps per graph are picked online: the first
// chunks try the candidates (results do not depend on them), the fastest per work
// item is kept.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <vector>

#include "sa/exchange.hpp"
#include "sa_internal.hpp"

namespace {

struct inst_desc {
  sa::sweep_args a;  // a.s: first spin array; tab: tables (per sweep: nsweeps x nsys, else nsys)
  uint64_t *s2;      // second spin array (instances that may take the fused 3D sweep)
  int kind, nsweeps, t, per_sweep, fast;
  int fused;  // swept by the fused kernel in this schedule (reads the current array, writes the other)
  int cur;    // 1: s2 holds the current spins
  int pad;
};

// t: the instance's sweep index (its index at the chunk start plus the sweep's
// offset in the chunk, a constant of the captured launch)
template <bool FAST>
__device__ __forceinline__ void batch_item(const inst_desc &in, int t, int color, int k, int sys) {
  sa::sweep_args a = in.a;
  if (in.cur) a.s = in.s2;  // half-sweeps update the current array in place
  if (in.per_sweep) a.tab += (int64_t)t * a.nsys;
  a.sweep = (uint64_t)t;
  if (a.L & 1) {  // odd L: k is a member lane (3D) or a wrap group (2D), sa_kernels.hpp
    if (in.kind == SA_KIND_3D) sa::odd3d_lane<true, FAST>(a, color, k, sys);
    else if (in.kind == SA_KIND_2D) sa::odd2d_group<true, FAST, false>(a, color, k, sys);
    else sa::odd2d_group<true, false, true>(a, color, k, sys);
    return;
  }
  if (in.kind == SA_KIND_3D) sa::half3d_column<true, FAST, 1, true>(a, color, k, sys);
  else if (in.kind == SA_KIND_2D) sa::half2d_site<true, FAST, false, 1>(a, color, k, sys);
  else sa::half2d_site<true, false, true, 1>(a, color, k, sys);
}

// work items of one system in a half-sweep: columns (3D) or sites (2D) of one
// colour for even L; for odd L the member lanes (3D, a multiple of 64 so that each
// system starts at a wave boundary) or the wrap groups (2D)
__host__ __device__ __forceinline__ int half_items(int kind, int L) {
  if (!(L & 1)) return L * (L >> 1);
  return kind == SA_KIND_3D ? sa::odd_width(L) : sa::odd_groups(L);
}

// ctrl[0] = blocks in the block map, ctrl[1] = descriptors; s = the sweep's offset in
// the chunk (descriptors hold the sweep index at the chunk start)
__global__ void __launch_bounds__(256) k_batch_half(const inst_desc *__restrict__ d, const int2 *__restrict__ bmap,
                                                    const int *__restrict__ ctrl, int color, int s) {
  const int nb = ctrl[0];
  for (int blk = blockIdx.x; blk < nb; blk += gridDim.x) {
    const int2 bm = bmap[blk];
    const inst_desc &in = d[bm.x];
    const int t = in.t + s;
    if (t >= in.nsweeps) continue;
    const int work = half_items(in.kind, in.a.L);
    const int item = bm.y + (int)threadIdx.x;
    if (item >= work * in.a.nsys) continue;
    const int sys = item / work, k = item - sys * work;
    if (in.fast) batch_item<true>(in, t, color, k, sys);
    else batch_item<false>(in, t, color, k, sys);
  }
}

// end of a chunk of k sweeps: sweep indices advance by the sweeps each instance did
__global__ void k_batch_advance(inst_desc *d, const int *__restrict__ ctrl, int k) {
  const int n = ctrl[1];
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const int done = min(k, max(0, d[i].nsweeps - d[i].t));
    d[i].t += done;
    if (d[i].fused) d[i].cur ^= done & 1;  // each fused sweep wrote the other array
  }
}

// Fused 3D sweeps of one group of instances (same L, band height TH and systems per
// workgroup MB, so one block size and LDS size): group g's blocks are entries
// ctrl[2+2g] .. +ctrl[3+2g] of the fused block map, each (descriptor, band index
// sb * (L/TH) + tile).  Each block runs the single-ladder fused walk on its band.
constexpr int kMaxFusedGroups = 24;
template <int MB, int TH>
__global__ void __launch_bounds__(640) k_batch_fused(const inst_desc *__restrict__ d, const int2 *__restrict__ fmap,
                                                     const int *__restrict__ ctrl, int g, int s) {
  extern __shared__ uint64_t lds[];
  const int off = ctrl[2 + 2 * g], n = ctrl[3 + 2 * g];
  for (int blk = blockIdx.x; blk < n; blk += gridDim.x) {
    const int2 bm = fmap[off + blk];
    const inst_desc &in = d[bm.x];
    const int t = in.t + s;
    if (t >= in.nsweeps) continue;  // the whole block skips together
    sa::sweep_args a = in.a;
    if (in.per_sweep) a.tab += (int64_t)t * a.nsys;
    a.sweep = (uint64_t)t;
    const int cur = in.cur ^ (s & 1);  // every earlier sweep of the chunk was fused too
    a.s = cur ? in.s2 : in.a.s;
    uint64_t *out = cur ? in.a.s : in.s2;
    const int ntile = a.L / TH;
    const int sb = bm.y / ntile, tile = bm.y - sb * ntile;
    sa::fused_band<true, true, MB, TH>(a, out, sb, tile, lds);
    __syncthreads();  // the next band reuses the LDS
  }
}

// ---- measurement of listed systems of many instances
// An item is up to kMeasSites consecutive sites of one system (one block of 256
// threads: 4 waves x 16 groups of 64 sites); counts go to cnt[slot*128 + lane]
// (satisfied -x,-y[,-z] bonds) and cnt[slot*128 + 64 + lane] (set spins), as in
// the single-ladder measurement.
constexpr int kMeasSites = 4096;
struct meas_item {
  const uint64_t *s, *jx, *jy, *jz;
  int64_t N, begin;
  int kind, L, slot, pad;
};

template <int KIND>
__device__ __forceinline__ void measure_item(const meas_item &it, int lane, int wv, int &sat, int &up) {
  uint64_t e[6] = {0, 0, 0, 0, 0, 0}, mc[5] = {0, 0, 0, 0, 0};
  const int64_t L = it.L, L2 = L * L, N = it.N;
  // wave wv takes groups wv, wv+4, ... (16 groups): E <= 48 < 64, M <= 16 < 32
  for (int g = wv; g < kMeasSites / 64; g += 4) {
    const int64_t i = it.begin + (int64_t)g * 64 + lane;
    uint64_t w = 0, bw = 0, bn = 0, bu = 0;
    if (i < N) {
      const int64_t x = i % L, y = (i / L) % L, z = i / L2;
      const int64_t iw = i - x + (x == 0 ? L - 1 : x - 1);
      const int64_t inn = (y == 0) ? i + L2 - L : i - L;
      w = it.s[i];
      bw = (KIND == sa::KIND_2P && x == 0) ? 0 : (it.jx[i] ^ w ^ it.s[iw]);
      bn = it.jy[i] ^ w ^ it.s[inn];
      if (KIND == sa::KIND_3D) {
        const int64_t iu = (z == 0) ? i + N - L2 : i - L2;
        bu = it.jz[i] ^ w ^ it.s[iu];
      }
    }
    const uint64_t sb = bw ^ bn ^ bu, cb = sa::maj3(bw, bn, bu);
    uint64_t t = e[0] & sb;
    e[0] ^= sb;
    uint64_t u = sa::maj3(e[1], cb, t);
    e[1] ^= cb ^ t;
#pragma unroll
    for (int b = 2; b < 6; ++b) {
      const uint64_t v = e[b] & u;
      e[b] ^= u;
      u = v;
    }
    t = w;
#pragma unroll
    for (int b = 0; b < 5; ++b) {
      const uint64_t v = mc[b] & t;
      mc[b] ^= t;
      t = v;
    }
  }
  sat = up = 0;
#pragma unroll
  for (int b = 0; b < 6; ++b) sat += __popcll(sa::wave_transpose64(e[b], lane)) << b;
#pragma unroll
  for (int b = 0; b < 5; ++b) up += __popcll(sa::wave_transpose64(mc[b], lane)) << b;
}

__global__ void __launch_bounds__(256) k_batch_measure(const meas_item *__restrict__ items, int *cnt) {
  __shared__ int red[2][64];
  const meas_item it = items[blockIdx.x];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (threadIdx.x < 128) (&red[0][0])[threadIdx.x] = 0;
  __syncthreads();
  int sat = 0, up = 0;
  if (it.kind == SA_KIND_3D) measure_item<sa::KIND_3D>(it, lane, wv, sat, up);
  else if (it.kind == SA_KIND_2D) measure_item<sa::KIND_2D>(it, lane, wv, sat, up);
  else measure_item<sa::KIND_2P>(it, lane, wv, sat, up);
  atomicAdd(&red[0][lane], sat);
  atomicAdd(&red[1][lane], up);
  __syncthreads();
  if (threadIdx.x < 128) atomicAdd(&cnt[it.slot * 128 + threadIdx.x], (&red[0][0])[threadIdx.x]);
}

// ---- masked swaps of many pairs: a ^= (a^b)&m, b ^= (a^b)&m (src/device.cuh:18-23)
constexpr int kSwapWords = 4096;
struct swap_item {
  uint64_t *a, *b;
  uint64_t mask;
  int64_t begin, end;
};

__global__ void __launch_bounds__(256) k_batch_swap(const swap_item *__restrict__ items) {
  const swap_item it = items[blockIdx.x];
  for (int64_t i = it.begin + threadIdx.x; i < it.end; i += blockDim.x) {
    const uint64_t x = it.a[i], y = it.b[i];
    const uint64_t d = (x ^ y) & it.mask;
    it.a[i] = x ^ d;
    it.b[i] = y ^ d;
  }
}

// pinned host staging that grows on demand (contents are not kept)
struct pinned {
  void *p = nullptr;
  size_t cap = 0;
  ~pinned() {
    if (p) (void)hipHostFree(p);
  }
  bool ensure(size_t bytes) {
    if (bytes <= cap) return true;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
    const size_t want = std::max<size_t>(bytes * 2, 4096);
    if (hipHostMalloc(&p, want, hipHostMallocDefault) != hipSuccess) {
      p = nullptr;
      return false;
    }
    cap = want;
    return true;
  }
};

// device array that grows on demand (contents are not kept)
template <class T>
struct dev_array {
  T *p = nullptr;
  size_t cap = 0;
  ~dev_array() {
    if (p) (void)hipFree(p);
  }
  // `moved`: the array has a new address (graphs that name it are stale)
  int ensure(size_t n, bool &moved) {
    moved = false;
    if (n <= cap) return SA_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    const size_t want = std::max<size_t>(n * 2, 64);
    if (hipMalloc(&p, want * sizeof(T)) != hipSuccess) {
      p = nullptr;
      return sa_fail(SA_ERR_NOMEM, "batch device array of %zu bytes", want * sizeof(T));
    }
    cap = want;
    moved = true;
    return SA_OK;
  }
};

struct pt_state {
  int every = 0;  // exchange every `every` sweeps; 0 = anneal instance
  std::vector<double> T;
  double h = 0;
  sa::mt19937 gen{};
  int uni_b = 0;
  int pend_i = 0;
  double pend_u = 0;
  std::vector<int32_t> pairs;
  std::vector<double> us;
  std::vector<uint64_t> masks;
  std::vector<float> E_min;
  std::vector<double> acc_sum;
  std::vector<int64_t> acc_n;
};

struct batch_inst {
  sa_ladder *lad = nullptr;
  uint64_t *s_first = nullptr;  // the ladder's spin array at attach (s_alt is the other)
  sa::sys_table *tabs = nullptr;
  int nsweeps = 0, per_sweep = 0, t = 0;
  bool fast = true, live = false;
  int64_t work = 0;   // work items per half-sweep (all systems)
  int64_t bytes = 0;  // device memory held
  pt_state pt;
};

// The fused 3D sweep inside the batch (k_batch_fused on forked graph branches) is
// switched off: one run of its test stopped making progress and no cause has been
// found in this code.  Off, 3D instances take the half-sweep kernels, no second
// spin array is allocated and SA_TUNE_BATCH_FUSED = 2 is refused.
constexpr bool kFusedBatch = false;

// online-tuning candidates: (fused 3D sweep 2 on / 1 off, threads per half-sweep
// block, sweeps per graph); the first chunks try them in this order
struct tune_cand {
  int fused, block, chunk;
};
const tune_cand kCands[2] = {{1, 128, 32}, {1, 64, 32}};
const int kNumCands = 2;

}  // namespace

struct sa_batch {
  hipStream_t stream = nullptr;
  int cus = 256;
  std::vector<batch_inst> inst;  // indexed by id
  // device state of the current schedule
  dev_array<inst_desc> d_desc;
  dev_array<int2> d_bmap;
  dev_array<int> d_ctrl;
  int ndesc = 0, nblocks = 0, bmap_bs = 0;
  std::vector<int> desc_id;  // instance id of each descriptor
  std::vector<char> desc_fused;  // descriptor swept by the fused kernel
  // fused 3D groups of the current schedule: (L, TH, MB) and block count each
  struct fgroup {
    int L, TH, MB, count;
  };
  std::vector<fgroup> groups;
  dev_array<int2> d_fmap;
  int sched_fused = -1;  // fused mode the schedule was built with
  bool dirty = true;
  std::map<int, hipGraphExec_t> graphs;  // chunk length -> graph over the current arrays, block size, groups
  int graph_bs = 0;
  std::vector<fgroup> graph_groups;
  std::vector<hipStream_t> side;  // side streams of the fused groups (fork/join inside the graph)
  std::vector<hipEvent_t> side_done;
  hipEvent_t fork = nullptr;
  // exchange staging
  pinned h_stage;
  dev_array<meas_item> d_meas;
  dev_array<swap_item> d_swap;
  dev_array<int> d_cnt;
  pinned h_cnt;
  // tuning: 0 = automatic; fused: 0 automatic, 1 off, 2 on
  int block = 0, chunk = 0, fused = 0;
  int tuned_block = 0, tuned_chunk = 0, tuned_fused = 0;
  std::vector<std::pair<int, float>> trials;  // (candidate, ms per work item)
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int64_t bytes_live = 0;
};

namespace {

void drop_graphs(sa_batch *b) {
  for (auto &g : b->graphs) (void)hipGraphExecDestroy(g.second);
  b->graphs.clear();
}

int check_id(const sa_batch *b, int id) {
  if (!b) return sa_fail(SA_ERR_ARG, "null batch");
  if (id < 0 || id >= (int)b->inst.size() || !b->inst[id].live) return sa_fail(SA_ERR_ARG, "unknown instance %d", id);
  return SA_OK;
}

int grid_for(const sa_batch *b, int bs) {
  // two blocks per wave slot at 4 waves/SIMD (the sweep kernel's register
  // budget): every CU filled, few enough blocks to loop
  return std::max(1, b->cus * 16 * 2 / (bs / 64));
}

// descriptors and block map of the unfinished instances, largest remaining work first
// fused 3D configuration of an instance: the tallest band (TH rows) that fits the
// kernel's launch bound and LDS, two systems per workgroup when nsys is even
bool fused_cfg(const sa_ladder *l, int &TH, int &MB) {
  if (l->kind != SA_KIND_3D || (int64_t)l->N * 8 >= ((int64_t)1 << 31)) return false;
  MB = (l->nsys % 2 == 0) ? 2 : 1;
  for (int th : {16, 8, 4}) {
    if (l->L % th || th > l->L || (th + 2) * (l->L / 2) > 640) continue;
    if ((size_t)3 * (th + 2) * MB * 2 * (l->L / 2) * 8 > 64 * 1024) continue;
    TH = th;
    return true;
  }
  return false;
}

// descriptors, the half-sweep block map and the fused groups' block maps of the
// unfinished instances, largest remaining work first
int rebuild(sa_batch *b, int bs, bool fused) {
  std::vector<int> ids;
  for (int i = 0; i < (int)b->inst.size(); ++i)
    if (b->inst[i].live && b->inst[i].t < b->inst[i].nsweeps) ids.push_back(i);
  std::stable_sort(ids.begin(), ids.end(), [&](int x, int y) {
    const auto &a = b->inst[x], &c = b->inst[y];
    return a.work * (int64_t)(a.nsweeps - a.t) > c.work * (int64_t)(c.nsweeps - c.t);
  });
  // which instances take the fused sweep, and their groups
  std::vector<sa_batch::fgroup> groups;
  std::vector<std::vector<int2>> gmaps;
  std::vector<int> gidx(ids.size(), -1);
  int64_t nblk = 0;
  for (size_t q = 0; q < ids.size(); ++q) {
    const batch_inst &in = b->inst[ids[q]];
    int TH = 0, MB = 0;
    if (fused && in.fast && in.lad->s_alt && fused_cfg(in.lad, TH, MB)) {
      int g = -1;
      for (size_t j = 0; j < groups.size(); ++j)
        if (groups[j].L == in.lad->L && groups[j].TH == TH && groups[j].MB == MB) g = (int)j;
      if (g < 0 && (int)groups.size() < kMaxFusedGroups) {
        g = (int)groups.size();
        groups.push_back({in.lad->L, TH, MB, 0});
        gmaps.emplace_back();
      }
      if (g >= 0) {
        gidx[q] = g;
        const int nb = (in.lad->L / TH) * (in.lad->nsys / MB);
        for (int k = 0; k < nb; ++k) gmaps[g].push_back(make_int2((int)q, k));
        groups[g].count += nb;
        continue;
      }
    }
    nblk += (in.work + bs - 1) / bs;
  }
  int64_t nf = 0;
  for (auto &m : gmaps) nf += (int64_t)m.size();
  if (nblk > INT32_MAX / 2 || nf > INT32_MAX / 2) return sa_fail(SA_ERR_ARG, "batch too large");
  // earlier launches may still read the old arrays and the staging buffer
  SA_HIP_TRY(hipStreamSynchronize(b->stream));
  bool m1 = false, m2 = false, m3 = false, m4 = false;
  int rc = b->d_desc.ensure(std::max<size_t>(ids.size(), 1), m1);
  if (!rc) rc = b->d_bmap.ensure(std::max<int64_t>(nblk, 1), m2);
  if (!rc) rc = b->d_ctrl.ensure(2 + 2 * kMaxFusedGroups, m3);
  if (!rc) rc = b->d_fmap.ensure(std::max<int64_t>(nf, 1), m4);
  if (rc) return rc;
  bool same_groups = groups.size() == b->graph_groups.size();
  for (size_t j = 0; same_groups && j < groups.size(); ++j)
    same_groups = groups[j].L == b->graph_groups[j].L && groups[j].TH == b->graph_groups[j].TH &&
                  groups[j].MB == b->graph_groups[j].MB && groups[j].count <= b->graph_groups[j].count;
  if (m1 || m2 || m3 || m4 || b->graph_bs != bs || !same_groups) {
    drop_graphs(b);
    b->graph_groups = groups;  // block counts at capture size the fused grids
  }
  b->graph_bs = bs;
  const int nctrl = 2 + 2 * kMaxFusedGroups;
  const size_t bytes = sizeof(int) * nctrl + sizeof(inst_desc) * ids.size() + sizeof(int2) * (nblk + nf) + 64;
  if (!b->h_stage.ensure(bytes)) return sa_fail(SA_ERR_NOMEM, "pinned staging");
  int *ctrl = (int *)b->h_stage.p;
  inst_desc *desc = (inst_desc *)((char *)b->h_stage.p + ((sizeof(int) * nctrl + 15) & ~(size_t)15));
  int2 *bmap = (int2 *)(desc + ids.size());
  int2 *fmap = bmap + nblk;
  int64_t q2 = 0;
  b->desc_fused.assign(ids.size(), 0);
  for (size_t q = 0; q < ids.size(); ++q) {
    const batch_inst &in = b->inst[ids[q]];
    const sa_ladder *l = in.lad;
    inst_desc d;
    std::memset(&d, 0, sizeof d);
    // a.s is the ladder's first array; the current one is l->s
    d.a.s = in.s_first;
    d.s2 = (l->s == in.s_first) ? l->s_alt : l->s;
    d.cur = (l->s == in.s_first) ? 0 : 1;
    d.a.jx = l->jx;
    d.a.jy = l->jy;
    d.a.jz = l->jz;
    d.a.tab = in.tabs;
    d.a.words = nullptr;
    d.a.seed = l->seed;
    d.a.sweep = 0;
    d.a.N = l->N;
    d.a.L = l->L;
    d.a.nsys = l->nsys;
    d.kind = l->kind;
    d.nsweeps = in.nsweeps;
    d.t = in.t;
    d.per_sweep = in.per_sweep;
    d.fast = in.fast ? 1 : 0;
    d.fused = gidx[q] >= 0 ? 1 : 0;
    desc[q] = d;
    b->desc_fused[q] = (char)d.fused;
    if (!d.fused)
      for (int64_t off = 0; off < in.work; off += bs) bmap[q2++] = make_int2((int)q, (int)off);
  }
  std::memset(ctrl, 0, sizeof(int) * nctrl);
  ctrl[0] = (int)nblk;
  ctrl[1] = (int)ids.size();
  int64_t fo = 0;
  for (size_t g = 0; g < gmaps.size(); ++g) {
    ctrl[2 + 2 * g] = (int)fo;
    ctrl[3 + 2 * g] = (int)gmaps[g].size();
    std::memcpy(fmap + fo, gmaps[g].data(), sizeof(int2) * gmaps[g].size());
    fo += (int64_t)gmaps[g].size();
  }
  SA_HIP_TRY(hipMemcpyAsync(b->d_ctrl.p, ctrl, sizeof(int) * nctrl, hipMemcpyHostToDevice, b->stream));
  if (!ids.empty()) {
    SA_HIP_TRY(hipMemcpyAsync(b->d_desc.p, desc, sizeof(inst_desc) * ids.size(), hipMemcpyHostToDevice, b->stream));
    if (nblk) SA_HIP_TRY(hipMemcpyAsync(b->d_bmap.p, bmap, sizeof(int2) * nblk, hipMemcpyHostToDevice, b->stream));
    if (nf) SA_HIP_TRY(hipMemcpyAsync(b->d_fmap.p, fmap, sizeof(int2) * nf, hipMemcpyHostToDevice, b->stream));
  }
  b->ndesc = (int)ids.size();
  b->desc_id = ids;
  b->nblocks = (int)nblk;
  b->bmap_bs = bs;
  b->groups = groups;
  b->sched_fused = fused ? 1 : 0;
  b->dirty = false;
  return SA_OK;
}

void launch_fused_group(sa_batch *b, int g, int s, hipStream_t st) {
  const sa_batch::fgroup &G = b->graph_groups[g];
  const dim3 grid((unsigned)std::max(1, G.count)), block((unsigned)((G.TH + 2) * (G.L / 2)));
  const size_t lds = (size_t)3 * (G.TH + 2) * G.MB * 2 * (G.L / 2) * 8;
#define SA_FUSED_CASE(MBV, THV)                                                                               \
  if (G.MB == MBV && G.TH == THV) {                                                                         \
    hipLaunchKernelGGL((k_batch_fused<MBV, THV>), grid, block, lds, st, b->d_desc.p, b->d_fmap.p,            \
                       b->d_ctrl.p, g, s);                                                                  \
    return;                                                                                                 \
  }
  SA_FUSED_CASE(1, 4)
  SA_FUSED_CASE(1, 8)
  SA_FUSED_CASE(1, 16)
  SA_FUSED_CASE(2, 4)
  SA_FUSED_CASE(2, 8)
  SA_FUSED_CASE(2, 16)
#undef SA_FUSED_CASE
}

// One sweep = the two half-sweep launches and one launch per fused group; the
// chunk ends with the advance.  The launches touch disjoint instances, so the fused groups go on side
// streams forked from and joined back to the batch stream (parallel branches of
// the captured graph) and overlap each other and the half-sweeps.
// side streams and events for ng fused groups (created outside any capture)
int ensure_side(sa_batch *b, int ng) {
  for (int g = (int)b->side.size(); g < ng; ++g) {
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr;
    SA_HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    SA_HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    b->side.push_back(st);
    b->side_done.push_back(ev);
  }
  if (!b->fork) SA_HIP_TRY(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
  return SA_OK;
}

int enqueue_sweeps(sa_batch *b, int k) {
  const int grid = grid_for(b, b->bmap_bs);
  const int ng = (int)b->graph_groups.size();
  for (int s = 0; s < k; ++s) {
    if (ng) SA_HIP_TRY(hipEventRecord(b->fork, b->stream));
    for (int g = 0; g < ng; ++g) {
      SA_HIP_TRY(hipStreamWaitEvent(b->side[g], b->fork, 0));
      launch_fused_group(b, g, s, b->side[g]);
      SA_HIP_TRY(hipEventRecord(b->side_done[g], b->side[g]));
    }
    for (int color = 0; color < 2; ++color)
      hipLaunchKernelGGL(k_batch_half, dim3(grid), dim3(b->bmap_bs), 0, b->stream, b->d_desc.p, b->d_bmap.p,
                         b->d_ctrl.p, color, s);
    for (int g = 0; g < ng; ++g) SA_HIP_TRY(hipStreamWaitEvent(b->stream, b->side_done[g], 0));
  }
  hipLaunchKernelGGL(k_batch_advance, dim3(1), dim3(256), 0, b->stream, b->d_desc.p, b->d_ctrl.p, k);
  return SA_OK;
}

// the graph of k sweeps over the current device arrays (captured on first use)
int graph_for(sa_batch *b, int k, hipGraphExec_t &out) {
  auto it = b->graphs.find(k);
  if (it != b->graphs.end()) {
    out = it->second;
    return SA_OK;
  }
  hipGraph_t g = nullptr;
  const int src = ensure_side(b, (int)b->graph_groups.size());
  if (src) return src;
  SA_HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  const int erc = enqueue_sweeps(b, k);
  const hipError_t ee = hipStreamEndCapture(b->stream, &g);
  if (erc) {
    if (g) (void)hipGraphDestroy(g);
    return erc;
  }
  SA_HIP_TRY(ee);
  hipGraphExec_t e = nullptr;
  const hipError_t rc = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (rc != hipSuccess) return sa_fail(SA_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(rc));
  b->graphs[k] = e;
  out = e;
  return SA_OK;
}

int64_t bonds(const sa_ladder *l) {
  return (l->kind == SA_KIND_3D ? 3 : 2) * l->N - (l->kind == SA_KIND_2P ? l->L : 0);
}

// measure systems sys[j] of instances ids[j]: E/M[j*64 + lane]
int measure_many(sa_batch *b, const std::vector<int> &ids, const std::vector<int> &sys, std::vector<int32_t> &E,
                 std::vector<int32_t> &M) {
  const int n = (int)ids.size();
  E.assign((size_t)n * 64, 0);
  M.assign((size_t)n * 64, 0);
  if (n == 0) return SA_OK;
  std::vector<meas_item> items;
  for (int j = 0; j < n; ++j) {
    const sa_ladder *l = b->inst[ids[j]].lad;
    for (int64_t beg = 0; beg < l->N; beg += kMeasSites) {
      meas_item it;
      std::memset(&it, 0, sizeof it);
      it.s = l->s + (int64_t)sys[j] * l->N;
      it.jx = l->jx;
      it.jy = l->jy;
      it.jz = l->jz;
      it.N = l->N;
      it.begin = beg;
      it.kind = l->kind;
      it.L = l->L;
      it.slot = j;
      items.push_back(it);
    }
  }
  SA_HIP_TRY(hipStreamSynchronize(b->stream));  // the staging buffer may still be read
  bool moved = false;
  int rc = b->d_meas.ensure(items.size(), moved);
  if (!rc) rc = b->d_cnt.ensure((size_t)n * 128, moved);
  if (rc) return rc;
  if (!b->h_stage.ensure(sizeof(meas_item) * items.size())) return sa_fail(SA_ERR_NOMEM, "pinned staging");
  if (!b->h_cnt.ensure(sizeof(int) * n * 128)) return sa_fail(SA_ERR_NOMEM, "pinned staging");
  std::memcpy(b->h_stage.p, items.data(), sizeof(meas_item) * items.size());
  SA_HIP_TRY(hipMemcpyAsync(b->d_meas.p, b->h_stage.p, sizeof(meas_item) * items.size(), hipMemcpyHostToDevice,
                            b->stream));
  hipLaunchKernelGGL(sa::k_fill_zero, dim3((n * 128 + 255) / 256), dim3(256), 0, b->stream, b->d_cnt.p, n * 128);
  hipLaunchKernelGGL(k_batch_measure, dim3((unsigned)items.size()), dim3(256), 0, b->stream, b->d_meas.p,
                     b->d_cnt.p);
  SA_HIP_TRY(hipGetLastError());
  SA_HIP_TRY(hipMemcpyAsync(b->h_cnt.p, b->d_cnt.p, sizeof(int) * n * 128, hipMemcpyDeviceToHost, b->stream));
  SA_HIP_TRY(hipStreamSynchronize(b->stream));
  const int *cnt = (const int *)b->h_cnt.p;
  for (int j = 0; j < n; ++j) {
    const sa_ladder *l = b->inst[ids[j]].lad;
    const int64_t nb = bonds(l);
    for (int k = 0; k < 64; ++k) {
      E[j * 64 + k] = (int32_t)(2 * (int64_t)cnt[j * 128 + k] - nb);
      M[j * 64 + k] = (int32_t)(2 * (int64_t)cnt[j * 128 + 64 + k] - l->N);
    }
  }
  return SA_OK;
}

// exchange step of every listed parallel-tempering instance (parallel_tempering::
// step after its sweeps, src/parallel_tempering.hpp:81-115)
int pt_exchange(sa_batch *b, const std::vector<int> &due) {
  if (due.empty()) return SA_OK;
  std::vector<int> ids, sys;
  for (int id : due) {
    batch_inst &in = b->inst[id];
    in.pt.pend_i = sa::uniform_int(in.pt.gen, 0, in.pt.uni_b);
    in.pt.pend_u = sa::canonical32(in.pt.gen);
    for (int s : {0, in.pt.pend_i, in.pt.pend_i + 1}) {
      ids.push_back(id);
      sys.push_back(s);
    }
  }
  std::vector<int32_t> E, M;
  int rc = measure_many(b, ids, sys, E, M);
  if (rc) return rc;
  std::vector<swap_item> items;
  for (size_t q = 0; q < due.size(); ++q) {
    batch_inst &in = b->inst[due[q]];
    pt_state &p = in.pt;
    float e[3][64];
    for (int r = 0; r < 3; ++r)
      for (int k = 0; k < 64; ++k)
        e[r][k] = (float)((double)E[(q * 3 + r) * 64 + k] + p.h * (double)M[(q * 3 + r) * 64 + k]);
    if (p.E_min.empty()) p.E_min.assign(e[0], e[0] + 64);
    for (int k = 0; k < 64; ++k) p.E_min[k] = std::max(e[0][k], p.E_min[k]);
    const int i = p.pend_i;
    const double dB = 1. / p.T[i] - 1. / p.T[i + 1];
    int acc = 0;
    const uint64_t mask = sa::exchange_word(e[1], e[2], dB, p.pend_u, 64, acc);
    p.acc_sum[i] += acc / 64.0;
    p.acc_n[i] += 1;
    p.pairs.push_back(i);
    p.us.push_back(p.pend_u);
    p.masks.push_back(mask);
    if (!mask) continue;
    const sa_ladder *l = in.lad;
    for (int64_t beg = 0; beg < l->N; beg += kSwapWords) {
      swap_item it;
      it.a = l->s + (int64_t)i * l->N;
      it.b = l->s + (int64_t)(i + 1) * l->N;
      it.mask = mask;
      it.begin = beg;
      it.end = std::min<int64_t>(l->N, beg + kSwapWords);
      items.push_back(it);
    }
  }
  if (items.empty()) return SA_OK;
  bool moved = false;
  rc = b->d_swap.ensure(items.size(), moved);
  if (rc) return rc;
  if (!b->h_stage.ensure(sizeof(swap_item) * items.size())) return sa_fail(SA_ERR_NOMEM, "pinned staging");
  std::memcpy(b->h_stage.p, items.data(), sizeof(swap_item) * items.size());
  SA_HIP_TRY(hipMemcpyAsync(b->d_swap.p, b->h_stage.p, sizeof(swap_item) * items.size(), hipMemcpyHostToDevice,
                            b->stream));
  hipLaunchKernelGGL(k_batch_swap, dim3((unsigned)items.size()), dim3(256), 0, b->stream, b->d_swap.p);
  SA_HIP_TRY(hipGetLastError());
  // every later writer of the staging buffer synchronises the stream first
  return SA_OK;
}

// host mirror of the device sweep indices after k more sweeps; collects the
// parallel-tempering instances whose exchange is due
void advance_host(sa_batch *b, int k, std::vector<int> &due) {
  for (size_t q = 0; q < b->desc_id.size(); ++q) {
    const int id = b->desc_id[q];
    batch_inst &in = b->inst[id];
    if (!in.live) continue;
    const int before = in.t;
    in.t = std::min(in.nsweeps, in.t + k);
    in.lad->t = (uint64_t)in.t;
    if (b->desc_fused[q] && ((in.t - before) & 1)) std::swap(in.lad->s, in.lad->s_alt);  // current array
    if (in.t >= in.nsweeps && before < in.nsweeps) b->dirty = true;
    if (in.pt.every && in.t > before && in.t % in.pt.every == 0) due.push_back(id);
  }
}

int sync_batch(sa_batch *b) {
  SA_HIP_TRY(hipStreamSynchronize(b->stream));
  return SA_OK;
}

// Everything an instance needs on the device, done without touching the batch (so
// it may run on any thread, while chunks execute): its ladder seeded like main,
// tables uploaded on the ladder's own stream.
int prepare_instance(int kind, int L, int nsys, uint64_t seed, int nsweeps, const double *T, int per_sweep, double h,
                     const pt_state *pt, batch_inst &in) {
  sa_ladder *l = nullptr;
  int rc = sa_ladder_create(kind, L, nsys, &l);
  if (rc) return rc;
  sa_ladder_set_rng(l, SA_RNG_COUNTER);
  // set_seed, init_J, set_seed, init_rand (src/main_gpu.cu:262-276)
  sa_ladder_seed(l, seed);
  rc = sa_ladder_init_couplings(l);
  if (!rc) rc = sa_ladder_seed(l, seed);
  if (!rc) rc = sa_ladder_init_spins(l);
  const int ntab = per_sweep ? std::max(1, nsweeps) * nsys : nsys;
  std::vector<sa::sys_table> tabs(ntab);
  bool fast = true;
  for (int q = 0; q < ntab && !rc; ++q) {
    bool f = false;
    sa_detail::make_table(kind, T[per_sweep && nsweeps == 0 ? q % nsys : q], h, tabs[q], f, nullptr);
    fast = fast && f;
  }
  // a second spin array for the fused 3D sweep (read one, write the other)
  int th_ = 0, mb_ = 0;
  if (!rc && kFusedBatch && fused_cfg(l, th_, mb_) && !l->s_alt && hipMalloc(&l->s_alt, (size_t)nsys * l->N * 8) != hipSuccess) {
    l->s_alt = nullptr;
    rc = sa_fail(SA_ERR_NOMEM, "second spin array");
  }
  in = batch_inst();
  in.lad = l;
  in.nsweeps = nsweeps;
  in.per_sweep = per_sweep ? 1 : 0;
  in.fast = fast;
  in.work = (int64_t)half_items(kind, L) * nsys;
  if (!rc && hipMalloc(&in.tabs, sizeof(sa::sys_table) * ntab) != hipSuccess) rc = sa_fail(SA_ERR_NOMEM, "tables");
  // on the ladder's stream (waited for below): admission may run while a chunk is
  // on the batch stream, and a null-stream copy would wait for it
  if (!rc && hipMemcpyAsync(in.tabs, tabs.data(), sizeof(sa::sys_table) * ntab, hipMemcpyHostToDevice,
                            l->stream) != hipSuccess)
    rc = sa_fail(SA_ERR_HIP, "table upload");
  // fixed temperatures also become the ladder's own tables, for plain sweeps later
  if (!rc && !per_sweep)
    for (int j = 0; j < nsys && !rc; ++j) rc = sa_ladder_set_temperature(l, j, T[j], h, nullptr);
  if (!rc) rc = sa_ladder_sync(l);
  if (rc) {
    if (in.tabs) (void)hipFree(in.tabs);
    sa_ladder_destroy(l);
    in = batch_inst();
    return rc;
  }
  in.bytes = sa_detail::batch_instance_bytes(kind, L, nsys, nsweeps, per_sweep);
  if (pt) in.pt = *pt;
  return SA_OK;
}

void attach_instance(sa_batch *b, batch_inst &&in, int *id) {
  in.live = true;
  in.s_first = in.lad->s;
  b->bytes_live += in.bytes;
  b->inst.push_back(std::move(in));
  *id = (int)b->inst.size() - 1;
  b->dirty = true;
}

int add_instance(sa_batch *b, int kind, int L, int nsys, uint64_t seed, int nsweeps, const double *T, int per_sweep,
                 double h, pt_state *pt, int *id) {
  batch_inst in;
  const int rc = prepare_instance(kind, L, nsys, seed, nsweeps, T, per_sweep, h, pt, in);
  if (rc) return rc;
  attach_instance(b, std::move(in), id);
  return SA_OK;
}

// parallel-tempering state as parallel_tempering's constructor sets it up
pt_state make_pt(int nT, const double *T, double h, int every) {
  pt_state p;
  p.every = every;
  p.T.assign(T, T + nT);
  std::sort(p.T.begin(), p.T.end());  // parallel_tempering sorts its temperatures
  p.h = h;
  p.gen.seed(123);
  p.uni_b = nT - 2;
  p.acc_sum.assign(nT - 1, 0.0);
  p.acc_n.assign(nT - 1, 0);
  return p;
}

}  // namespace

struct sa_detail::prepared_instance {
  batch_inst in;
};

namespace sa_detail {
int batch_geometry(int kind, int L, int nsys) {
  if (kind < SA_KIND_2P || kind > SA_KIND_3D) return sa_fail(SA_ERR_ARG, "unknown kind %d", kind);
  if (nsys < 1) return sa_fail(SA_ERR_ARG, "nsys must be >= 1");
  if (L < 3) return sa_fail(SA_ERR_UNSUPPORTED, "device path needs L >= 3 (got %d)", L);
  if ((kind == SA_KIND_3D && L > 1024) || (kind != SA_KIND_3D && L > 32768))
    return sa_fail(SA_ERR_ARG, "L=%d too large", L);
  // k_batch_half and the block map carry an instance's work item in an int
  if ((int64_t)half_items(kind, L) * nsys >= ((int64_t)1 << 31))
    return sa_fail(SA_ERR_UNSUPPORTED, "instance too large for the batch: %lld work items per half-sweep (limit 2^31)",
                   (long long)half_items(kind, L) * nsys);
  return SA_OK;
}

int batch_prepare(int kind, int L, int nsys, uint64_t seed, int nsweeps, const double *T, int per_sweep, double h,
                  int pt_every, prepared_instance **out) {
  *out = nullptr;
  int rc = batch_geometry(kind, L, nsys);
  if (rc) return rc;
  std::unique_ptr<prepared_instance> p(new prepared_instance());
  if (pt_every > 0) {
    if (nsys < 2 || nsweeps % pt_every) return sa_fail(SA_ERR_ARG, "bad parallel-tempering instance");
    pt_state pt = make_pt(nsys, T, h, pt_every);
    rc = prepare_instance(kind, L, nsys, seed, nsweeps, pt.T.data(), 0, h, &pt, p->in);
  } else {
    rc = prepare_instance(kind, L, nsys, seed, nsweeps, T, per_sweep, h, nullptr, p->in);
  }
  if (rc) return rc;
  *out = p.release();
  return SA_OK;
}

int batch_attach(sa_batch *b, prepared_instance *p, int *id) {
  attach_instance(b, std::move(p->in), id);
  delete p;
  return SA_OK;
}

void batch_discard(prepared_instance *p) {
  if (!p) return;
  if (p->in.tabs) (void)hipFree(p->in.tabs);
  sa_ladder_destroy(p->in.lad);
  delete p;
}

int64_t batch_instance_bytes(int kind, int L, int nsys, int nsweeps, int per_sweep) {
  const int ntab = per_sweep ? std::max(1, nsweeps) * nsys : nsys;
  const int64_t N = (int64_t)L * L * (kind == SA_KIND_3D ? L : 1);
  const bool second = kFusedBatch && kind == SA_KIND_3D && L % 4 == 0;  // fused-sweep ping-pong array
  return ladder_bytes(kind, L, nsys) + (int64_t)sizeof(sa::sys_table) * ntab + (second ? (int64_t)nsys * N * 8 : 0);
}

int batch_launch_chunk(sa_batch *b, int max_sweeps, int *k_out) {
  *k_out = 0;
  // fixed by the caller, tuned, or the next trial
  int trial = -1;
  if ((!b->block || !b->chunk || !b->fused) && !(b->tuned_block && b->tuned_chunk && b->tuned_fused)) {
    trial = (int)b->trials.size();
    if (trial >= kNumCands) {
      int best = 0;
      for (int q = 1; q < kNumCands; ++q)
        if (b->trials[q].second < b->trials[best].second) best = q;
      b->tuned_fused = kCands[best].fused;
      b->tuned_block = kCands[best].block;
      b->tuned_chunk = kCands[best].chunk;
      trial = -1;
    }
  }
  const int bs = b->block ? b->block : (trial >= 0 ? kCands[trial].block : b->tuned_block);
  const int kfull = b->chunk ? b->chunk : (trial >= 0 ? kCands[trial].chunk : b->tuned_chunk);
  const int fmode = b->fused ? b->fused : (trial >= 0 ? kCands[trial].fused : b->tuned_fused);
  const bool fz = fmode == 2;
  if (b->dirty || b->bmap_bs != bs || b->sched_fused != (fz ? 1 : 0)) {
    const int rc = rebuild(b, bs, fz);
    if (rc) return rc;
  }
  if (b->ndesc == 0 || max_sweeps <= 0) return SA_OK;
  // the chunk ends at the next exchange point of any parallel-tempering instance
  int64_t k = std::min(kfull, max_sweeps);
  int remaining = 0;
  for (int id : b->desc_id) {
    const batch_inst &in = b->inst[id];
    remaining = std::max(remaining, in.nsweeps - in.t);
    if (in.pt.every) k = std::min<int64_t>(k, in.pt.every - in.t % in.pt.every);
  }
  k = std::min<int64_t>(k, remaining);
  if (k <= 0) return SA_OK;
  hipGraphExec_t g = nullptr;
  int rc = graph_for(b, (int)k, g);
  if (rc) return rc;
  if (trial >= 0) SA_HIP_TRY(hipEventRecord(b->ev0, b->stream));
  SA_HIP_TRY(hipGraphLaunch(g, b->stream));
  if (trial >= 0) {
    SA_HIP_TRY(hipEventRecord(b->ev1, b->stream));
    SA_HIP_TRY(hipEventSynchronize(b->ev1));
    float ms = 0;
    SA_HIP_TRY(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    // time per work item, so trials on different schedules compare
    double work = 0;
    for (int id : b->desc_id)
      work += (double)b->inst[id].work * std::min<int64_t>(k, b->inst[id].nsweeps - b->inst[id].t);
    b->trials.push_back({trial, (float)(ms / std::max(work, 1.0))});
  }
  *k_out = (int)k;
  return SA_OK;
}

int batch_finish_chunk(sa_batch *b, int k) {
  std::vector<int> due;
  advance_host(b, k, due);
  return pt_exchange(b, due);
}

int batch_wait(sa_batch *b) { return sync_batch(b); }

sa_ladder *batch_instance_ladder(sa_batch *b, int id) {
  return (id >= 0 && id < (int)b->inst.size() && b->inst[id].live) ? b->inst[id].lad : nullptr;
}

bool batch_done(const sa_batch *b, int id) {
  return id >= 0 && id < (int)b->inst.size() && b->inst[id].live && b->inst[id].t >= b->inst[id].nsweeps;
}
}  // namespace sa_detail

extern "C" {

int sa_batch_create(sa_batch **out) {
  if (!out) return sa_fail(SA_ERR_ARG, "null output handle");
  sa_batch *b = new sa_batch();
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess &&
      prop.multiProcessorCount > 0)
    b->cus = prop.multiProcessorCount;
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&b->ev0) != hipSuccess ||
      hipEventCreate(&b->ev1) != hipSuccess) {
    delete b;
    return sa_fail(SA_ERR_HIP, "stream/event creation failed");
  }
  *out = b;
  return SA_OK;
}

int sa_batch_destroy(sa_batch *b) {
  if (!b) return SA_OK;
  (void)hipStreamSynchronize(b->stream);
  drop_graphs(b);
  for (auto &in : b->inst)
    if (in.live) {
      if (in.tabs) (void)hipFree(in.tabs);
      sa_ladder_destroy(in.lad);
    }
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->fork) (void)hipEventDestroy(b->fork);
  for (auto e : b->side_done) (void)hipEventDestroy(e);
  for (auto st : b->side) (void)hipStreamDestroy(st);
  hipStream_t s = b->stream;
  delete b;  // device arrays and staging buffers free themselves
  (void)hipStreamDestroy(s);
  return SA_OK;
}

int sa_batch_add(sa_batch *b, int kind, int L, int nsys, uint64_t seed, int nsweeps, const double *T, int per_sweep,
                 double h, int *id) {
  const int rc = sa_detail::batch_geometry(kind, L, nsys);  // before anything touches the device
  if (rc) return rc;
  if (!b || !T || !id || nsweeps < 0) return sa_fail(SA_ERR_ARG, "bad argument");
  return add_instance(b, kind, L, nsys, seed, nsweeps, T, per_sweep, h, nullptr, id);
}

int sa_batch_add_pt(sa_batch *b, int kind, int L, int nT, const double *T, double h, uint64_t seed, int every,
                    int nsteps, int *id) {
  const int rc = sa_detail::batch_geometry(kind, L, nT);  // before anything touches the device
  if (rc) return rc;
  if (!b || !T || !id || nT < 2 || every < 1 || nsteps < 0) return sa_fail(SA_ERR_ARG, "bad argument");
  if ((int64_t)every * nsteps > INT32_MAX / 2) return sa_fail(SA_ERR_ARG, "too many sweeps");
  pt_state p = make_pt(nT, T, h, every);
  return add_instance(b, kind, L, nT, seed, every * nsteps, p.T.data(), 0, h, &p, id);
}

int sa_batch_run(sa_batch *b, int sweeps) {
  if (!b) return sa_fail(SA_ERR_ARG, "null batch");
  int64_t left = sweeps > 0 ? sweeps : INT64_MAX;
  while (left > 0) {
    int k = 0;
    int rc = sa_detail::batch_launch_chunk(b, (int)std::min<int64_t>(left, INT32_MAX), &k);
    if (rc) return rc;
    if (k == 0) break;
    rc = sa_detail::batch_finish_chunk(b, k);
    if (rc) return rc;
    left -= k;
  }
  int rc = sync_batch(b);
  if (rc) return rc;
  int unfinished = 0;
  for (auto &in : b->inst)
    if (in.live && in.t < in.nsweeps) ++unfinished;
  return unfinished;
}

int sa_batch_sweeps_done(sa_batch *b, int id, int *t) {
  int rc = check_id(b, id);
  if (rc) return rc;
  if (!t) return sa_fail(SA_ERR_ARG, "null argument");
  *t = b->inst[id].t;
  return SA_OK;
}

int sa_batch_get_spins(sa_batch *b, int id, int sys, uint64_t *s) {
  int rc = check_id(b, id);
  if (rc) return rc;
  rc = sync_batch(b);
  if (rc) return rc;
  return sa_ladder_get_spins(b->inst[id].lad, sys, s);
}

int sa_batch_measure(sa_batch *b, int id, int32_t *E, int32_t *M) {
  int rc = check_id(b, id);
  if (rc) return rc;
  rc = sync_batch(b);
  if (rc) return rc;
  return sa_ladder_measure(b->inst[id].lad, E, M);
}

int sa_batch_pt_result(sa_batch *b, int id, int64_t capacity, int32_t *pairs, double *us, uint64_t *masks,
                       float *E_min, double *acceptance) {
  int rc = check_id(b, id);
  if (rc) return rc;
  const pt_state &p = b->inst[id].pt;
  if (!p.every) return sa_fail(SA_ERR_STATE, "instance %d is not a parallel-tempering instance", id);
  const int64_t n = std::min<int64_t>(capacity, (int64_t)p.pairs.size());
  for (int64_t k = 0; k < n; ++k) {
    if (pairs) pairs[k] = p.pairs[k];
    if (us) us[k] = p.us[k];
    if (masks) masks[k] = p.masks[k];
  }
  if (E_min)
    for (int k = 0; k < 64; ++k) E_min[k] = p.E_min.empty() ? 0.f : p.E_min[k];
  if (acceptance)
    for (size_t i = 0; i < p.acc_sum.size(); ++i) acceptance[i] = p.acc_n[i] ? p.acc_sum[i] / p.acc_n[i] : 0.0;
  return (int)std::min<int64_t>((int64_t)p.pairs.size(), INT32_MAX);
}

sa_ladder *sa_batch_ladder(sa_batch *b, int id) {
  if (check_id(b, id)) return nullptr;
  if (sync_batch(b)) return nullptr;
  b->dirty = true;  // the caller may sweep it (the fused 3D sweep swaps its spin arrays)
  return b->inst[id].lad;
}

int sa_batch_remove(sa_batch *b, int id) {
  int rc = check_id(b, id);
  if (rc) return rc;
  rc = sync_batch(b);
  if (rc) return rc;
  batch_inst &in = b->inst[id];
  if (in.tabs) (void)hipFree(in.tabs);
  sa_ladder_destroy(in.lad);
  b->bytes_live -= in.bytes;
  in = batch_inst();
  b->dirty = true;
  return SA_OK;
}

int64_t sa_batch_bytes(const sa_batch *b) { return b ? b->bytes_live : 0; }

void *sa_batch_stream(sa_batch *b) { return b ? (void *)b->stream : nullptr; }

int sa_batch_set_tuning(sa_batch *b, int key, int value) {
  if (!b) return sa_fail(SA_ERR_ARG, "null batch");
  if (key == SA_TUNE_BATCH_BLOCK) {
    if (value != 0 && value != 64 && value != 128 && value != 256)
      return sa_fail(SA_ERR_ARG, "batch block must be 0 (auto), 64, 128 or 256");
    b->block = value;
  } else if (key == SA_TUNE_BATCH_CHUNK) {
    if (value < 0 || value > 4096) return sa_fail(SA_ERR_ARG, "batch chunk must be 0 (auto) .. 4096");
    b->chunk = value;
  } else if (key == SA_TUNE_BATCH_FUSED) {
    if (value < 0 || value > 2) return sa_fail(SA_ERR_ARG, "batch fused must be 0 (auto), 1 (off) or 2 (on)");
    if (value == 2 && !kFusedBatch) return sa_fail(SA_ERR_ARG, "the fused batch sweep is switched off in this build");
    b->fused = value;
  } else {
    return sa_fail(SA_ERR_ARG, "unknown batch tuning key %d", key);
  }
  b->dirty = true;
  return SA_OK;
}

int sa_batch_get_tuning(sa_batch *b, int key, int *value) {
  if (!b || !value) return sa_fail(SA_ERR_ARG, "null argument");
  if (key == SA_TUNE_BATCH_BLOCK) *value = b->block ? b->block : b->tuned_block;
  else if (key == SA_TUNE_BATCH_CHUNK) *value = b->chunk ? b->chunk : b->tuned_chunk;
  else if (key == SA_TUNE_BATCH_FUSED) *value = b->fused ? b->fused : b->tuned_fused;
  else return sa_fail(SA_ERR_ARG, "unknown batch tuning key %d", key);
  return SA_OK;
}

}  // extern "C"
