// sa_batch.hip -- many independent instances of uneven size in one stream of
// launches (the C ABI sa_batch_* of include/sa_hip.h).
//
// Each instance is a ladder (couplings, spins, nsys systems) with its own
// temperature schedule and its own counter-RNG seed.  A half-sweep launch covers
// every unfinished instance: the grid is the concatenation of the instances' work
// items (3D: (column, system), 2D: (site, system)); a block belongs to one
// instance (block map), reads the instance's descriptor and runs the same
// per-column / per-site code as the single-instance kernels, so an instance's
// spins are bit for bit those of running it alone.  The instance's sweep index
// lives in its device descriptor and is advanced by a one-thread-per-instance
// kernel after the second half, so a chunk of sweeps is a fixed sequence of
// launches, captured once as a hipGraph and replayed until an instance finishes
// and the block map changes.  Block size and chunk length are tuned online: the
// first chunks try the candidates (results do not depend on them) and the
// fastest is kept.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <vector>

#include "sa_internal.hpp"

namespace {

struct inst_desc {
  sa::sweep_args a;  // tab: the instance's tables (per sweep: nsweeps x nsys, else nsys)
  int kind, nsweeps, t, per_sweep, fast, pad;
};

template <bool FAST>
__device__ __forceinline__ void batch_item(const inst_desc &in, int color, int k, int sys) {
  sa::sweep_args a = in.a;
  if (in.per_sweep) a.tab += (int64_t)in.t * a.nsys;
  a.sweep = (uint64_t)in.t;
  if (in.kind == SA_KIND_3D) sa::half3d_column<true, FAST, 1>(a, color, k, sys);
  else if (in.kind == SA_KIND_2D) sa::half2d_site<true, FAST, false, 1>(a, color, k, sys);
  else sa::half2d_site<true, false, true, 1>(a, color, k, sys);
}

__global__ void __launch_bounds__(256) k_batch_half(const inst_desc *__restrict__ d, const int2 *__restrict__ bmap,
                                                    int color) {
  const int2 bm = bmap[blockIdx.x];
  const inst_desc &in = d[bm.x];
  if (in.t >= in.nsweeps) return;
  const int L = in.a.L, work = L * (L >> 1);
  const int item = bm.y + (int)threadIdx.x;
  if (item >= work * in.a.nsys) return;
  const int sys = item / work, k = item - sys * work;
  if (in.fast) batch_item<true>(in, color, k, sys);
  else batch_item<false>(in, color, k, sys);
}

__global__ void k_batch_advance(inst_desc *d, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && d[i].t < d[i].nsweeps) d[i].t += 1;
}

struct batch_inst {
  sa_ladder *lad = nullptr;
  sa::sys_table *tabs = nullptr;
  int nsweeps = 0, per_sweep = 0, t = 0;
  bool fast = true, live = false;
  int64_t work = 0;  // work items per half-sweep (all systems)
};

const int kBlocks[3] = {64, 128, 256};
const int kChunks[2] = {8, 32};

}  // namespace

struct sa_batch {
  hipStream_t stream = nullptr;
  std::vector<batch_inst> inst;  // indexed by id
  // device state of the current schedule
  inst_desc *d_desc = nullptr;
  int desc_cap = 0, ndesc = 0;
  std::vector<int> desc_id;  // instance id of each descriptor
  int2 *d_bmap = nullptr;
  int bmap_cap = 0, nblocks = 0, bmap_bs = 0;
  bool dirty = true;
  hipGraphExec_t exec = nullptr;
  int exec_k = 0;
  // tuning: 0 = automatic
  int block = 0, chunk = 0;
  int tuned_block = 0, tuned_chunk = 0;
  std::vector<std::pair<int, float>> trials;  // (candidate, ms per sweep)
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

namespace {

void drop_graph(sa_batch *b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  b->exec = nullptr;
  b->exec_k = 0;
}

int check_id(const sa_batch *b, int id) {
  if (!b) return sa_fail(SA_ERR_ARG, "null batch");
  if (id < 0 || id >= (int)b->inst.size() || !b->inst[id].live) return sa_fail(SA_ERR_ARG, "unknown instance %d", id);
  return SA_OK;
}

// descriptors and block map of the unfinished instances, largest first
int rebuild(sa_batch *b, int bs) {
  std::vector<int> ids;
  for (int i = 0; i < (int)b->inst.size(); ++i)
    if (b->inst[i].live && b->inst[i].t < b->inst[i].nsweeps) ids.push_back(i);
  std::stable_sort(ids.begin(), ids.end(), [&](int x, int y) {
    const auto &a = b->inst[x], &c = b->inst[y];
    return a.work * (int64_t)(a.nsweeps - a.t) > c.work * (int64_t)(c.nsweeps - c.t);
  });
  // earlier launches still read the old descriptors and block map
  SA_HIP_TRY(hipStreamSynchronize(b->stream));
  std::vector<inst_desc> desc(ids.size());
  std::vector<int2> bmap;
  for (size_t q = 0; q < ids.size(); ++q) {
    const batch_inst &in = b->inst[ids[q]];
    const sa_ladder *l = in.lad;
    inst_desc &d = desc[q];
    d.a.s = l->s;
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
    d.pad = 0;
    for (int64_t off = 0; off < in.work; off += bs) bmap.push_back(make_int2((int)q, (int)off));
  }
  if ((int)desc.size() > b->desc_cap) {
    if (b->d_desc) (void)hipFree(b->d_desc);
    b->d_desc = nullptr;
    const int cap = std::max<int>(64, (int)desc.size() * 2);
    if (hipMalloc(&b->d_desc, sizeof(inst_desc) * cap) != hipSuccess) return sa_fail(SA_ERR_NOMEM, "batch descriptors");
    b->desc_cap = cap;
  }
  if ((int)bmap.size() > b->bmap_cap) {
    if (b->d_bmap) (void)hipFree(b->d_bmap);
    b->d_bmap = nullptr;
    const int cap = std::max<int>(1024, (int)bmap.size() * 2);
    if (hipMalloc(&b->d_bmap, sizeof(int2) * cap) != hipSuccess) return sa_fail(SA_ERR_NOMEM, "batch block map");
    b->bmap_cap = cap;
  }
  if (!desc.empty()) {
    SA_HIP_TRY(hipMemcpyAsync(b->d_desc, desc.data(), sizeof(inst_desc) * desc.size(), hipMemcpyHostToDevice,
                              b->stream));
    SA_HIP_TRY(hipMemcpyAsync(b->d_bmap, bmap.data(), sizeof(int2) * bmap.size(), hipMemcpyHostToDevice, b->stream));
    SA_HIP_TRY(hipStreamSynchronize(b->stream));  // the host vectors end here
  }
  b->ndesc = (int)desc.size();
  b->desc_id = ids;
  b->nblocks = (int)bmap.size();
  b->bmap_bs = bs;
  b->dirty = false;
  drop_graph(b);
  return SA_OK;
}

void enqueue_sweeps(sa_batch *b, int k) {
  for (int s = 0; s < k; ++s) {
    for (int color = 0; color < 2; ++color)
      hipLaunchKernelGGL(k_batch_half, dim3(b->nblocks), dim3(b->bmap_bs), 0, b->stream, b->d_desc, b->d_bmap, color);
    hipLaunchKernelGGL(k_batch_advance, dim3((b->ndesc + 255) / 256), dim3(256), 0, b->stream, b->d_desc, b->ndesc);
  }
}

// the graph of k sweeps of the current schedule (captured on first use)
int prepare_graph(sa_batch *b, int k) {
  if (b->exec && b->exec_k == k) return SA_OK;
  drop_graph(b);
  hipGraph_t g = nullptr;
  SA_HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  enqueue_sweeps(b, k);
  SA_HIP_TRY(hipStreamEndCapture(b->stream, &g));
  const hipError_t e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    b->exec = nullptr;
    return sa_fail(SA_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
  }
  b->exec_k = k;
  return SA_OK;
}

// host mirror of the device sweep indices after k more sweeps
void advance_host(sa_batch *b, int k) {
  for (int id : b->desc_id) {
    batch_inst &in = b->inst[id];
    const int before = in.t;
    in.t = std::min(in.nsweeps, in.t + k);
    in.lad->t = (uint64_t)in.t;
    if (in.t >= in.nsweeps && before < in.nsweeps) b->dirty = true;
  }
}

int sync_batch(sa_batch *b) {
  SA_HIP_TRY(hipStreamSynchronize(b->stream));
  return SA_OK;
}

}  // namespace

extern "C" {

int sa_batch_create(sa_batch **out) {
  if (!out) return sa_fail(SA_ERR_ARG, "null output handle");
  sa_batch *b = new sa_batch();
  if (hipStreamCreate(&b->stream) != hipSuccess || hipEventCreate(&b->ev0) != hipSuccess ||
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
  drop_graph(b);
  for (auto &in : b->inst)
    if (in.live) {
      if (in.tabs) (void)hipFree(in.tabs);
      sa_ladder_destroy(in.lad);
    }
  if (b->d_desc) (void)hipFree(b->d_desc);
  if (b->d_bmap) (void)hipFree(b->d_bmap);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  (void)hipStreamDestroy(b->stream);
  delete b;
  return SA_OK;
}

int sa_batch_add(sa_batch *b, int kind, int L, int nsys, uint64_t seed, int nsweeps, const double *T, int per_sweep,
                 double h, int *id) {
  if (!b || !T || !id || nsweeps < 0) return sa_fail(SA_ERR_ARG, "bad argument");
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
  batch_inst in;
  in.lad = l;
  in.nsweeps = nsweeps;
  in.per_sweep = per_sweep ? 1 : 0;
  in.fast = fast;
  in.work = (int64_t)L * (L / 2) * nsys;
  if (!rc && hipMalloc(&in.tabs, sizeof(sa::sys_table) * ntab) != hipSuccess) rc = sa_fail(SA_ERR_NOMEM, "tables");
  if (!rc && hipMemcpy(in.tabs, tabs.data(), sizeof(sa::sys_table) * ntab, hipMemcpyHostToDevice) != hipSuccess)
    rc = sa_fail(SA_ERR_HIP, "table upload");
  if (!rc) rc = sa_ladder_sync(l);
  if (rc) {
    if (in.tabs) (void)hipFree(in.tabs);
    sa_ladder_destroy(l);
    return rc;
  }
  // fixed temperatures also become the ladder's own tables, for plain sweeps later
  if (!per_sweep)
    for (int j = 0; j < nsys; ++j) sa_ladder_set_temperature(l, j, T[j], h, nullptr);
  in.live = true;
  b->inst.push_back(in);
  *id = (int)b->inst.size() - 1;
  b->dirty = true;
  return SA_OK;
}

int sa_batch_run(sa_batch *b, int sweeps) {
  if (!b) return sa_fail(SA_ERR_ARG, "null batch");
  const int ncand = 3 * 2;
  int64_t left = sweeps > 0 ? sweeps : INT64_MAX;
  while (left > 0) {
    // block size and chunk length: fixed by the caller, tuned, or the next trial
    int trial = -1;
    if ((!b->block || !b->chunk) && !(b->tuned_block && b->tuned_chunk)) {
      trial = (int)b->trials.size();
      if (trial >= ncand) {
        int best = 0;
        for (int q = 1; q < ncand; ++q)
          if (b->trials[q].second < b->trials[best].second) best = q;
        b->tuned_block = kBlocks[best / 2];
        b->tuned_chunk = kChunks[best % 2];
        continue;
      }
    }
    const int bs = b->block ? b->block : (trial >= 0 ? kBlocks[trial / 2] : b->tuned_block);
    const int kfull = b->chunk ? b->chunk : (trial >= 0 ? kChunks[trial % 2] : b->tuned_chunk);
    if (b->dirty || b->bmap_bs != bs) {
      const int rc = rebuild(b, bs);
      if (rc) return rc;
    }
    if (b->ndesc == 0) break;
    int remaining = 0;
    for (int id : b->desc_id) remaining = std::max(remaining, b->inst[id].nsweeps - b->inst[id].t);
    const int k = (int)std::min<int64_t>({(int64_t)kfull, left, (int64_t)remaining});
    if (k == kfull) {
      const int rc = prepare_graph(b, k);
      if (rc) return rc;
    }
    if (trial >= 0) SA_HIP_TRY(hipEventRecord(b->ev0, b->stream));
    if (k == kfull) {
      SA_HIP_TRY(hipGraphLaunch(b->exec, b->stream));
    } else {  // a short tail: plain launches
      enqueue_sweeps(b, k);
      SA_HIP_TRY(hipGetLastError());
    }
    if (trial >= 0) {
      SA_HIP_TRY(hipEventRecord(b->ev1, b->stream));
      SA_HIP_TRY(hipEventSynchronize(b->ev1));
      float ms = 0;
      SA_HIP_TRY(hipEventElapsedTime(&ms, b->ev0, b->ev1));
      // time per work item, so trials on different schedules compare
      double work = 0;
      for (int id : b->desc_id) work += (double)b->inst[id].work * std::min(k, b->inst[id].nsweeps - b->inst[id].t);
      b->trials.push_back({trial, (float)(ms / std::max(work, 1.0))});
    }
    advance_host(b, k);
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
  in = batch_inst();
  b->dirty = true;
  return SA_OK;
}

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
  } else {
    return sa_fail(SA_ERR_ARG, "unknown batch tuning key %d", key);
  }
  drop_graph(b);
  b->dirty = true;
  return SA_OK;
}

int sa_batch_get_tuning(sa_batch *b, int key, int *value) {
  if (!b || !value) return sa_fail(SA_ERR_ARG, "null argument");
  if (key == SA_TUNE_BATCH_BLOCK) *value = b->block ? b->block : b->tuned_block;
  else if (key == SA_TUNE_BATCH_CHUNK) *value = b->chunk ? b->chunk : b->tuned_chunk;
  else return sa_fail(SA_ERR_ARG, "unknown batch tuning key %d", key);
  return SA_OK;
}

}  // extern "C"
