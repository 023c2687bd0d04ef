// sa_noalloc.hip -- sa_measure_noalloc (include/sa_hip.h): the measurement of
// sa_ladder_measure on caller-owned device arrays, with no allocation of any kind.
//
// The measurement kernels need a counter buffer (kMeasureCopies x nsys x 128 ints of
// partial sums).  Here it is `scratch`, a file-scope device array of fixed size:
// each GPU that loads the library holds its own copy until the process ends.  Host
// code never passes the array's name to a kernel (on the host it is only a stand-in
// for the symbol); the device address comes from hipGetSymbolAddress, taken once per
// device on the first call there, with that device current.  The counters are
// cleared by a fill kernel (no memset node, so the call can be captured in a graph)
// and summed by a second kernel that writes E and M to device memory.
#include <hip/hip_runtime.h>

#include <mutex>

#include "sa_internal.hpp"

namespace {

constexpr int kScratchInts = sa::kMeasureCopies * SA_NOALLOC_MAX_SYSTEMS * 128;
static __device__ __align__(256) int scratch[kScratchInts];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
int *g_scratch[kMaxDevices] = {};  // device address of `scratch` per device (nullptr: not looked up yet)

// E = 2*satisfied - bonds, M = 2*set - N per listed system and lane, summing the copies
__global__ void __launch_bounds__(128) k_noalloc_finish(const int *__restrict__ cnt, int nsys, int64_t bonds, int64_t N,
                                                        int32_t *__restrict__ E, int32_t *__restrict__ M) {
  const int j = blockIdx.x, q = threadIdx.x;  // q < 64: bond count of lane q; q >= 64: spin count of lane q-64
  int64_t sum = 0;
  for (int c = 0; c < sa::kMeasureCopies; ++c) sum += cnt[((int64_t)c * nsys + j) * 128 + q];
  if (q < 64)
    E[j * 64 + q] = (int32_t)(2 * sum - bonds);
  else
    M[j * 64 + q - 64] = (int32_t)(2 * sum - N);
}

}  // namespace

extern "C" int sa_measure_noalloc(int kind, int L, int nsys, const uint64_t *s, const uint64_t *jx, const uint64_t *jy,
                                  const uint64_t *jz, int32_t *E, int32_t *M, void *stream) {
  if (kind < SA_KIND_2P || kind > SA_KIND_3D) return sa_fail(SA_ERR_ARG, "unknown kind %d", kind);
  if (L < 3) return sa_fail(SA_ERR_UNSUPPORTED, "device path needs L >= 3 (got %d)", L);
  if (nsys < 1 || nsys > SA_NOALLOC_MAX_SYSTEMS)
    return sa_fail(SA_ERR_ARG, "sa_measure_noalloc: nsys=%d outside [1, %d] (the fixed counter buffer's size)", nsys,
                   SA_NOALLOC_MAX_SYSTEMS);
  if (!s || !jx || !jy || (kind == SA_KIND_3D && !jz) || !E || !M)
    return sa_fail(SA_ERR_ARG, "sa_measure_noalloc: null array");
  const int64_t N = (int64_t)L * L * (kind == SA_KIND_3D ? L : 1);
  const int64_t bonds = (kind == SA_KIND_3D ? 3 : 2) * N - (kind == SA_KIND_2P ? L : 0);
  // per-lane counts are 32-bit: every bond count must stay below 2^31
  if (bonds >= ((int64_t)1 << 31)) return sa_fail(SA_ERR_ARG, "sa_measure_noalloc: L=%d too large", L);
  int dev = 0;
  SA_HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return sa_fail(SA_ERR_ARG, "device %d beyond %d", dev, kMaxDevices);
  int *cnt = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
      void *p = nullptr;
      const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !p)
        return sa_fail(SA_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
      g_scratch[dev] = static_cast<int *>(p);
    }
    cnt = g_scratch[dev];
  }
  const hipStream_t st = static_cast<hipStream_t>(stream);
  sa_detail::launch_measure(kind, L, N, s, jx, jy, jz, nsys, nullptr, sa::sys4{{-1, -1, -1, -1}}, cnt, st);
  hipLaunchKernelGGL(k_noalloc_finish, dim3(nsys), dim3(128), 0, st, cnt, nsys, bonds, N, E, M);
  SA_HIP_TRY(hipGetLastError());
  return SA_OK;
}
