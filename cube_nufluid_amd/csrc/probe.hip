// HBM streaming probe: the copy bandwidth a plain, well-formed streaming
// kernel reaches on this GPU, used as the practical ceiling beside the
// 8 TB/s peak when the step's kernels are scored against HBM.  Each thread
// moves UNR 16-byte vectors per grid-stride iteration, the loads issued
// before the stores; the nontemporal variant marks both as streaming.
#include "cube_internal.hpp"

typedef float v4f __attribute__((ext_vector_type(4)));

template <int UNR, bool NT>
__global__ __launch_bounds__(256) void k_stream_copy(const v4f* __restrict__ src, v4f* __restrict__ dst, int64_t n4) {
  const int64_t stride = (int64_t)gridDim.x * 256 * UNR;
  for (int64_t base = (int64_t)blockIdx.x * 256 * UNR + threadIdx.x; base < n4; base += stride) {
    v4f v[UNR];
#pragma unroll
    for (int u = 0; u < UNR; u++) {
      const int64_t i = base + (int64_t)u * 256;
      if (i < n4) v[u] = NT ? __builtin_nontemporal_load(src + i) : src[i];
    }
#pragma unroll
    for (int u = 0; u < UNR; u++) {
      const int64_t i = base + (int64_t)u * 256;
      if (i < n4) {
        if (NT)
          __builtin_nontemporal_store(v[u], dst + i);
        else
          dst[i] = v[u];
      }
    }
  }
}

extern "C" int cube_hbm_copy_probe(int device, int64_t bytes, int iters, double* gbps) {
  CUBE_EXCLUSIVE;
  if (bytes < (1 << 20) || bytes % 16 || iters < 1 || !gbps) return CUBE_ERR_ARG;
  if (hipSetDevice(device) != hipSuccess) return CUBE_ERR_HIP;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return CUBE_ERR_HIP;
  void *src = nullptr, *dst = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = CUBE_OK;
  const int64_t n4 = bytes / 16;
  if (hipMalloc(&src, bytes) != hipSuccess || hipMalloc(&dst, bytes) != hipSuccess || hipStreamCreate(&st) != hipSuccess ||
      hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess ||
      hipMemsetAsync(src, 1, bytes, st) != hipSuccess || hipMemsetAsync(dst, 0, bytes, st) != hipSuccess) {
    rc = CUBE_ERR_HIP;
  } else {
    // gbps[0..3]: best copy rate (read + write bytes per second / 1e9) of
    // {plain, nontemporal} x {8 and 32 workgroups per CU}
    const int cus = prop.multiProcessorCount;
    for (int var = 0; var < 4 && rc == CUBE_OK; var++) {
      const unsigned grid = (unsigned)(cus * (var & 1 ? 32 : 8));
      double best = 0;
      for (int it = 0; it < iters + 1 && rc == CUBE_OK; it++) {
        (void)hipEventRecord(e0, st);
        if (var < 2)
          k_stream_copy<4, false><<<grid, 256, 0, st>>>((const v4f*)src, (v4f*)dst, n4);
        else
          k_stream_copy<4, true><<<grid, 256, 0, st>>>((const v4f*)src, (v4f*)dst, n4);
        (void)hipEventRecord(e1, st);
        if (hipEventSynchronize(e1) != hipSuccess) rc = CUBE_ERR_HIP;
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (it > 0 && ms > 0) best = fmax(best, 2.0 * (double)bytes / (ms * 1e-3) / 1e9);
      }
      gbps[var] = best;
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  if (src) (void)hipFree(src);
  if (dst) (void)hipFree(dst);
  return rc;
}
