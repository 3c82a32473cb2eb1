// Device-wide exclusive scan of per-cell counts (the role of spine_tile /
// spine_image, variables.f90:103-218) and deterministic reductions.
#include "cube_internal.hpp"
#include <vector>

#define SCAN_NT 256
#define SCAN_PER 8
#define SCAN_BLK (SCAN_NT * SCAN_PER)

__device__ __forceinline__ int64_t block_excl_scan(int64_t v, int64_t* tmp, int64_t* total) {
  // Hillis-Steele over SCAN_NT entries in LDS (tmp has SCAN_NT slots)
  int t = threadIdx.x;
  tmp[t] = v;
  __syncthreads();
  for (int off = 1; off < SCAN_NT; off <<= 1) {
    int64_t a = t >= off ? tmp[t - off] : 0;
    __syncthreads();
    tmp[t] += a;
    __syncthreads();
  }
  int64_t incl = tmp[t];
  *total = tmp[SCAN_NT - 1];
  __syncthreads();
  return incl - v;
}

__global__ void k_scan_block_sums(const int32_t* cnt, int64_t n, int64_t* bsum) {
  __shared__ int64_t tmp[SCAN_NT];
  int64_t b0 = (int64_t)blockIdx.x * SCAN_BLK;
  int64_t s = 0;
  for (int q = 0; q < SCAN_PER; q++) {
    int64_t i = b0 + (int64_t)threadIdx.x * SCAN_PER + q;
    if (i < n) s += cnt[i];
  }
  int64_t total;
  block_excl_scan(s, tmp, &total);
  if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}
// single workgroup: exclusive scan of nb block sums in place (chunked)
__global__ void k_scan_top(int64_t* bsum, int64_t nb) {
  __shared__ int64_t tmp[SCAN_NT];
  int64_t carry = 0;
  for (int64_t c0 = 0; c0 < nb; c0 += SCAN_NT) {
    int64_t i = c0 + threadIdx.x;
    int64_t v = i < nb ? bsum[i] : 0;
    int64_t total;
    int64_t ex = block_excl_scan(v, tmp, &total);
    if (i < nb) bsum[i] = ex + carry;
    carry += total;
    __syncthreads();
  }
  if (threadIdx.x == 0) bsum[nb] = carry;
}
__global__ void k_scan_apply(const int32_t* cnt, int64_t n, const int64_t* bsum, int64_t* out) {
  __shared__ int64_t tmp[SCAN_NT];
  int64_t b0 = (int64_t)blockIdx.x * SCAN_BLK;
  int64_t loc[SCAN_PER];
  int64_t s = 0;
  for (int q = 0; q < SCAN_PER; q++) {
    int64_t i = b0 + (int64_t)threadIdx.x * SCAN_PER + q;
    loc[q] = s;
    if (i < n) s += cnt[i];
  }
  int64_t total;
  int64_t ex = block_excl_scan(s, tmp, &total);
  int64_t base = bsum[blockIdx.x] + ex;
  for (int q = 0; q < SCAN_PER; q++) {
    int64_t i = b0 + (int64_t)threadIdx.x * SCAN_PER + q;
    if (i < n) out[i] = base + loc[q];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = bsum[gridDim.x];
}

// cstart[0..n] = exclusive prefix of cnt[0..n-1], cstart[n] = total
int scan_counts(cube_handle* h, const int32_t* cnt, int64_t n, int64_t* cstart) {
  int64_t nb = (n + SCAN_BLK - 1) / SCAN_BLK;
  if (h->scan_tmp_n < nb + 1) {
    if (h->scan_tmp) (void)hipFree(h->scan_tmp);
    HIP_TRY(h, hipMalloc(&h->scan_tmp, sizeof(int64_t) * (nb + 1)));
    h->scan_tmp_n = nb + 1;
  }
  k_scan_block_sums<<<(unsigned)nb, SCAN_NT, 0, h->stream>>>(cnt, n, h->scan_tmp);
  k_scan_top<<<1, SCAN_NT, 0, h->stream>>>(h->scan_tmp, nb);
  k_scan_apply<<<(unsigned)nb, SCAN_NT, 0, h->stream>>>(cnt, n, h->scan_tmp, cstart);
  HIP_TRY(h, hipGetLastError());
  return 0;
}

// Sum nparts x nfields doubles (field-major within a part) in a fixed order:
// one workgroup per field, fixed stride partition, fixed tree.
__global__ void k_reduce_fields(const double* part, int64_t nparts, int nfields, double* out) {
  __shared__ double tmp[SCAN_NT];
  int f = blockIdx.x;
  double s = 0;
  for (int64_t i = threadIdx.x; i < nparts; i += SCAN_NT) s += part[i * nfields + f];
  tmp[threadIdx.x] = s;
  __syncthreads();
  for (int off = SCAN_NT / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) tmp[threadIdx.x] += tmp[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[f] = tmp[0];
}
// Large inputs in two levels: RL_BLOCKS workgroups per field each sum a
// fixed contiguous chunk, then one workgroup per field sums the chunks; the
// partition depends only on nparts, so the order is fixed
#define RL_BLOCKS 256
__global__ void k_reduce_chunks(const double* part, int64_t nparts, int nfields, double* l1) {
  __shared__ double tmp[SCAN_NT];
  const int f = blockIdx.y;
  const int64_t chunk = (nparts + RL_BLOCKS - 1) / RL_BLOCKS;
  const int64_t i0 = (int64_t)blockIdx.x * chunk, i1 = min(nparts, i0 + chunk);
  double s = 0;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += SCAN_NT) s += part[i * nfields + f];
  tmp[threadIdx.x] = s;
  __syncthreads();
  for (int off = SCAN_NT / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) tmp[threadIdx.x] += tmp[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) l1[(int64_t)blockIdx.x * nfields + f] = tmp[0];
}
int reduce_doubles_dev(cube_handle* h, const double* part, int64_t nparts, int nfields, double* dev_out) {
  if (nparts > 64 * 1024 && nfields <= 8 && h->dred_l1) {
    k_reduce_chunks<<<dim3(RL_BLOCKS, nfields), SCAN_NT, 0, h->stream>>>(part, nparts, nfields, h->dred_l1);
    k_reduce_fields<<<nfields, SCAN_NT, 0, h->stream>>>(h->dred_l1, RL_BLOCKS, nfields, dev_out);
  } else {
    k_reduce_fields<<<nfields, SCAN_NT, 0, h->stream>>>(part, nparts, nfields, dev_out);
  }
  HIP_TRY(h, hipGetLastError());
  return 0;
}
int reduce_doubles(cube_handle* h, const double* part, int64_t nparts, int nfields, double* host_out) {
  double* dout = h->dred + h->dred_n - 16;  // tail of the partials buffer
  int r = reduce_doubles_dev(h, part, nparts, nfields, dout);
  if (r) return r;
  HIP_TRY(h, hipMemcpyAsync(host_out, dout, sizeof(double) * nfields, hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  return 0;
}
