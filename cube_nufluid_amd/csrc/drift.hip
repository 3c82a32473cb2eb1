// Fixed-point drift, re-binning and velocity re-quantisation
// (update_particle.f90:13-259).
//
// The reference processes each tile's extended region (physical cells plus
// an ncb-cell ghost shell) in the order ilayer -> k -> j -> i -> l
// (update_particle.f90:57-63, 99-105), counts arrivals per destination cell,
// accumulates the new coarse-cell mean velocity in that order, and places
// each particle by its arrival rank.  Ghost particles are periodic copies, so
// a particle that ends in physical cell d of tile T is the copy that T sees
// at tile-local source cell d - delta.  This file reproduces that result
// without copies and for any displacement up to 7 coarse cells:
//   pass 1  (per particle): destination cell and displacement delta, counted
//           per destination cell;
//   scan    : new cell offsets in checkpoint order;
//   scatter (per particle): a sort key (source layer/plane/row/column in the
//           destination tile's frame, then particle index) into the
//           destination cell's segment;
//   sort    : each segment sorted by key (per thread, or per workgroup for
//           crowded cells), which is the reference's arrival order;
//   write   (per destination cell): vfield_new accumulated in that order,
//           then xp/vp/pid at their arrival rank.
// All arithmetic on positions and velocities is the reference's: double xq,
// single-precision tan through an exact host table, double atan through
// host-exact thresholds, nint half away from zero, narrowing with wrap.
#include "cube_internal.hpp"
#include <string.h>
#include <algorithm>

#define DR_NT 256
#define DR_CH 256
#define DMAX 7

__device__ __forceinline__ uint16_t pack_disp(int dx, int dy, int dz) {
  return (uint16_t)((dx + 8) | ((dy + 8) << 4) | ((dz + 8) << 8));
}

__global__ __launch_bounds__(DR_NT) void k_drift_pass1(Geo g, const void* __restrict__ xp, const void* __restrict__ vp,
                                                       const int64_t* __restrict__ cstart, int64_t n3c,
                                                       const float* __restrict__ vfield, float dt_mid, double c_old,
                                                       const float* __restrict__ vdec, int dmax, uint16_t* disp,
                                                       int32_t* cnt_new, uint32_t* flags) {
  __shared__ int64_t cs[DR_CH + 1];
  __shared__ int smax[3];
  __shared__ int sbad, sconf;
  const int64_t r0 = (int64_t)blockIdx.x * DR_CH;
  for (int i = threadIdx.x; i <= DR_CH; i += DR_NT) {
    int64_t r = r0 + i;
    cs[i] = cstart[r < n3c ? r : n3c];
  }
  if (threadIdx.x < 3) smax[threadIdx.x] = 0;
  if (threadIdx.x == 0) { sbad = 0; sconf = 0; }
  __syncthreads();
  int lmax[3] = {0, 0, 0}, bad = 0, conf = 0;
  for (int64_t p = cs[0] + threadIdx.x; p < cs[DR_CH]; p += DR_NT) {
    int lo = 0, hi = DR_CH;
    while (hi - lo > 1) {
      int mid = (lo + hi) >> 1;
      if (cs[mid] <= p) lo = mid; else hi = mid;
    }
    int64_t r = r0 + lo;
    int32_t c[3];
    cell_of_rank(g, r, c);
    int del[3];
    for (int e = 0; e < 3; e++) {
      int il = c[e] % g.nt + 1;  // tile-local 1-based source index
      double xo = xoff(load_xu(xp, 3 * p + e, g.izipx), g.izipx);
      double vreal = vdecode(vdec, load_v(vp, 3 * p + e, g.izipv), g.izipv, c_old) + (double)vfield[r * 3 + e];
      double deltax = ((double)dt_mid * vreal) / (double)CUBE_NCELL;
      int ig = (int)ceil(((double)il - 1.0) + xo + deltax);
      int d = ig - il;
      if (ig < 1 || ig > g.nt) {
        // the destination lies in the neighbouring tile: that tile computes
        // this particle from its own frame (source index il -+ nt)
        int sl = ig < 1 ? il + g.nt : il - g.nt;
        int ig2 = (int)ceil(((double)sl - 1.0) + xo + deltax);
        int d2 = ig2 - sl;
        if (d2 != d) {
          conf++;
          if (ig2 >= 1 && ig2 <= g.nt) d = d2;
        }
      }
      del[e] = d;
      int ad = d < 0 ? -d : d;
      if (ad > dmax) bad = 1;
      lmax[e] = max(lmax[e], ad);
    }
    if (bad) continue;
    disp[p] = pack_disp(del[0], del[1], del[2]);
    int64_t rd = rank_of(g, wrapc(c[0] + del[0], g.nc), wrapc(c[1] + del[1], g.nc), wrapc(c[2] + del[2], g.nc));
    atomicAdd(&cnt_new[rd], 1);
  }
  for (int e = 0; e < 3; e++) atomicMax(&smax[e], lmax[e]);
  if (bad) atomicOr(&sbad, 1);
  if (conf) atomicAdd(&sconf, conf);
  __syncthreads();
  if (threadIdx.x < 3) atomicMax(&flags[threadIdx.x], (uint32_t)smax[threadIdx.x]);
  if (threadIdx.x == 0) {
    if (sbad) atomicOr(&flags[3], 1u);
    if (sconf) atomicAdd(&flags[4], (uint32_t)sconf);
  }
}

// ---------------------------------------------------------------- scatter
// Every arrival gets a 64-bit key in its destination cell's segment of
// `keys` (cs_new order):
//   bits 44..63  ilayer of the source z-plane in the destination tile's frame
//   bits 32..43  (7-dz, 7-dy, 7-dx), 4 bits each: source plane, row, column
//                ascending in that frame
//   bits  0..31  source particle index (its position l inside the source cell)
// Sorting a segment by key gives the reference's arrival order
// ilayer -> k -> j -> i -> l (update_particle.f90:99-105).  Slots come from
// atomics and so are in no particular order before the sort.
__device__ __forceinline__ uint64_t arrival_key(int ilayer, int dx, int dy, int dz, int64_t p) {
  return ((uint64_t)ilayer << 44) | ((uint64_t)(7 - dz) << 40) | ((uint64_t)(7 - dy) << 36) | ((uint64_t)(7 - dx) << 32) |
         (uint64_t)(uint32_t)p;
}
__device__ __forceinline__ void key_disp(uint64_t k, int d[3]) {
  d[0] = 7 - (int)((k >> 32) & 15);
  d[1] = 7 - (int)((k >> 36) & 15);
  d[2] = 7 - (int)((k >> 40) & 15);
}

#define DR_BIG 256  // cells with more arrivals are sorted by a whole workgroup

__global__ __launch_bounds__(DR_NT) void k_drift_scatter(Geo g, const int64_t* __restrict__ cstart, int64_t n3c,
                                                         const uint16_t* __restrict__ disp, const int64_t* __restrict__ cs_new,
                                                         int nlayer, int32_t* cnt, uint64_t* keys, int32_t* big,
                                                         uint32_t* nbig) {
  __shared__ int64_t cs[DR_CH + 1];
  const int64_t r0 = (int64_t)blockIdx.x * DR_CH;
  for (int i = threadIdx.x; i <= DR_CH; i += DR_NT) {
    int64_t r = r0 + i;
    cs[i] = cstart[r < n3c ? r : n3c];
  }
  __syncthreads();
  for (int64_t p = cs[0] + threadIdx.x; p < cs[DR_CH]; p += DR_NT) {
    int lo = 0, hi = DR_CH;
    while (hi - lo > 1) {
      int mid = (lo + hi) >> 1;
      if (cs[mid] <= p) lo = mid; else hi = mid;
    }
    int32_t c[3];
    cell_of_rank(g, r0 + lo, c);
    const uint32_t dc = disp[p];
    const int d[3] = {(int)(dc & 15) - 8, (int)((dc >> 4) & 15) - 8, (int)((dc >> 8) & 15) - 8};
    const int cz = wrapc(c[2] + d[2], g.nc);
    const int64_t rd = rank_of(g, wrapc(c[0] + d[0], g.nc), wrapc(c[1] + d[1], g.nc), cz);
    const int sz = cz % g.nt + 1 - d[2];  // source plane in the destination tile's frame
    const int ilayer = (sz - 1 + g.ncb) % nlayer;
    const int32_t slot = atomicSub(&cnt[rd], 1) - 1;
    keys[cs_new[rd] + slot] = arrival_key(ilayer, d[0], d[1], d[2], p);
    if (slot == DR_BIG) big[atomicAdd(nbig, 1u)] = (int32_t)rd;
  }
}

// ---------------------------------------------------------------- big cells
// One workgroup per listed cell (grid-stride over the list).  Segments up to
// DR_SORT_LDS keys are bitonic-sorted in LDS; longer ones are sorted in
// DR_SORT_LDS chunks and then merged pairwise between `keys` and `tmp`
// (each element's output position is its index in its own run plus its rank
// in the partner run, so a pass is one parallel step).
#define DR_SORT_NT 1024
#define DR_SORT_LDS 8192
__device__ void block_bitonic(uint64_t* s, int n, int npow2) {
  for (int i = n + threadIdx.x; i < npow2; i += DR_SORT_NT) s[i] = ~0ull;
  __syncthreads();
  for (int k = 2; k <= npow2; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < npow2; i += DR_SORT_NT) {
        const int l = i ^ j;
        if (l > i) {
          const uint64_t a = s[i], b = s[l];
          const bool up = (i & k) == 0;
          if ((a > b) == up) { s[i] = b; s[l] = a; }
        }
      }
      __syncthreads();
    }
}

__global__ __launch_bounds__(DR_SORT_NT) void k_drift_sort_big(const int64_t* __restrict__ cs_new, const int32_t* __restrict__ big,
                                                               const uint32_t* __restrict__ nbig, uint64_t* keys, uint64_t* tmp) {
  __shared__ uint64_t s[DR_SORT_LDS];
  const uint32_t nb = *nbig;
  for (uint32_t bi = blockIdx.x; bi < nb; bi += gridDim.x) {
    const int64_t r = big[bi];
    const int64_t base = cs_new[r];
    const int64_t n = cs_new[r + 1] - base;
    // sort runs of DR_SORT_LDS in LDS
    for (int64_t c0 = 0; c0 < n; c0 += DR_SORT_LDS) {
      const int m = (int)min((int64_t)DR_SORT_LDS, n - c0);
      int np2 = 1;
      while (np2 < m) np2 <<= 1;
      for (int i = threadIdx.x; i < m; i += DR_SORT_NT) s[i] = keys[base + c0 + i];
      __syncthreads();
      block_bitonic(s, m, np2);
      for (int i = threadIdx.x; i < m; i += DR_SORT_NT) keys[base + c0 + i] = s[i];
      __syncthreads();
    }
    // merge passes (keys are unique: the low word is the particle index)
    uint64_t* src = keys + base;
    uint64_t* dst = tmp + base;
    for (int64_t run = DR_SORT_LDS; run < n; run <<= 1) {
      __threadfence();
      __syncthreads();
      for (int64_t i = threadIdx.x; i < n; i += DR_SORT_NT) {
        const int64_t a0 = (i / (2 * run)) * (2 * run);
        const int64_t b0 = min(a0 + run, n), b1 = min(a0 + 2 * run, n);
        const uint64_t v = src[i];
        int64_t lo, hi;  // rank of v in the partner run
        if (i < b0) { lo = b0; hi = b1; } else { lo = a0; hi = b0; }
        const int64_t p0 = lo;
        while (lo < hi) {
          const int64_t m = (lo + hi) >> 1;
          if (src[m] < v) lo = m + 1; else hi = m;
        }
        const int64_t own = i < b0 ? i - a0 : i - b0;
        dst[a0 + own + (lo - p0)] = v;
      }
      uint64_t* t = src; src = dst; dst = t;
    }
    __threadfence();
    __syncthreads();
    if (src != keys + base)
      for (int64_t i = threadIdx.x; i < n; i += DR_SORT_NT) keys[base + i] = src[i];
    __threadfence();
    __syncthreads();
  }
}

// ---------------------------------------------------------------- write
// One thread per destination cell; a workgroup's DR_NT consecutive cells own
// one contiguous key range, staged in LDS when it fits.  Each thread sorts
// its own segment (cells above DR_BIG arrive sorted), sums vfield_new in
// arrival order, then writes xp/vp/pid at their slots and the sigma_v sums.
#define DR_WCAP 6144
__device__ __forceinline__ void seg_sort(uint64_t* a, int n) {
  for (int i = 1; i < n; i++) {
    const uint64_t v = a[i];
    int j = i - 1;
    while (j >= 0 && a[j] > v) { a[j + 1] = a[j]; j--; }
    a[j + 1] = v;
  }
}

__global__ __launch_bounds__(DR_NT) void k_drift_write(Geo g, const void* __restrict__ xp, const void* __restrict__ vp,
                                                       const int32_t* __restrict__ pid, const float* __restrict__ vfield,
                                                       const int64_t* __restrict__ cs_new, uint64_t* keys, int64_t n3c,
                                                       float dt_mid, double c_old, const float* __restrict__ vdec,
                                                       const double* __restrict__ venc, int32_t kmax, void* xp_o, void* vp_o,
                                                       int32_t* pid_o, float* vfield_o, int32_t* rhoc_o, double* part) {
  __shared__ uint64_t sk[DR_WCAP];
  __shared__ double red[3][DR_NT / 64];
  const int64_t r0 = (int64_t)blockIdx.x * DR_NT;
  const int64_t rdst = r0 + threadIdx.x;
  const int64_t kb0 = cs_new[r0], kb1 = cs_new[min(r0 + DR_NT, n3c)];
  const bool staged = kb1 - kb0 <= DR_WCAP;
  if (staged)
    for (int64_t i = kb0 + threadIdx.x; i < kb1; i += DR_NT) sk[i - kb0] = keys[i];
  __syncthreads();
  double sum_c = 0, sum_r = 0, sum_s = 0;
  const double wv = (double)0.1f;  // weight_v=0.1 is a default-real literal
  const double xscale = (g.izipx == 1 ? 1.0 / 256.0 : 1.0 / 65536.0) * (double)CUBE_NCELL;
  if (rdst < n3c) {
    const int64_t base = cs_new[rdst];
    const int32_t nnew = (int32_t)(cs_new[rdst + 1] - base);
    uint64_t* kk = staged ? sk + (base - kb0) : keys + base;
    if (nnew <= DR_BIG) seg_sort(kk, nnew);
    int32_t c[3];
    cell_of_rank(g, rdst, c);
    float vf[3];
    for (int e = 0; e < 3; e++) vf[e] = (float)((double)vfield[rdst * 3 + e] * wv);
    for (int32_t k = 0; k < nnew; k++) {
      const uint64_t key = kk[k];
      const int64_t p = (int64_t)(uint32_t)key;
      int d[3];
      key_disp(key, d);
      const int64_t rs = rank_of(g, wrapc(c[0] - d[0], g.nc), wrapc(c[1] - d[1], g.nc), wrapc(c[2] - d[2], g.nc));
      for (int e = 0; e < 3; e++) {
        double vreal = vdecode(vdec, load_v(vp, 3 * p + e, g.izipv), g.izipv, c_old) + (double)vfield[rs * 3 + e];
        vf[e] = (float)((double)vf[e] + vreal);
      }
    }
    float vfn[3];
    for (int e = 0; e < 3; e++) vfn[e] = (float)((double)vf[e] / ((double)nnew + wv));
    for (int32_t k = 0; k < nnew; k++) {
      const uint64_t key = kk[k];
      const int64_t p = (int64_t)(uint32_t)key, q = base + k;
      int d[3];
      key_disp(key, d);
      const int64_t rs = rank_of(g, wrapc(c[0] - d[0], g.nc), wrapc(c[1] - d[1], g.nc), wrapc(c[2] - d[2], g.nc));
      double vr[3];
      for (int e = 0; e < 3; e++) {
        double vreal = vdecode(vdec, load_v(vp, 3 * p + e, g.izipv), g.izipv, c_old) + (double)vfield[rs * 3 + e];
        int64_t xn = (int64_t)load_xu(xp, 3 * p + e, g.izipx);
        xn = g.izipx == 1 ? (int64_t)(int8_t)xn : (int64_t)(int16_t)xn;
        xn += (int64_t)nint_d((double)dt_mid * vreal / xscale);
        store_x(xp_o, 3 * q + e, g.izipx, xn);
        int32_t code_v = vencode(venc, kmax, g.izipv, c_old, vreal - (double)vfn[e]);
        store_v(vp_o, 3 * q + e, g.izipv, code_v);
        vr[e] = vdecode(vdec, code_v, g.izipv, c_old);
      }
      pid_o[q] = pid[p];
      sum_r += vr[0] * vr[0] + vr[1] * vr[1] + vr[2] * vr[2];
      for (int e = 0; e < 3; e++) vr[e] += (double)vfn[e];
      sum_s += vr[0] * vr[0] + vr[1] * vr[1] + vr[2] * vr[2];
    }
    for (int e = 0; e < 3; e++) vfield_o[rdst * 3 + e] = vfn[e];
    rhoc_o[rdst] = nnew;
    sum_c = (double)(vfn[0] * vfn[0] + vfn[1] * vfn[1] + vfn[2] * vfn[2]);
  }
  double v3[3] = {sum_c, sum_r, sum_s};
  for (int f = 0; f < 3; f++) {
    double v = v3[f];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if ((threadIdx.x & 63) == 0) red[f][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    double t = 0;
    for (int i = 0; i < DR_NT / 64; i++) t += red[threadIdx.x][i];
    part[(int64_t)blockIdx.x * 3 + threadIdx.x] = t;
  }
}

int drift_run(cube_handle* h, float dt_old, float dt, cube_drift_out* out) {
  Geo g = geo_of(h);
  const float dt_mid = (dt_old + dt) / 2;
  const int nlayer = 2 * (int)ceilf(dt_mid * h->vz_max / (float)CUBE_NCELL) + 1;
  const double c_old = vcoef(h->sigma_vi);
  const int dmax = (int)(h->ncb < DMAX ? h->ncb : DMAX);
  uint32_t* flags = h->uflags;  // [0..2] max disp, [3] too far, [4] frame conflicts, [6] cells above DR_BIG
  HIP_TRY(h, hipMemsetAsync(flags, 0, sizeof(uint32_t) * 8, h->stream));
  HIP_TRY(h, hipMemsetAsync(h->rhoc2, 0, sizeof(int32_t) * h->n3c, h->stream));
  int64_t nb1 = (h->n3c + DR_CH - 1) / DR_CH;
  int pid1 = prof_begin(h, PS_DRIFT1);
  k_drift_pass1<<<(unsigned)nb1, DR_NT, 0, h->stream>>>(g, h->xp, h->vp, h->cstart, h->n3c, h->vfield, dt_mid, c_old, h->vdec, dmax,
                                                         h->disp, h->rhoc2, flags);
  prof_end(h, pid1);
  HIP_TRY(h, hipGetLastError());
  uint32_t hf[8];
  HIP_TRY(h, hipMemcpyAsync(hf, flags, sizeof(hf), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  if (hf[3])
    return cube_fail(h, CUBE_ERR_CAPACITY, "a particle moved more than %d coarse cells in one drift (buffer ncb=%lld)", dmax,
                     (long long)h->ncb);
  int pid2 = prof_begin(h, PS_DRIFT_SCAN);
  int r = scan_counts(h, h->rhoc2, h->n3c, h->cstart2);
  prof_end(h, pid2);
  if (r) return r;
  int64_t nb3 = (h->n3c + DR_NT - 1) / DR_NT;
  if (nb3 * 3 > h->dred_n - 16) return cube_fail(h, CUBE_ERR_ARG, "reduction buffer too small");
  uint64_t* keys = (uint64_t*)h->order;
  uint32_t* nbig = flags + 6;
  int pid3 = prof_begin(h, PS_DRIFT_SCATTER);
  k_drift_scatter<<<(unsigned)nb1, DR_NT, 0, h->stream>>>(g, h->cstart, h->n3c, h->disp, h->cstart2, nlayer, h->rhoc2, keys,
                                                           h->dbig, nbig);
  prof_end(h, pid3);
  HIP_TRY(h, hipGetLastError());
  int pid4 = prof_begin(h, PS_DRIFT_SORT);
  k_drift_sort_big<<<1024, DR_SORT_NT, 0, h->stream>>>(h->cstart2, h->dbig, nbig, keys, (uint64_t*)h->dsort_tmp);
  prof_end(h, pid4);
  HIP_TRY(h, hipGetLastError());
  int pid5 = prof_begin(h, PS_DRIFT_WRITE);
  k_drift_write<<<(unsigned)nb3, DR_NT, 0, h->stream>>>(g, h->xp, h->vp, h->pid, h->vfield, h->cstart2, keys, h->n3c, dt_mid, c_old,
                                                         h->vdec, h->venc, h->vkmax, h->xp2, h->vp2, h->pid2, h->vfield2,
                                                         h->rhoc2, h->dred);
  prof_end(h, pid5);
  HIP_TRY(h, hipGetLastError());
  double sums[3];
  if ((r = reduce_doubles(h, h->dred, nb3, 3, sums))) return r;
  int64_t npnew;
  HIP_TRY(h, hipMemcpy(&npnew, h->cstart2 + h->n3c, sizeof(int64_t), hipMemcpyDeviceToHost));
  HIP_TRY(h, hipMemcpy(hf, flags, sizeof(hf), hipMemcpyDeviceToHost));
  if (npnew != h->np) return cube_fail(h, CUBE_ERR_CHECK, "npcheck %lld != nplocal %lld", (long long)npnew, (long long)h->np);
  std::swap(h->xp, h->xp2);
  std::swap(h->vp, h->vp2);
  std::swap(h->pid, h->pid2);
  std::swap(h->rhoc, h->rhoc2);
  std::swap(h->vfield, h->vfield2);
  std::swap(h->cstart, h->cstart2);
  const double npg = (double)h->np;
  double std_vsim = sqrt(sums[2] / npg);
  double std_vsim_c = sqrt(sums[0] / (double)h->nc / (double)h->nc / (double)h->nc);
  double std_vsim_res = sqrt(sums[1] / npg);
  h->sigma_vi_new = (float)(std_vsim_res / (double)sqrtf(3.0f));
  if (out) {
    out->sigma_vi_new = h->sigma_vi_new;
    out->nlayer = nlayer;
    out->std_vsim = std_vsim;
    out->std_vsim_c = std_vsim_c;
    out->std_vsim_res = std_vsim_res;
    out->npcheck = npnew;
    for (int e = 0; e < 3; e++) out->max_disp[e] = (int32_t)hf[e];
    out->frame_conflicts = (int32_t)hf[4];
  }
  return 0;
}
