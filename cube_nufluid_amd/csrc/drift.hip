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
#define DR_LBITS 20  // particles per source cell < 2^20 (sort key field)
#define DR_RBITS 27  // coarse cells < 2^27 (nc <= 512)

__device__ __forceinline__ uint16_t pack_disp(int dx, int dy, int dz) {
  return (uint16_t)((dx + 8) | ((dy + 8) << 4) | ((dz + 8) << 8));
}

// Lanes of a wave hold consecutive particles, so equal destination cells in
// neighbouring lanes form runs; one atomic per run instead of per particle.
// key < 0 marks an idle lane (made unique per lane by the caller).
struct WaveRun { int head, len; };
__device__ __forceinline__ WaveRun wave_run(int64_t key) {
  const int lane = threadIdx.x & 63;
  const int64_t prev = __shfl_up(key, 1);
  const bool is_head = lane == 0 || key != prev;
  const unsigned long long H = __ballot(is_head);
  const unsigned long long upto = ~0ull >> (63 - lane);  // bits 0..lane
  const unsigned long long above = H & ~upto;
  WaveRun w;
  w.head = 63 - __clzll(H & upto);
  w.len = (above ? __ffsll((long long)above) - 1 : 64) - lane;  // meaningful in the head lane
  return w;
}

#define DP_SL 2
__global__ __launch_bounds__(DR_NT) void k_drift_pass1(Geo g, const void* __restrict__ xp, const void* __restrict__ vp,
                                                       const int64_t* __restrict__ cstart, int64_t n3c,
                                                       const float* __restrict__ vfield, float dt_mid, double c_old,
                                                       const double* __restrict__ vdt, int dmax, uint16_t* disp,
                                                       int32_t* cnt_new, uint32_t* mslot) {
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
  const int lane = threadIdx.x & 63;
  int lmax[3] = {0, 0, 0}, bad = 0, conf = 0;
  // DP_SL particles (DR_NT apart) per thread and batch: their codes and
  // decoded velocities are loaded together
  for (int64_t pb = cs[0]; pb < cs[DR_CH]; pb += DR_NT * DP_SL) {
    int64_t pp[DP_SL], rr[DP_SL];
    bool ok[DP_SL];
    double xo[DP_SL][3], vd[DP_SL][3];
    float vf[DP_SL][3];
#pragma unroll
    for (int j = 0; j < DP_SL; j++) {
      pp[j] = pb + j * DR_NT + threadIdx.x;
      ok[j] = pp[j] < cs[DR_CH];
      int lo = 0, hi = DR_CH;
      while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (cs[mid] <= pp[j]) lo = mid; else hi = mid;
      }
      if (ok[j] && cs[lo + 1] - cs[lo] > ((int64_t)1 << DR_LBITS)) bad |= 2;
      rr[j] = r0 + lo;
#pragma unroll
      for (int e = 0; e < 3; e++) {
        xo[j][e] = ok[j] ? xoff(load_xu(xp, 3 * pp[j] + e, g.izipx), g.izipx) : 0.0;
        vd[j][e] = ok[j] ? (double)load_v(vp, 3 * pp[j] + e, g.izipv) : 0.0;
        vf[j][e] = ok[j] ? vfield[rr[j] * 3 + e] : 0.f;
      }
    }
#pragma unroll
    for (int j = 0; j < DP_SL; j++)
#pragma unroll
      for (int e = 0; e < 3; e++) vd[j][e] = ok[j] ? vdecode(vdt, (int32_t)vd[j][e], g.izipv) : 0.0;
#pragma unroll
    for (int j = 0; j < DP_SL; j++) {
      const int64_t p = pp[j];
      int64_t rd = -1 - lane;
      if (ok[j]) {
        const int64_t r = rr[j];
        int32_t c[3];
        cell_of_rank(g, r, c);
        int del[3];
        bool far = false;
        for (int e = 0; e < 3; e++) {
          int il = (c[e] & (g.nt - 1)) + 1;  // tile-local 1-based source index
          if (g.lnt < 0) il = c[e] % g.nt + 1;
          double vreal = vd[j][e] + (double)vf[j][e];
          double deltax = ((double)dt_mid * vreal) / (double)CUBE_NCELL;
          int ig = (int)ceil(((double)il - 1.0) + xo[j][e] + deltax);
          int d = ig - il;
          if (ig < 1 || ig > g.nt) {
            // the destination lies in the neighbouring tile: that tile computes
            // this particle from its own frame (source index il -+ nt)
            int sl = ig < 1 ? il + g.nt : il - g.nt;
            int ig2 = (int)ceil(((double)sl - 1.0) + xo[j][e] + deltax);
            int d2 = ig2 - sl;
            if (d2 != d) {
              conf++;
              if (ig2 >= 1 && ig2 <= g.nt) d = d2;
            }
          }
          del[e] = d;
          int ad = d < 0 ? -d : d;
          if (ad > dmax) far = true;
          lmax[e] = max(lmax[e], ad);
        }
        if (far) {
          // flagged: the drift fails and its result is discarded; the
          // particle is counted in its own cell so that the later passes
          // of the same launch sequence stay in bounds
          bad |= 1;
          disp[p] = pack_disp(0, 0, 0);
          rd = r;
        } else {
          disp[p] = pack_disp(del[0], del[1], del[2]);
          rd = rank_of(g, wrapc(c[0] + del[0], g.nc), wrapc(c[1] + del[1], g.nc), wrapc(c[2] + del[2], g.nc));
        }
      }
      const WaveRun w = wave_run(rd);
      if (rd >= 0 && w.head == lane) atomicAdd(&cnt_new[rd], w.len);
    }
  }
  for (int e = 0; e < 3; e++) atomicMax(&smax[e], lmax[e]);
  if (bad) atomicOr(&sbad, bad);
  if (conf) atomicAdd(&sconf, conf);
  __syncthreads();
  // slot fields 0..2 max displacement, 3 too far / crowded, 4 frame conflicts
  uint32_t* sl = mslot + (blockIdx.x & (MSLOTS - 1)) * MSLOT_FIELDS;
  if (threadIdx.x < 3 && smax[threadIdx.x]) atomicMax(&sl[threadIdx.x], (uint32_t)smax[threadIdx.x]);
  if (threadIdx.x == 0) {
    if (sbad) atomicOr(&sl[3], (uint32_t)sbad);
    if (sconf) atomicAdd(&sl[4], (uint32_t)sconf);
  }
}

// ---------------------------------------------------------------- scatter
// Every arrival gets a 64-bit key in its destination cell's segment of
// `keys` (cs_new order):
//   bits 59..63  ilayer of the source z-plane in the destination tile's frame
//   bits 47..58  (7-dz, 7-dy, 7-dx), 4 bits each: source plane, row, column
//                ascending in that frame
//   bits 27..46  l, the particle's position inside its source cell
//   bits  0..26  source cell rank (fixed by the destination and dz,dy,dx)
// Sorting a segment by key gives the reference's arrival order
// ilayer -> k -> j -> i -> l (update_particle.f90:99-105).  Slots come from
// atomics and so are in no particular order before the sort.
__device__ __forceinline__ uint64_t arrival_key(int ilayer, int dx, int dy, int dz, int64_t l, int64_t rs) {
  return ((uint64_t)ilayer << 59) | ((uint64_t)(7 - dz) << 55) | ((uint64_t)(7 - dy) << 51) | ((uint64_t)(7 - dx) << 47) |
         ((uint64_t)l << DR_RBITS) | (uint64_t)rs;
}
__device__ __forceinline__ int64_t key_rank(uint64_t k) { return (int64_t)(k & ((1ull << DR_RBITS) - 1)); }
__device__ __forceinline__ int64_t key_l(uint64_t k) { return (int64_t)((k >> DR_RBITS) & ((1ull << DR_LBITS) - 1)); }

#define DR_SMALL 16   // cells up to this many arrivals: one thread sorts
#define DR_MED 1024   // up to this many: one wavefront; above: a 1024-thread workgroup

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
  const int lane = threadIdx.x & 63;
  for (int64_t pb = cs[0]; pb < cs[DR_CH]; pb += DR_NT) {
    const int64_t p = pb + threadIdx.x;
    int64_t rd = -1 - lane;
    uint64_t key = 0;
    if (p < cs[DR_CH]) {
      int lo = 0, hi = DR_CH;
      while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (cs[mid] <= p) lo = mid; else hi = mid;
      }
      const int64_t rs = r0 + lo;
      int32_t c[3];
      cell_of_rank(g, rs, c);
      const uint32_t dc = disp[p];
      const int d[3] = {(int)(dc & 15) - 8, (int)((dc >> 4) & 15) - 8, (int)((dc >> 8) & 15) - 8};
      const int cz = wrapc(c[2] + d[2], g.nc);
      rd = rank_of(g, wrapc(c[0] + d[0], g.nc), wrapc(c[1] + d[1], g.nc), cz);
      const int sz = cz % g.nt + 1 - d[2];  // source plane in the destination tile's frame
      const int ilayer = (sz - 1 + g.ncb) % nlayer;
      key = arrival_key(ilayer, d[0], d[1], d[2], p - cs[lo], rs);
    }
    const WaveRun w = wave_run(rd);
    int64_t base = 0;
    if (rd >= 0 && w.head == lane) {
      const int32_t old = atomicSub(&cnt[rd], w.len);
      base = cs_new[rd] + old - w.len;
      // the run holding slot DR_MED lists a crowded cell once (rare)
      if (old > DR_MED && old - w.len <= DR_MED) big[atomicAdd(nbig, 1u)] = (int32_t)rd;
    }
    base = __shfl(base, w.head);
    if (rd >= 0) keys[base + (lane - w.head)] = key;
  }
}

// ---------------------------------------------------------------- sorting
// Cells up to DR_SMALL arrivals: one thread each, insertion sort in LDS (the
// small cells of a workgroup's DR_NT cells are packed into LDS, at most
// DR_NT*DR_SMALL keys).  Up to DR_MED: the workgroup's wavefronts take them
// in turn, bitonic in LDS.  Larger (listed by the scatter): one 1024-thread
// workgroup per cell; segments up to DR_SORT_LDS keys are bitonic-sorted in
// LDS, longer ones sorted in DR_SORT_LDS chunks and then merged pairwise
// between `keys` and `tmp` (an element's output position is its index in its
// own run plus its rank in the partner run).
__device__ __forceinline__ void wave_sync() {
  // a wavefront's LDS operations execute in order; this keeps the compiler
  // from moving them across the point
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
// compare-exchange network over npow2 keys by NT cooperating lanes; sync()
// separates the stages
template <int NT, typename Sync>
__device__ __forceinline__ void bitonic(uint64_t* s, int n, int npow2, int tid, Sync sync) {
  for (int i = n + tid; i < npow2; i += NT) s[i] = ~0ull;
  sync();
  for (int k = 2; k <= npow2; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int t = tid; t < (npow2 >> 1); t += NT) {
        const int i = ((t & ~(j - 1)) << 1) | (t & (j - 1)), l = i + j;
        const uint64_t a = s[i], b = s[l];
        const bool up = (i & k) == 0;
        if ((a > b) == up) { s[i] = b; s[l] = a; }
      }
      sync();
    }
}

#define DR_MED_LDS 1024
__global__ __launch_bounds__(DR_NT) void k_drift_sort_cells(const int64_t* __restrict__ cs_new, int64_t n3c, uint64_t* keys) {
  __shared__ uint64_t sk[DR_NT * DR_SMALL];  // small cells; then DR_NT/64 wavefront areas of DR_MED_LDS
  __shared__ int32_t soff[DR_NT / 64];
  __shared__ int32_t med[DR_NT];
  __shared__ int32_t nmed;
  __shared__ int32_t loff[DR_NT + 1];
  const int64_t r = (int64_t)blockIdx.x * DR_NT + threadIdx.x;
  int64_t base = 0;
  int n = 0;
  if (r < n3c) {
    base = cs_new[r];
    n = (int)(cs_new[r + 1] - base);
  }
  // Usual case: the workgroup's contiguous key range fits in LDS.  It is
  // loaded coalesced, and each key is written to its cell's start plus its
  // rank in the cell (the keys are distinct), one thread per key.  Keys of
  // cells above DR_MED are left for k_drift_sort_big.
  {
    const int64_t rb = (int64_t)blockIdx.x * DR_NT;
    const int ncell = (int)min((int64_t)DR_NT, n3c - rb);
    const int64_t k0 = cs_new[rb], tot = cs_new[rb + ncell] - k0;
    if (tot <= DR_NT * DR_SMALL) {
      if ((int)threadIdx.x < ncell) loff[threadIdx.x] = (int32_t)(base - k0);
      if (threadIdx.x == 0) loff[ncell] = (int32_t)tot;
      constexpr int G = 8;
      for (int64_t i0 = 0; i0 < tot; i0 += G * DR_NT) {
        uint64_t v[G];
#pragma unroll
        for (int q = 0; q < G; q++) {
          const int64_t i = i0 + q * DR_NT + threadIdx.x;
          v[q] = i < tot ? keys[k0 + i] : 0;
        }
#pragma unroll
        for (int q = 0; q < G; q++) {
          const int64_t i = i0 + q * DR_NT + threadIdx.x;
          if (i < tot) sk[i] = v[q];
        }
      }
      __syncthreads();
      for (int i = threadIdx.x; i < (int)tot; i += DR_NT) {
        int lo = 0, hi = ncell;  // last cell starting at or before i
        while (hi - lo > 1) {
          const int m = (lo + hi) >> 1;
          if (loff[m] <= i) lo = m; else hi = m;
        }
        const int cb = loff[lo], ce = loff[lo + 1];
        const uint64_t v = sk[i];
        int rank = i - cb;
        if (ce - cb <= DR_MED) {
          rank = 0;
          for (int j = cb; j < ce; j++) rank += sk[j] < v ? 1 : 0;
        }
        keys[k0 + cb + rank] = v;
      }
      return;
    }
  }
  if (threadIdx.x == 0) nmed = 0;
  const int m = n <= DR_SMALL ? n : 0;  // keys this thread stages
  int incl = m;
  for (int o = 1; o < 64; o <<= 1) {
    const int t = __shfl_up(incl, o);
    if ((threadIdx.x & 63) >= o) incl += t;
  }
  if ((threadIdx.x & 63) == 63) soff[threadIdx.x >> 6] = incl;
  __syncthreads();
  if (n > DR_SMALL && n <= DR_MED) med[atomicAdd(&nmed, 1)] = threadIdx.x;
  int wbase = 0;
  for (int w = 0; w < (int)(threadIdx.x >> 6); w++) wbase += soff[w];
  if (m > 1) {
    uint64_t* a = sk + wbase + incl - m;
    for (int i = 0; i < m; i++) a[i] = keys[base + i];
    for (int i = 1; i < m; i++) {
      const uint64_t v = a[i];
      int j = i - 1;
      while (j >= 0 && a[j] > v) { a[j + 1] = a[j]; j--; }
      a[j + 1] = v;
    }
    for (int i = 0; i < m; i++) keys[base + i] = a[i];
  }
  __syncthreads();
  // medium cells: wavefront w sorts list entries w, w + DR_NT/64, ...
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  uint64_t* ws = sk + wv * DR_MED_LDS;
  const int nm = nmed;
  for (int q = wv; q < nm; q += DR_NT / 64) {
    const int64_t rr = (int64_t)blockIdx.x * DR_NT + med[q];
    const int64_t b0 = cs_new[rr];
    const int nn = (int)(cs_new[rr + 1] - b0);
    int np2 = 64;
    while (np2 < nn) np2 <<= 1;
    for (int i = lane; i < nn; i += 64) ws[i] = keys[b0 + i];
    bitonic<64>(ws, nn, np2, lane, [] __device__() { wave_sync(); });
    for (int i = lane; i < nn; i += 64) keys[b0 + i] = ws[i];
    wave_sync();
  }
}

#define DR_SORT_NT 1024
#define DR_SORT_LDS 8192
__global__ __launch_bounds__(DR_SORT_NT) void k_drift_sort_big(const int64_t* __restrict__ cs_new, const int32_t* __restrict__ big,
                                                               const uint32_t* __restrict__ nbig, uint64_t* keys, uint64_t* tmp) {
  __shared__ uint64_t s[DR_SORT_LDS];
  const uint32_t nb = *nbig;
  for (uint32_t bi = blockIdx.x; bi < nb; bi += gridDim.x) {
    const int64_t r = big[bi];
    const int64_t base = cs_new[r];
    const int64_t n = cs_new[r + 1] - base;
    for (int64_t c0 = 0; c0 < n; c0 += DR_SORT_LDS) {
      const int m = (int)min((int64_t)DR_SORT_LDS, n - c0);
      int np2 = 1;
      while (np2 < m) np2 <<= 1;
      for (int i = threadIdx.x; i < m; i += DR_SORT_NT) s[i] = keys[base + c0 + i];
      __syncthreads();
      bitonic<DR_SORT_NT>(s, m, np2, (int)threadIdx.x, [] __device__() { __syncthreads(); });
      for (int i = threadIdx.x; i < m; i += DR_SORT_NT) keys[base + c0 + i] = s[i];
      __syncthreads();
    }
    // merge passes (keys are unique: the low bits hold the source cell and l)
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

// ---------------------------------------------------------- gather + cells
// One workgroup per DW_CELLS destination cells; their slots (checkpoint
// order after the drift) are contiguous.
//   gather (one thread per slot, DG_SL slots per batch with each stage's
//     loads issued together): the arrival's full velocity vreal
//     (update_particle.f90:69), its new position code (:117) and PID.
//     vreal stays in LDS when the workgroup's slots fit (else in HBM).
//   cells (one thread per cell): vfield_new summed in arrival order
//     (update_particle.f90:70-83).
//   slots (DC_SL per batch): the velocity residual re-quantised with the old
//     sigma (:119) and the sigma_v sums (:174-200).
// part: 3 doubles per workgroup, at its first cell's index / DW_CELLS.
#define DW_CELLS 128
#define DW_STAGE 1536
#define DG_SL 4
#define DC_SL 2
template <bool STAGED>
__device__ __forceinline__ void drift_write_body(Geo g, const uint64_t* __restrict__ keys, int64_t r0, int ncell,
                                                 const int64_t* __restrict__ cstart, const void* __restrict__ xp,
                                                 const void* __restrict__ vp, const int32_t* __restrict__ pid,
                                                 const float* __restrict__ vfield, float dt_mid, double c_old,
                                                 const double* __restrict__ vdt, const double* __restrict__ venc,
                                                 int32_t kmax, void* xp_o, int32_t* pid_o, double* vr, void* vp_o,
                                                 float* vfield_o, int32_t* rhoc_o, double* part, const int64_t* scs,
                                                 float* svf, double (*red)[DR_NT / 64], double* svr, const double* sat) {
  const int64_t qa = scs[0], q1 = scs[ncell];
  // vreal of slot q: LDS when the workgroup's slots fit, else the HBM scratch
  auto vr_at = [&](int64_t q, int e) -> double& { return STAGED ? svr[3 * (q - qa) + e] : vr[3 * q + e]; };
  // ---- gather
  const double xscale = (g.izipx == 1 ? 1.0 / 256.0 : 1.0 / 65536.0) * (double)CUBE_NCELL;
  for (int64_t q0 = qa + threadIdx.x; q0 < q1; q0 += (int64_t)DR_NT * DG_SL) {
    uint64_t key[DG_SL];
    bool ok[DG_SL];
#pragma unroll
    for (int j = 0; j < DG_SL; j++) {
      ok[j] = q0 + j * DR_NT < q1;
      key[j] = ok[j] ? keys[q0 + j * DR_NT] : 0;
    }
    int64_t rs[DG_SL], p[DG_SL];
#pragma unroll
    for (int j = 0; j < DG_SL; j++) {
      rs[j] = key_rank(key[j]);
      p[j] = ok[j] ? cstart[rs[j]] + key_l(key[j]) : 0;
    }
    int32_t vc[DG_SL][3], pv[DG_SL];
    uint32_t xu[DG_SL][3];
    float vf[DG_SL][3];
#pragma unroll
    for (int j = 0; j < DG_SL; j++) {
      pv[j] = ok[j] ? pid[p[j]] : 0;
#pragma unroll
      for (int e = 0; e < 3; e++) {
        vc[j][e] = ok[j] ? load_v(vp, 3 * p[j] + e, g.izipv) : 0;
        xu[j][e] = ok[j] ? load_xu(xp, 3 * p[j] + e, g.izipx) : 0;
        vf[j][e] = ok[j] ? vfield[rs[j] * 3 + e] : 0.f;
      }
    }
    double vd[DG_SL][3];
#pragma unroll
    for (int j = 0; j < DG_SL; j++)
#pragma unroll
      for (int e = 0; e < 3; e++) vd[j][e] = ok[j] ? vdecode(vdt, vc[j][e], g.izipv) : 0.0;
#pragma unroll
    for (int j = 0; j < DG_SL; j++) {
      if (!ok[j]) continue;
      const int64_t q = q0 + j * DR_NT;
#pragma unroll
      for (int e = 0; e < 3; e++) {
        const double vreal = vd[j][e] + (double)vf[j][e];
        int64_t xn = (int64_t)xu[j][e];
        xn = g.izipx == 1 ? (int64_t)(int8_t)xn : (int64_t)(int16_t)xn;
        xn += (int64_t)nint_d((double)dt_mid * vreal / xscale);
        store_x(xp_o, 3 * q + e, g.izipx, xn);
        vr_at(q, e) = vreal;
      }
      pid_o[q] = pv[j];
    }
  }
  if (!STAGED) __threadfence_block();
  __syncthreads();
  // ---- cells
  const double wv = (double)0.1f;  // weight_v=0.1 is a default-real literal
  double sum_c = 0, sum_r = 0, sum_s = 0;
  if ((int)threadIdx.x < ncell) {
    const int64_t r = r0 + threadIdx.x;
    const int64_t b = scs[threadIdx.x], n = scs[threadIdx.x + 1] - b;
    float vf[3];
    for (int e = 0; e < 3; e++) vf[e] = (float)((double)vfield[r * 3 + e] * wv);
    // in arrival order; the loads of U arrivals are issued together
    constexpr int U = 4;
    for (int64_t k0 = 0; k0 < n; k0 += U) {
      double a[U][3];
#pragma unroll
      for (int u = 0; u < U; u++)
#pragma unroll
        for (int e = 0; e < 3; e++) a[u][e] = k0 + u < n ? vr_at(b + k0 + u, e) : 0.0;
#pragma unroll
      for (int u = 0; u < U; u++)
        if (k0 + u < n)
#pragma unroll
          for (int e = 0; e < 3; e++) vf[e] = (float)((double)vf[e] + a[u][e]);
    }
    float vfn[3];
    for (int e = 0; e < 3; e++) {
      vfn[e] = (float)((double)vf[e] / ((double)n + wv));
      svf[threadIdx.x * 3 + e] = vfn[e];
      vfield_o[r * 3 + e] = vfn[e];
    }
    rhoc_o[r] = (int32_t)n;
    sum_c = (double)(vfn[0] * vfn[0] + vfn[1] * vfn[1] + vfn[2] * vfn[2]);
  }
  __syncthreads();
  // ---- slots: re-quantisation and sigma sums, DC_SL per batch
  for (int64_t q0 = qa + threadIdx.x; q0 < q1; q0 += (int64_t)DR_NT * DC_SL) {
    bool ok[DC_SL];
    int cl[DC_SL];
    double x[DC_SL][3];
#pragma unroll
    for (int j = 0; j < DC_SL; j++) {
      const int64_t q = q0 + j * DR_NT;
      ok[j] = q < q1;
      int lo = 0, hi = ncell;
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (scs[mid] <= q) lo = mid; else hi = mid;
      }
      cl[j] = lo;
#pragma unroll
      for (int e = 0; e < 3; e++) x[j][e] = ok[j] ? vr_at(q, e) : 0.0;
    }
    int32_t n[DC_SL][3];
#pragma unroll
    for (int j = 0; j < DC_SL; j++)
#pragma unroll
      for (int e = 0; e < 3; e++) {
        x[j][e] = c_old * (x[j][e] - (double)svf[cl[j] * 3 + e]);
        n[j][e] = venc_fast(venc, sat, kmax, g.izipv, x[j][e]);
      }
    double v[DC_SL][3];
#pragma unroll
    for (int j = 0; j < DC_SL; j++)
#pragma unroll
      for (int e = 0; e < 3; e++) {
        if (ok[j]) store_v(vp_o, 3 * (q0 + j * DR_NT) + e, g.izipv, n[j][e]);
        v[j][e] = vdecode(vdt, n[j][e], g.izipv);
      }
#pragma unroll
    for (int j = 0; j < DC_SL; j++) {
      if (!ok[j]) continue;
      double w[3] = {v[j][0], v[j][1], v[j][2]};
      sum_r += w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
      for (int e = 0; e < 3; e++) w[e] += (double)svf[cl[j] * 3 + e];
      sum_s += w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
    }
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
    part[r0 / DW_CELLS * 3 + threadIdx.x] = t;
  }
}

__global__ __launch_bounds__(DR_NT) void k_drift_write(Geo g, const uint64_t* __restrict__ keys, const int64_t* __restrict__ cs_new,
                                                       int64_t n3c, const int64_t* __restrict__ cstart, const void* __restrict__ xp,
                                                       const void* __restrict__ vp, const int32_t* __restrict__ pid,
                                                       const float* __restrict__ vfield, float dt_mid, double c_old,
                                                       const double* __restrict__ vdt, const double* __restrict__ venc,
                                                       int32_t kmax, void* xp_o, int32_t* pid_o, double* vr, void* vp_o,
                                                       float* vfield_o, int32_t* rhoc_o, double* part) {
  __shared__ int64_t scs[DW_CELLS + 1];
  __shared__ float svf[DW_CELLS * 3];
  __shared__ double red[3][DR_NT / 64];
  __shared__ double svr[3 * DW_STAGE];
  __shared__ double sat[33];
  // XCD-contiguous block order: neighbouring cell rows gather from the same
  // source cells, so they run on the same XCD (the same L2)
  const int64_t r0 = (int64_t)xcd_remap(blockIdx.x, gridDim.x) * DW_CELLS;
  const int ncell = (int)min((int64_t)DW_CELLS, n3c - r0);
  for (int i = threadIdx.x; i <= DW_CELLS; i += DR_NT) scs[i] = cs_new[r0 + min(i, ncell)];
  atan_tab_to_lds(sat);
  __syncthreads();
  if (scs[ncell] - scs[0] <= DW_STAGE)
    drift_write_body<true>(g, keys, r0, ncell, cstart, xp, vp, pid, vfield, dt_mid, c_old, vdt, venc, kmax, xp_o, pid_o, vr, vp_o,
                           vfield_o, rhoc_o, part, scs, svf, red, svr, sat);
  else
    drift_write_body<false>(g, keys, r0, ncell, cstart, xp, vp, pid, vfield, dt_mid, c_old, vdt, venc, kmax, xp_o, pid_o, vr,
                            vp_o, vfield_o, rhoc_o, part, scs, svf, red, svr, sat);
}

// sigma_vi_new = sqrt(sum v_res^2 / N) / sqrt(3) (update_particle.f90:213,
// in the host's operation order) and the new velocity coefficient
// vcoef(sigma_vi_new) = sqrt(pi/2) / (sigma * 2.5), formed on the device so
// that the kick can read them without a host round trip
__global__ void k_sigma_new(const double* __restrict__ dsum, double npg, float sqrt3f, float sqrtpi2f, float* sigma_out,
                            double* cnew_out) {
  const float sg = (float)(sqrt(dsum[1] / npg) / (double)sqrt3f);
  *sigma_out = sg;
  *cnew_out = (double)sqrtpi2f / ((double)sg * 2.5);
}
__global__ void k_vtab_fill_dev(const float* __restrict__ dec, int n, const double* __restrict__ c, double* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (double)dec[i] / *c;
}
__global__ void k_set_double(double* p, double v) { *p = v; }

int sigma_new_to_device(cube_handle* h) {
  const int n = h->p.izipv == 1 ? 256 : 65536;
  k_set_double<<<1, 1, 0, h->stream>>>(h->dscal, vcoef(h->sigma_vi_new));
  k_vtab_fill_dev<<<(n + 255) / 256, 256, 0, h->stream>>>(h->vdec, n, h->dscal, h->vdt_new);
  HIP_TRY(h, hipGetLastError());
  return 0;
}

// update_particle.f90:57-226 as one launch sequence.  The results the host
// needs (flags, npcheck, the sigma sums) are gathered in h->drep; the state
// buffers are swapped at once and swapped back by drift_finish if the drift
// failed.  Unless the handle is inside a deferred step, the report is
// copied and checked before returning.
static int drift_launch(cube_handle* h, float dt_old, float dt, int* nlayer_out) {
  Geo g = geo_of(h);
  const float dt_mid = (dt_old + dt) / 2;
  const int nlayer = 2 * (int)ceilf(dt_mid * h->vz_max / (float)CUBE_NCELL) + 1;
  *nlayer_out = nlayer;
  if (nlayer > 31) return cube_fail(h, CUBE_ERR_CAPACITY, "nlayer=%d above 31 (dt*vz_max too large)", nlayer);
  const double c_old = vcoef(h->sigma_vi);
  const double* vdt = vtab_for(h, c_old);
  const int dmax = (int)(h->ncb < DMAX ? h->ncb : DMAX);
  uint32_t* flags = h->uflags;  // [0..2] max disp, [3] too far / crowded, [4] frame conflicts, [6] crowded cells
  const int64_t nb1 = (h->n3c + DR_CH - 1) / DR_CH;
  const int64_t nb3 = (h->n3c + DR_NT - 1) / DR_NT;
  const int64_t nbw = (h->n3c + DW_CELLS - 1) / DW_CELLS;
  if (nb3 * 3 > h->dred_n - 16 || nbw * 3 > h->dred_n - 16) return cube_fail(h, CUBE_ERR_ARG, "reduction buffer too small");
  int r;
  if ((r = fill_u32(h, flags, 8, 0u)) || (r = fill_u32(h, h->rhoc2, h->n3c, 0u)) ||
      (r = fill_u32(h, h->mslots, MSLOTS * MSLOT_FIELDS, 0u)))
    return r;
  int pid1 = prof_begin(h, PS_DRIFT1);
  k_drift_pass1<<<(unsigned)nb1, DR_NT, 0, h->stream>>>(g, h->xp, h->vp, h->cstart, h->n3c, h->vfield, dt_mid, c_old, vdt, dmax,
                                                         h->disp, h->rhoc2, h->mslots);
  prof_end(h, pid1);
  HIP_TRY(h, hipGetLastError());
  if ((r = slots_reduce(h, 5, 1u << 3, 1u << 4, flags))) return r;
  int pid2 = prof_begin(h, PS_DRIFT_SCAN);
  r = scan_counts(h, h->rhoc2, h->n3c, h->cstart2);
  prof_end(h, pid2);
  if (r) return r;
  uint64_t* keys = (uint64_t*)h->order;
  uint32_t* nbig = flags + 6;
  int pid3 = prof_begin(h, PS_DRIFT_SCATTER);
  k_drift_scatter<<<(unsigned)nb1, DR_NT, 0, h->stream>>>(g, h->cstart, h->n3c, h->disp, h->cstart2, nlayer, h->rhoc2, keys,
                                                           h->dbig, nbig);
  prof_end(h, pid3);
  HIP_TRY(h, hipGetLastError());
  int pid4 = prof_begin(h, PS_DRIFT_SORT);
  k_drift_sort_cells<<<(unsigned)nb3, DR_NT, 0, h->stream>>>(h->cstart2, h->n3c, keys);
  k_drift_sort_big<<<512, DR_SORT_NT, 0, h->stream>>>(h->cstart2, h->dbig, nbig, keys, (uint64_t*)h->dscratch);
  prof_end(h, pid4);
  HIP_TRY(h, hipGetLastError());
  int pid5 = prof_begin(h, PS_DRIFT_WRITE);
  k_drift_write<<<(unsigned)nbw, DR_NT, 0, h->stream>>>(g, keys, h->cstart2, h->n3c, h->cstart, h->xp, h->vp, h->pid, h->vfield,
                                                         dt_mid, c_old, vdt, h->venc, h->vkmax, h->xp2, h->pid2,
                                                         (double*)h->dscratch, h->vp2, h->vfield2, h->rhoc2, h->dred);
  prof_end(h, pid5);
  HIP_TRY(h, hipGetLastError());
  StepReport* d = h->drep;
  if ((r = reduce_doubles_dev(h, h->dred, nbw, 3, d->dsum))) return r;
  const int n = h->p.izipv == 1 ? 256 : 65536;
  k_sigma_new<<<1, 1, 0, h->stream>>>(d->dsum, (double)h->np, sqrtf(3.0f), sqrtf(PI_F / 2), &d->sigma_new, h->dscal);
  k_vtab_fill_dev<<<(n + 255) / 256, 256, 0, h->stream>>>(h->vdec, n, h->dscal, h->vdt_new);
  HIP_TRY(h, hipGetLastError());
  HIP_TRY(h, hipMemcpyAsync(&d->npnew, h->cstart2 + h->n3c, sizeof(int64_t), hipMemcpyDeviceToDevice, h->stream));
  HIP_TRY(h, hipMemcpyAsync(d->dflags, flags, sizeof(d->dflags), hipMemcpyDeviceToDevice, h->stream));
  std::swap(h->xp, h->xp2);
  std::swap(h->vp, h->vp2);
  std::swap(h->pid, h->pid2);
  std::swap(h->rhoc, h->rhoc2);
  std::swap(h->vfield, h->vfield2);
  std::swap(h->cstart, h->cstart2);
  return 0;
}

static void drift_unswap(cube_handle* h) {
  std::swap(h->xp, h->xp2);
  std::swap(h->vp, h->vp2);
  std::swap(h->pid, h->pid2);
  std::swap(h->rhoc, h->rhoc2);
  std::swap(h->vfield, h->vfield2);
  std::swap(h->cstart, h->cstart2);
}

// checks of update_particle.f90:90-96 and the npcheck of :150, on the
// copied report; a failed drift leaves the state as it was
int drift_finish(cube_handle* h, const StepReport* rp, int nlayer, cube_drift_out* out) {
  const int dmax = (int)(h->ncb < DMAX ? h->ncb : DMAX);
  const char* why = nullptr;
  int code = 0;
  char buf[160];
  if (rp->dflags[3] & 1) {
    snprintf(buf, sizeof buf, "a particle moved more than %d coarse cells in one drift (buffer ncb=%lld)", dmax,
             (long long)h->ncb);
    why = buf;
    code = CUBE_ERR_CAPACITY;
  } else if (rp->dflags[3] & 2) {
    snprintf(buf, sizeof buf, "a coarse cell holds more than 2^%d particles", DR_LBITS);
    why = buf;
    code = CUBE_ERR_CAPACITY;
  } else if (rp->npnew != h->np) {
    snprintf(buf, sizeof buf, "npcheck %lld != nplocal %lld", (long long)rp->npnew, (long long)h->np);
    why = buf;
    code = CUBE_ERR_CHECK;
  }
  if (code) {
    drift_unswap(h);
    int r = sigma_new_to_device(h);  // the failed drift overwrote them
    (void)r;
    return cube_fail(h, code, "%s", why);
  }
  const double npg = (double)h->np;
  double std_vsim = sqrt(rp->dsum[2] / npg);
  double std_vsim_c = sqrt(rp->dsum[0] / (double)h->nc / (double)h->nc / (double)h->nc);
  double std_vsim_res = sqrt(rp->dsum[1] / npg);
  h->sigma_vi_new = rp->sigma_new;
  if (out) {
    out->sigma_vi_new = h->sigma_vi_new;
    out->nlayer = nlayer;
    out->std_vsim = std_vsim;
    out->std_vsim_c = std_vsim_c;
    out->std_vsim_res = std_vsim_res;
    out->npcheck = rp->npnew;
    for (int e = 0; e < 3; e++) out->max_disp[e] = (int32_t)rp->dflags[e];
    out->frame_conflicts = (int32_t)rp->dflags[4];
  }
  return 0;
}

int drift_launch_step(cube_handle* h, float dt_old, float dt, int* nlayer) { return drift_launch(h, dt_old, dt, nlayer); }

int drift_run(cube_handle* h, float dt_old, float dt, cube_drift_out* out) {
  int nlayer = 0;
  int r = drift_launch(h, dt_old, dt, &nlayer);
  if (r) return r;
  HIP_TRY(h, hipMemcpyAsync(h->rep, h->drep, sizeof(StepReport), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  return drift_finish(h, h->rep, nlayer, out);
}
