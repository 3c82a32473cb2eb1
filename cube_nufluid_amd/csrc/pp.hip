// Short-range particle-particle force (ext_pp_force.f90:9-160).
//
// The reference links particles into per-fine-cell lists and, for every
// particle, sums the tapered pair force over all particles in the fine cells
// within +-pp_range.  Here one workgroup owns a BP^3 block of coarse cells
// (a 4BP fine-cell cube).  Candidate partners -- particles of the block and
// of the surrounding coarse cells whose fine cell lies within pp_range of
// the block -- are staged in LDS in chunks; each thread owns one particle of
// the block and tests fine-cell adjacency exactly as the lists do, so the
// pair set is identical.  Positions relative to the block are exact in
// FP32 (they carry at most 19 significant bits), the pair force is FP32 and
// the per-particle sum is a 64-bit fixed-point integer in 2^-32 units, so it
// does not depend on the order in which candidates landed in LDS and the
// result is the same on every run.  The kick and re-quantisation with the
// current sigma (ext_pp_force.f90:90, 114-119) are applied in place.
#include "cube_internal.hpp"
#include <string.h>
#include <algorithm>

#define PP_NT 64
#define PP_BP 2
#define PP_CAP 1024
#define PP_FIX 4294967296.0  // 2^32

struct PPcand { float x, y, z; int32_t fc; };

__device__ __forceinline__ int32_t pack_fc(int a, int b, int c) {  // each in [-8, 23]
  return ((a + 8) & 31) | (((b + 8) & 31) << 5) | (((c + 8) & 31) << 10);
}

__global__ __launch_bounds__(PP_NT) void k_pp(Geo g, const void* __restrict__ xp, void* vp, const int64_t* __restrict__ cstart,
                                              float mass, float rsoft, int pp_range, float a_mid, float dt, double c_old,
                                              const float* __restrict__ vdec, const double* __restrict__ venc, int32_t kmax,
                                              uint32_t* f2max_bits, unsigned long long* npairs_out,
                                              unsigned long long* nupd_out, double* probe, const int32_t* blist, int bscale) {
  constexpr int SPAN = PP_BP + 2, NCELL = SPAN * SPAN * SPAN, FB = 4 * PP_BP;
  __shared__ int64_t beg[NCELL];
  __shared__ int32_t off[NCELL + 1];
  __shared__ PPcand cand[PP_CAP];
  __shared__ int ncand;
  __shared__ int32_t ioff[PP_BP * PP_BP * PP_BP + 1];
  __shared__ int64_t ibeg[PP_BP * PP_BP * PP_BP];
  // blocks of PP_BP^3 cells; with a block list, entry blockIdx.x/bscale^3 names
  // a coarser block of (bscale*PP_BP)^3 cells and blockIdx.x%bscale^3 the sub-block
  int c0x, c0y, c0z;
  if (blist) {
    const int nsub = bscale * bscale * bscale, big = blist[blockIdx.x / nsub], sub = blockIdx.x % nsub;
    const int nbb = g.nc / (PP_BP * bscale);
    c0x = (big % nbb) * PP_BP * bscale + (sub % bscale) * PP_BP;
    c0y = ((big / nbb) % nbb) * PP_BP * bscale + ((sub / bscale) % bscale) * PP_BP;
    c0z = (big / (nbb * nbb)) * PP_BP * bscale + (sub / (bscale * bscale)) * PP_BP;
  } else {
    const int nbx = g.nc / PP_BP;
    c0x = (blockIdx.x % nbx) * PP_BP;
    c0y = ((blockIdx.x / nbx) % nbx) * PP_BP;
    c0z = (blockIdx.x / (nbx * nbx)) * PP_BP;
  }
  // region cells [c0-1, c0+PP_BP], inner cells [c0, c0+PP_BP)
  for (int i = threadIdx.x; i < NCELL; i += PP_NT) {
    int rx = i % SPAN - 1, ry = (i / SPAN) % SPAN - 1, rz = i / (SPAN * SPAN) - 1;
    int64_t r = rank_of(g, wrapc(c0x + rx, g.nc), wrapc(c0y + ry, g.nc), wrapc(c0z + rz, g.nc));
    beg[i] = cstart[r];
    off[i + 1] = (int32_t)(cstart[r + 1] - cstart[r]);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    off[0] = 0;
    for (int i = 0; i < NCELL; i++) off[i + 1] += off[i];
    ioff[0] = 0;
    int q = 0;
    for (int z = 0; z < PP_BP; z++)
      for (int y = 0; y < PP_BP; y++)
        for (int x = 0; x < PP_BP; x++) {
          int ci = ((z + 1) * SPAN + (y + 1)) * SPAN + (x + 1);
          ibeg[q] = beg[ci];
          ioff[q + 1] = ioff[q] + (off[ci + 1] - off[ci]);
          q++;
        }
  }
  __syncthreads();
  const int32_t ntot = off[NCELL], ninner = ioff[PP_BP * PP_BP * PP_BP];
  const int lo_fc = -pp_range, hi_fc = FB - 1 + pp_range;
  float f2loc = 0.f;
  unsigned long long npl = 0;
  for (int32_t ib = 0; ib < ninner; ib += PP_NT) {
    const int32_t it = ib + threadIdx.x;
    const bool act = it < ninner;
    int64_t pi = 0;
    float xi[3] = {0, 0, 0};
    int fi[3] = {0, 0, 0};
    if (act) {
      int q = 0;
      while (ioff[q + 1] <= it) q++;
      pi = ibeg[q] + (it - ioff[q]);
      int rel[3] = {q % PP_BP, (q / PP_BP) % PP_BP, q / (PP_BP * PP_BP)};
      for (int e = 0; e < 3; e++) {
        uint32_t u = load_xu(xp, 3 * pi + e, g.izipx);
        xi[e] = (float)((double)(CUBE_NCELL * rel[e]) + (double)CUBE_NCELL * xoff(u, g.izipx));
        fi[e] = CUBE_NCELL * rel[e] + xsub(u, g.izipx);
      }
    }
    long long fa[3] = {0, 0, 0};
    for (int32_t j0 = 0; j0 < ntot; j0 += PP_CAP) {
      if (threadIdx.x == 0) ncand = 0;
      __syncthreads();
      const int32_t j1 = min(ntot, j0 + PP_CAP);
      for (int32_t t = j0 + threadIdx.x; t < j1; t += PP_NT) {
        int lo = 0, hi = NCELL;
        while (hi - lo > 1) {
          int mid = (lo + hi) >> 1;
          if (off[mid] <= t) lo = mid; else hi = mid;
        }
        int64_t pj = beg[lo] + (t - off[lo]);
        int rel[3] = {lo % SPAN - 1, (lo / SPAN) % SPAN - 1, lo / (SPAN * SPAN) - 1};
        float x[3];
        int fc[3];
        bool keep = true;
        for (int e = 0; e < 3; e++) {
          uint32_t u = load_xu(xp, 3 * pj + e, g.izipx);
          x[e] = (float)((double)(CUBE_NCELL * rel[e]) + (double)CUBE_NCELL * xoff(u, g.izipx));
          fc[e] = CUBE_NCELL * rel[e] + xsub(u, g.izipx);
          keep = keep && fc[e] >= lo_fc && fc[e] <= hi_fc;
        }
        if (keep) {
          int slot = atomicAdd(&ncand, 1);
          cand[slot] = PPcand{x[0], x[1], x[2], pack_fc(fc[0], fc[1], fc[2])};
        }
      }
      __syncthreads();
      const int nc_ = ncand;
      if (act) {
        for (int j = 0; j < nc_; j++) {
          PPcand cj = cand[j];
          int ax = (cj.fc & 31) - 8, ay = ((cj.fc >> 5) & 31) - 8, az = ((cj.fc >> 10) & 31) - 8;
          if (abs(ax - fi[0]) > pp_range || abs(ay - fi[1]) > pp_range || abs(az - fi[2]) > pp_range) continue;
          npl++;
          float dx = cj.x - xi[0], dy = cj.y - xi[1], dz = cj.z - xi[2];
          float r2 = dx * dx + dy * dy + dz * dz;
          float r = sqrtf(r2);
          if (!(r > rsoft)) continue;
          float rc = r * (1.0f / 16.0f);
          float rc3 = rc * rc * rc;
          float pc = 1.0f - 1.75f * rc3 + 0.75f * rc3 * rc * rc;
          float s = mass * pc / (r2 * r);
          fa[0] += llrint((double)(s * dx) * PP_FIX);
          fa[1] += llrint((double)(s * dy) * PP_FIX);
          fa[2] += llrint((double)(s * dz) * PP_FIX);
        }
      }
      __syncthreads();
    }
    const double f[3] = {(double)fa[0] * (1.0 / PP_FIX), (double)fa[1] * (1.0 / PP_FIX), (double)fa[2] * (1.0 / PP_FIX)};
    if (act) {
      if (probe) {
        for (int e = 0; e < 3; e++) probe[3 * pi + e] = f[e];
      } else {
        for (int e = 0; e < 3; e++) {
          double v = vdecode(vdec, load_v(vp, 3 * pi + e, g.izipv), g.izipv, c_old);
          v = v + f[e] * (double)a_mid * (double)dt / 6.0 / (double)PI_F;
          store_v(vp, 3 * pi + e, g.izipv, vencode(venc, kmax, g.izipv, c_old, v));
        }
      }
      f2loc = fmaxf(f2loc, (float)(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]));
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    f2loc = fmaxf(f2loc, __shfl_xor(f2loc, o));
    npl += __shfl_xor(npl, o);
  }
  if (threadIdx.x == 0) {
    atomicMax(f2max_bits, __float_as_uint(f2loc));
    atomicAdd(npairs_out, npl);
    atomicAdd(nupd_out, (unsigned long long)ninner);
  }
}

// ------------------------------------------------------------ binned PP
// One workgroup per PB^3 coarse cells.  Candidates (particles of the block
// and its neighbour cells whose fine cell is within pp_range of the block)
// are counting-sorted by fine cell in LDS; each block particle then visits
// only the (2 pp_range+1)^3 neighbouring fine-cell runs, which is exactly the
// linked-list neighbourhood of ext_pp_force.f90:92-110.  Blocks whose
// candidates exceed PB_CAP are appended to an overflow list and handled by
// k_pp above.
#define PB 4
#define PB_NT 256
#define PB_CAP 2048
#define PB_WMAX 20  // window width 4*PB + 2*pp_range for pp_range <= 2
struct PPent { float x, y, z; int32_t p; };
__global__ __launch_bounds__(PB_NT) void k_pp_binned(Geo g, const void* __restrict__ xp, void* vp,
                                                     const int64_t* __restrict__ cstart, float mass, float rsoft, int pr,
                                                     float a_mid, float dt, double c_old, const float* __restrict__ vdec,
                                                     const double* __restrict__ venc, int32_t kmax, uint32_t* f2max_bits,
                                                     unsigned long long* npairs_out, unsigned long long* nupd_out,
                                                     double* probe, int32_t* ovf_list, uint32_t* ovf_n) {
  constexpr int SPAN = PB + 2, NCELL = SPAN * SPAN * SPAN, FB = 4 * PB;
  __shared__ int64_t beg[NCELL];
  __shared__ int32_t off[NCELL + 1];
  __shared__ int32_t wcnt[PB_WMAX * PB_WMAX * PB_WMAX + 1];
  __shared__ PPent ent[PB_CAP];
  __shared__ int32_t ncand, ninner_s;
  __shared__ float redf[PB_NT / 64];
  __shared__ unsigned long long redn[PB_NT / 64];
  const int nbx = g.nc / PB;
  const int c0x = (blockIdx.x % nbx) * PB, c0y = ((blockIdx.x / nbx) % nbx) * PB, c0z = (blockIdx.x / (nbx * nbx)) * PB;
  const int W = FB + 2 * pr, W3 = W * W * W, lo = -pr;
  const int tid = threadIdx.x;
  for (int i = tid; i < NCELL; i += PB_NT) {
    int rx = i % SPAN - 1, ry = (i / SPAN) % SPAN - 1, rz = i / (SPAN * SPAN) - 1;
    int64_t r = rank_of(g, wrapc(c0x + rx, g.nc), wrapc(c0y + ry, g.nc), wrapc(c0z + rz, g.nc));
    beg[i] = cstart[r];
    off[i + 1] = (int32_t)(cstart[r + 1] - cstart[r]);
  }
  for (int i = tid; i <= W3; i += PB_NT) wcnt[i] = 0;
  if (tid == 0) { ncand = 0; ninner_s = 0; }
  __syncthreads();
  if (tid == 0) {
    off[0] = 0;
    for (int i = 0; i < NCELL; i++) off[i + 1] += off[i];
  }
  __syncthreads();
  const int32_t ntot = off[NCELL];
  // pass 1: count candidates per window fine cell
  for (int32_t t = tid; t < ntot; t += PB_NT) {
    int a = 0, b = NCELL;
    while (b - a > 1) { int m = (a + b) >> 1; if (off[m] <= t) a = m; else b = m; }
    const int64_t p = beg[a] + (t - off[a]);
    const int rel[3] = {a % SPAN - 1, (a / SPAN) % SPAN - 1, a / (SPAN * SPAN) - 1};
    int w[3];
    bool keep = true;
    for (int e = 0; e < 3; e++) {
      w[e] = CUBE_NCELL * rel[e] + xsub(load_xu(xp, 3 * p + e, g.izipx), g.izipx) - lo;
      keep = keep && w[e] >= 0 && w[e] < W;
    }
    if (keep) {
      atomicAdd(&wcnt[(w[2] * W + w[1]) * W + w[0]], 1);
      atomicAdd(&ncand, 1);
    }
  }
  __syncthreads();
  const int32_t nc_ = ncand;
  if (nc_ > PB_CAP) {  // too many candidates for LDS: hand the block to the chunked kernel
    if (tid == 0) {
      uint32_t slot = atomicAdd(ovf_n, 1u);
      ovf_list[slot] = blockIdx.x;
    }
    return;
  }
  // exclusive scan of wcnt (W3 <= 13824) by one wavefront: serial chunks + shuffle scan
  if (tid < 64) {
    const int per = (W3 + 63) / 64;
    const int i0 = min(W3, tid * per), i1 = min(W3, i0 + per);
    int32_t s = 0;
    for (int i = i0; i < i1; i++) s += wcnt[i];
    int32_t incl = s;
    for (int o = 1; o < 64; o <<= 1) {
      int32_t v = __shfl_up(incl, o);
      if (tid >= o) incl += v;
    }
    int32_t run = incl - s;
    for (int i = i0; i < i1; i++) {
      int32_t c = wcnt[i];
      wcnt[i] = run;
      run += c;
    }
    if (tid == 63) wcnt[W3] = incl;
  }
  __syncthreads();
  // pass 2: scatter candidates into fine-cell order (wcnt[i] becomes the end of run i)
  for (int32_t t = tid; t < ntot; t += PB_NT) {
    int a = 0, b = NCELL;
    while (b - a > 1) { int m = (a + b) >> 1; if (off[m] <= t) a = m; else b = m; }
    const int64_t p = beg[a] + (t - off[a]);
    const int rel[3] = {a % SPAN - 1, (a / SPAN) % SPAN - 1, a / (SPAN * SPAN) - 1};
    int w[3];
    float x[3];
    bool keep = true;
    for (int e = 0; e < 3; e++) {
      uint32_t u = load_xu(xp, 3 * p + e, g.izipx);
      w[e] = CUBE_NCELL * rel[e] + xsub(u, g.izipx) - lo;
      x[e] = (float)((double)(CUBE_NCELL * rel[e]) + (double)CUBE_NCELL * xoff(u, g.izipx));
      keep = keep && w[e] >= 0 && w[e] < W;
    }
    if (keep) {
      int slot = atomicAdd(&wcnt[(w[2] * W + w[1]) * W + w[0]], 1);
      ent[slot] = PPent{x[0], x[1], x[2], (int32_t)p};
    }
  }
  __syncthreads();
  // after pass 2, wcnt[i] = end of run i = start of run i+1; start of run 0 is 0
  float f2loc = 0.f;
  unsigned long long npl = 0;
  int32_t ninner = 0;
  for (int32_t t = tid; t < nc_; t += PB_NT) {
    const PPent pi = ent[t];
    const int fi[3] = {(int)floorf(pi.x) - lo, (int)floorf(pi.y) - lo, (int)floorf(pi.z) - lo};
    if (fi[0] < pr || fi[0] >= pr + FB || fi[1] < pr || fi[1] >= pr + FB || fi[2] < pr || fi[2] >= pr + FB) continue;
    ninner++;
    long long fa[3] = {0, 0, 0};
    for (int kz = fi[2] - pr; kz <= fi[2] + pr; kz++)
      for (int ky = fi[1] - pr; ky <= fi[1] + pr; ky++) {
        const int rowc = (kz * W + ky) * W;
        const int j0 = (rowc + fi[0] - pr) == 0 ? 0 : wcnt[rowc + fi[0] - pr - 1];
        const int j1 = wcnt[rowc + fi[0] + pr];  // runs of consecutive x cells are contiguous
        for (int j = j0; j < j1; j++) {
          const PPent pj = ent[j];
          npl++;
          float dx = pj.x - pi.x, dy = pj.y - pi.y, dz = pj.z - pi.z;
          float r2 = dx * dx + dy * dy + dz * dz;
          float r = sqrtf(r2);
          if (!(r > rsoft)) continue;
          float rc = r * (1.0f / 16.0f);
          float rc3 = rc * rc * rc;
          float pc = 1.0f - 1.75f * rc3 + 0.75f * rc3 * rc * rc;
          float sc = mass * pc / (r2 * r);
          fa[0] += llrint((double)(sc * dx) * PP_FIX);
          fa[1] += llrint((double)(sc * dy) * PP_FIX);
          fa[2] += llrint((double)(sc * dz) * PP_FIX);
        }
      }
    const double f[3] = {(double)fa[0] * (1.0 / PP_FIX), (double)fa[1] * (1.0 / PP_FIX), (double)fa[2] * (1.0 / PP_FIX)};
    const int64_t p = pi.p;
    if (probe) {
      for (int e = 0; e < 3; e++) probe[3 * p + e] = f[e];
    } else {
      for (int e = 0; e < 3; e++) {
        double v = vdecode(vdec, load_v(vp, 3 * p + e, g.izipv), g.izipv, c_old);
        v = v + f[e] * (double)a_mid * (double)dt / 6.0 / (double)PI_F;
        store_v(vp, 3 * p + e, g.izipv, vencode(venc, kmax, g.izipv, c_old, v));
      }
    }
    f2loc = fmaxf(f2loc, (float)(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]));
  }
  for (int o = 32; o > 0; o >>= 1) {
    f2loc = fmaxf(f2loc, __shfl_xor(f2loc, o));
    npl += __shfl_xor(npl, o);
  }
  atomicAdd(&ninner_s, ninner);
  if ((tid & 63) == 0) { redf[tid >> 6] = f2loc; redn[tid >> 6] = npl; }
  __syncthreads();
  if (tid == 0) {
    float m = 0.f;
    unsigned long long nn = 0;
    for (int i = 0; i < PB_NT / 64; i++) { m = fmaxf(m, redf[i]); nn += redn[i]; }
    atomicMax(f2max_bits, __float_as_uint(m));
    atomicAdd(npairs_out, nn);
    atomicAdd(nupd_out, (unsigned long long)ninner_s);
  }
}

int pp_run(cube_handle* h, float a_mid, float dt, double c_old, cube_pp_out* out, double* probe) {
  Geo g = geo_of(h);
  if (h->nc % PB) return cube_fail(h, CUBE_ERR_ARG, "nc must be a multiple of %d for PP", PB);
  if (h->p.pp_range < 1 || h->p.pp_range > 4)
    return cube_fail(h, CUBE_ERR_ARG, "pp_range=%d outside 1..4 (parameters.f90:62)", h->p.pp_range);
  uint32_t* f2 = h->uflags + 8;
  unsigned long long* cnt = (unsigned long long*)(h->uflags + 16);
  uint32_t* novf = h->uflags + 20;
  HIP_TRY(h, hipMemsetAsync(h->uflags + 8, 0, sizeof(uint32_t) * 16, h->stream));
  const int64_t nb = (h->nc / PB) * (h->nc / PB) * (h->nc / PB);
  int32_t* ovf = (int32_t*)h->order;  // idle outside the drift
  uint32_t nov = 0;
  int pr = prof_begin(h, PS_PP);
  if (h->p.pp_range > 2) {  // window beyond the binned kernel's LDS table: chunked kernel for every block
    const int nbs = (int)((h->nc / PP_BP) * (h->nc / PP_BP) * (h->nc / PP_BP));
    k_pp<<<(unsigned)nbs, PP_NT, 0, h->stream>>>(g, h->xp, h->vp, h->cstart, h->mass_p, h->p.rsoft, h->p.pp_range, a_mid, dt,
                                                 c_old, h->vdec, h->venc, h->vkmax, f2, cnt, cnt + 1, probe, nullptr, 1);
    HIP_TRY(h, hipGetLastError());
  } else {
  k_pp_binned<<<(unsigned)nb, PB_NT, 0, h->stream>>>(g, h->xp, h->vp, h->cstart, h->mass_p, h->p.rsoft, h->p.pp_range, a_mid, dt,
                                                      c_old, h->vdec, h->venc, h->vkmax, f2, cnt, cnt + 1, probe, ovf, novf);
  HIP_TRY(h, hipGetLastError());
  HIP_TRY(h, hipMemcpyAsync(&nov, novf, sizeof(uint32_t), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  if (nov) {
    const int bscale = PB / PP_BP, nsub = bscale * bscale * bscale;
    k_pp<<<(unsigned)((int64_t)nov * nsub), PP_NT, 0, h->stream>>>(g, h->xp, h->vp, h->cstart, h->mass_p, h->p.rsoft, h->p.pp_range,
                                                                   a_mid, dt, c_old, h->vdec, h->venc, h->vkmax, f2, cnt, cnt + 1,
                                                                   probe, ovf, bscale);
    HIP_TRY(h, hipGetLastError());
  }
  }
  prof_end(h, pr);
  uint32_t hv[12];
  HIP_TRY(h, hipMemcpyAsync(hv, h->uflags + 8, sizeof(uint32_t) * 12, hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  float f2max;
  memcpy(&f2max, &hv[0], 4);
  unsigned long long c2[2];
  memcpy(c2, &hv[8], sizeof(c2));
  if (out) {
    out->f2_max_pp = f2max;
    out->npairs = (int64_t)c2[0];
    out->itest1 = (int64_t)c2[1];
  }
  h->pp_overflow_blocks = nov;
  return 0;
}
