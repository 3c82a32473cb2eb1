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
// does not depend on the order in which candidates landed in LDS.  This
// chunked kernel handles only blocks the faster kernels below cannot hold.  The kick and re-quantisation with the
// current sigma (ext_pp_force.f90:90, 114-119) are applied in place.
#include "cube_internal.hpp"
#include <string.h>
#include <algorithm>
#include <vector>
#include <type_traits>

#define PP_NT 64
#define PP_BP 2
#define PP_CAP 1024
#define PP_FIX 4294967296.0  // 2^32

struct PPcand { float x, y, z; int32_t fc; };

// The r > rsoft test of ext_pp_force.f90:101-106 is decided as the
// reference decides it, sqrt(sum(xvec21**2)) > rsoft in double.  The float
// r2 is within 3 ulps of the exact sum of the squared (exact FP32) offsets,
// so only pairs that close to rsoft^2 are re-tested in double.
__device__ __forceinline__ bool pf_outside_rsoft(float r2, float dx, float dy, float dz, float rsoft, float rsoft2) {
  if (fabsf(r2 - rsoft2) > 6e-7f * rsoft2) return r2 > rsoft2;
  const double e2 = (double)dx * (double)dx + (double)dy * (double)dy + (double)dz * (double)dz;
  return sqrt(e2) > (double)rsoft;
}

__device__ __forceinline__ int32_t pack_fc(int a, int b, int c) {  // each in [-8, 23]
  return ((a + 8) & 31) | (((b + 8) & 31) << 5) | (((c + 8) & 31) << 10);
}

__global__ __launch_bounds__(PP_NT) void k_pp(Geo g, const void* __restrict__ xp, void* vp, const int64_t* __restrict__ cstart,
                                              float mass, float rsoft, int pp_range, float a_mid, float dt, double c_old,
                                              const double* __restrict__ vdt, const double* __restrict__ venc, int32_t kmax,
                                              uint32_t* f2max_bits, unsigned long long* npairs_out,
                                              unsigned long long* nupd_out, double* probe, const int32_t* blist,
                                              const uint32_t* __restrict__ nblist, int bscale) {
  constexpr int SPAN = PP_BP + 2, NCELL = SPAN * SPAN * SPAN, FB = 4 * PP_BP;
  __shared__ int64_t beg[NCELL];
  __shared__ int32_t off[NCELL + 1];
  __shared__ PPcand cand[PP_CAP];
  __shared__ int ncand;
  __shared__ int32_t ioff[PP_BP * PP_BP * PP_BP + 1];
  __shared__ int64_t ibeg[PP_BP * PP_BP * PP_BP];
  // blocks of PP_BP^3 cells: entry vb/bscale^3 of the block list (its length
  // *nblist is read on the device, so the launch needs no host round trip)
  // names a coarser block of (bscale*PP_BP)^3 cells and vb%bscale^3 the
  // sub-block; the workgroups stride over the list
  const int nsub = bscale * bscale * bscale;
  const uint32_t nvb = *nblist * (uint32_t)nsub;
  float f2loc = 0.f;
  unsigned long long npl = 0, ntarg = 0;
  for (uint32_t vb = blockIdx.x; vb < nvb; vb += gridDim.x) {
  int c0x, c0y, c0z;
  {
    const int big = blist[vb / nsub], sub = (int)(vb % nsub);
    const int nbb = g.nc / (PP_BP * bscale);
    c0x = (big % nbb) * PP_BP * bscale + (sub % bscale) * PP_BP;
    c0y = ((big / nbb) % nbb) * PP_BP * bscale + ((sub / bscale) % bscale) * PP_BP;
    c0z = (big / (nbb * nbb)) * PP_BP * bscale + (sub / (bscale * bscale)) * PP_BP;
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
  ntarg += (unsigned long long)ninner;
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
          if (!pf_outside_rsoft(r2, dx, dy, dz, rsoft, rsoft * rsoft)) continue;
          float r = sqrtf(r2);
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
          double v = vdecode(vdt, load_v(vp, 3 * pi + e, g.izipv), g.izipv);
          v = v + f[e] * (double)a_mid * (double)dt / 6.0 / (double)PI_F;
          store_v(vp, 3 * pi + e, g.izipv, vencode(venc, kmax, g.izipv, c_old, v));
        }
      }
      f2loc = fmaxf(f2loc, (float)(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]));
    }
  }
  __syncthreads();  // the next block overwrites the staged tables
  }
  for (int o = 32; o > 0; o >>= 1) {
    f2loc = fmaxf(f2loc, __shfl_xor(f2loc, o));
    npl += __shfl_xor(npl, o);
  }
  if (threadIdx.x == 0 && (npl || ntarg)) {
    const int sl = blockIdx.x & 1023;
    atomicMax(f2max_bits + sl, __float_as_uint(f2loc));
    atomicAdd(npairs_out + sl, npl);
    atomicAdd(nupd_out + sl, ntarg);
  }
}

// ------------------------------------------------------------ fine-cell PP
// One workgroup per PB^3 coarse cells (a 4PB-wide fine-cell block).
// Candidate partners -- particles of the block and of its neighbour coarse
// cells whose fine cell lies within pp_range of the block, a W^3 window with
// W = 4PB + 2 pp_range -- are counting-sorted by window fine cell, so each
// block particle visits exactly the (2 pp_range+1)^3 neighbouring fine cells
// of ext_pp_force.f90:92-110 as (2 pp_range+1)^2 contiguous x-runs.
//   fast path : the sorted candidates sit in LDS (up to the cap);
//   dense path: blocks with more candidates are listed; a prep kernel sorts
//               their candidates into a global pool (the fine work mesh, idle
//               between drift and PM) and cuts their targets into work items
//               of PD_NT, which a pairs kernel walks grid-stride.  A block the
//               pool cannot hold goes to the chunked kernel above.
// Every sum has a fixed value whatever the thread timing: the fast path
// sorts each fine cell's candidates by particle index and sums in double in
// that order; the dense path sums in 2^-32 fixed point.  Pair forces are
// FP32 with exact FP32 relative positions (FP64 in the oracle precision).
// PB = 4 serves pp_range <= 2, PB = 2 pp_range 3..4 (window <= 20 cells).
#define PD_NT 256
#define PP_SLOTS 1024
#define PF_OWN 4096  // region particles covered by the LDS owner table
struct PPent { float x, y, z; int32_t p; };

struct PPctx {  // per-call constants
  float mass, rsoft, a_mid, dt;
  double c_old;
  const double* vdt;  // decoded velocities for c_old
  const double* venc;
  int32_t kmax, pr;
};

template <int PB>
struct PBgeo {
  static constexpr int SPAN = PB + 2, NCELL = SPAN * SPAN * SPAN, FB = 4 * PB;
};

template <int PB>
__device__ __forceinline__ void pf_origin(const Geo& g, int64_t b, int& x, int& y, int& z) {
  const int nbx = g.nc / PB;
  x = (int)(b % nbx) * PB;
  y = (int)((b / nbx) % nbx) * PB;
  z = (int)(b / ((int64_t)nbx * nbx)) * PB;
}

// inclusive block scan of one int per thread (NT threads); wsum has NT/64 slots
template <int NT>
__device__ __forceinline__ int32_t block_incl_scan(int32_t v, int32_t* wsum) {
  const int lane = threadIdx.x & 63;
  int32_t incl = v;
  for (int o = 1; o < 64; o <<= 1) {
    const int32_t t = __shfl_up(incl, o);
    if (lane >= o) incl += t;
  }
  if (lane == 63) wsum[threadIdx.x >> 6] = incl;
  __syncthreads();
  int32_t wb = 0;
  for (int w = 0; w < (int)(threadIdx.x >> 6); w++) wb += wsum[w];
  __syncthreads();
  return wb + incl;
}

// region = SPAN^3 coarse cells [c0-1, c0+PB]; beg/off staged for the workgroup
template <int PB, int NT>
__device__ __forceinline__ void pf_region(const Geo& g, const int64_t* cstart, int c0x, int c0y, int c0z, int64_t* beg,
                                          int32_t* off, int32_t* wsum) {
  using G = PBgeo<PB>;
  static_assert(G::NCELL <= NT, "one thread per region cell");
  const int i = threadIdx.x;
  int32_t c = 0;
  if (i < G::NCELL) {
    const int rx = i % G::SPAN - 1, ry = (i / G::SPAN) % G::SPAN - 1, rz = i / (G::SPAN * G::SPAN) - 1;
    const int64_t r = rank_of(g, wrapc(c0x + rx, g.nc), wrapc(c0y + ry, g.nc), wrapc(c0z + rz, g.nc));
    beg[i] = cstart[r];
    c = (int32_t)(cstart[r + 1] - cstart[r]);
  }
  const int32_t incl = block_incl_scan<NT>(c, wsum);
  if (i < G::NCELL) off[i + 1] = incl;
  if (i == 0) off[0] = 0;
}

// window cell of region particle t (false if outside the window) and its
// block-local fine coordinates (exact in FP32: at most 20 significant bits)
// region cell of region particle t: from the owner table when given, else
// by binary search of the offsets
template <int PB>
__device__ __forceinline__ bool pf_cand(const Geo& g, const void* xp, const int64_t* beg, const int32_t* off, int32_t t, int pr,
                                        int W, int& wi, PPent& e, const uint8_t* owner = nullptr) {
  using G = PBgeo<PB>;
  int a = 0;
  if (owner) {
    a = owner[t];
  } else {
    int b = G::NCELL;
    while (b - a > 1) {
      const int m = (a + b) >> 1;
      if (off[m] <= t) a = m; else b = m;
    }
  }
  const int64_t p = beg[a] + (t - off[a]);
  const int rel[3] = {a % G::SPAN - 1, (a / G::SPAN) % G::SPAN - 1, a / (G::SPAN * G::SPAN) - 1};
  int w[3];
  float x[3];
  bool keep = true;
  // 4*(u+0.5)/2^(8 izipx) = (2u+1)*2^(1-8 izipx): exact in FP32, as is the sum
  const float res = g.izipx == 1 ? 0.0078125f : 3.0517578125e-05f;
  for (int d = 0; d < 3; d++) {
    const uint32_t u = load_xu(xp, 3 * p + d, g.izipx);
    w[d] = CUBE_NCELL * rel[d] + xsub(u, g.izipx) + pr;
    x[d] = (float)(CUBE_NCELL * rel[d]) + (float)(2 * u + 1) * res;
    keep = keep && w[d] >= 0 && w[d] < W;
  }
  wi = (w[2] * W + w[1]) * W + w[0];
  e = PPent{x[0], x[1], x[2], (int32_t)p};
  return keep;
}

// exclusive scan of cnt[0..n) in place, cnt[n] = total
template <int NT>
__device__ void pf_scan(int32_t* cnt, int n, int32_t* wsum) {
  const int per = (n + NT - 1) / NT;
  const int i0 = min(n, (int)threadIdx.x * per), i1 = min(n, i0 + per);
  int32_t s = 0;
  for (int i = i0; i < i1; i++) s += cnt[i];
  int32_t run = block_incl_scan<NT>(s, wsum) - s;
  for (int i = i0; i < i1; i++) {
    const int32_t c = cnt[i];
    cnt[i] = run;
    run += c;
  }
  if (threadIdx.x == NT - 1) cnt[n] = run;
  __syncthreads();
}

// Pair-force accumulators: FP32 pairs sum in 2^-32 fixed point (exact and
// order-independent); the FP64 oracle path sums doubles.
template <typename RT>
struct PPacc;
template <>
struct PPacc<float> {
  long long a[3] = {0, 0, 0};
  __device__ __forceinline__ void add(int e, float v) { a[e] += llrint((double)v * PP_FIX); }
  __device__ __forceinline__ double get(int e) const { return (double)a[e] * (1.0 / PP_FIX); }
};
// FP32 pairs summed in double in a fixed candidate order (runs sorted by
// particle index), which is as reproducible as the fixed-point sum
struct PPaccOrd {
  double a[3] = {0, 0, 0};
  __device__ __forceinline__ void add(int e, float v) { a[e] += (double)v; }
  __device__ __forceinline__ double get(int e) const { return a[e]; }
};
template <>
struct PPacc<double> {
  double a[3] = {0, 0, 0};
  __device__ __forceinline__ void add(int e, double v) { a[e] += v; }
  __device__ __forceinline__ double get(int e) const { return a[e]; }
};

// pair term of ext_pp_force.f90:99-108 for r > rsoft: m pc(r/16) dx/r^3,
// pc = 1 - 7/4 rc^3 + 3/4 rc^5 as 1 + rc^3 (3/4 rc^2 - 7/4).  FP32 with
// explicit multiply-adds (the pair force has a tolerance bar, not a bit
// bar); r2 from multiply-adds is within 2 ulps of the exact sum of the
// squared offsets, inside the margin of the r > rsoft test
__device__ __forceinline__ float pf_r2(float dx, float dy, float dz) { return fmaf(dz, dz, fmaf(dy, dy, dx * dx)); }
__device__ __forceinline__ float pf_scale(float r2, float ri, float mass) {
  const float rc = r2 * ri * (1.0f / 16.0f);
  const float rc2 = rc * rc;
  const float pc = fmaf(rc2 * rc, fmaf(0.75f, rc2, -1.75f), 1.0f);
  return mass * pc * (ri * ri * ri);
}
__device__ __forceinline__ bool pf_pair(float dx, float dy, float dz, float mass, float rsoft, float rsoft2, float& sc) {
  const float r2 = pf_r2(dx, dy, dz);
  if (!pf_outside_rsoft(r2, dx, dy, dz, rsoft, rsoft2)) return false;
  sc = pf_scale(r2, rsqrtf(r2), mass);
  return true;
}
// the same pair term with the r > rsoft test as a select; an included pair
// has r2 > rsoft2 > 0 to rounding, which keeps the hardware reciprocal
// square root away from denormal inputs
__device__ __forceinline__ float pf_pair_sel(float dx, float dy, float dz, float mass, float rsoft, float rsoft2) {
  const float r2 = pf_r2(dx, dy, dz);
  const bool ok = pf_outside_rsoft(r2, dx, dy, dz, rsoft, rsoft2);
  const float sc = pf_scale(r2, __builtin_amdgcn_rsqf(ok ? r2 : 1.0f), mass);
  return ok ? sc : 0.0f;
}
__device__ __forceinline__ bool pf_pair(double dx, double dy, double dz, double mass, double rsoft, double rsoft2, double& sc) {
  const double r2 = dx * dx + dy * dy + dz * dz;
  (void)rsoft2;
  const double r = sqrt(r2);
  if (!(r > rsoft)) return false;
  const double rc = r * (1.0 / 16.0);
  const double rc3 = rc * rc * rc;
  const double pc = 1.0 - 1.75 * rc3 + 0.75 * rc3 * rc * rc;
  sc = mass * pc / (r2 * r);
  return true;
}
__device__ __forceinline__ double pf_pair_sel(double dx, double dy, double dz, double mass, double rsoft, double rsoft2) {
  double sc;
  return pf_pair(dx, dy, dz, mass, rsoft, rsoft2, sc) ? sc : 0.0;
}

// run-start tables: int32 entries (dense pool), or two 16-bit entries per
// word in LDS (the binned kernel: values never exceed its candidate cap)
struct RunI32 {
  const int32_t* p;
  __device__ __forceinline__ int get(int i) const { return p[i]; }
};
struct RunU16 {
  uint32_t* w;
  __device__ __forceinline__ int get(int i) const { return (int)((w[i >> 1] >> ((i & 1) * 16)) & 0xFFFFu); }
  // add one to entry i, returning its previous value
  __device__ __forceinline__ int inc(int i) const {
    return (int)((atomicAdd(&w[i >> 1], 1u << ((i & 1) * 16)) >> ((i & 1) * 16)) & 0xFFFFu);
  }
};
__host__ __device__ constexpr int pf_wc16_bytes(int W3) { return ((W3 + 3) / 2 * 4 + 15) / 16 * 16; }

// exclusive scan of the 16-bit counts at entries 1..n of t in place, entry
// n+1 = total; each thread owns whole words, so no two threads write one
template <int NT>
__device__ void pf_scan16(RunU16 t, int n, int32_t* wsum) {
  const int nw = (n + 3) / 2;  // words covering entries 0..n+1
  const int per = (nw + NT - 1) / NT;
  const int w0 = min(nw, (int)threadIdx.x * per), w1 = min(nw, w0 + per);
  int32_t s = 0;
  for (int w = w0; w < w1; w++)
    for (int h = 0; h < 2; h++) {
      const int e = 2 * w + h;
      if (e >= 1 && e <= n) s += t.get(e);
    }
  int32_t run = block_incl_scan<NT>(s, wsum) - s;
  for (int w = w0; w < w1; w++) {
    uint32_t word = t.w[w];
    for (int h = 0; h < 2; h++) {
      const int e = 2 * w + h;
      if (e >= 1 && e <= n + 1) {
        const uint32_t c = (word >> (16 * h)) & 0xFFFFu;
        word = (word & ~(0xFFFFu << (16 * h))) | ((uint32_t)run << (16 * h));
        if (e <= n) run += (int32_t)c;
      }
    }
    t.w[w] = word;
  }
  __syncthreads();
}

// force on one target from the candidate runs (start[] exclusive offsets);
// candidate coordinates are exact in FP32, so RT only sets the pair
// arithmetic.  ROWSUM: candidates are in a fixed order, so each x-run is
// summed in RT and added to the accumulator once per run.
template <typename RT, bool ROWSUM, typename Acc, typename EntT, typename RunT>
__device__ __forceinline__ void pf_force(const EntT* ent, const RunT start, int W, int pr, const PPent& ti, float mass_f,
                                         float rsoft_f, Acc& acc, unsigned long long& npl) {
  const int wx = (int)floorf(ti.x) + pr, wy = (int)floorf(ti.y) + pr, wz = (int)floorf(ti.z) + pr;
  const RT mass = (RT)mass_f, rsoft = (RT)rsoft_f, rsoft2 = (RT)rsoft_f * (RT)rsoft_f, tx = (RT)ti.x, ty = (RT)ti.y,
           tz = (RT)ti.z;
  for (int kz = wz - pr; kz <= wz + pr; kz++)
    for (int ky = wy - pr; ky <= wy + pr; ky++) {
      const int row = (kz * W + ky) * W;
      const int j0 = start.get(row + wx - pr), j1 = start.get(row + wx + pr + 1);
      npl += (unsigned long long)(j1 - j0);
      RT sx = 0, sy = 0, sz = 0;
      for (int j = j0; j < j1; j++) {
        const PPent pj = ent[j];
        const RT dx = (RT)pj.x - tx, dy = (RT)pj.y - ty, dz = (RT)pj.z - tz;
        RT sc;
        if (ROWSUM) {
          // branch-free: an excluded pair adds sc = 0 (a signed zero, which
          // leaves the running sum unchanged)
          sc = pf_pair_sel(dx, dy, dz, mass, rsoft, rsoft2);
          sx = fma(sc, dx, sx);
          sy = fma(sc, dy, sy);
          sz = fma(sc, dz, sz);
        } else {
          if (!pf_pair(dx, dy, dz, mass, rsoft, rsoft2, sc)) continue;
          acc.add(0, sc * dx);
          acc.add(1, sc * dy);
          acc.add(2, sc * dz);
        }
      }
      if (ROWSUM && j1 > j0) {
        acc.add(0, sx);
        acc.add(1, sy);
        acc.add(2, sz);
      }
    }
}

// kick (ext_pp_force.f90:114-119) from prefetched codes, or probe write;
// returns |F|^2 in float
template <typename Acc>
__device__ __forceinline__ float pf_apply(const Geo& g, const PPctx& c, void* vp, double* probe, int64_t p, const int32_t vc[3],
                                          const Acc& acc) {
  const double f[3] = {acc.get(0), acc.get(1), acc.get(2)};
  if (probe) {
    for (int e = 0; e < 3; e++) probe[3 * p + e] = f[e];
  } else {
    for (int e = 0; e < 3; e++) {
      double v = vdecode(c.vdt, vc[e], g.izipv);
      v = v + f[e] * (double)c.a_mid * (double)c.dt / 6.0 / (double)PI_F;
      store_v(vp, 3 * p + e, g.izipv, vencode(c.venc, c.kmax, g.izipv, c.c_old, v));
    }
  }
  return (float)(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
}

// per-workgroup totals: f2 max, pair count, targets
template <int NT>
__device__ void pf_reduce(float f2loc, unsigned long long npl, int32_t ninner, float* redf, unsigned long long* redn,
                          int32_t* redi, uint32_t* f2max_bits, unsigned long long* npairs_out, unsigned long long* nupd_out) {
  for (int o = 32; o > 0; o >>= 1) {
    f2loc = fmaxf(f2loc, __shfl_xor(f2loc, o));
    npl += __shfl_xor(npl, o);
    ninner += __shfl_xor(ninner, o);
  }
  if ((threadIdx.x & 63) == 0) {
    redf[threadIdx.x >> 6] = f2loc;
    redn[threadIdx.x >> 6] = npl;
    redi[threadIdx.x >> 6] = ninner;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float m = 0.f;
    unsigned long long nn = 0, ni = 0;
    for (int i = 0; i < NT / 64; i++) { m = fmaxf(m, redf[i]); nn += redn[i]; ni += (unsigned long long)redi[i]; }
    // one of PP_SLOTS slots per workgroup: millions of workgroups must not
    // meet on one address (max and integer sums are order-independent)
    const int sl = blockIdx.x & (PP_SLOTS - 1);
    if (m > 0.f) atomicMax(f2max_bits + sl, __float_as_uint(m));
    if (nn) atomicAdd(npairs_out + sl, nn);
    if (ni) atomicAdd(nupd_out + sl, ni);
  }
  __syncthreads();
}

template <int PB>
struct PFcfg;
template <>
struct PFcfg<4> { static constexpr int NT = 512, K = 4, CAP = 2048; };
template <>
struct PFcfg<2> { static constexpr int NT = 128, K = 4, CAP = 1024; };

__host__ __device__ constexpr int pf_wc_bytes(int W3) { return ((W3 + 2) * 4 + 15) / 16 * 16; }

// LDS (dynamic): the run table, W^3+2 16-bit entries (entry 1+i: count,
// then start, then end of run i; entry 0 = 0), then CAP entries
template <int PB, typename RT>
__global__ __launch_bounds__(PFcfg<PB>::NT) void k_pp_fast(Geo g, PPctx c, const void* __restrict__ xp, void* vp,
                                                          const int64_t* __restrict__ cstart, uint32_t* f2max_bits,
                                                          unsigned long long* npairs_out, unsigned long long* nupd_out,
                                                          double* probe, int32_t* dense, uint32_t* ndense) {
  using G = PBgeo<PB>;
  constexpr int NT = PFcfg<PB>::NT, K = PFcfg<PB>::K, CAP = PFcfg<PB>::CAP;
  extern __shared__ __align__(16) uint8_t pf_dyn[];
  __shared__ int64_t beg[G::NCELL];
  __shared__ int32_t off[G::NCELL + 1];
  __shared__ int32_t ncand_s, wsum[NT / 64], redi[NT / 64];
  __shared__ float redf[NT / 64];
  __shared__ unsigned long long redn[NT / 64];
  __shared__ uint8_t owner[PF_OWN];
  __shared__ int32_t inoff[PB * PB * PB + 1];
  __shared__ uint8_t ia[PB * PB * PB];
  int c0x, c0y, c0z;
  pf_origin<PB>(g, blockIdx.x, c0x, c0y, c0z);
  const int pr = c.pr, W = G::FB + 2 * pr, W3 = W * W * W;
  const RunU16 wc{(uint32_t*)pf_dyn};
  PPent* ent = (PPent*)(pf_dyn + pf_wc16_bytes(W3));
  for (int i = threadIdx.x; i < (W3 + 3) / 2; i += NT) wc.w[i] = 0u;
  if (threadIdx.x == 0) ncand_s = 0;
  pf_region<PB, NT>(g, cstart, c0x, c0y, c0z, beg, off, wsum);
  __syncthreads();
  const int32_t ntot = off[G::NCELL];
  // owner table: region cell of each of the first PF_OWN region particles
  static_assert(G::NCELL <= 256, "owner entries are bytes");
  const bool use_owner = ntot <= PF_OWN;
  if (use_owner && (int)threadIdx.x < G::NCELL)
    for (int32_t t = off[threadIdx.x]; t < off[threadIdx.x + 1]; t++) owner[t] = (uint8_t)threadIdx.x;
  __syncthreads();
  const uint8_t* own = use_owner ? owner : nullptr;
  // pass 1: the first K*NT region particles are kept in registers
  PPent re[K];
  int rw[K];
  bool rk[K];
  int32_t mine = 0;
#pragma unroll
  for (int k = 0; k < K; k++) {
    const int32_t t = threadIdx.x + k * NT;
    rk[k] = t < ntot && pf_cand<PB>(g, xp, beg, off, t, pr, W, rw[k], re[k], own);
  }
#pragma unroll
  for (int k = 0; k < K; k++)
    if (rk[k]) {
      wc.inc(1 + rw[k]);
      mine++;
    }
  for (int32_t t = threadIdx.x + K * NT; t < ntot; t += NT) {
    int wi;
    PPent e;
    if (pf_cand<PB>(g, xp, beg, off, t, pr, W, wi, e)) {
      wc.inc(1 + wi);
      mine++;
    }
  }
  if (mine) atomicAdd(&ncand_s, mine);
  __syncthreads();
  const int32_t ncand = ncand_s;
  if (ncand > CAP) {
    if (threadIdx.x == 0) dense[atomicAdd(ndense, 1u)] = (int32_t)blockIdx.x;
    return;
  }
  pf_scan16<NT>(wc, W3, wsum);
  // scatter; afterwards entry 1+i is the end of run i, so entry i is the start of run i
#pragma unroll
  for (int k = 0; k < K; k++)
    if (rk[k]) ent[wc.inc(1 + rw[k])] = re[k];
  for (int32_t t = threadIdx.x + K * NT; t < ntot; t += NT) {
    int wi;
    PPent e;
    if (pf_cand<PB>(g, xp, beg, off, t, pr, W, wi, e)) ent[wc.inc(1 + wi)] = e;
  }
  __syncthreads();
  // runs in particle-index order, so the pair sums below have a fixed
  // order: each entry moves to its run's start plus its rank by particle
  // index within the run (entries are held in registers across the barrier)
  {
    constexpr int PE = (CAP + NT - 1) / NT;
    PPent me[PE];
    int dst[PE];
#pragma unroll
    for (int k = 0; k < PE; k++) {
      const int m = threadIdx.x + k * NT;
      dst[k] = -1;
      if (m < ncand) {
        me[k] = ent[m];
        const int wi = (((int)floorf(me[k].z) + pr) * W + ((int)floorf(me[k].y) + pr)) * W + (int)floorf(me[k].x) + pr;
        const int a = wc.get(wi), b = wc.get(wi + 1);
        int rank = 0;
        if (b - a > 1)
          for (int j = a; j < b; j++) rank += ent[j].p < me[k].p ? 1 : 0;
        dst[k] = a + rank;
      }
    }
    __syncthreads();
    // the owner table is dead now; its bytes hold tpos[t], the sorted
    // position of region particle t < PF_OWN/2, for the target lookup below
    uint16_t* tpos = reinterpret_cast<uint16_t*>(owner);
#pragma unroll
    for (int k = 0; k < PE; k++)
      if (dst[k] >= 0) {
        ent[dst[k]] = me[k];
        const int rx = (int)floorf(me[k].x * 0.25f), ry = (int)floorf(me[k].y * 0.25f), rz = (int)floorf(me[k].z * 0.25f);
        const int a = ((rz + 1) * G::SPAN + (ry + 1)) * G::SPAN + (rx + 1);
        const int32_t t = off[a] + (int32_t)((int64_t)me[k].p - beg[a]);
        if (t < PF_OWN / 2) tpos[t] = (uint16_t)dst[k];
      }
  }
  __syncthreads();
  // targets: the particles of the workgroup's own PB^3 coarse cells, dealt
  // out densely to the threads (inner cell ib <-> region cell ia[ib])
  if (threadIdx.x < 64) {
    const int ib = threadIdx.x;
    int32_t cnt = 0;
    if (ib < PB * PB * PB) {
      const int a = ((ib / (PB * PB)) + 1) * G::SPAN * G::SPAN + ((ib / PB) % PB + 1) * G::SPAN + (ib % PB) + 1;
      ia[ib] = (uint8_t)a;
      cnt = off[a + 1] - off[a];
    }
    int32_t incl = cnt;
    for (int o = 1; o < 64; o <<= 1) {
      const int32_t v = __shfl_up(incl, o);
      if (ib >= o) incl += v;
    }
    if (ib < PB * PB * PB) inoff[ib + 1] = incl;
    if (ib == 0) inoff[0] = 0;
  }
  __syncthreads();
  const int32_t ntarget = inoff[PB * PB * PB];
  float f2loc = 0.f;
  unsigned long long npl = 0;
  int32_t ninner = 0;
  for (int32_t k = threadIdx.x; k < ntarget; k += NT) {
    int lo = 0, hi = PB * PB * PB;
    while (hi - lo > 1) {
      const int m = (lo + hi) >> 1;
      if (inoff[m] <= k) lo = m; else hi = m;
    }
    const int a = ia[lo];
    const int32_t t = off[a] + (k - inoff[lo]);
    PPent ti;
    if (t < PF_OWN / 2) {
      ti = ent[reinterpret_cast<const uint16_t*>(owner)[t]];
    } else {
      int wi_unused;
      (void)pf_cand<PB>(g, xp, beg, off, t, pr, W, wi_unused, ti);
    }
    ninner++;
    int32_t vc[3] = {0, 0, 0};
    if (!probe)
      for (int e = 0; e < 3; e++) vc[e] = load_v(vp, 3 * (int64_t)ti.p + e, g.izipv);
    typename std::conditional<sizeof(RT) == 4, PPaccOrd, PPacc<double>>::type acc;
    pf_force<RT, true>(ent, wc, W, pr, ti, c.mass, c.rsoft, acc, npl);
    f2loc = fmaxf(f2loc, pf_apply(g, c, vp, probe, ti.p, vc, acc));
  }
  pf_reduce<NT>(f2loc, npl, ninner, redf, redn, redi, f2max_bits, npairs_out, nupd_out);
}

// Dense blocks, step 1 (one workgroup per listed block, grid-stride): sort
// the candidates into the pool as [start table: W^3+1 int32][entries], and
// cut the block's targets into work items of PD_NT.
template <int PB>
__global__ __launch_bounds__(PD_NT) void k_pp_dense_prep(Geo g, int pr, const void* __restrict__ xp,
                                                         const int64_t* __restrict__ cstart, const int32_t* __restrict__ dense,
                                                         const uint32_t* __restrict__ ndense, PPent* pool, uint64_t pool_cap,
                                                         unsigned long long* pool_used, int64_t* info, int64_t* items,
                                                         uint32_t* nitems, int32_t* fall, uint32_t* nfall) {
  using G = PBgeo<PB>;
  constexpr int NR = G::FB * G::FB;  // inner rows
  extern __shared__ __align__(16) uint8_t pf_dyn[];
  __shared__ int64_t beg[G::NCELL];
  __shared__ int32_t off[G::NCELL + 1];
  __shared__ int32_t wsum[PD_NT / 64];
  __shared__ int32_t rowc[NR + 1];
  __shared__ unsigned long long base_s;
  const int W = G::FB + 2 * pr, W3 = W * W * W;
  int32_t* wcnt = (int32_t*)pf_dyn;
  const uint64_t tbl_units = (uint64_t)(W3 + 1 + 3) / 4;  // the table in 16-byte units
  const uint32_t nd = *ndense;
  for (uint32_t di = blockIdx.x; di < nd; di += gridDim.x) {
    const int32_t b = dense[di];
    int c0x, c0y, c0z;
    pf_origin<PB>(g, b, c0x, c0y, c0z);
    for (int i = threadIdx.x; i <= W3; i += PD_NT) wcnt[i] = 0;
    pf_region<PB, PD_NT>(g, cstart, c0x, c0y, c0z, beg, off, wsum);
    __syncthreads();
    const int32_t ntot = off[G::NCELL];
    for (int32_t t = threadIdx.x; t < ntot; t += PD_NT) {
      int wi;
      PPent e;
      if (pf_cand<PB>(g, xp, beg, off, t, pr, W, wi, e)) atomicAdd(&wcnt[wi], 1);
    }
    __syncthreads();
    pf_scan<PD_NT>(wcnt, W3, wsum);
    const int32_t ncand = wcnt[W3];
    // targets: inner rows (y, z in [pr, pr+FB)), x-runs [pr, pr+FB)
    int32_t n = 0;
    if ((int)threadIdx.x < NR) {
      const int y = (threadIdx.x % G::FB) + pr, z = (threadIdx.x / G::FB) + pr;
      const int row = (z * W + y) * W;
      n = wcnt[row + pr + G::FB] - wcnt[row + pr];
    }
    const int32_t incl = block_incl_scan<PD_NT>(n, wsum);
    if ((int)threadIdx.x < NR) rowc[threadIdx.x + 1] = incl;
    if (threadIdx.x == 0) {
      rowc[0] = 0;
      const uint64_t units = tbl_units + (uint64_t)ncand;
      const unsigned long long o = atomicAdd(pool_used, (unsigned long long)units);
      base_s = o + units <= pool_cap ? o : ~0ull;
    }
    __syncthreads();
    const unsigned long long base = base_s;
    const int32_t ntarg = rowc[NR];
    if (base == ~0ull) {
      if (threadIdx.x == 0) fall[atomicAdd(nfall, 1u)] = b;
      __syncthreads();
      continue;
    }
    int32_t* tbl = (int32_t*)(pool + base);
    PPent* ent = pool + base + tbl_units;
    for (int i = threadIdx.x; i <= W3; i += PD_NT) tbl[i] = wcnt[i];
    __syncthreads();
    for (int32_t t = threadIdx.x; t < ntot; t += PD_NT) {
      int wi;
      PPent e;
      if (pf_cand<PB>(g, xp, beg, off, t, pr, W, wi, e)) ent[atomicAdd(&wcnt[wi], 1)] = e;
    }
    const int nch = (ntarg + PD_NT - 1) / PD_NT;
    if (threadIdx.x == 0) {
      info[2 * di] = (int64_t)base;
      info[2 * di + 1] = ((int64_t)b << 32) | (int64_t)ntarg;
      const uint32_t i0 = atomicAdd(nitems, (uint32_t)nch);
      for (int ch = 0; ch < nch; ch++) items[i0 + ch] = ((int64_t)di << 32) | ch;
    }
    __syncthreads();
  }
}

// Dense blocks, step 2: one work item (PD_NT targets of one block) per
// workgroup iteration; candidates are read from the pool.
template <int PB, typename RT>
__global__ __launch_bounds__(PD_NT) void k_pp_dense_pairs(Geo g, PPctx c, void* vp, const PPent* __restrict__ pool,
                                                          const int64_t* __restrict__ info, const int64_t* __restrict__ items,
                                                          const uint32_t* __restrict__ nitems, uint32_t* f2max_bits,
                                                          unsigned long long* npairs_out, unsigned long long* nupd_out,
                                                          double* probe) {
  using G = PBgeo<PB>;
  constexpr int NR = G::FB * G::FB;
  extern __shared__ __align__(16) uint8_t pf_dyn[];
  __shared__ int32_t rowc[NR + 1];
  __shared__ int32_t wsum[PD_NT / 64];
  __shared__ float redf[PD_NT / 64];
  __shared__ unsigned long long redn[PD_NT / 64];
  __shared__ int32_t redi[PD_NT / 64];
  const int pr = c.pr, W = G::FB + 2 * pr, W3 = W * W * W;
  int32_t* st = (int32_t*)pf_dyn;
  const uint64_t tbl_units = (uint64_t)(W3 + 1 + 3) / 4;
  const uint32_t ni = *nitems;
  for (uint32_t it = blockIdx.x; it < ni; it += gridDim.x) {
    const int64_t item = items[it];
    const int64_t di = item >> 32, ch = item & 0xffffffff;
    const uint64_t base = (uint64_t)info[2 * di];
    const int32_t ntarg = (int32_t)(info[2 * di + 1] & 0xffffffff);
    const int32_t* tbl = (const int32_t*)(pool + base);
    const PPent* ent = pool + base + tbl_units;
    for (int i = threadIdx.x; i <= W3; i += PD_NT) st[i] = tbl[i];
    __syncthreads();
    int32_t n = 0;
    if ((int)threadIdx.x < NR) {
      const int y = (threadIdx.x % G::FB) + pr, z = (threadIdx.x / G::FB) + pr;
      const int row = (z * W + y) * W;
      n = st[row + pr + G::FB] - st[row + pr];
    }
    const int32_t incl = block_incl_scan<PD_NT>(n, wsum);
    if ((int)threadIdx.x < NR) rowc[threadIdx.x + 1] = incl;
    if (threadIdx.x == 0) rowc[0] = 0;
    __syncthreads();
    float f2loc = 0.f;
    unsigned long long npl = 0;
    int32_t ninner = 0;
    const int32_t k = (int32_t)ch * PD_NT + threadIdx.x;
    if (k < ntarg) {
      int a = 0, b = NR;
      while (b - a > 1) {
        const int m = (a + b) >> 1;
        if (rowc[m] <= k) a = m; else b = m;
      }
      const int y = (a % G::FB) + pr, z = (a / G::FB) + pr;
      const int32_t e = st[(z * W + y) * W + pr] + (k - rowc[a]);
      const PPent ti = ent[e];
      int32_t vc[3] = {0, 0, 0};
      if (!probe)
        for (int q = 0; q < 3; q++) vc[q] = load_v(vp, 3 * (int64_t)ti.p + q, g.izipv);
      PPacc<RT> acc;
      pf_force<RT, false>(ent, RunI32{st}, W, pr, ti, c.mass, c.rsoft, acc, npl);
      f2loc = pf_apply(g, c, vp, probe, ti.p, vc, acc);
      ninner = 1;
    }
    pf_reduce<PD_NT>(f2loc, npl, ninner, redf, redn, redi, f2max_bits, npairs_out, nupd_out);
  }
}

template <int PB, typename RT>
static int pp_launch(cube_handle* h, const Geo& g, const PPctx& c, uint32_t* f2, unsigned long long* npr, unsigned long long* nup,
                     double* probe, int32_t* dense, uint32_t* ndense, int64_t* info, int64_t* items, uint32_t* nitems,
                     int32_t* fall, uint32_t* nfall, unsigned long long* pool_used) {
  using G = PBgeo<PB>;
  const int W = G::FB + 2 * c.pr, W3 = W * W * W;
  const int64_t nbx = h->nc / PB, nb = nbx * nbx * nbx;
  const size_t lds_fast = pf_wc16_bytes(W3) + sizeof(PPent) * PFcfg<PB>::CAP;
  const size_t lds_tab = sizeof(int32_t) * (W3 + 1);
  const uint64_t pool_cap = h->fine.bytes_R() / sizeof(PPent);
  HIP_TRY(h, hipFuncSetAttribute((const void*)k_pp_fast<PB, RT>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_fast));
  k_pp_fast<PB, RT><<<(unsigned)nb, PFcfg<PB>::NT, lds_fast, h->stream>>>(g, c, h->xp, h->vp, h->cstart, f2, npr, nup, probe, dense,
                                                                      ndense);
  HIP_TRY(h, hipGetLastError());
  k_pp_dense_prep<PB><<<2048, PD_NT, lds_tab, h->stream>>>(g, c.pr, h->xp, h->cstart, dense, ndense, (PPent*)h->fine.W, pool_cap,
                                                            pool_used, info, items, nitems, fall, nfall);
  HIP_TRY(h, hipGetLastError());
  k_pp_dense_pairs<PB, RT><<<8192, PD_NT, lds_tab, h->stream>>>(g, c, h->vp, (const PPent*)h->fine.W, info, items, nitems, f2, npr, nup,
                                                             probe);
  HIP_TRY(h, hipGetLastError());
  return 0;
}

// per-slot totals summed on the device into the step report: pairs and
// targets (integer sums) and the f2 maximum (positive floats order as
// their bit patterns)
__global__ __launch_bounds__(256) void k_pp_totals(const unsigned long long* __restrict__ npr,
                                                   const unsigned long long* __restrict__ nup,
                                                   const uint32_t* __restrict__ f2, StepReport* d) {
  __shared__ unsigned long long sp[256], st[256];
  __shared__ uint32_t sf[256];
  unsigned long long a = 0, b = 0;
  uint32_t m = 0;
  for (int i = threadIdx.x; i < PP_SLOTS; i += 256) {
    a += npr[i];
    b += nup[i];
    m = max(m, f2[i]);
  }
  sp[threadIdx.x] = a;
  st[threadIdx.x] = b;
  sf[threadIdx.x] = m;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      sp[threadIdx.x] += sp[threadIdx.x + o];
      st[threadIdx.x] += st[threadIdx.x + o];
      sf[threadIdx.x] = max(sf[threadIdx.x], sf[threadIdx.x + o]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    d->pp_pairs = sp[0];
    d->pp_targets = st[0];
    d->pp_f2 = sf[0];
  }
}

// ext_pp_force as one launch sequence: fast kernel, dense prep and pairs,
// and the chunked kernel over whatever the pool could not hold (its list
// length is read on the device); totals and flags into h->drep
int pp_launch_step(cube_handle* h, float a_mid, float dt, double c_old, double* probe) {
  Geo g = geo_of(h);
  const int prange = h->p.pp_range;
  if (prange < 1 || prange > 4)
    return cube_fail(h, CUBE_ERR_ARG, "pp_range=%d outside 1..4 (parameters.f90:62)", prange);
  const int PBsel = prange <= 2 ? 4 : 2;
  if (h->nc % PBsel) return cube_fail(h, CUBE_ERR_ARG, "nc must be a multiple of %d for PP", PBsel);
  // uflags[9] dense blocks; [10] work items; [11] pool fallbacks; [16..17] pool units used.
  // Per-slot totals in pp_red: pairs and targets [PP_SLOTS] u64 each, f2 max bits [PP_SLOTS] u32.
  uint32_t* ndense = h->uflags + 9;
  uint32_t* nitems = h->uflags + 10;
  uint32_t* nfall = h->uflags + 11;
  unsigned long long* pool_used = (unsigned long long*)(h->uflags + 16);
  unsigned long long* npr = (unsigned long long*)h->pp_red;
  unsigned long long* nup = npr + PP_SLOTS;
  uint32_t* f2 = (uint32_t*)(nup + PP_SLOTS);
  int rc;
  if ((rc = fill_u32(h, h->uflags + 8, 12, 0u))) return rc;
  if ((rc = fill_u32(h, h->pp_red, PP_SLOTS * (2 * sizeof(unsigned long long) + sizeof(uint32_t)) / 4, 0u))) return rc;
  // pp_meta: per 2^3-cell block (the finer of the two block sizes) 2 int64 of
  // info and 2 int32 list entries, plus the work items
  const int64_t nb2 = (h->nc / 2) * (h->nc / 2) * (h->nc / 2);
  int64_t* info = (int64_t*)h->pp_meta;
  int64_t* items = info + 2 * nb2;
  int32_t* dense = (int32_t*)(items + h->p.np_max / PD_NT + nb2 + 1);
  int32_t* fall = dense + nb2;
  PPctx c{h->mass_p, h->p.rsoft, a_mid, dt, c_old, vtab_for(h, c_old), h->venc, h->vkmax, prange};
  int pr = prof_begin(h, PS_PP);
#define PP_ARGS h, g, c, f2, npr, nup, probe, dense, ndense, info, items, nitems, fall, nfall, pool_used
  if (h->prec == 4) rc = PBsel == 4 ? pp_launch<4, float>(PP_ARGS) : pp_launch<2, float>(PP_ARGS);
  else rc = PBsel == 4 ? pp_launch<4, double>(PP_ARGS) : pp_launch<2, double>(PP_ARGS);
#undef PP_ARGS
  if (rc) return rc;
  // blocks the pool could not hold: chunked kernel on their PP_BP^3 sub-blocks
  const int bscale = PBsel / PP_BP;
  k_pp<<<256, PP_NT, 0, h->stream>>>(g, h->xp, h->vp, h->cstart, h->mass_p, h->p.rsoft, prange, a_mid, dt, c_old, c.vdt, h->venc,
                                     h->vkmax, f2, npr, nup, probe, fall, nfall, bscale);
  HIP_TRY(h, hipGetLastError());
  prof_end(h, pr);
  k_pp_totals<<<1, 256, 0, h->stream>>>(npr, nup, f2, h->drep);
  HIP_TRY(h, hipGetLastError());
  HIP_TRY(h, hipMemcpyAsync(h->drep->ppflags, h->uflags + 8, sizeof(h->drep->ppflags), hipMemcpyDeviceToDevice, h->stream));
  return 0;
}

int pp_finish(cube_handle* h, const StepReport* rp, cube_pp_out* out) {
  float f2max;
  memcpy(&f2max, &rp->pp_f2, 4);
  if (out) {
    out->f2_max_pp = f2max;
    out->npairs = (int64_t)rp->pp_pairs;
    out->itest1 = (int64_t)rp->pp_targets;
    out->dense_blocks = (int64_t)rp->ppflags[1];
  }
  h->pp_overflow_blocks = rp->ppflags[3];
  return 0;
}

int pp_run(cube_handle* h, float a_mid, float dt, double c_old, cube_pp_out* out, double* probe) {
  int r = pp_launch_step(h, a_mid, dt, c_old, probe);
  if (r) return r;
  HIP_TRY(h, hipMemcpyAsync(h->rep, h->drep, sizeof(StepReport), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  return pp_finish(h, h->rep, out);
}
