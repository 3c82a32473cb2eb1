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
// the per-particle sum is FP64.  The kick and re-quantisation with the
// current sigma (ext_pp_force.f90:90, 114-119) are applied in place.
#include "cube_internal.hpp"
#include <string.h>
#include <algorithm>

#define PP_NT 64
#define PP_BP 2
#define PP_CAP 1024

struct PPcand { float x, y, z; int32_t fc; };

__device__ __forceinline__ int32_t pack_fc(int a, int b, int c) {  // each in [-8, 23]
  return ((a + 8) & 31) | (((b + 8) & 31) << 5) | (((c + 8) & 31) << 10);
}

__global__ __launch_bounds__(PP_NT) void k_pp(Geo g, const void* __restrict__ xp, void* vp, const int64_t* __restrict__ cstart,
                                              float mass, float rsoft, int pp_range, float a_mid, float dt, double c_old,
                                              const float* __restrict__ vdec, const double* __restrict__ venc, int32_t kmax,
                                              uint32_t* f2max_bits, unsigned long long* npairs_out,
                                              unsigned long long* nupd_out, double* probe) {
  constexpr int SPAN = PP_BP + 2, NCELL = SPAN * SPAN * SPAN, FB = 4 * PP_BP;
  __shared__ int64_t beg[NCELL];
  __shared__ int32_t off[NCELL + 1];
  __shared__ PPcand cand[PP_CAP];
  __shared__ int ncand;
  __shared__ int32_t ioff[PP_BP * PP_BP * PP_BP + 1];
  __shared__ int64_t ibeg[PP_BP * PP_BP * PP_BP];
  const int nbx = g.nc / PP_BP;
  const int bx = blockIdx.x % nbx, by = (blockIdx.x / nbx) % nbx, bz = blockIdx.x / (nbx * nbx);
  const int c0x = bx * PP_BP, c0y = by * PP_BP, c0z = bz * PP_BP;
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
    double f[3] = {0, 0, 0};
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
          f[0] += (double)(s * dx);
          f[1] += (double)(s * dy);
          f[2] += (double)(s * dz);
        }
      }
      __syncthreads();
    }
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

int pp_run(cube_handle* h, float a_mid, float dt, double c_old, cube_pp_out* out, double* probe) {
  Geo g = geo_of(h);
  if (h->nc % PP_BP) return cube_fail(h, CUBE_ERR_ARG, "nc must be even for PP");
  if (h->p.pp_range < 1 || h->p.pp_range > 4)
    return cube_fail(h, CUBE_ERR_ARG, "pp_range=%d outside 1..4 (parameters.f90:62)", h->p.pp_range);
  uint32_t* f2 = h->uflags + 8;
  unsigned long long* cnt = (unsigned long long*)(h->uflags + 16);
  HIP_TRY(h, hipMemsetAsync(h->uflags + 8, 0, sizeof(uint32_t) * 16, h->stream));
  int64_t nb = (h->nc / PP_BP) * (h->nc / PP_BP) * (h->nc / PP_BP);
  int pr = prof_begin(h, PS_PP);
  k_pp<<<(unsigned)nb, PP_NT, 0, h->stream>>>(g, h->xp, h->vp, h->cstart, h->mass_p, h->p.rsoft, h->p.pp_range, a_mid, dt, c_old,
                                               h->vdec, h->venc, h->vkmax, f2, cnt, cnt + 1, probe);
  prof_end(h, pr);
  HIP_TRY(h, hipGetLastError());
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
  return 0;
}
