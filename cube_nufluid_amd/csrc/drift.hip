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
// without copies:
//   pass 1 (per particle): destination cell and displacement delta, counted
//            per destination cell;
//   scan   : new cell offsets in checkpoint order;
//   pass 3 (per destination cell): walk the source cells d - delta in the
//            reference's processing order (ilayer of the source k, then k, j,
//            i), take the particles whose delta matches, accumulate vfield_new
//            in that order, then write xp/vp/pid at their arrival rank.
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

// One thread per destination cell (rank r).  part: 3 doubles per block
// (sum_c, sum_r, sum_s of update_particle.f90:174-200).
__global__ __launch_bounds__(DR_NT) void k_drift_pass3(Geo g, const void* __restrict__ xp, const void* __restrict__ vp,
                                                       const int32_t* __restrict__ pid, const int64_t* __restrict__ cstart,
                                                       const float* __restrict__ vfield, const uint16_t* __restrict__ disp,
                                                       const int64_t* __restrict__ cs_new, int64_t n3c, int rx, int ry, int rz,
                                                       int nlayer, float dt_mid, double c_old, const float* __restrict__ vdec,
                                                       const double* __restrict__ venc, int32_t kmax, void* xp_o, void* vp_o,
                                                       int32_t* pid_o, float* vfield_o, int32_t* rhoc_o, double* part,
                                                       uint32_t* flags) {
  __shared__ double red[3][DR_NT / 64];
  const int64_t rdst = (int64_t)blockIdx.x * DR_NT + threadIdx.x;
  double sum_c = 0, sum_r = 0, sum_s = 0;
  const double wv = (double)0.1f;  // weight_v=0.1 is a default-real literal
  const double xscale = (g.izipx == 1 ? 1.0 / 256.0 : 1.0 / 65536.0) * (double)CUBE_NCELL;
  if (rdst < n3c) {
    int32_t c[3];
    cell_of_rank(g, rdst, c);
    int dl[3] = {c[0] % g.nt + 1, c[1] % g.nt + 1, c[2] % g.nt + 1};
    const int64_t base = cs_new[rdst];
    const int32_t nnew = (int32_t)(cs_new[rdst + 1] - base);
    float vf[3];
    for (int e = 0; e < 3; e++) vf[e] = (float)((double)vfield[rdst * 3 + e] * wv);
    float vfn[3] = {0, 0, 0};
    for (int sweep = 0; sweep < 2; sweep++) {
      int32_t k = 0;
      for (int L = 0; L < nlayer; L++) {
        for (int sz = dl[2] - rz; sz <= dl[2] + rz; sz++) {
          if ((sz - 1 + g.ncb) % nlayer != L) continue;
          int dz = dl[2] - sz;
          for (int sy = dl[1] - ry; sy <= dl[1] + ry; sy++) {
            int dy = dl[1] - sy;
            for (int sx = dl[0] - rx; sx <= dl[0] + rx; sx++) {
              int dx = dl[0] - sx;
              uint16_t code = pack_disp(dx, dy, dz);
              int64_t rs = rank_of(g, wrapc(c[0] - dx, g.nc), wrapc(c[1] - dy, g.nc), wrapc(c[2] - dz, g.nc));
              const int64_t p1 = cstart[rs + 1];
              for (int64_t p = cstart[rs]; p < p1; p++) {
                if (disp[p] != code) continue;
                double vreal[3];
                for (int e = 0; e < 3; e++)
                  vreal[e] = vdecode(vdec, load_v(vp, 3 * p + e, g.izipv), g.izipv, c_old) + (double)vfield[rs * 3 + e];
                if (sweep == 0) {
                  for (int e = 0; e < 3; e++) vf[e] = (float)((double)vf[e] + vreal[e]);
                } else if (k < nnew) {
                  const int64_t q = base + k;
                  double vr[3];
                  for (int e = 0; e < 3; e++) {
                    int64_t xn = (int64_t)load_xu(xp, 3 * p + e, g.izipx);
                    if (g.izipx == 1) xn = (int64_t)(int8_t)xn; else xn = (int64_t)(int16_t)xn;
                    xn += (int64_t)nint_d((double)dt_mid * vreal[e] / xscale);
                    store_x(xp_o, 3 * q + e, g.izipx, xn);
                    int32_t code_v = vencode(venc, kmax, g.izipv, c_old, vreal[e] - (double)vfn[e]);
                    store_v(vp_o, 3 * q + e, g.izipv, code_v);
                    vr[e] = vdecode(vdec, code_v, g.izipv, c_old);
                  }
                  pid_o[q] = pid[p];
                  sum_r += vr[0] * vr[0] + vr[1] * vr[1] + vr[2] * vr[2];
                  for (int e = 0; e < 3; e++) vr[e] += (double)vfn[e];
                  sum_s += vr[0] * vr[0] + vr[1] * vr[1] + vr[2] * vr[2];
                }
                k++;
              }
            }
          }
        }
      }
      if (sweep == 0) {
        for (int e = 0; e < 3; e++) vfn[e] = (float)((double)vf[e] / ((double)nnew + wv));
      }
      if (k != nnew) atomicOr(&flags[5], 1u);
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
    double s = 0;
    for (int i = 0; i < DR_NT / 64; i++) s += red[threadIdx.x][i];
    part[(int64_t)blockIdx.x * 3 + threadIdx.x] = s;
  }
}

// ---------------------------------------------------------------- order
// One workgroup per OBX x OBY x OBZ block of destination cells (one thread
// each).  The displacement codes of the source region (block grown by the
// largest displacement) are staged in LDS; each thread walks its source
// cells in the reference's processing order and records, for every arrival,
// its particle index and source cell at the arrival's final slot.
#define OBX 8
#define OBY 8
#define OBZ 4
#define ORD_MAXC 2048
#define ORD_CAP 16384
__global__ __launch_bounds__(OBX * OBY * OBZ) void k_drift_order(Geo g, const int64_t* __restrict__ cstart,
                                                                  const uint16_t* __restrict__ disp,
                                                                  const int64_t* __restrict__ cs_new, int rx, int ry,
                                                                  int rz, int nlayer, int32_t* order, int32_t* srank,
                                                                  uint32_t* flags) {
  __shared__ int64_t s_beg[ORD_MAXC];
  __shared__ int32_t s_rank[ORD_MAXC];
  __shared__ int32_t s_cnt[ORD_MAXC];
  __shared__ int32_t s_off[ORD_MAXC + 1];
  __shared__ uint16_t s_disp[ORD_CAP];
  __shared__ int32_t s_tmp[OBX * OBY * OBZ];
  const int SX = OBX + 2 * rx, SY = OBY + 2 * ry, SZ = OBZ + 2 * rz, NC = SX * SY * SZ;
  const int nbx = g.nc / OBX, nby = g.nc / OBY;
  const int b = blockIdx.x;
  const int b0x = (b % nbx) * OBX, b0y = ((b / nbx) % nby) * OBY, b0z = (b / (nbx * nby)) * OBZ;
  const int tid = threadIdx.x, NT = OBX * OBY * OBZ;
  // source region: cells b0 - r .. b0 + OB + r - 1 (periodic)
  for (int i = tid; i < NC; i += NT) {
    int ix = i % SX, iy = (i / SX) % SY, iz = i / (SX * SY);
    int64_t rr = rank_of(g, wrapc(b0x - rx + ix, g.nc), wrapc(b0y - ry + iy, g.nc), wrapc(b0z - rz + iz, g.nc));
    s_beg[i] = cstart[rr];
    s_rank[i] = (int32_t)rr;
    s_cnt[i] = (int32_t)(cstart[rr + 1] - cstart[rr]);
  }
  __syncthreads();
  // exclusive prefix of s_cnt: contiguous chunk per thread, then a block scan of chunk sums
  const int per = (NC + NT - 1) / NT;
  const int i0 = min(NC, tid * per), i1 = min(NC, i0 + per);
  int32_t local_sum = 0;
  for (int i = i0; i < i1; i++) local_sum += s_cnt[i];
  s_tmp[tid] = local_sum;
  __syncthreads();
  for (int o = 1; o < NT; o <<= 1) {
    int32_t a = tid >= o ? s_tmp[tid - o] : 0;
    __syncthreads();
    s_tmp[tid] += a;
    __syncthreads();
  }
  int32_t run = s_tmp[tid] - local_sum;
  for (int i = i0; i < i1; i++) {
    s_off[i] = run;
    run += s_cnt[i];
  }
  const int32_t total = s_tmp[NT - 1];
  if (tid == 0) s_off[NC] = total;
  __syncthreads();
  const bool staged = total <= ORD_CAP;
  if (staged) {
    for (int i = tid; i < NC; i += NT) {
      const int32_t o0 = s_off[i], cnt = s_off[i + 1] - o0;
      const int64_t p0 = s_beg[i];
      for (int32_t q = 0; q < cnt; q++) s_disp[o0 + q] = disp[p0 + q];
    }
  }
  __syncthreads();
  // destination cell of this thread
  const int tx = tid % OBX, ty = (tid / OBX) % OBY, tz = tid / (OBX * OBY);
  const int c[3] = {b0x + tx, b0y + ty, b0z + tz};
  const int64_t rdst = rank_of(g, c[0], c[1], c[2]);
  const int dl[3] = {c[0] % g.nt + 1, c[1] % g.nt + 1, c[2] % g.nt + 1};
  const int64_t base = cs_new[rdst];
  const int32_t nnew = (int32_t)(cs_new[rdst + 1] - base);
  int32_t k = 0;
  for (int L = 0; L < nlayer; L++) {
    for (int sz = dl[2] - rz; sz <= dl[2] + rz; sz++) {
      if ((sz - 1 + g.ncb) % nlayer != L) continue;
      const int dz = dl[2] - sz;
      for (int sy = dl[1] - ry; sy <= dl[1] + ry; sy++) {
        const int dy = dl[1] - sy;
        for (int sx = dl[0] - rx; sx <= dl[0] + rx; sx++) {
          const int dx = dl[0] - sx;
          const uint16_t code = pack_disp(dx, dy, dz);
          // region index of the source cell (c - d) relative to origin b0 - r
          const int ri = ((tz - dz + rz) * SY + (ty - dy + ry)) * SX + (tx - dx + rx);
          const int32_t q0 = s_off[ri], q1 = s_off[ri + 1];
          const int64_t p0 = s_beg[ri];
          for (int32_t q = q0; q < q1; q++) {
            const uint16_t dc = staged ? s_disp[q] : disp[p0 + (q - q0)];
            if (dc != code) continue;
            if (k < nnew) {
              order[base + k] = (int32_t)(p0 + (q - q0));
              srank[base + k] = s_rank[ri];
            }
            k++;
          }
        }
      }
    }
  }
  if (k != nnew) atomicOr(&flags[5], 1u);
}

// One thread per destination cell: ordered vfield_new sum, then xp/vp/pid in
// arrival order, and the sigma_v sums.
__global__ __launch_bounds__(DR_NT) void k_drift_write(Geo g, const void* __restrict__ xp, const void* __restrict__ vp,
                                                       const int32_t* __restrict__ pid, const float* __restrict__ vfield,
                                                       const int64_t* __restrict__ cs_new, const int32_t* __restrict__ order,
                                                       const int32_t* __restrict__ srank, int64_t n3c, float dt_mid,
                                                       double c_old, const float* __restrict__ vdec,
                                                       const double* __restrict__ venc, int32_t kmax, void* xp_o, void* vp_o,
                                                       int32_t* pid_o, float* vfield_o, int32_t* rhoc_o, double* part) {
  __shared__ double red[3][DR_NT / 64];
  const int64_t rdst = (int64_t)blockIdx.x * DR_NT + threadIdx.x;
  double sum_c = 0, sum_r = 0, sum_s = 0;
  const double wv = (double)0.1f;  // weight_v=0.1 is a default-real literal
  const double xscale = (g.izipx == 1 ? 1.0 / 256.0 : 1.0 / 65536.0) * (double)CUBE_NCELL;
  if (rdst < n3c) {
    const int64_t base = cs_new[rdst];
    const int32_t nnew = (int32_t)(cs_new[rdst + 1] - base);
    float vf[3];
    for (int e = 0; e < 3; e++) vf[e] = (float)((double)vfield[rdst * 3 + e] * wv);
    for (int32_t k = 0; k < nnew; k++) {
      const int64_t p = order[base + k], rs = srank[base + k];
      for (int e = 0; e < 3; e++) {
        double vreal = vdecode(vdec, load_v(vp, 3 * p + e, g.izipv), g.izipv, c_old) + (double)vfield[rs * 3 + e];
        vf[e] = (float)((double)vf[e] + vreal);
      }
    }
    float vfn[3];
    for (int e = 0; e < 3; e++) vfn[e] = (float)((double)vf[e] / ((double)nnew + wv));
    for (int32_t k = 0; k < nnew; k++) {
      const int64_t p = order[base + k], rs = srank[base + k], q = base + k;
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
  uint32_t* flags = h->uflags;  // [0..2] max disp, [3] too far, [4] frame conflicts, [5] count mismatch
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
  const int rx = (int)hf[0], ry = (int)hf[1], rz = (int)hf[2];
  const bool ordered = (OBX + 2 * rx) * (OBY + 2 * ry) * (OBZ + 2 * rz) <= ORD_MAXC && h->nc % OBX == 0;
  int pid3 = prof_begin(h, PS_DRIFT3);
  if (ordered) {
    const int64_t nbo = (h->nc / OBX) * (h->nc / OBY) * (h->nc / OBZ);
    k_drift_order<<<(unsigned)nbo, OBX * OBY * OBZ, 0, h->stream>>>(g, h->cstart, h->disp, h->cstart2, rx, ry, rz, nlayer,
                                                                    h->order, h->order + h->p.np_max, flags);
    k_drift_write<<<(unsigned)nb3, DR_NT, 0, h->stream>>>(g, h->xp, h->vp, h->pid, h->vfield, h->cstart2, h->order,
                                                           h->order + h->p.np_max, h->n3c, dt_mid, c_old, h->vdec, h->venc,
                                                           h->vkmax, h->xp2, h->vp2, h->pid2, h->vfield2, h->rhoc2, h->dred);
  } else {
    k_drift_pass3<<<(unsigned)nb3, DR_NT, 0, h->stream>>>(g, h->xp, h->vp, h->pid, h->cstart, h->vfield, h->disp, h->cstart2, h->n3c,
                                                           rx, ry, rz, nlayer, dt_mid, c_old, h->vdec,
                                                           h->venc, h->vkmax, h->xp2, h->vp2, h->pid2, h->vfield2, h->rhoc2,
                                                           h->dred, flags);
  }
  prof_end(h, pid3);
  HIP_TRY(h, hipGetLastError());
  double sums[3];
  if ((r = reduce_doubles(h, h->dred, nb3, 3, sums))) return r;
  int64_t npnew;
  HIP_TRY(h, hipMemcpy(&npnew, h->cstart2 + h->n3c, sizeof(int64_t), hipMemcpyDeviceToHost));
  HIP_TRY(h, hipMemcpy(hf, flags, sizeof(hf), hipMemcpyDeviceToHost));
  if (hf[5]) return cube_fail(h, CUBE_ERR_CHECK, "drift re-binning lost particles");
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
