// Particle <-> mesh kernels of particle_mesh (pm.f90): CIC mass assignment on
// the fine and coarse meshes and the per-component kick that applies the
// fine kick (pm.f90:131-156) and the coarse kick (pm.f90:335-372) with their
// intermediate re-quantisation.
//
// Coordinates follow the reference's tile frames: positions are formed in
// double relative to the tile that owns the destination node and rounded to
// float there (pm.f90:87, 138, 255, 348), so every CIC weight equals the
// reference's weight for the same particle.
#include "cube_internal.hpp"

#define CIC_NT 256
#define CIC_U 4  // particles per thread and batch in the CIC kernels

// --------------------------------------------------------------- helpers
// Load cstart for the cells of a SPAN^3 region around block origin c0 (lo
// cells below it) into LDS as per-cell begin/offset arrays (one segment per
// cell).  The offsets are an exclusive scan over the whole workgroup:
// each thread takes PER consecutive cells.
template <int SPAN>
__device__ void load_region(const Geo& g, const int64_t* cstart, int c0x, int c0y, int c0z, int lo,
                            int64_t* beg, int32_t* off) {
  constexpr int NCELL = SPAN * SPAN * SPAN, PER = (NCELL + CIC_NT - 1) / CIC_NT;
  __shared__ int32_t wsum[CIC_NT / 64];
  int32_t cnt[PER];
  int32_t tsum = 0;
#pragma unroll
  for (int k = 0; k < PER; k++) {
    const int i = threadIdx.x * PER + k;
    cnt[k] = 0;
    if (i < NCELL) {
      const int rx = i % SPAN - lo, ry = (i / SPAN) % SPAN - lo, rz = i / (SPAN * SPAN) - lo;
      const int64_t r = rank_of(g, wrapc(c0x + rx, g.nc), wrapc(c0y + ry, g.nc), wrapc(c0z + rz, g.nc));
      const int64_t b = cstart[r];
      beg[i] = b;
      cnt[k] = (int32_t)(cstart[r + 1] - b);
    }
    tsum += cnt[k];
  }
  const int lane = threadIdx.x & 63;
  int32_t incl = tsum;
  for (int o = 1; o < 64; o <<= 1) {
    const int32_t t = __shfl_up(incl, o);
    if (lane >= o) incl += t;
  }
  if (lane == 63) wsum[threadIdx.x >> 6] = incl;
  __syncthreads();
  int32_t run = incl - tsum;
  for (int w = 0; w < (int)(threadIdx.x >> 6); w++) run += wsum[w];
#pragma unroll
  for (int k = 0; k < PER; k++) {
    const int i = threadIdx.x * PER + k;
    if (i < NCELL) off[i] = run;
    run += cnt[k];
  }
  if (threadIdx.x == CIC_NT - 1) off[NCELL] = run;
  __syncthreads();
}
__device__ __forceinline__ int seg_find(const int32_t* off, int n, int32_t t) {
  int lo = 0, hi = n;  // off[lo] <= t < off[hi]
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (off[mid] <= t) lo = mid; else hi = mid;
  }
  return lo;
}

// ------------------------------------------------------------- fine CIC
// One workgroup per BC^3 coarse cells = (4BC)^3 fine nodes, binned in LDS.
// Nodes sit on fine-cell left edges (pm.f90:87 has no -0.5); a particle in
// coarse cell c touches fine nodes 4c..4c+4, so the block gathers cells
// [c0-1, c0+BC) and keeps taps on its own nodes: every node is written once.
// Each weight (the reference's float product) is accumulated as a 64-bit
// fixed-point integer (2^-40 units), so node sums are exact and independent
// of the order in which threads add them; the node value is the sum rounded
// once to float.
#define CIC_FIX 1099511627776.0  // 2^40
// CENTRE: the halo finder's density (halofind.f90:159), whose nodes sit at
// fine-cell centres ('-0.5'): a particle in cell c then touches nodes
// 4c-1..4c+4 and the block gathers cells [c0-1, c0+BC].  R has row pitch
// `pitch`.
template <int BC, typename T, bool CENTRE = false>
__global__ __launch_bounds__(CIC_NT) void k_cic_fine(Geo g, const void* __restrict__ xp, const int64_t* __restrict__ cstart,
                                                     float mass_f, T* R, int64_t N, int64_t pitch, double* part) {
  const T mass = (T)mass_f;
  constexpr int F = 4 * BC, SPAN = BC + (CENTRE ? 2 : 1), NCELL = SPAN * SPAN * SPAN;
  __shared__ unsigned long long s[F * F * F];
  __shared__ int64_t beg[NCELL];
  __shared__ int32_t off[NCELL + 1];
  __shared__ double red[CIC_NT / 64];
  const int nbx = g.nc / BC;
  // XCD-contiguous block order: x neighbours, which write the two halves of
  // shared 128-byte lines, run on one L2
  const int lb = (int)xcd_remap(blockIdx.x, gridDim.x);
  const int bx = lb % nbx, by = (lb / nbx) % nbx, bz = lb / (nbx * nbx);
  const int c0x = bx * BC, c0y = by * BC, c0z = bz * BC;
  // block-local tile frame: 0-based offset of c0 inside its tile
  const int l0x = c0x % g.nt, l0y = c0y % g.nt, l0z = c0z % g.nt;
  for (int i = threadIdx.x; i < F * F * F; i += CIC_NT) s[i] = 0ull;
  load_region<SPAN>(g, cstart, c0x, c0y, c0z, 1, beg, off);
  const int32_t total = off[NCELL];
  // CIC_U particles per thread and batch: their cell searches and position
  // loads are issued together, then their taps added
  for (int32_t t0 = threadIdx.x; t0 < total; t0 += CIC_NT * CIC_U) {
    int cib[CIC_U];
    uint32_t ub[CIC_U][3];
#pragma unroll
    for (int j = 0; j < CIC_U; j++) {
      const int32_t t = t0 + j * CIC_NT;
      cib[j] = t < total ? seg_find(off, NCELL, t) : 0;
      const int64_t p = beg[cib[j]] + (t - off[cib[j]]);
#pragma unroll
      for (int e = 0; e < 3; e++) ub[j][e] = t < total ? load_xu(xp, 3 * p + e, g.izipx) : 0u;
    }
#pragma unroll
    for (int j = 0; j < CIC_U; j++) {
      if (t0 + j * CIC_NT >= total) break;
      const int ci = cib[j];
      int rel[3] = {ci % SPAN - 1, (ci / SPAN) % SPAN - 1, ci / (SPAN * SPAN) - 1};
      int l0[3] = {l0x, l0y, l0z};
      T dx1[3], dx2[3];
      int n1[3];
      for (int e = 0; e < 3; e++) {
        int il = l0[e] + rel[e] + 1;  // tile-local 1-based cell index (0 = left ghost)
        uint32_t u = ub[j][e];
        T tx;
        if constexpr (CENTRE)
          tx = (T)((double)(CUBE_NCELL * (il - 1)) + (double)CUBE_NCELL * xoff(u, g.izipx) - 0.5);
        else
          tx = (T)((double)(CUBE_NCELL * (il - 1)) + (double)CUBE_NCELL * xoff(u, g.izipx));
        int a1 = (int)floor(tx) + 1;
        dx1[e] = (T)a1 - tx;
        dx2[e] = 1 - dx1[e];
        n1[e] = (a1 - 1) - CUBE_NCELL * l0[e];  // block-local node of tap 1
      }
      const T w[8] = {dx1[0] * dx1[1] * dx1[2] * mass, dx2[0] * dx1[1] * dx1[2] * mass,
                      dx1[0] * dx2[1] * dx1[2] * mass, dx1[0] * dx1[1] * dx2[2] * mass,
                      dx1[0] * dx2[1] * dx2[2] * mass, dx2[0] * dx1[1] * dx2[2] * mass,
                      dx2[0] * dx2[1] * dx1[2] * mass, dx2[0] * dx2[1] * dx2[2] * mass};
      const int o[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 1, 1}, {1, 0, 1}, {1, 1, 0}, {1, 1, 1}};
#pragma unroll
      for (int q = 0; q < 8; q++) {
        int x = n1[0] + o[q][0], y = n1[1] + o[q][1], z = n1[2] + o[q][2];
        if ((unsigned)x < (unsigned)F && (unsigned)y < (unsigned)F && (unsigned)z < (unsigned)F)
          atomicAdd(&s[(z * F + y) * F + x], (unsigned long long)llrint((double)w[q] * CIC_FIX));
      }
    }
  }
  __syncthreads();
  double ms = 0;
  for (int i = threadIdx.x; i < F * F * F; i += CIC_NT) {
    int x = i % F, y = (i / F) % F, z = i / (F * F);
    int64_t gx = 4 * c0x + x, gy = 4 * c0y + y, gz = 4 * c0z + z;
    double vd = (double)(long long)s[i] * (1.0 / CIC_FIX);
    R[(gz * N + gy) * pitch + gx] = (T)vd;
    ms += vd;
  }
  for (int off2 = 32; off2 > 0; off2 >>= 1) ms += __shfl_xor(ms, off2);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = ms;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0;
    for (int i = 0; i < CIC_NT / 64; i++) m += red[i];
    part[blockIdx.x] = m;
  }
}

// ------------------------------------------------------------ coarse CIC
// Nodes at coarse-cell centres (pm.f90:255 '-0.5'); a particle in cell c
// touches nodes c-1..c+1, so the block gathers cells [c0-1, c0+BC].
template <int BC, typename T>
__global__ __launch_bounds__(CIC_NT) void k_cic_coarse(Geo g, const void* __restrict__ xp, const int64_t* __restrict__ cstart,
                                                       float mass_f, T* R, double* part) {
  const T mass = (T)mass_f;
  constexpr int SPAN = BC + 2, NCELL = SPAN * SPAN * SPAN;
  __shared__ unsigned long long s[BC * BC * BC];
  __shared__ int64_t beg[NCELL];
  __shared__ int32_t off[NCELL + 1];
  __shared__ double red[CIC_NT / 64];
  const int nbx = g.nc / BC;
  // XCD-contiguous block order: x neighbours, which write the two halves of
  // shared 128-byte lines, run on one L2
  const int lb = (int)xcd_remap(blockIdx.x, gridDim.x);
  const int bx = lb % nbx, by = (lb / nbx) % nbx, bz = lb / (nbx * nbx);
  const int c0x = bx * BC, c0y = by * BC, c0z = bz * BC;
  const int l0x = c0x % g.nt, l0y = c0y % g.nt, l0z = c0z % g.nt;
  for (int i = threadIdx.x; i < BC * BC * BC; i += CIC_NT) s[i] = 0ull;
  load_region<SPAN>(g, cstart, c0x, c0y, c0z, 1, beg, off);
  const int32_t total = off[NCELL];
  for (int32_t t0 = threadIdx.x; t0 < total; t0 += CIC_NT * CIC_U) {
    int cib[CIC_U];
    uint32_t ub[CIC_U][3];
#pragma unroll
    for (int j = 0; j < CIC_U; j++) {
      const int32_t t = t0 + j * CIC_NT;
      cib[j] = t < total ? seg_find(off, NCELL, t) : 0;
      const int64_t p = beg[cib[j]] + (t - off[cib[j]]);
#pragma unroll
      for (int e = 0; e < 3; e++) ub[j][e] = t < total ? load_xu(xp, 3 * p + e, g.izipx) : 0u;
    }
#pragma unroll
    for (int j = 0; j < CIC_U; j++) {
      if (t0 + j * CIC_NT >= total) break;
      const int ci = cib[j];
      int rel[3] = {ci % SPAN - 1, (ci / SPAN) % SPAN - 1, ci / (SPAN * SPAN) - 1};
      int l0[3] = {l0x, l0y, l0z};
      T dx1[3], dx2[3];
      int n1[3];
      for (int e = 0; e < 3; e++) {
        int il = l0[e] + rel[e] + 1;
        uint32_t u = ub[j][e];
        T tx = (T)(((double)(il - 1) + xoff(u, g.izipx)) - 0.5);
        int a1 = (int)floor(tx) + 1;  // 1-based tile-local node
        dx1[e] = (T)a1 - tx;
        dx2[e] = 1 - dx1[e];
        n1[e] = (a1 - 1) - l0[e];
      }
      const T w[8] = {dx1[0] * dx1[1] * dx1[2] * mass, dx2[0] * dx1[1] * dx1[2] * mass,
                      dx1[0] * dx2[1] * dx1[2] * mass, dx1[0] * dx1[1] * dx2[2] * mass,
                      dx1[0] * dx2[1] * dx2[2] * mass, dx2[0] * dx1[1] * dx2[2] * mass,
                      dx2[0] * dx2[1] * dx1[2] * mass, dx2[0] * dx2[1] * dx2[2] * mass};
      const int o[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 1, 1}, {1, 0, 1}, {1, 1, 0}, {1, 1, 1}};
#pragma unroll
      for (int q = 0; q < 8; q++) {
        int x = n1[0] + o[q][0], y = n1[1] + o[q][1], z = n1[2] + o[q][2];
        if ((unsigned)x < (unsigned)BC && (unsigned)y < (unsigned)BC && (unsigned)z < (unsigned)BC)
          atomicAdd(&s[(z * BC + y) * BC + x], (unsigned long long)llrint((double)w[q] * CIC_FIX));
      }
    }
  }
  __syncthreads();
  const int64_t N = g.nc;
  double ms = 0;
  for (int i = threadIdx.x; i < BC * BC * BC; i += CIC_NT) {
    int x = i % BC, y = (i / BC) % BC, z = i / (BC * BC);
    double vd = (double)(long long)s[i] * (1.0 / CIC_FIX);
    R[((int64_t)(c0z + z) * N + (c0y + y)) * (N + 2) + (c0x + x)] = (T)vd;
    ms += vd;
  }
  for (int off2 = 32; off2 > 0; off2 >>= 1) ms += __shfl_xor(ms, off2);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = ms;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0;
    for (int i = 0; i < CIC_NT / 64; i++) m += red[i];
    part[blockIdx.x] = m;
  }
}

int cic_fine(cube_handle* h, double* mass, double* mass_dev) {
  Geo g = geo_of(h);
  if (h->nt % 4) return cube_fail(h, CUBE_ERR_ARG, "nt=%lld must be a multiple of 4", (long long)h->nt);
  int64_t nb = (h->nc / 4) * (h->nc / 4) * (h->nc / 4);
  if (nb > h->dred_n - 16) return cube_fail(h, CUBE_ERR_ARG, "reduction buffer too small");
  int pr = prof_begin(h, PS_CIC_F);
  if (h->prec == 4)
    k_cic_fine<4, float><<<(unsigned)nb, CIC_NT, 0, h->stream>>>(g, h->xp, h->cstart, h->mass_p, (float*)h->fine.R, h->fine.N,
                                                                h->fine.N + 2, h->dred);
  else
    k_cic_fine<4, double><<<(unsigned)nb, CIC_NT, 0, h->stream>>>(g, h->xp, h->cstart, h->mass_p, (double*)h->fine.R, h->fine.N,
                                                                 h->fine.N + 2, h->dred);
  prof_end(h, pr);
  HIP_TRY(h, hipGetLastError());
  if (mass_dev) return reduce_doubles_dev(h, h->dred, nb, 1, mass_dev);
  if (mass) return reduce_doubles(h, h->dred, nb, 1, mass);
  return 0;
}
// halo-finder density at fine-cell centres into out (nf^3 float, pitch nf)
int cic_halo_density(cube_handle* h, float* out) {
  Geo g = geo_of(h);
  if (h->nt % 4) return cube_fail(h, CUBE_ERR_ARG, "nt=%lld must be a multiple of 4", (long long)h->nt);
  int64_t nb = (h->nc / 4) * (h->nc / 4) * (h->nc / 4);
  if (nb > h->dred_n - 16) return cube_fail(h, CUBE_ERR_ARG, "reduction buffer too small");
  k_cic_fine<4, float, true><<<(unsigned)nb, CIC_NT, 0, h->stream>>>(g, h->xp, h->cstart, h->mass_p, out, h->fine.N, h->fine.N,
                                                                    h->dred);
  HIP_TRY(h, hipGetLastError());
  return 0;
}
int cic_coarse(cube_handle* h, double* mass, double* mass_dev) {
  Geo g = geo_of(h);
  int64_t nb = (h->nc / 8) * (h->nc / 8) * (h->nc / 8);
  int pr = prof_begin(h, PS_CIC_C);
  if (h->prec == 4)
    k_cic_coarse<8, float><<<(unsigned)nb, CIC_NT, 0, h->stream>>>(g, h->xp, h->cstart, h->mass_p, (float*)h->coarse.R, h->dred);
  else
    k_cic_coarse<8, double><<<(unsigned)nb, CIC_NT, 0, h->stream>>>(g, h->xp, h->cstart, h->mass_p, (double*)h->coarse.R, h->dred);
  prof_end(h, pr);
  HIP_TRY(h, hipGetLastError());
  if (mass_dev) return reduce_doubles_dev(h, h->dred, nb, 1, mass_dev);
  if (mass) return reduce_doubles(h, h->dred, nb, 1, mass);
  return 0;
}

// ------------------------------------------------------------------ kick
// The fine kick (decode with sigma_old, encode with sigma_new;
// pm.f90:145-154) and the coarse kick (decode/encode with sigma_new;
// pm.f90:353-363), all three components per particle, as two passes over
// the particles like the reference's: the fine pass stores each particle's
// re-quantised code in vp, and the coarse pass decodes it from there.
//
// A workgroup of either pass takes BX consecutive coarse cells of one
// x-row of a tile.  The fine pass stages the fine force nodes its cells'
// particles can touch in LDS, 4*c0 + [0, 4*BX] x [0, 4] x [0, 4] of the
// three components; the coarse pass the coarse nodes c0 - 1 + [0, BX+1] x
// [0, 2] x [0, 2].  Slab rows are dealt to the wavefronts (a row index is
// uniform, so its address is scalar) and a row's nodes to the lanes.  The
// particles then gather their 8 taps per component from LDS.  A float
// in-cell coordinate can round up onto the next cell's far node; those
// rare particles read their fine taps from HBM instead.
//
// Each tap adds (double)(F*k*w1*w2*w3) with k = a_mid*dt/6/pi formed once
// in the mesh precision (the reference forms F*a_mid*dt/6/pi per tap; the
// two agree to rounding).  With a probe array set, a pass writes its
// interpolated forces instead of kicking.  Mesh sizes are powers of two, so
// node indices wrap with masks.
struct KickFine {
  const void* f[3];
  int64_t pf[3];
};

// the cell of particle p among the workgroup's BX cells (offsets cs)
template <int BX>
__device__ __forceinline__ int cell_of(const int64_t* cs, int64_t p) {
  int lo = 0, hi = BX;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (cs[mid] <= p) lo = mid; else hi = mid;
  }
  return lo;
}

// maximum over the wavefront of non-negative floats as their bit patterns
// (which order as the floats do), by DPP: pairs and quads, half rows, rows,
// then rows 0-1 into row 1 and rows 0-3 into row 3; lane 63 holds it
__device__ __forceinline__ uint32_t wave_max_bits(uint32_t v) {
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xb1, 0xf, 0xf, false));   // quad_perm 1,0,3,2
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4e, 0xf, 0xf, false));   // quad_perm 2,3,0,1
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x141, 0xf, 0xf, false));  // row_half_mirror
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x140, 0xf, 0xf, false));  // row_mirror
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false));  // row_bcast:15
  v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false));  // row_bcast:31
  return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

// block reduction of NV non-negative float maxima into slot fields
// f0..f0+NV-1
template <int NV, int W>
__device__ __forceinline__ void slot_max(const float (&mv)[NV], uint32_t (*red)[W], uint32_t* mslot, int f0) {
  const int tid = threadIdx.x, wv = tid >> 6;
#pragma unroll
  for (int d = 0; d < NV; d++) {
    const uint32_t m = wave_max_bits(__float_as_uint(mv[d]));
    if ((tid & 63) == 0) red[d][wv] = m;
  }
  __syncthreads();
  if (tid < NV) {
    uint32_t m = red[tid][0];
    for (int i = 1; i < W; i++) m = max(m, red[tid][i]);
    atomicMax(mslot + (blockIdx.x & (MSLOTS - 1)) * MSLOT_FIELDS + f0 + tid, m);
  }
}

template <typename T, int BX, int NT>
__global__ __launch_bounds__(NT) void k_kick_fine(Geo g, const void* __restrict__ xp, void* vp,
                                                  const int64_t* __restrict__ cstart, KickFine F, int64_t Nf, T kf,
                                                  const double* __restrict__ cnew_p,
                                                  const double* __restrict__ vdt_old, const double* __restrict__ venc,
                                                  int32_t kmax, uint32_t* mslot, double* probe_f) {
  constexpr int W = NT / 64;                               // wavefronts
  constexpr int XF = 4 * BX, SX = XF + 1;                  // slab [3][5][5][SX]
  constexpr int FROW = 25 * SX, NFS = 3 * FROW;
  constexpr int NRF = 75;                                  // rows (d, z, y)
  constexpr int NCH = (XF + 63) / 64;                      // 64-node chunks of a row's first XF nodes
  __shared__ T sF[NFS];
  __shared__ int64_t cs[BX + 1];
  __shared__ uint32_t red[1][W];
  __shared__ double sAt[33];
  const int nbx = g.nc / BX;
  // XCD-contiguous block order: x-adjacent segments share the line that
  // holds a slab row's far node, y-adjacent ones four of their five rows
  const int b = (int)xcd_remap(blockIdx.x, gridDim.x);
  const int32_t c0[3] = {(b % nbx) * BX, (b / nbx) % g.nc, b / (nbx * g.nc)};
  const int tid = threadIdx.x, lane = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const double c_new = *cnew_p;  // vcoef(sigma_vi_new), formed on the device by the drift
  const uint32_t mf = (uint32_t)Nf - 1;
  const int64_t r0 = rank_of(g, c0[0], c0[1], c0[2]);
  if (tid <= BX) cs[tid] = cstart[r0 + tid];
  atan_tab_to_lds(sAt);
  __syncthreads();
  const int64_t p0 = cs[0];
  const int32_t ntot = (int32_t)(cs[BX] - p0);
  // this thread's first particle: its position and velocity codes are
  // requested before the slab and its old velocities decoded while the
  // slab arrives
  int64_t pa = p0 + tid;
  uint32_t xa[3] = {0u, 0u, 0u};
  int32_t va[3] = {0, 0, 0};
  double vo_a[3] = {0.0, 0.0, 0.0};
  if (tid < ntot) {
#pragma unroll
    for (int e = 0; e < 3; e++) xa[e] = load_xu(xp, 3 * pa + e, g.izipx);
#pragma unroll
    for (int e = 0; e < 3; e++) va[e] = load_v(vp, 3 * pa + e, g.izipv);
  }
  // plane q = 5*d + z of the slab (component d, z row) holds the 5 rows
  // y = 0..4; a wavefront takes planes wv, wv + W, ... (uniform, so plane
  // and row addresses are scalar) and its lanes consecutive nodes of a row,
  // at byte offsets that are the same for every row.  The far face x = XF
  // of all 75 rows is loaded by the first 75 threads.  All loads are issued
  // before the LDS stores.
  {
    constexpr int NPL = 15, KPL = (NPL + W - 1) / W;
    T v[KPL][5][NCH], vfar = (T)0;
    const uint32_t xb = 4u * (uint32_t)c0[0];
    uint32_t xo[NCH];
#pragma unroll
    for (int ch = 0; ch < NCH; ch++) xo[ch] = ((xb + (uint32_t)(ch * 64 + lane)) & mf) * (uint32_t)sizeof(T);
#pragma unroll
    for (int k = 0; k < KPL; k++) {
      const int q = wv + k * W;
      if (q < NPL) {
        const int d = q / 5, z = q % 5;
        const T* f = (const T*)(d == 0 ? F.f[0] : (d == 1 ? F.f[1] : F.f[2]));
        const int64_t P = d == 0 ? F.pf[0] : (d == 1 ? F.pf[1] : F.pf[2]);
        const uint32_t gz = (4u * (uint32_t)c0[2] + z) & mf;
        const T* plane = f + (int64_t)gz * Nf * P;
#pragma unroll
        for (int y = 0; y < 5; y++) {
          const char* row = (const char*)(plane + (int64_t)((4u * (uint32_t)c0[1] + y) & mf) * P);
#pragma unroll
          for (int ch = 0; ch < NCH; ch++)
            if (ch * 64 + lane < XF) v[k][y][ch] = *(const T*)(row + xo[ch]);
        }
      }
    }
    if (tid < NRF) {
      const int d = tid / 25, z = (tid / 5) % 5, y = tid % 5;
      const T* f = (const T*)(d == 0 ? F.f[0] : (d == 1 ? F.f[1] : F.f[2]));
      const int64_t P = d == 0 ? F.pf[0] : (d == 1 ? F.pf[1] : F.pf[2]);
      const uint32_t gy = (4u * (uint32_t)c0[1] + y) & mf, gz = (4u * (uint32_t)c0[2] + z) & mf;
      vfar = f[(int64_t)(gz * (uint32_t)Nf + gy) * P + ((xb + XF) & mf)];
    }
    if (tid < ntot) {
#pragma unroll
      for (int e = 0; e < 3; e++) vo_a[e] = vdecode(vdt_old, va[e], g.izipv);
    }
#pragma unroll
    for (int k = 0; k < KPL; k++) {
      const int q = wv + k * W;
      if (q < NPL) {
#pragma unroll
        for (int y = 0; y < 5; y++)
#pragma unroll
          for (int ch = 0; ch < NCH; ch++) {
            const int x = ch * 64 + lane;
            if (x < XF) sF[(q * 5 + y) * SX + x] = v[k][y][ch];
          }
      }
    }
    if (tid < NRF) sF[tid * SX + XF] = vfar;
  }
  __syncthreads();
  // max |F|^2 over the fine nodes this workgroup owns (the slab without its
  // far faces; the workgroups' owned nodes tile the mesh), summed in float
  // in component order: the dt_fine input of pm.f90.  Thread tid takes x =
  // tid % XF of the 16 (z, y) rows tid / XF, tid / XF + NT / XF, ...
  static_assert(NT % XF == 0, "whole rows per pass");
  float f2loc = 0.f;
  {
    const int x = tid % XF;
#pragma unroll
    for (int row = tid / XF; row < 16; row += NT / XF) {
      const int o = ((row >> 2) * 5 + (row & 3)) * SX + x;
      const float a = (float)sF[o], bq = (float)sF[FROW + o], c = (float)sF[2 * FROW + o];
      float s2 = a * a;
      s2 = fmaf(bq, bq, s2);
      s2 = fmaf(c, c, s2);
      f2loc = fmaxf(f2loc, s2);
    }
  }
  for (int32_t k = tid; k < ntot; k += NT) {
    // this particle's codes and old velocities; the next particle's codes
    // are requested now and decoded at the end of this one
    const int64_t p = pa;
    uint32_t xu[3];
    double vo[3];
#pragma unroll
    for (int e = 0; e < 3; e++) {
      xu[e] = xa[e];
      vo[e] = vo_a[e];
    }
    const int32_t kn = k + NT;
    if (kn < ntot) {
      pa = p0 + kn;
#pragma unroll
      for (int e = 0; e < 3; e++) xa[e] = load_xu(xp, 3 * pa + e, g.izipx);
#pragma unroll
      for (int e = 0; e < 3; e++) va[e] = load_v(vp, 3 * pa + e, g.izipv);
    }
    const int lo = cell_of<BX>(cs, p);
    const int32_t c[3] = {c0[0] + lo, c0[1], c0[2]};
    T dx1[3], dx2[3];
    uint32_t f1[3], f2[3];
    int lf[3];
    for (int e = 0; e < 3; e++) {
      const int il = (g.lnt >= 0 ? (c[e] & (g.nt - 1)) : c[e] % g.nt) + 1;
      const int torg = c[e] - (il - 1);
      const double xo = xoff(xu[e], g.izipx);
      const T tx = (T)((double)(CUBE_NCELL * (il - 1)) + (double)CUBE_NCELL * xo);
      const int a1 = (int)floor(tx) + 1;
      dx1[e] = (T)a1 - tx;
      dx2[e] = 1 - dx1[e];
      // a tile-local float coordinate may round up to 4*nt; the mask wraps it
      f1[e] = (uint32_t)(CUBE_NCELL * torg + a1 - 1) & mf;
      f2[e] = (f1[e] + 1) & mf;
      lf[e] = CUBE_NCELL * (c[e] - c0[e]) + (a1 - 1 - CUBE_NCELL * (il - 1));  // 0..4*BX (x), 0..4 (y, z)
    }
    const bool in_slab = lf[0] < XF && lf[1] < 4 && lf[2] < 4;
    // the 8 taps: bit 0 picks x, bit 1 y, bit 2 z (order of pm.f90:146-153)
    const int ord[8] = {0, 1, 2, 4, 6, 5, 3, 7};
    int sf[8];
    T Fw[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int o = ord[q], ox = o & 1, oy = (o >> 1) & 1, oz = o >> 2;
      sf[q] = ((lf[2] + oz) * 5 + lf[1] + oy) * SX + lf[0] + ox;
      Fw[q] = (ox ? dx2[0] : dx1[0]) * (oy ? dx2[1] : dx1[1]) * (oz ? dx2[2] : dx1[2]);
    }
    // each tap's F*k*w in the mesh precision added to the old velocity in
    // tap order (a probe: F*w summed in double), then re-quantised with the
    // new sigma
    int32_t n[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
      double v = probe_f ? 0.0 : vo[d];
      const T kd = probe_f ? (T)1 : kf;
      if (in_slab) {
#pragma unroll
        for (int q = 0; q < 8; q++) v = v + (double)(sF[d * FROW + sf[q]] * kd * Fw[q]);
      } else {
        const T* Ff = (const T*)F.f[d];
        const int64_t P = F.pf[d];
#pragma unroll
        for (int q = 0; q < 8; q++) {
          const int o = ord[q], ox = o & 1, oy = (o >> 1) & 1, oz = o >> 2;
          const uint32_t rowf = (oz ? f2[2] : f1[2]) * (uint32_t)Nf + (oy ? f2[1] : f1[1]);
          v = v + (double)(Ff[(int64_t)rowf * P + (ox ? f2[0] : f1[0])] * kd * Fw[q]);
        }
      }
      if (probe_f) probe_f[3 * p + d] = v;
      else n[d] = venc_fast(venc, sAt, kmax, g.izipv, c_new * v);
    }
    if (!probe_f) {
#pragma unroll
      for (int d = 0; d < 3; d++) store_v(vp, 3 * p + d, g.izipv, n[d]);
    }
    if (kn < ntot) {
#pragma unroll
      for (int e = 0; e < 3; e++) vo_a[e] = vdecode(vdt_old, va[e], g.izipv);
    }
  }
  const float mv[1] = {f2loc};
  slot_max<1, W>(mv, red, mslot, 0);  // slot field 0: max |F|^2
}

template <typename T, int BX, int NT>
__global__ __launch_bounds__(NT) void k_kick_coarse(Geo g, const void* __restrict__ xp, void* vp,
                                                    const int64_t* __restrict__ cstart, int64_t n3c,
                                                    const float* __restrict__ vfield, const T* __restrict__ Fc, T kf,
                                                    const double* __restrict__ cnew_p,
                                                    const double* __restrict__ vdt_new, const double* __restrict__ venc,
                                                    int32_t kmax, uint32_t* mslot, double* probe_c) {
  constexpr int W = NT / 64;
  constexpr int CX = BX + 2, CROW = 9 * CX, NCS = 3 * CROW;  // slab [3][3][3][CX]
  static_assert(CX <= 64, "a wavefront per coarse row");
  __shared__ T sC[NCS];
  __shared__ int64_t cs[BX + 1];
  __shared__ uint32_t red[3][W];
  __shared__ double sAt[33];
  const int nbx = g.nc / BX;
  const int b = blockIdx.x;
  const int32_t c0[3] = {(b % nbx) * BX, (b / nbx) % g.nc, b / (nbx * g.nc)};
  const int tid = threadIdx.x, lane = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const double c_new = *cnew_p;
  const uint32_t Nc = (uint32_t)g.nc, mc = Nc - 1;
  const int64_t r0 = rank_of(g, c0[0], c0[1], c0[2]);
  if (tid <= BX) cs[tid] = cstart[r0 + tid];
  atan_tab_to_lds(sAt);
  __syncthreads();
  const int64_t p0 = cs[0];
  const int32_t ntot = (int32_t)(cs[BX] - p0);
  int64_t pa = p0 + tid;
  uint32_t xa[3] = {0u, 0u, 0u};
  int32_t va[3] = {0, 0, 0};
  double vo_a[3] = {0.0, 0.0, 0.0};
  if (tid < ntot) {
#pragma unroll
    for (int e = 0; e < 3; e++) xa[e] = load_xu(xp, 3 * pa + e, g.izipx);
#pragma unroll
    for (int e = 0; e < 3; e++) va[e] = load_v(vp, 3 * pa + e, g.izipv);
  }
  // plane q = 3*d + z (component d, z row) holds the 3 rows y = 0..2; a
  // wavefront takes planes wv, wv + W, ..., its first CX lanes the nodes
  {
    constexpr int NPL = 9, KPL = (NPL + W - 1) / W;
    T vc[KPL][3];
    const uint32_t gx = (uint32_t)(c0[0] - 1 + lane) & mc;
#pragma unroll
    for (int k = 0; k < KPL; k++) {
      const int q = wv + k * W;
      if (q < NPL) {
        const int d = q / 3, z = q % 3;
        const uint32_t gz = (uint32_t)(c0[2] - 1 + z) & mc;
        const T* plane = Fc + (size_t)d * n3c + (size_t)gz * Nc * Nc;
#pragma unroll
        for (int y = 0; y < 3; y++) {
          const T* row = plane + (size_t)(((uint32_t)(c0[1] - 1 + y) & mc) * Nc);
          if (lane < CX) vc[k][y] = row[gx];
        }
      }
    }
    if (tid < ntot) {
#pragma unroll
      for (int e = 0; e < 3; e++) vo_a[e] = vdecode(vdt_new, va[e], g.izipv);
    }
    // the slab holds F*k (F for a probe): each tap is then (F*k)*w
    const T kd = probe_c ? (T)1 : kf;
#pragma unroll
    for (int k = 0; k < KPL; k++) {
      const int q = wv + k * W;
      if (q < NPL) {
#pragma unroll
        for (int y = 0; y < 3; y++)
          if (lane < CX) sC[(q * 3 + y) * CX + lane] = vc[k][y] * kd;
      }
    }
  }
  __syncthreads();
  float vmax[3] = {0.f, 0.f, 0.f};
  for (int32_t k = tid; k < ntot; k += NT) {
    const int64_t p = pa;
    uint32_t xu[3];
    double vo[3];
#pragma unroll
    for (int e = 0; e < 3; e++) {
      xu[e] = xa[e];
      vo[e] = vo_a[e];
    }
    const int32_t kn = k + NT;
    if (kn < ntot) {
      pa = p0 + kn;
#pragma unroll
      for (int e = 0; e < 3; e++) xa[e] = load_xu(xp, 3 * pa + e, g.izipx);
#pragma unroll
      for (int e = 0; e < 3; e++) va[e] = load_v(vp, 3 * pa + e, g.izipv);
    }
    const int lo = cell_of<BX>(cs, p);
    float vf[3];
#pragma unroll
    for (int e = 0; e < 3; e++) vf[e] = vfield[(r0 + lo) * 3 + e];
    const int32_t c[3] = {c0[0] + lo, c0[1], c0[2]};
    T cx1[3], cx2[3];
    int lc[3];
    for (int e = 0; e < 3; e++) {
      const double xo = xoff(xu[e], g.izipx);
      const T tc = (T)(((double)c[e] + xo) - 0.5);
      const int b1 = (int)floor(tc) + 1;  // 1-based global node: b1-1 is c-1 or c
      cx1[e] = (T)b1 - tc;
      cx2[e] = 1 - cx1[e];
      lc[e] = b1 - c0[e];  // local node of the lower tap: 0..BX (x), 0..1 (y, z)
    }
    const int ord[8] = {0, 1, 2, 4, 6, 5, 3, 7};
    int sc[8];
    T Cw[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int o = ord[q], ox = o & 1, oy = (o >> 1) & 1, oz = o >> 2;
      sc[q] = ((lc[2] + oz) * 3 + lc[1] + oy) * CX + lc[0] + ox;
      Cw[q] = (ox ? cx2[0] : cx1[0]) * (oy ? cx2[1] : cx1[1]) * (oz ? cx2[2] : cx1[2]);
    }
    // the velocity decoded from the fine pass's code, the coarse taps in
    // tap order, vmax of the kicked velocity plus the cell's vfield, and the
    // final code (a probe: F*w summed in double)
    int32_t n[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
      double v = probe_c ? 0.0 : vo[d];
#pragma unroll
      for (int q = 0; q < 8; q++) v = v + (double)(sC[d * CROW + sc[q]] * Cw[q]);
      if (probe_c) {
        probe_c[3 * p + d] = v;
      } else {
        vmax[d] = fmaxf(vmax[d], (float)fabs(v + (double)vf[d]));
        n[d] = venc_fast(venc, sAt, kmax, g.izipv, c_new * v);
      }
    }
    if (!probe_c) {
#pragma unroll
      for (int d = 0; d < 3; d++) store_v(vp, 3 * p + d, g.izipv, n[d]);
    }
    if (kn < ntot) {
#pragma unroll
      for (int e = 0; e < 3; e++) vo_a[e] = vdecode(vdt_new, va[e], g.izipv);
    }
  }
  slot_max<3, W>(vmax, red, mslot, 1);  // slot fields 1..3: vmax
}

template <typename T, int BX, int NT>
static void kick_launch(cube_handle* h, const void* const Ff[3], const int64_t pf[3], float a_mid, float dt,
                        const double* c_new, const double* vo, const double* vn, double* probe_f, double* probe_c) {
  Geo g = geo_of(h);
  KickFine F;
  for (int d = 0; d < 3; d++) {
    F.f[d] = Ff[d];
    F.pf[d] = pf[d];
  }
  const unsigned nb = (unsigned)(h->n3c / BX);
  // k = a_mid*dt/6/pi in the mesh precision, in this operation order
  const T kf = (T)a_mid * (T)dt / (T)6 / (T)PI_F;
  int pr = prof_begin(h, PS_KICK);
  k_kick_fine<T, BX, NT><<<nb, NT, 0, h->stream>>>(g, h->xp, h->vp, h->cstart, F, h->fine.N, kf, c_new, vo, h->venc,
                                                   h->vkmax, h->mslots, probe_f);
  prof_end(h, pr);
  pr = prof_begin(h, PS_KICK_C);
  k_kick_coarse<T, BX, NT><<<nb, NT, 0, h->stream>>>(g, h->xp, h->vp, h->cstart, h->n3c, h->vfield, (const T*)h->fc, kf,
                                                     c_new, vn, h->venc, h->vkmax, h->mslots, probe_c);
  prof_end(h, pr);
}

template <typename T>
static int kick_pick(cube_handle* h, const void* const Ff[3], const int64_t pf[3], float a_mid, float dt,
                     const double* c_new, const double* vo, const double* vn, double* probe_f, double* probe_c) {
  if (h->nt % 32 == 0)
    kick_launch<T, 32, 256>(h, Ff, pf, a_mid, dt, c_new, vo, vn, probe_f, probe_c);
  else if (h->nt % 16 == 0)
    kick_launch<T, 16, 256>(h, Ff, pf, a_mid, dt, c_new, vo, vn, probe_f, probe_c);
  else if (h->nt % 4 == 0)
    kick_launch<T, 4, 256>(h, Ff, pf, a_mid, dt, c_new, vo, vn, probe_f, probe_c);
  else
    return cube_fail(h, CUBE_ERR_ARG, "nt=%lld must be a multiple of 4", (long long)h->nt);
  return 0;
}

int kick_all(cube_handle* h, const void* const Ff[3], const int64_t pf[3], float a_mid, float dt, double c_old,
             uint32_t* vmax_bits, uint32_t* f2max_bits, double* probe_f, double* probe_c) {
  const double* vo = vtab_for(h, c_old);
  const double* c_new = h->dscal;  // and its table h->vdt_new, both written by the drift
  const double* vn = h->vdt_new;
  if (vmax_bits != f2max_bits + 1) return cube_fail(h, CUBE_ERR_ARG, "kick: vmax must follow f2max");
  if (!probe_f != !probe_c) return cube_fail(h, CUBE_ERR_ARG, "kick: probe both force arrays or neither");
  int r0 = fill_u32(h, h->mslots, MSLOTS * MSLOT_FIELDS, 0u);
  if (r0) return r0;
  int r = h->prec == 4
              ? kick_pick<float>(h, Ff, pf, a_mid, dt, c_new, vo, vn, probe_f, probe_c)
              : kick_pick<double>(h, Ff, pf, a_mid, dt, c_new, vo, vn, probe_f, probe_c);
  if (r) return r;
  HIP_TRY(h, hipGetLastError());
  // max |F|^2 always; vmax only for a kick (a probe leaves it unset)
  return slots_reduce(h, probe_f ? 1 : 4, 0u, 0u, f2max_bits);
}
