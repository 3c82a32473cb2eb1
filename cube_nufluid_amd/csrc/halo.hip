// Runtime halo finder (halofind.f90:18-598) on the GPU.
//
// The reference runs tile by tile on the tile's extended arrays, whose
// ghost particles are separate copies with their own p_tag; peaks in one
// tile never see the tags of another.  Here every tile is worked at once:
//   1. cic_halo_density: fine-mesh CIC at cell centres (halofind.f90:145-
//      177) into fine.S, each node in its owning tile's frame, summed in
//      2^-40 fixed point;
//   2. k_halo_peaks: 27-point maxima above den_peak_cutoff (:180-183), one
//      candidate per tile whose extended node range 1-ncell..nft+ncell holds
//      the node;
//   3. k_halo_walk: the mesh-mass walk over the radius-sorted offset table
//      (:186-199) and the parabolic peak position (:203-208), one thread per
//      candidate;
//   4. peaks sorted by (tile in the reference's order, density descending,
//      loop index) with two stable radix sorts (:216-219, :230);
//   5. k_halo_process: a persistent kernel; each workgroup takes the next
//      peak in that order, waits until every earlier peak of its tile whose
//      search sphere overlaps its own is finished (only those can have
//      tagged particles it would gather), then gathers the untagged
//      particles within rsearch (:268-327), builds the refined grid and its
//      parabolic maximum (:293-356), sorts by radius, finds the overdensity
//      radius (:406-418) and tags the members (:421).  The result equals
//      processing the tile's peaks one after another;
//   6. catalogue kernels: halo_info of the physical halos in output order
//      (:422-506), with the running sums the reference never resets
//      (inertia, dx_mean, inert) carried by single-thread scans.
// A particle's tag is a 27-bit mask, one bit per tile holding a copy of it
// (the tile's displacement -1..1 per axis from the particle's own tile).
#include <hipcub/hipcub.hpp>

#include "cube_internal.hpp"

#define HNT 256
#define HFIX 1099511627776.0  // 2^40
#define HMAXCELLS 2048        // search-box cells a workgroup stages in LDS
#define HSPIN_LIMIT (1u << 26)

int cic_halo_density(cube_handle* h, float* out);

static_assert(sizeof(cube_halo_params) == 80, "cube_halo_params layout");
static_assert(sizeof(cube_halo_catalog) == 80, "cube_halo_catalog layout");

namespace {

struct HalopCfg {
  int32_t nc, nt, nnt, ncb, N, nft, nfb;
  float den_odc, cutoff, mass_p;
  int32_t min_np, search_ratio, refine_ratio, newbox_max;
  int64_t nq, capn, gcap, mcap;
  float newgrid;
};

// per-peak data in processing order
struct HPk {
  float h[3];     // parabolic node maximum, tile frame
  float rp, rs;   // r_proxy, rsearch
  int32_t t;      // tile
  int32_t cs0[3], cs1[3];  // search box, tile-local 1-based cells
  int32_t fr0[3], nb[3];   // refined box origin (fine) and newbox
  int32_t bin;
};

struct HRes {
  float hfin[3], radius;
  int32_t i_odc, phys;
  int64_t moff;
};

__device__ __forceinline__ int32_t wrapn(int32_t a, int32_t n) { return ((a % n) + n) % n; }
__device__ __forceinline__ float norm3(float x, float y, float z) { return sqrtf((x * x + y * y) + z * z); }
__device__ __forceinline__ float fmodulo(float a, float p) { return a - floorf(a / p) * p; }
__device__ __forceinline__ float cbrt_ref(float x) { return (float)exp(log((double)x) * (double)(1.0f / 3.0f)); }
__device__ __forceinline__ float log10_ref(float x) { return (float)log10((double)x); }
__device__ __forceinline__ float exp10_ref(float x) { return (float)pow(10.0, (double)x); }

// ----------------------------------------------------------- offset table
// key = r2 << 25 | generation index ((i+R)*D + (j+R))*D + (k+R), i outer
__global__ void k_hoff_gen(int32_t R, uint64_t* keys) {
  const int32_t D = 2 * R + 1;
  const int64_t n = (int64_t)D * D * D;
  for (int64_t gi = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; gi < n; gi += (int64_t)gridDim.x * blockDim.x) {
    const int32_t i = (int32_t)(gi / ((int64_t)D * D)) - R, j = (int32_t)((gi / D) % D) - R, k = (int32_t)(gi % D) - R;
    const int64_t r2 = (int64_t)i * i + (int64_t)j * j + (int64_t)k * k;
    keys[gi] = r2 > (int64_t)R * R ? ~0ull : ((uint64_t)r2 << 25) | (uint64_t)gi;
  }
}
__device__ __forceinline__ void hoff_decode(uint64_t key, int32_t R, int32_t& dx, int32_t& dy, int32_t& dz, int32_t& r2) {
  const int32_t D = 2 * R + 1;
  const int64_t gi = (int64_t)(key & ((1ull << 25) - 1));
  dx = (int32_t)(gi / ((int64_t)D * D)) - R;
  dy = (int32_t)((gi / D) % D) - R;
  dz = (int32_t)(gi % D) - R;
  r2 = (int32_t)(key >> 25);
}

// ------------------------------------------------------------------ peaks
__global__ void k_halo_peaks(const float* __restrict__ rho, HalopCfg c, uint32_t* ncand, int64_t cap, int4* cand) {
  const int32_t N = c.N;
  const int64_t n3 = (int64_t)N * N * N;
  for (int64_t gi = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; gi < n3; gi += (int64_t)gridDim.x * blockDim.x) {
    const float v = rho[gi];
    if (!(v > c.cutoff)) continue;
    const int32_t x = (int32_t)(gi % N), y = (int32_t)((gi / N) % N), z = (int32_t)(gi / ((int64_t)N * N));
    float m = v;
    for (int dz = -1; dz <= 1; dz++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++)
          m = fmaxf(m, rho[((int64_t)wrapn(z + dz, N) * N + wrapn(y + dy, N)) * N + wrapn(x + dx, N)]);
    if (m != v) continue;
    // (tile, tile-local 1-based node) options per axis
    int32_t tt[3][2], ll[3][2], no[3];
    const int32_t co[3] = {x, y, z};
    for (int a = 0; a < 3; a++) {
      const int32_t th = co[a] / c.nft, i0 = co[a] % c.nft + 1;
      tt[a][0] = th; ll[a][0] = i0; no[a] = 1;
      if (i0 <= CUBE_NCELL) { tt[a][1] = wrapn(th - 1, c.nnt); ll[a][1] = i0 + c.nft; no[a] = 2; }
      else if (i0 >= c.nft - CUBE_NCELL + 1) { tt[a][1] = wrapn(th + 1, c.nnt); ll[a][1] = i0 - c.nft; no[a] = 2; }
    }
    for (int a2 = 0; a2 < no[2]; a2++)
      for (int a1 = 0; a1 < no[1]; a1++)
        for (int a0 = 0; a0 < no[0]; a0++) {
          const uint32_t s = atomicAdd(ncand, 1u);
          if ((int64_t)s < cap) {
            const int32_t t = (tt[2][a2] * c.nnt + tt[1][a1]) * c.nnt + tt[0][a0];
            cand[s] = make_int4(t, ll[0][a0], ll[1][a1], ll[2][a2]);
          }
        }
  }
}

// ------------------------------------------------------------ mesh walk
__global__ void k_halo_walk(const float* __restrict__ rho, HalopCfg c, const int4* __restrict__ cand, int64_t ncand,
                            const uint64_t* __restrict__ tab, int64_t ntab, int32_t R, uint32_t* nvalid, int32_t* vidx,
                            float* cden, float* cmesh, float* chpos, uint32_t* cloop) {
  const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (s >= ncand) return;
  const int4 cd = cand[s];
  const int32_t N = c.N, tx = cd.x % c.nnt, ty = (cd.x / c.nnt) % c.nnt, tz = cd.x / (c.nnt * c.nnt);
  const int32_t T0[3] = {tx * c.nft, ty * c.nft, tz * c.nft};
  const int32_t i0 = cd.y, j0 = cd.z, k0 = cd.w;
  auto R3 = [&](int32_t ix, int32_t iy, int32_t iz) {  // tile-local 1-based node
    return rho[((int64_t)wrapn(T0[2] + iz - 1, N) * N + wrapn(T0[1] + iy - 1, N)) * N + wrapn(T0[0] + ix - 1, N)];
  };
  float m_mesh = 0.f;
  const int32_t lo = 1 - c.nfb, hi = c.nft + c.nfb - 1;
  int32_t dx, dy, dz, r2;
  if (ntab > 0) hoff_decode(tab[0], R, dx, dy, dz, r2);
  for (int64_t ii = 1; ii <= ntab; ii++) {
    int32_t nx = 0, ny = 0, nz = 0, nr2 = -1;
    if (ii < ntab) hoff_decode(tab[ii], R, nx, ny, nz, nr2);
    const int32_t ix = i0 + dx, iy = j0 + dy, iz = k0 + dz;
    const bool skip = ix <= lo || ix >= hi || iy <= lo || iy >= hi || iz <= lo || iz >= hi;
    if (!skip) {
      m_mesh = m_mesh + R3(ix, iy, iz);
      if (!(ii < ntab && nr2 == r2) && ii >= 19 && m_mesh / (float)ii < c.den_odc) break;
    }
    dx = nx; dy = ny; dz = nz; r2 = nr2;
  }
  if (!(m_mesh > c.mass_p * (float)c.min_np / 2.0f)) return;
  const uint32_t v = atomicAdd(nvalid, 1u);
  const float cc = R3(i0, j0, k0);
  const float xm = R3(i0 - 1, j0, k0), xp = R3(i0 + 1, j0, k0);
  const float ym = R3(i0, j0 - 1, k0), yp = R3(i0, j0 + 1, k0);
  const float zm = R3(i0, j0, k0 - 1), zp = R3(i0, j0, k0 + 1);
  vidx[v] = (int32_t)s;
  cden[v] = cc;
  cmesh[v] = m_mesh;
  chpos[3 * v] = ((float)i0 - 0.5f) + 0.5f * (xp - xm) / (-xm + 2 * cc - xp);
  chpos[3 * v + 1] = ((float)j0 - 0.5f) + 0.5f * (yp - ym) / (-ym + 2 * cc - yp);
  chpos[3 * v + 2] = ((float)k0 - 0.5f) + 0.5f * (zp - zm) / (-zm + 2 * cc - zp);
  const int32_t E = c.nft + 2 * CUBE_NCELL, b = CUBE_NCELL - 1;
  cloop[v] = (uint32_t)(((k0 + b) * E + (j0 + b)) * E + (i0 + b));
}

__global__ void k_halo_key_lo(int64_t nv, const uint32_t* __restrict__ cloop, uint64_t* key, int32_t* perm) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nv) return;
  key[i] = cloop[i];
  perm[i] = (int32_t)i;
}
__global__ void k_halo_key_hi(int64_t nv, const int32_t* __restrict__ perm, const int32_t* __restrict__ vidx,
                              const int4* __restrict__ cand, const float* __restrict__ cden, int32_t ntiles,
                              uint64_t* key) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nv) return;
  const int32_t v = perm[i];
  const uint32_t db = __float_as_uint(cden[v]);  // den > cutoff >= 0
  key[i] = ((uint64_t)(ntiles - 1 - cand[vidx[v]].x) << 31) | (uint64_t)(0x7FFFFFFFu - db);
}

// per-peak boxes, in processing order
__global__ void k_halo_prep(int64_t npk, const int32_t* __restrict__ order, const int32_t* __restrict__ vidx,
                            const int4* __restrict__ cand, const float* __restrict__ cmesh,
                            const float* __restrict__ chpos, HalopCfg c, float gs, int32_t nbin, HPk* pk,
                            uint32_t* bincnt, uint32_t* errf, unsigned long long* gmax) {
  const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (s >= npk) return;
  const int32_t v = order[s];
  HPk p;
  for (int d = 0; d < 3; d++) p.h[d] = chpos[3 * v + d];
  p.t = cand[vidx[v]].x;
  const float mp = cmesh[v];
  p.rp = cbrt_ref(0.75f / PI_F / c.den_odc * mp);
  p.rs = fminf((float)c.search_ratio * p.rp, (float)(c.ncb - 1) * (float)CUBE_NCELL);
  uint64_t gc = 1, ncell = 1;
  int32_t b[3];
  for (int d = 0; d < 3; d++) {
    p.cs0[d] = (int32_t)floorf((p.h[d] - p.rs) / (float)CUBE_NCELL) + 1;
    p.cs1[d] = (int32_t)floorf((p.h[d] + p.rs) / (float)CUBE_NCELL) + 1;
    const int32_t cr0 = (int32_t)floorf((p.h[d] - p.rp) / (float)CUBE_NCELL) + 1;
    const int32_t cr1 = (int32_t)floorf((p.h[d] + p.rp) / (float)CUBE_NCELL) + 1;
    p.fr0[d] = CUBE_NCELL * (cr0 - 1);
    p.nb[d] = c.refine_ratio * (CUBE_NCELL * cr1 - p.fr0[d]);
    if (p.cs0[d] < 1 - c.ncb || p.cs1[d] > c.nt + c.ncb) atomicOr(errf, 1u);
    if (p.nb[d] > c.newbox_max) atomicOr(errf, 2u);
    gc *= (uint64_t)(p.nb[d] + 2);
    ncell *= (uint64_t)(p.cs1[d] - p.cs0[d] + 1);
    const float q = floorf((p.h[d] + 2.0f * CUBE_NCELL) / gs);
    b[d] = q < 0 ? 0 : (q >= nbin ? nbin - 1 : (int32_t)q);
  }
  if (ncell > HMAXCELLS) atomicOr(errf, 4u);
  atomicMax(gmax, (unsigned long long)gc);
  p.bin = ((p.t * nbin + b[2]) * nbin + b[1]) * nbin + b[0];
  pk[s] = p;
  atomicAdd(bincnt + p.bin, 1u);
}
__global__ void k_halo_bin_fill(int64_t npk, const HPk* __restrict__ pk, const int64_t* __restrict__ binoff,
                                uint32_t* binfill, int32_t* binlist) {
  const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (s >= npk) return;
  const int32_t b = pk[s].bin;
  binlist[binoff[b] + atomicAdd(binfill + b, 1u)] = (int32_t)s;
}

// ------------------------------------------------------- block utilities
__device__ __forceinline__ uint64_t block_scan_excl(uint64_t v, uint64_t* wsum, uint64_t& total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint64_t incl = v;
  for (int o = 1; o < 64; o <<= 1) {
    const uint64_t t = __shfl_up(incl, o);
    if (lane >= o) incl += t;
  }
  if (lane == 63) wsum[w] = incl;
  __syncthreads();
  uint64_t run = 0, tot = 0;
  for (int i = 0; i < HNT / 64; i++) {
    if (i < w) run += wsum[i];
    tot += wsum[i];
  }
  __syncthreads();
  total = tot;
  return run + incl - v;
}
template <typename V, typename Op>
__device__ __forceinline__ V block_reduce(V v, V* buf, Op op) {
  for (int o = 32; o > 0; o >>= 1) v = op(v, __shfl_xor(v, o));
  if ((threadIdx.x & 63) == 0) buf[threadIdx.x >> 6] = v;
  __syncthreads();
  V r = buf[0];
  for (int i = 1; i < HNT / 64; i++) r = op(r, buf[i]);
  __syncthreads();
  return r;
}

// ------------------------------------------------------- peak processing
struct HWs {  // one workgroup's scratch in global memory
  int32_t* q;       // gathered particle index
  int32_t* rk;      // its cell rank
  uint32_t* bit;    // its copy bit
  float* x;         // [3] tile-frame position
  uint32_t* key[2]; // radius bits, ping-pong
  uint32_t* idx[2]; // gather index, ping-pong
  unsigned long long* grid;
};

__global__ __launch_bounds__(HNT) void k_halo_process(Geo g, HalopCfg c, const void* __restrict__ xp,
                                                      const int64_t* __restrict__ cstart, int64_t npk,
                                                      const HPk* __restrict__ pk, int32_t nbin,
                                                      const int64_t* __restrict__ binoff, const int32_t* __restrict__ binlist,
                                                      uint32_t* tag, int32_t* done, uint32_t* next, HRes* res,
                                                      int32_t* mq, int32_t* mrk, uint32_t* mbit, unsigned long long* mcount,
                                                      uint32_t* errf, char* wsbase, int64_t wsbytes) {
  __shared__ int32_t s_p;
  __shared__ int64_t beg[HMAXCELLS];
  __shared__ int32_t off[HMAXCELLS + 1];
  __shared__ int32_t crk[HMAXCELLS];
  __shared__ uint64_t wsum[HNT / 64];
  __shared__ long long rbuf[HNT / 64];
  __shared__ int64_t s_moff;
  const int tid = threadIdx.x;
  HWs ws;
  {
    char* b = wsbase + (int64_t)blockIdx.x * wsbytes;
    const int64_t n = c.capn;
    ws.q = (int32_t*)b; b += 4 * n;
    ws.rk = (int32_t*)b; b += 4 * n;
    ws.bit = (uint32_t*)b; b += 4 * n;
    ws.x = (float*)b; b += 12 * n;
    ws.key[0] = (uint32_t*)b; b += 4 * n;
    ws.key[1] = (uint32_t*)b; b += 4 * n;
    ws.idx[0] = (uint32_t*)b; b += 4 * n;
    ws.idx[1] = (uint32_t*)b; b += 4 * n;
    b = (char*)(((uintptr_t)b + 255) & ~(uintptr_t)255);
    ws.grid = (unsigned long long*)b;
  }
  for (;;) {
    if (tid == 0) s_p = (int32_t)atomicAdd(next, 1u);
    __syncthreads();
    const int32_t P = s_p;
    __syncthreads();
    if ((int64_t)P >= npk) break;
    const HPk p = pk[P];
    // wait for the earlier peaks of this tile whose spheres overlap
    {
      const int32_t bx = p.bin % nbin, by = (p.bin / nbin) % nbin, bz = (p.bin / (nbin * nbin)) % nbin;
      bool ok = true;
      for (int nb = 0; nb < 27; nb++) {
        const int32_t x = bx + nb % 3 - 1, y = by + (nb / 3) % 3 - 1, z = bz + nb / 9 - 1;
        if (x < 0 || y < 0 || z < 0 || x >= nbin || y >= nbin || z >= nbin) continue;
        const int32_t b = ((p.t * nbin + z) * nbin + y) * nbin + x;
        for (int64_t e = binoff[b] + tid; e < binoff[b + 1]; e += HNT) {
          const int32_t Q = binlist[e];
          if (Q >= P) continue;
          const HPk q = pk[Q];
          const float d0 = q.h[0] - p.h[0], d1 = q.h[1] - p.h[1], d2 = q.h[2] - p.h[2];
          const float rr = p.rs + q.rs + 1e-3f;
          if (d0 * d0 + d1 * d1 + d2 * d2 >= rr * rr) continue;
          uint32_t spins = 0;
          while (__hip_atomic_load(done + Q, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == 0) {
            __builtin_amdgcn_s_sleep(2);
            if (++spins > HSPIN_LIMIT) { ok = false; break; }
          }
          if (!ok) break;
        }
        if (!ok) break;
      }
      if (!ok) atomicOr(errf, 8u);
      __threadfence();
    }
    __syncthreads();
    const int32_t T[3] = {p.t % c.nnt, (p.t / c.nnt) % c.nnt, p.t / (c.nnt * c.nnt)};
    // stage the search box: cells, their ranks and particle offsets
    const int32_t ncx = p.cs1[0] - p.cs0[0] + 1, ncy = p.cs1[1] - p.cs0[1] + 1, ncz = p.cs1[2] - p.cs0[2] + 1;
    const int32_t ncell = ncx * ncy * ncz;
    {
      constexpr int PER = (HMAXCELLS + HNT - 1) / HNT;
      int32_t cnt[PER];
      uint64_t tsum = 0;
#pragma unroll
      for (int k = 0; k < PER; k++) {
        const int i = tid * PER + k;
        cnt[k] = 0;
        if (i < ncell) {
          const int32_t ci = p.cs0[0] + i % ncx, cj = p.cs0[1] + (i / ncx) % ncy, ck = p.cs0[2] + i / (ncx * ncy);
          const int64_t r = rank_of(g, wrapn(T[0] * c.nt + ci - 1, c.nc), wrapn(T[1] * c.nt + cj - 1, c.nc),
                                    wrapn(T[2] * c.nt + ck - 1, c.nc));
          const int64_t b0 = cstart[r];
          beg[i] = b0;
          crk[i] = (int32_t)r;
          cnt[k] = (int32_t)(cstart[r + 1] - b0);
        }
        tsum += (uint64_t)cnt[k];
      }
      uint64_t tot;
      uint64_t run = block_scan_excl(tsum, wsum, tot);
#pragma unroll
      for (int k = 0; k < PER; k++) {
        const int i = tid * PER + k;
        if (i < ncell) off[i] = (int32_t)run;
        run += (uint64_t)cnt[k];
      }
      if (tid == 0) off[ncell] = (int32_t)tot;
      __syncthreads();
    }
    const int32_t total = off[ncell];
    // gather untagged particles within rsearch, in the reference's order
    // (cells k, j, i; particles in cell order)
    uint64_t n_odc = 0;
    bool over = false;
    for (int32_t t0 = 0; t0 < total; t0 += HNT) {
      const int32_t t = t0 + tid;
      uint32_t flag = 0, bt = 0;
      int64_t q = 0;
      int32_t ci = 0;
      float pos[3] = {0.f, 0.f, 0.f};
      if (t < total) {
        int lo = 0, hi = ncell;
        while (hi - lo > 1) {
          const int mid = (lo + hi) >> 1;
          if (off[mid] <= t) lo = mid; else hi = mid;
        }
        ci = lo;
        q = beg[ci] + (t - off[ci]);
        const int32_t cc[3] = {p.cs0[0] + ci % ncx, p.cs0[1] + (ci / ncx) % ncy, p.cs0[2] + ci / (ncx * ncy)};
        int32_t td[3];
        for (int d = 0; d < 3; d++) {
          pos[d] = (float)((double)(CUBE_NCELL * (cc[d] - 1)) + (double)CUBE_NCELL * xoff(load_xu(xp, 3 * q + d, g.izipx), g.izipx));
          td[d] = cc[d] >= 1 ? (cc[d] - 1) / c.nt : -1;
        }
        bt = 1u << (uint32_t)(((td[2] + 1) * 3 + (td[1] + 1)) * 3 + (td[0] + 1));
        const float rr = norm3(pos[0] - p.h[0], pos[1] - p.h[1], pos[2] - p.h[2]);
        if (rr < p.rs) {
          const uint32_t tg = __hip_atomic_load(tag + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          flag = (tg & bt) ? 0u : 1u;
        }
      }
      uint64_t cnt;
      const uint64_t pre = block_scan_excl(flag, wsum, cnt);
      if (flag) {
        const uint64_t slot = n_odc + pre;
        if ((int64_t)slot < c.capn) {
          ws.q[slot] = (int32_t)q;
          ws.rk[slot] = crk[ci];
          ws.bit[slot] = bt;
          ws.x[3 * slot] = pos[0];
          ws.x[3 * slot + 1] = pos[1];
          ws.x[3 * slot + 2] = pos[2];
        }
      }
      n_odc += cnt;
    }
    if ((int64_t)n_odc > c.capn) over = true;
    HRes out;
    out.i_odc = 0;
    out.phys = 0;
    out.moff = -1;
    out.radius = 0.f;
    for (int d = 0; d < 3; d++) out.hfin[d] = p.h[d];
    if (over) {
      if (tid == 0) atomicOr(errf, 16u);
    } else {
      // refined grid (halofind.f90:293-314): clear, then CIC of the gathered
      // particles within r_proxy in 2^-40 fixed point
      const int32_t D0 = p.nb[0] + 2, D1 = p.nb[1] + 2, D2 = p.nb[2] + 2;
      const int64_t gcn = (int64_t)D0 * D1 * D2;
      for (int64_t i = tid; i < gcn; i += HNT) ws.grid[i] = 0ull;
      __syncthreads();
      const float ng = c.newgrid;
      for (int64_t m = tid; m < (int64_t)n_odc; m += HNT) {
        const float x0 = ws.x[3 * m], x1 = ws.x[3 * m + 1], x2 = ws.x[3 * m + 2];
        const float rr = norm3(x0 - p.h[0], x1 - p.h[1], x2 - p.h[2]);
        if (!(rr < p.rp)) continue;
        const float xx[3] = {x0, x1, x2};
        int32_t i1[3];
        float dx1[3], dx2[3];
        for (int d = 0; d < 3; d++) {
          const float pb = xx[d] - 0.5f * ng;
          i1[d] = (int32_t)floorf((pb - (float)p.fr0[d]) / ng) + 1;
          dx1[d] = (float)i1[d] - (pb - (float)p.fr0[d]) / ng;
          dx2[d] = 1 - dx1[d];
        }
        const float w[8] = {dx1[0] * dx1[1] * dx1[2], dx2[0] * dx1[1] * dx1[2], dx1[0] * dx2[1] * dx1[2],
                            dx1[0] * dx1[1] * dx2[2], dx1[0] * dx2[1] * dx2[2], dx2[0] * dx1[1] * dx2[2],
                            dx2[0] * dx2[1] * dx1[2], dx2[0] * dx2[1] * dx2[2]};
        const int o[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 1, 1}, {1, 0, 1}, {1, 1, 0}, {1, 1, 1}};
#pragma unroll
        for (int k = 0; k < 8; k++) {
          const int32_t a0 = i1[0] + o[k][0], a1 = i1[1] + o[k][1], a2 = i1[2] + o[k][2];
          if (a0 < 0 || a0 >= D0 || a1 < 0 || a1 >= D1 || a2 < 0 || a2 >= D2) { atomicOr(errf, 32u); continue; }
          atomicAdd(ws.grid + ((int64_t)a2 * D1 + a1) * D0 + a0, (unsigned long long)llrint((double)w[k] * HFIX));
        }
      }
      __threadfence();
      __syncthreads();
      // maxloc over indices 1.. (column-major first maximum; all zero -> 1,1,1)
      long long best = -1;
      int64_t bl = INT64_MAX;
      for (int64_t i = tid; i < gcn; i += HNT) {
        const int32_t a0 = (int32_t)(i % D0), a1 = (int32_t)((i / D0) % D1), a2 = (int32_t)(i / ((int64_t)D0 * D1));
        if (a0 < 1 || a1 < 1 || a2 < 1) continue;
        const long long v = (long long)__hip_atomic_load(ws.grid + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (v > best || (v == best && i < bl)) { best = v; bl = i; }
      }
      const long long bmax = block_reduce<long long>(best, rbuf, [](long long a, long long b) { return a > b ? a : b; });
      const long long lmin = block_reduce<long long>(best == bmax ? (long long)bl : (long long)INT64_MAX, rbuf,
                                                     [](long long a, long long b) { return a < b ? a : b; });
      int32_t bi = 1, bj = 1, bk = 1;
      if (bmax > 0) {
        bi = (int32_t)(lmin % D0);
        bj = (int32_t)((lmin / D0) % D1);
        bk = (int32_t)(lmin / ((int64_t)D0 * D1));
      }
      auto FG = [&](int32_t i, int32_t j, int32_t k) -> float {
        if (i < 0 || j < 0 || k < 0 || i >= D0 || j >= D1 || k >= D2) return 0.f;
        return (float)((double)(long long)__hip_atomic_load(ws.grid + ((int64_t)k * D1 + j) * D0 + i, __ATOMIC_RELAXED,
                                                            __HIP_MEMORY_SCOPE_AGENT) * (1.0 / HFIX));
      };
      float hf[3];
      {
        const float cc = FG(bi, bj, bk);
        const float xm = FG(bi - 1, bj, bk), xq = FG(bi + 1, bj, bk);
        const float ym = FG(bi, bj - 1, bk), yq = FG(bi, bj + 1, bk);
        const float zm = FG(bi, bj, bk - 1), zq = FG(bi, bj, bk + 1);
        hf[0] = ((float)bi - 0.5f) + 0.5f * (xq - xm) / (-xm + 2 * cc - xq);
        hf[1] = ((float)bj - 0.5f) + 0.5f * (yq - ym) / (-ym + 2 * cc - yq);
        hf[2] = ((float)bk - 0.5f) + 0.5f * (zq - zm) / (-zm + 2 * cc - zm);  // halofind.f90:356 as written
      }
      for (int d = 0; d < 3; d++) hf[d] = (float)p.fr0[d] + ng * hf[d];
      const float nftf = (float)c.nft;
      const int phys = hf[0] > 0 && hf[1] > 0 && hf[2] > 0 && hf[0] < nftf && hf[1] < nftf && hf[2] < nftf;
      // radii from the refined centre, sorted with gathering order kept
      // for equal radii: stable LSD radix sort, 2 bits per pass
      for (int64_t m = tid; m < (int64_t)n_odc; m += HNT) {
        ws.key[0][m] = __float_as_uint(norm3(ws.x[3 * m] - hf[0], ws.x[3 * m + 1] - hf[1], ws.x[3 * m + 2] - hf[2]));
        ws.idx[0][m] = (uint32_t)m;
      }
      __syncthreads();
      int cur = 0;
      for (int bitpos = 0; bitpos < 32; bitpos += 2) {
        // bucket totals; a pass whose digits are all equal keeps the order
        uint32_t bc[4];
        {
          uint64_t c0 = 0, c1 = 0, c2 = 0;
          for (int64_t m = tid; m < (int64_t)n_odc; m += HNT) {
            const uint32_t dgt = (ws.key[cur][m] >> bitpos) & 3u;
            c0 += dgt == 0; c1 += dgt == 1; c2 += dgt == 2;
          }
          uint64_t t0, t1, t2;
          block_scan_excl(c0, wsum, t0);
          block_scan_excl(c1, wsum, t1);
          block_scan_excl(c2, wsum, t2);
          const uint64_t t3 = n_odc - t0 - t1 - t2;
          if (t0 == n_odc || t1 == n_odc || t2 == n_odc || t3 == n_odc) continue;
          bc[0] = 0; bc[1] = (uint32_t)t0; bc[2] = (uint32_t)(t0 + t1); bc[3] = (uint32_t)(t0 + t1 + t2);
        }
        uint32_t run[4] = {bc[0], bc[1], bc[2], bc[3]};
        for (int64_t m0 = 0; m0 < (int64_t)n_odc; m0 += HNT) {
          const int64_t m = m0 + tid;
          uint32_t dgt = 4;
          if (m < (int64_t)n_odc) dgt = (ws.key[cur][m] >> bitpos) & 3u;
          const uint64_t onehot = dgt < 4 ? 1ull << (16 * dgt) : 0ull;
          uint64_t tot;
          const uint64_t pre = block_scan_excl(onehot, wsum, tot);
          if (dgt < 4) {
            const uint32_t dst = run[dgt] + (uint32_t)((pre >> (16 * dgt)) & 0xFFFFu);
            ws.key[cur ^ 1][dst] = ws.key[cur][m];
            ws.idx[cur ^ 1][dst] = ws.idx[cur][m];
          }
          for (int d = 0; d < 4; d++) run[d] += (uint32_t)((tot >> (16 * d)) & 0xFFFFu);
        }
        __syncthreads();
        cur ^= 1;
      }
      // first ii >= 2 with enclosed density below den_odc
      const float kd = 0.75f / PI_F;
      int64_t found = INT64_MAX;
      for (int64_t m0 = 1; m0 < (int64_t)n_odc && found == INT64_MAX; m0 += HNT) {
        const int64_t m = m0 + tid;
        long long f = INT64_MAX;
        if (m < (int64_t)n_odc) {
          const float r = __uint_as_float(ws.key[cur][m]);
          const float den2 = kd * (float)(m + 1) * c.mass_p / (r * r * r);
          if (den2 < c.den_odc) f = m + 1;
        }
        found = block_reduce<long long>(f, rbuf, [](long long a, long long b) { return a < b ? a : b; });
      }
      if (found != INT64_MAX) {
        const int64_t ii = found;
        const float r_i = __uint_as_float(ws.key[cur][ii - 1]), r_m = __uint_as_float(ws.key[cur][ii - 2]);
        const float den2 = kd * (float)ii * c.mass_p / (r_i * r_i * r_i);
        const float den1 = kd * (float)(ii - 1) * c.mass_p / (r_m * r_m * r_m);
        const float r2 = log10_ref(r_i), r1 = log10_ref(r_m), d2 = log10_ref(den2), d1 = log10_ref(den1);
        const float w1 = log10_ref(c.den_odc) - d2, w2 = d1 - log10_ref(c.den_odc);
        out.radius = exp10_ref((w1 * r1 + w2 * r2) / (d1 - d2));
        out.i_odc = (int32_t)(ii - 1);
      }
      for (int d = 0; d < 3; d++) out.hfin[d] = hf[d];
      out.phys = phys;
      if (out.i_odc >= c.min_np) {
        for (int64_t m = tid; m < out.i_odc; m += HNT) atomicOr(tag + ws.q[ws.idx[cur][m]], ws.bit[ws.idx[cur][m]]);
        if (phys) {
          if (tid == 0) {
            const unsigned long long o = atomicAdd(mcount, (unsigned long long)out.i_odc);
            s_moff = (int64_t)o;
            if ((int64_t)(o + out.i_odc) > c.mcap) atomicOr(errf, 64u);
          }
          __syncthreads();
          out.moff = s_moff;
          if (out.moff + out.i_odc <= c.mcap)
            for (int64_t m = tid; m < out.i_odc; m += HNT) {
              const uint32_t e = ws.idx[cur][m];
              mq[out.moff + m] = ws.q[e];
              mrk[out.moff + m] = ws.rk[e];
              mbit[out.moff + m] = ws.bit[e];
            }
        }
      }
    }
    if (tid == 0) res[P] = out;
    __threadfence();
    __syncthreads();
    if (tid == 0) __hip_atomic_store(done + P, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// ---------------------------------------------------------------- catalogue
__global__ void k_halo_flags(int64_t npk, const HRes* __restrict__ res, int32_t min_np, int32_t* is_halo, int32_t* words) {
  const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (s >= npk) return;
  const bool hh = res[s].i_odc >= min_np && res[s].phys;
  is_halo[s] = hh ? 1 : 0;
  words[s] = hh ? res[s].i_odc + 2 : 0;
}

// member position / velocity in the peak's tile frame
__device__ __forceinline__ void member_xv(const Geo& g, const HalopCfg& c, int32_t tile, int32_t q, int32_t rk, uint32_t bit,
                                          const void* xp, const void* vp, const float* vfield, const double* vdt,
                                          float x[3], float v[3]) {
  // cell of rank rk: tile-major, x fastest inside the tile
  const int32_t nt = c.nt, nnt = c.nnt, n3t = nt * nt * nt;
  const int32_t tr = rk / n3t, lr = rk % n3t;
  const int32_t lc[3] = {lr % nt, (lr / nt) % nt, lr / (nt * nt)};
  (void)tr;
  const int32_t b = 31 - __clz(bit);
  const int32_t td[3] = {b % 3 - 1, (b / 3) % 3 - 1, b / 9 - 1};
  (void)tile; (void)nnt;
  for (int d = 0; d < 3; d++) {
    const int32_t cc = td[d] * nt + lc[d] + 1;
    x[d] = (float)((double)(CUBE_NCELL * (cc - 1)) + (double)CUBE_NCELL * xoff(load_xu(xp, 3 * (int64_t)q + d, g.izipx), g.izipx));
    v[d] = (float)(vdecode(vdt, load_v(vp, 3 * (int64_t)q + d, g.izipv), g.izipv) + (double)vfield[3 * (int64_t)rk + d]);
  }
}

struct HSums {
  double in[6];   // inertia terms of this halo
  double sdx[3];  // q-space dx sums
  int64_t nptemp;
  float xw[3];
};

// per halo: everything but the running totals (halofind.f90:426-444, 459-483)
__global__ __launch_bounds__(HNT) void k_halo_cat1(Geo g, HalopCfg c, int64_t npk, const HPk* __restrict__ pk,
                                                   const HRes* __restrict__ res, const int32_t* __restrict__ is_halo,
                                                   const int64_t* __restrict__ hidx, const int64_t* __restrict__ woff,
                                                   const int32_t* __restrict__ mq, const int32_t* __restrict__ mrk,
                                                   const uint32_t* __restrict__ mbit, const void* __restrict__ xp,
                                                   const void* __restrict__ vp, const float* __restrict__ vfield,
                                                   const double* __restrict__ vdt, const int32_t* __restrict__ pid,
                                                   float box, float h0, float a, float vsim2phys, float cen0, float cen1,
                                                   float cen2, float* rec, HSums* sums, int32_t* words, int16_t* hid,
                                                   int64_t nhid) {
  __shared__ double dbuf[HNT / 64];
  __shared__ long long lbuf[HNT / 64];
  for (int64_t s = blockIdx.x; s < npk; s += gridDim.x) {
    if (!is_halo[s]) continue;
    const HRes r = res[s];
    const HPk p = pk[s];
    const int64_t hh = hidx[s];
    const int32_t n = r.i_odc;
    const int32_t T[3] = {p.t % c.nnt, (p.t / c.nnt) % c.nnt, p.t / (c.nnt * c.nnt)};
    const float off[3] = {(float)(T[0] * c.nft), (float)(T[1] * c.nft), (float)(T[2] * c.nft)};
    const float nfr = (float)c.N;
    auto sum = [&](double v) { return block_reduce<double>(v, dbuf, [](double x, double y) { return x + y; }); };
    double sx[3] = {0, 0, 0}, sv[3] = {0, 0, 0};
    for (int32_t m = threadIdx.x; m < n; m += HNT) {
      float x[3], v[3];
      member_xv(g, c, p.t, mq[r.moff + m], mrk[r.moff + m], mbit[r.moff + m], xp, vp, vfield, vdt, x, v);
      for (int d = 0; d < 3; d++) { sx[d] += x[d]; sv[d] += v[d]; }
    }
    float xm[3], vm[3];
    for (int d = 0; d < 3; d++) {
      xm[d] = (float)(sum(sx[d]) / (double)n);
      vm[d] = (float)(sum(sv[d]) / (double)n);
    }
    float xw[3];
    for (int d = 0; d < 3; d++) xw[d] = fmodulo(xm[d] + off[d], nfr);
    double vx[3] = {0, 0, 0}, vd = 0, am[3] = {0, 0, 0}, in[6] = {0, 0, 0, 0, 0, 0}, sdx[3] = {0, 0, 0};
    long long npt = 0;
    const int64_t nq = c.nq;
    const float sp = (float)c.N / (float)nq, half = (float)(c.N / 2);
    for (int32_t m = threadIdx.x; m < n; m += HNT) {
      float x[3], v[3];
      member_xv(g, c, p.t, mq[r.moff + m], mrk[r.moff + m], mbit[r.moff + m], xp, vp, vfield, vdt, x, v);
      for (int d = 0; d < 3; d++) {
        const double e = (double)(x[d] - xm[d]), f = (double)(v[d] - vm[d]);
        vx[d] += e * e;
        vd += f * f;
      }
      const double e0 = x[0] - (xw[0] - off[0]), e1 = x[1] - (xw[1] - off[1]), e2 = x[2] - (xw[2] - off[2]);
      am[0] += e1 * (double)v[2] - e2 * (double)v[1];
      am[1] += e2 * (double)v[0] - e0 * (double)v[2];
      am[2] += e0 * (double)v[1] - e1 * (double)v[0];
      in[0] += e0 * e0; in[1] += e1 * e1; in[2] += e2 * e2;
      in[3] += e0 * e1; in[4] += e1 * e2; in[5] += e2 * e0;
      const int64_t qt = (int64_t)pid[mq[r.moff + m]] - 1;
      if (qt >= 0) {
        npt++;
        const int64_t qi[3] = {qt % nq, (qt / nq) % nq, qt / (nq * nq)};
        for (int d = 0; d < 3; d++) {
          const float qp = (float)qi[d] * sp + 0.5f;
          sdx[d] += (double)(fmodulo(qp - xw[d] + half, nfr) - half);
        }
      }
    }
    float* H = rec + 42 * hh;
    for (int d = 0; d < 3; d++) vx[d] = sum(vx[d]);
    vd = sum(vd);
    for (int d = 0; d < 3; d++) am[d] = sum(am[d]);
    for (int k = 0; k < 6; k++) in[k] = sum(in[k]);
    for (int d = 0; d < 3; d++) sdx[d] = sum(sdx[d]);
    npt = block_reduce<long long>(npt, lbuf, [](long long x, long long y) { return x + y; });
    if (threadIdx.x == 0) {
      for (int d = 0; d < 3; d++) H[d] = fmodulo(r.hfin[d] + off[d], nfr);
      H[3] = (float)n;
      H[4] = r.radius;
      H[5] = sqrtf((float)(vd / (double)(n - 1)));
      for (int d = 0; d < 3; d++) {
        H[6 + d] = xw[d];
        H[9 + d] = vm[d];
        H[12 + d] = (float)am[d];
        H[15 + d] = (float)(vx[d] / (double)(n - 1));
      }
      float los[3];
      const float cen[3] = {cen0, cen1, cen2};
      for (int d = 0; d < 3; d++) los[d] = xw[d] - (cen[d] / 500.0f) * (float)c.nc;
      const float ln = norm3(los[0], los[1], los[2]);
      for (int d = 0; d < 3; d++) los[d] = los[d] / ln;
      const float dot = (vm[0] * los[0] + vm[1] * los[1]) + vm[2] * los[2];
      for (int d = 0; d < 3; d++) {
        float sxd = dot * los[d];
        sxd = sxd * vsim2phys / a / (100 * h0);
        sxd = sxd / (h0 * box / nfr);
        H[39 + d] = fmodulo(xw[d] + sxd, nfr);
      }
      HSums S;
      for (int k = 0; k < 6; k++) S.in[k] = in[k];
      for (int d = 0; d < 3; d++) { S.sdx[d] = sdx[d]; S.xw[d] = xw[d]; }
      S.nptemp = npt;
      sums[hh] = S;
    }
    // halo_pid words and hid
    const int64_t w0 = woff[s];
    if (threadIdx.x == 0) {
      words[w0] = n;
      words[w0 + n + 1] = 0;
    }
    for (int32_t m = threadIdx.x; m < n; m += HNT) {
      const int32_t pv = pid[mq[r.moff + m]];
      words[w0 + 1 + m] = pv;
      if (hid && pv >= 1 && pv <= nhid) hid[pv - 1] = (int16_t)(hh + 1);
    }
    __syncthreads();
  }
}

// running totals in catalogue order: inertia (x), dx_mean -> q_mean
__global__ void k_halo_seq1(int64_t nhalo, const HSums* __restrict__ sums, int32_t N, float* rec, float* qmean) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double run[6] = {0, 0, 0, 0, 0, 0};
  float dxm[3] = {0.f, 0.f, 0.f};
  const float nfr = (float)N;
  for (int64_t hh = 0; hh < nhalo; hh++) {
    const HSums S = sums[hh];
    float* H = rec + 42 * hh;
    for (int k = 0; k < 6; k++) run[k] += S.in[k];
    const float sym[9] = {(float)run[0], (float)run[3], (float)run[5], (float)run[3], (float)run[1],
                          (float)run[4], (float)run[5], (float)run[4], (float)run[2]};
    for (int k = 0; k < 9; k++) H[18 + k] = sym[k];
    for (int d = 0; d < 3; d++) {
      dxm[d] = (float)(((double)dxm[d] + S.sdx[d]) / (double)S.nptemp);
      const float qm = fmodulo(S.xw[d] + dxm[d], nfr);
      H[27 + d] = qm;
      qmean[3 * hh + d] = qm;
    }
  }
}

__global__ __launch_bounds__(HNT) void k_halo_cat2(HalopCfg c, int64_t npk, const HRes* __restrict__ res,
                                                   const int32_t* __restrict__ is_halo, const int64_t* __restrict__ hidx,
                                                   const int32_t* __restrict__ mq, const int32_t* __restrict__ pid,
                                                   const float* __restrict__ qmean, double* iq_out) {
  __shared__ double dbuf[HNT / 64];
  for (int64_t s = blockIdx.x; s < npk; s += gridDim.x) {
    if (!is_halo[s]) continue;
    const HRes r = res[s];
    const int64_t hh = hidx[s];
    const int64_t nq = c.nq;
    const float nfr = (float)c.N, sp = (float)c.N / (float)nq, half = (float)(c.N / 2);
    const float qm[3] = {qmean[3 * hh], qmean[3 * hh + 1], qmean[3 * hh + 2]};
    double iq[6] = {0, 0, 0, 0, 0, 0};
    for (int32_t m = threadIdx.x; m < r.i_odc; m += HNT) {
      const int64_t qt = (int64_t)pid[mq[r.moff + m]] - 1;
      if (qt < 0) continue;
      const int64_t qi[3] = {qt % nq, (qt / nq) % nq, qt / (nq * nq)};
      double e[3];
      for (int d = 0; d < 3; d++) {
        const float qp = (float)qi[d] * sp + 0.5f;
        e[d] = (double)(fmodulo(qp + 0.5f - qm[d] + half, nfr) - half);
      }
      iq[0] += e[0] * e[0]; iq[1] += e[1] * e[1]; iq[2] += e[2] * e[2];
      iq[3] += e[0] * e[1]; iq[4] += e[1] * e[2]; iq[5] += e[2] * e[0];
    }
    for (int k = 0; k < 6; k++) {
      const double t = block_reduce<double>(iq[k], dbuf, [](double x, double y) { return x + y; });
      if (threadIdx.x == 0) iq_out[6 * hh + k] = t;
    }
  }
}
__global__ void k_halo_seq2(int64_t nhalo, const double* __restrict__ iq, float* rec) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double run[6] = {0, 0, 0, 0, 0, 0};
  for (int64_t hh = 0; hh < nhalo; hh++) {
    for (int k = 0; k < 6; k++) run[k] += iq[6 * hh + k];
    float* H = rec + 42 * hh;
    const float sym[9] = {(float)run[0], (float)run[3], (float)run[5], (float)run[3], (float)run[1],
                          (float)run[4], (float)run[5], (float)run[4], (float)run[2]};
    for (int k = 0; k < 9; k++) H[30 + k] = sym[k];
  }
}

// one device allocation carved into aligned pieces
struct Arena {
  char* base = nullptr;
  size_t used = 0, cap = 0;
  template <typename T>
  T* take(size_t n) {
    used = (used + 255) & ~(size_t)255;
    T* p = (T*)(base + used);
    used += n * sizeof(T);
    return p;
  }
};

}  // namespace

extern "C" int cube_halofind(cube_handle* h, const cube_halo_params* hp, cube_halo_catalog* out) {
  CUBE_EXCLUSIVE;
  if (!h || !hp || !out) return CUBE_ERR_ARG;
  if (hipSetDevice(h->p.device) != hipSuccess) return cube_fail(h, CUBE_ERR_HIP, "hipSetDevice failed");
  if (!out->halos || !out->pid_words) return cube_fail(h, CUBE_ERR_ARG, "halo catalogue buffers missing");
  if (h->ncb > h->nt) return cube_fail(h, CUBE_ERR_ARG, "halo finder needs ncb <= nt");
  if (hp->nc_halo_max < 1 || hp->nc_halo_max > 160) return cube_fail(h, CUBE_ERR_ARG, "nc_halo_max must lie in 1..160");
  if (h->nt < 4) return cube_fail(h, CUBE_ERR_ARG, "halo finder needs nt >= 4");
  if (!(hp->den_peak_cutoff >= 0)) return cube_fail(h, CUBE_ERR_ARG, "den_peak_cutoff must be >= 0");
  if (!(hp->a > 0)) return cube_fail(h, CUBE_ERR_ARG, "halo finder needs the scale factor a > 0");
  if (hp->refine_ratio < 1 || hp->search_ratio < 1 || hp->nq < 1 || hp->max_halo_np < 2)
    return cube_fail(h, CUBE_ERR_ARG, "invalid halo finder parameters");
  if (!h->fine.S) return cube_fail(h, CUBE_ERR_ARG, "halo finder needs the fine mesh");
  hipStream_t st = h->stream;
  HalopCfg c{};
  c.nc = (int32_t)h->nc; c.nt = (int32_t)h->nt; c.nnt = (int32_t)h->nnt; c.ncb = (int32_t)h->ncb;
  c.N = (int32_t)h->fine.N; c.nft = CUBE_NCELL * c.nt; c.nfb = CUBE_NCELL * c.ncb;
  c.den_odc = hp->den_odc; c.cutoff = hp->den_peak_cutoff; c.mass_p = h->mass_p;
  c.min_np = hp->min_halo_particles; c.search_ratio = hp->search_ratio; c.refine_ratio = hp->refine_ratio;
  c.newbox_max = hp->newbox_max; c.nq = hp->nq; c.newgrid = 1.0f / (float)hp->refine_ratio;
  const Geo g = geo_of(h);
  const int64_t np = h->np, N = c.N, n3 = N * N * N;

  // 1. density into fine.S
  float* rho = (float*)h->fine.S;
  int rc = cic_halo_density(h, rho);
  if (rc) return rc;

  // offset table
  const int32_t R = hp->nc_halo_max, Dt = 2 * R + 1;
  const int64_t ngen = (int64_t)Dt * Dt * Dt;
  int64_t ntab = 0;
  for (int64_t i = -R; i <= R; i++)
    for (int64_t j = -R; j <= R; j++) {
      const int64_t rem = (int64_t)R * R - i * i - j * j;
      if (rem < 0) continue;
      int64_t k = 0;
      while ((k + 1) * (k + 1) <= rem) k++;
      ntab += 2 * k + 1;
    }
  const int64_t cand_cap = std::max<int64_t>(1 << 16, std::min<int64_t>(n3 / 8, 1 << 23));
  const int64_t mcap = std::max<int64_t>(1 << 16, np + (np >> 2));
  // first arena: everything whose size does not depend on the peaks
  size_t sort_tmp = 0;
  {
    size_t a = 0, b = 0;
    HIP_TRY(h, hipcub::DeviceRadixSort::SortKeys(nullptr, a, (uint64_t*)nullptr, (uint64_t*)nullptr, (int)ngen, 0, 41, st));
    HIP_TRY(h, hipcub::DeviceRadixSort::SortPairs(nullptr, b, (uint64_t*)nullptr, (uint64_t*)nullptr, (int32_t*)nullptr,
                                                  (int32_t*)nullptr, (int)cand_cap, 0, 43, st));
    sort_tmp = std::max(a, b);
  }
  Arena A;
  A.cap = 2 * sizeof(uint64_t) * ngen + sort_tmp + cand_cap * (sizeof(int4) + 4 * sizeof(int32_t) + 7 * sizeof(float) +
                                                               2 * sizeof(uint64_t) + sizeof(HPk) + sizeof(HRes) + 32) +
          (size_t)np * 4 + (size_t)mcap * 12 + (1 << 20);
  Arena B;
  char* wsbase = nullptr;
  auto cleanup = [&]() {
    (void)hipFree(A.base);
    (void)hipFree(B.base);
    (void)hipFree(wsbase);
  };
  HIP_TRY(h, hipMalloc(&A.base, A.cap));
  uint64_t* tabk = A.take<uint64_t>(ngen);
  uint64_t* tabs = A.take<uint64_t>(ngen);
  void* tmp = A.take<char>(sort_tmp);
  int4* cand = A.take<int4>(cand_cap);
  uint32_t* ctr = A.take<uint32_t>(16);  // 0 ncand, 1 nvalid, 2 next, 3 errf
  unsigned long long* uctr = A.take<unsigned long long>(4);  // 0 members, 1 grid max
  int32_t* vidx = A.take<int32_t>(cand_cap);
  float* cden = A.take<float>(cand_cap);
  float* cmesh = A.take<float>(cand_cap);
  float* chpos = A.take<float>(3 * cand_cap);
  uint32_t* cloop = A.take<uint32_t>(cand_cap);
  int32_t* perm0 = A.take<int32_t>(cand_cap);
  int32_t* perm1 = A.take<int32_t>(cand_cap);
  uint64_t* k0 = A.take<uint64_t>(cand_cap);
  uint64_t* k1 = A.take<uint64_t>(cand_cap);
  uint32_t* tag = A.take<uint32_t>(np);
  int32_t* mq = A.take<int32_t>(mcap);
  int32_t* mrk = A.take<int32_t>(mcap);
  uint32_t* mbit = A.take<uint32_t>(mcap);
  if (A.used > A.cap) { cleanup(); return cube_fail(h, CUBE_ERR_ARG, "halo arena sizing"); }
#define HTRY(call)                                                                          \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess) {                                                                 \
      cleanup();                                                                            \
      return cube_fail(h, CUBE_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_));     \
    }                                                                                       \
  } while (0)
  HTRY(hipMemsetAsync(ctr, 0, 16 * sizeof(uint32_t), st));
  HTRY(hipMemsetAsync(uctr, 0, 4 * sizeof(unsigned long long), st));
  HTRY(hipMemsetAsync(tag, 0, (size_t)np * 4, st));
  k_hoff_gen<<<4096, 256, 0, st>>>(R, tabk);
  {
    size_t tb = sort_tmp;
    HTRY(hipcub::DeviceRadixSort::SortKeys(tmp, tb, tabk, tabs, (int)ngen, 0, 41, st));
  }
  // 2. peaks
  k_halo_peaks<<<8192, 256, 0, st>>>(rho, c, ctr + 0, cand_cap, cand);
  HTRY(hipGetLastError());
  uint32_t hc[4];
  HTRY(hipMemcpyAsync(hc, ctr, sizeof(hc), hipMemcpyDeviceToHost, st));
  HTRY(hipStreamSynchronize(st));
  if ((int64_t)hc[0] > cand_cap) {
    cleanup();
    return cube_fail(h, CUBE_ERR_CAPACITY, "halo finder: %u candidate peaks exceed the table (%lld)", hc[0], (long long)cand_cap);
  }
  const int64_t ncand = hc[0];
  // 3. mesh walk
  if (ncand > 0)
    k_halo_walk<<<(unsigned)((ncand + 127) / 128), 128, 0, st>>>(rho, c, cand, ncand, tabs, ntab, R, ctr + 1, vidx, cden,
                                                                  cmesh, chpos, cloop);
  HTRY(hipGetLastError());
  HTRY(hipMemcpyAsync(hc, ctr, sizeof(hc), hipMemcpyDeviceToHost, st));
  HTRY(hipStreamSynchronize(st));
  const int64_t npk = hc[1];
  out->npeaks = ncand;
  out->nprocessed = npk;
  out->nhalo = 0;
  out->npid_words = 0;
  if (npk == 0) { cleanup(); return 0; }
  // 4. order: stable by loop index, then by (tile order, density descending)
  {
    size_t tb = sort_tmp;
    k_halo_key_lo<<<(unsigned)((npk + 255) / 256), 256, 0, st>>>(npk, cloop, k0, perm0);
    HTRY(hipcub::DeviceRadixSort::SortPairs(tmp, tb, k0, k1, perm0, perm1, (int)npk, 0, 32, st));
    const int32_t ntiles = c.nnt * c.nnt * c.nnt;
    k_halo_key_hi<<<(unsigned)((npk + 255) / 256), 256, 0, st>>>(npk, perm1, vidx, cand, cden, ntiles, k0);
    tb = sort_tmp;
    HTRY(hipcub::DeviceRadixSort::SortPairs(tmp, tb, k0, k1, perm1, perm0, (int)npk, 0, 43, st));
  }
  // perm0[s] = valid slot of the s-th peak in processing order
  const float gs = std::max(2.0f * (float)(c.ncb - 1) * (float)CUBE_NCELL, 1.0f);
  const int32_t nbin = (int32_t)std::ceil((c.nft + 4.0f * CUBE_NCELL) / gs) + 1;
  const int64_t nbins = (int64_t)c.nnt * c.nnt * c.nnt * nbin * nbin * nbin;
  B.cap = (size_t)npk * (sizeof(HPk) + sizeof(HRes) + 4 * sizeof(int32_t) + 2 * sizeof(int64_t) + 64) +
          (size_t)(nbins + 1) * (sizeof(int64_t) + 2 * sizeof(uint32_t)) + (1 << 20);
  HTRY(hipMalloc(&B.base, B.cap));
  HPk* pk = B.take<HPk>(npk);
  HRes* res = B.take<HRes>(npk);
  int32_t* done = B.take<int32_t>(npk);
  int32_t* binlist = B.take<int32_t>(npk);
  uint32_t* bincnt = B.take<uint32_t>(nbins + 1);
  uint32_t* binfill = B.take<uint32_t>(nbins + 1);
  int64_t* binoff = B.take<int64_t>(nbins + 1);
  int32_t* is_halo = B.take<int32_t>(npk);
  int32_t* wcount = B.take<int32_t>(npk);
  int64_t* hidx = B.take<int64_t>(npk + 1);
  int64_t* woff = B.take<int64_t>(npk + 1);
  HTRY(hipMemsetAsync(done, 0, sizeof(int32_t) * npk, st));
  HTRY(hipMemsetAsync(bincnt, 0, sizeof(uint32_t) * (nbins + 1), st));
  HTRY(hipMemsetAsync(binfill, 0, sizeof(uint32_t) * (nbins + 1), st));
  k_halo_prep<<<(unsigned)((npk + 255) / 256), 256, 0, st>>>(npk, perm0, vidx, cand, cmesh, chpos, c, gs, nbin, pk, bincnt,
                                                             ctr + 3, uctr + 1);
  HTRY(hipGetLastError());
  {
    const int r2 = scan_counts(h, (const int32_t*)bincnt, nbins, binoff);
    if (r2) { cleanup(); return r2; }
  }
  k_halo_bin_fill<<<(unsigned)((npk + 255) / 256), 256, 0, st>>>(npk, pk, binoff, binfill, binlist);
  HTRY(hipGetLastError());
  unsigned long long gmax = 0;
  HTRY(hipMemcpyAsync(hc, ctr, sizeof(hc), hipMemcpyDeviceToHost, st));
  HTRY(hipMemcpyAsync(&gmax, uctr + 1, sizeof(gmax), hipMemcpyDeviceToHost, st));
  HTRY(hipStreamSynchronize(st));
  if (hc[3] & 1u) { cleanup(); return cube_fail(h, CUBE_ERR_CHECK, "halo search box beyond the tile buffer (halofind.f90:245-252)"); }
  if (hc[3] & 2u) { cleanup(); return cube_fail(h, CUBE_ERR_CHECK, "newbox>newbox_max (halofind.f90:259)"); }
  if (hc[3] & 4u) { cleanup(); return cube_fail(h, CUBE_ERR_CAPACITY, "halo search box holds more than %d cells", HMAXCELLS); }
  // 5. persistent processing: workgroup scratch sized for the largest
  // refined grid and the search-list capacity
  int dev = 0, ncu = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  c.capn = std::min<int64_t>(hp->max_halo_np, std::max<int64_t>(4096, std::min<int64_t>(np, 1 << 20)));
  c.gcap = (int64_t)gmax;
  c.mcap = mcap;
  const int64_t wsbytes = ((int64_t)c.capn * 40 + 512 + (int64_t)gmax * 8 + 255) & ~(int64_t)255;
  size_t freeb = 0, totb = 0;
  (void)hipMemGetInfo(&freeb, &totb);
  int64_t nwg = std::min<int64_t>(std::max(ncu, 1), npk);
  const int64_t budget = (int64_t)(freeb / 2);
  if (nwg * wsbytes > budget) nwg = std::max<int64_t>(1, budget / wsbytes);
  HTRY(hipMalloc(&wsbase, (size_t)(nwg * wsbytes)));
  const double* vdt = vtab_for(h, vcoef(h->sigma_vi));
  k_halo_process<<<(unsigned)nwg, HNT, 0, st>>>(g, c, h->xp, h->cstart, npk, pk, nbin, binoff, binlist, tag, done, ctr + 2,
                                                 res, mq, mrk, mbit, uctr + 0, ctr + 3, wsbase, wsbytes);
  hipError_t pe = hipGetLastError();
  if (pe == hipSuccess) pe = hipMemcpyAsync(hc, ctr, sizeof(hc), hipMemcpyDeviceToHost, st);
  if (pe == hipSuccess) pe = hipStreamSynchronize(st);
  if (pe != hipSuccess) { cleanup(); return cube_fail(h, CUBE_ERR_HIP, "halo processing failed: %s", hipGetErrorString(pe)); }
  if (hc[3] & 8u) { cleanup(); return cube_fail(h, CUBE_ERR_CHECK, "halo finder: dependency wait exceeded its bound"); }
  if (hc[3] & 16u) { cleanup(); return cube_fail(h, CUBE_ERR_CAPACITY, "halo search list longer than %lld particles", (long long)c.capn); }
  if (hc[3] & 32u) { cleanup(); return cube_fail(h, CUBE_ERR_CHECK, "refined grid index outside 0..newbox+1"); }
  if (hc[3] & 64u) { cleanup(); return cube_fail(h, CUBE_ERR_CAPACITY, "halo member list capacity"); }
  // 6. catalogue
  k_halo_flags<<<(unsigned)((npk + 255) / 256), 256, 0, st>>>(npk, res, c.min_np, is_halo, wcount);
  if ((rc = scan_counts(h, is_halo, npk, hidx)) || (rc = scan_counts(h, wcount, npk, woff))) { cleanup(); return rc; }
  int64_t nh = 0, nw = 0;
  HTRY(hipMemcpyAsync(&nh, hidx + npk, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  HTRY(hipMemcpyAsync(&nw, woff + npk, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  HTRY(hipStreamSynchronize(st));
  if (nh > out->max_halos || nw > out->max_pid_words) {
    cleanup();
    return cube_fail(h, CUBE_ERR_CAPACITY, "halo catalogue: %lld halos / %lld pid words exceed the buffers", (long long)nh,
                     (long long)nw);
  }
  out->nhalo = nh;
  out->npid_words = nw;
  if (nh > 0) {
    float* rec = nullptr;
    HSums* sums = nullptr;
    float* qmean = nullptr;
    double* iq = nullptr;
    int32_t* words = nullptr;
    int16_t* hidd = nullptr;
    hipError_t e = hipMalloc(&rec, sizeof(float) * 42 * nh);
    if (e == hipSuccess) e = hipMalloc(&sums, sizeof(HSums) * nh);
    if (e == hipSuccess) e = hipMalloc(&qmean, sizeof(float) * 3 * nh);
    if (e == hipSuccess) e = hipMalloc(&iq, sizeof(double) * 6 * nh);
    if (e == hipSuccess) e = hipMalloc(&words, sizeof(int32_t) * nw);
    if (e == hipSuccess && out->hid && out->nhid > 0) {
      e = hipMalloc(&hidd, sizeof(int16_t) * out->nhid);
      if (e == hipSuccess) e = hipMemsetAsync(hidd, 0, sizeof(int16_t) * out->nhid, st);
    }
    if (e == hipSuccess) {
      const unsigned nb1 = (unsigned)std::min<int64_t>(npk, 65535);
      k_halo_cat1<<<nb1, HNT, 0, st>>>(g, c, npk, pk, res, is_halo, hidx, woff, mq, mrk, mbit, h->xp, h->vp, h->vfield, vdt,
                                       h->pid, hp->box, hp->h0, hp->a, hp->vsim2phys, hp->los_centre[0], hp->los_centre[1],
                                       hp->los_centre[2], rec, sums, words, hidd, hidd ? out->nhid : 0);
      k_halo_seq1<<<1, 1, 0, st>>>(nh, sums, c.N, rec, qmean);
      k_halo_cat2<<<nb1, HNT, 0, st>>>(c, npk, res, is_halo, hidx, mq, h->pid, qmean, iq);
      k_halo_seq2<<<1, 1, 0, st>>>(nh, iq, rec);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out->halos, rec, sizeof(float) * 42 * nh, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(out->pid_words, words, sizeof(int32_t) * nw, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess && hidd) e = hipMemcpyAsync(out->hid, hidd, sizeof(int16_t) * out->nhid, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(rec); (void)hipFree(sums); (void)hipFree(qmean); (void)hipFree(iq); (void)hipFree(words);
    if (hidd) (void)hipFree(hidd);
    if (e != hipSuccess) { cleanup(); return cube_fail(h, CUBE_ERR_HIP, "halo catalogue failed: %s", hipGetErrorString(e)); }
  } else if (out->hid && out->nhid > 0) {
    memset(out->hid, 0, sizeof(int16_t) * out->nhid);
  }
  cleanup();
  return 0;
#undef HTRY
}
