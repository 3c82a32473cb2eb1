// Seeded Zel'dovich initial conditions generated in HBM (the role of the
// reference's IC tool, util/initial_conditions.f90:214-954, which writes
// CUBE's first checkpoint).
//
//  1. unit Gaussian noise on the particle lattice (ng = np_nc*nc per dim),
//     one counter-based SplitMix64 draw + Box-Muller per site;
//  2. R2C, then per component psi_e(k) = i k_e/k^2 * sqrt(ng^3 P(k)/V) * W(k),
//     zero on the component's own Nyquist plane, C2R;
//  3. lattice point q = i/np_nc + 1/(2 ncell) (coarse units), x = q + psi/ncell,
//     v = psi * vfac; cell = floor(x), in-cell code floor(frac(x)*2^(8 izipx));
//  4. particles binned per cell in lattice (pid) order, vfield = cell mean,
//     sigma_vi = rms residual / sqrt(3), velocities encoded relative to vfield.
// The result is a complete resident state in checkpoint order.
#include "cube_internal.hpp"
#include <math.h>
#include <string.h>
#include <algorithm>
#include <vector>

#define IC_NT 256

__host__ __device__ inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

__global__ void k_ic_noise(float* R, int64_t ng, uint64_t seed0) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= ng * ng * ng) return;
  uint64_t hsh = splitmix64(seed0 + (uint64_t)i);
  double u1 = ((double)(hsh >> 32) + 0.5) * (1.0 / 4294967296.0);
  double u2 = ((double)(hsh & 0xffffffffull) + 0.5) * (1.0 / 4294967296.0);
  double gv = sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
  int64_t x = i % ng, yz = i / ng;
  R[yz * (ng + 2) + x] = (float)gv;
}

// W = R * i k_e/k^2 * amp(k), amp = sqrt(ng^3 P(k)/V)/ng^3 * (nf/box)
__global__ void k_ic_mult(const float2* C, float2* W, int64_t ng, int e, double kf, double scale, const double* lnk,
                          const double* lnp, int nk) {
  int64_t H = ng / 2 + 1;
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= H * ng * ng) return;
  int64_t kx = i % H, ky = (i / H) % ng, kz = i / (H * ng);
  int64_t sx = kx, sy = ky < ng / 2 ? ky : ky - ng, sz = kz < ng / 2 ? kz : kz - ng;
  int64_t se = e == 0 ? sx : (e == 1 ? sy : sz);
  float2 out{0.f, 0.f};
  bool nyq = (se == ng / 2) || (se == -ng / 2);
  if ((sx | sy | sz) != 0 && !nyq) {
    double k2 = kf * kf * (double)(sx * sx + sy * sy + sz * sz);
    double k = sqrt(k2);
    double lk = log(k);
    double lp;
    if (lk <= lnk[0]) lp = lnp[0];
    else if (lk >= lnk[nk - 1]) lp = lnp[nk - 1];
    else {
      int lo = 0, hi = nk - 1;
      while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (lnk[mid] <= lk) lo = mid; else hi = mid;
      }
      double t = (lk - lnk[lo]) / (lnk[hi] - lnk[lo]);
      lp = lnp[lo] + t * (lnp[hi] - lnp[lo]);
    }
    double s = (kf * (double)se) / k2 * sqrt(exp(lp)) * scale;
    float2 c = C[i];
    out = float2{(float)(-(double)c.y * s), (float)((double)c.x * s)};
  }
  W[i] = out;
}

// per lattice site: cell, in-cell code and velocity for component e
__global__ void k_ic_particles(const float* psi, int64_t ng, int e, int np_nc, int nc, int izipx, double vfac,
                               int16_t* cellc, uint16_t* ucode, float* vel) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int64_t n = ng * ng * ng;
  if (i >= n) return;
  int64_t x = i % ng, y = (i / ng) % ng, z = i / (ng * ng);
  int64_t ie = e == 0 ? x : (e == 1 ? y : z);
  double p = (double)psi[(z * ng + y) * (ng + 2) + x];
  double q = (double)ie / np_nc + 0.5 / CUBE_NCELL;
  double xc = q + p / CUBE_NCELL;
  xc = fmod(xc, (double)nc);
  if (xc < 0) xc += nc;
  double fl = floor(xc);
  int c = (int)fl;
  if (c >= nc) c -= nc;
  const int64_t nbin = (int64_t)1 << (8 * izipx);
  int64_t u = (int64_t)floor((xc - fl) * (double)nbin);
  u = u < 0 ? 0 : (u >= nbin ? nbin - 1 : u);
  cellc[e * n + i] = (int16_t)c;
  ucode[e * n + i] = (uint16_t)u;
  vel[e * n + i] = (float)(p * vfac);
}

__global__ void k_ic_count(Geo g, const int16_t* cellc, int64_t n, int32_t* cnt) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  atomicAdd(&cnt[rank_of(g, cellc[i], cellc[n + i], cellc[2 * n + i])], 1);
}
__global__ void k_ic_scatter(Geo g, const int16_t* cellc, int64_t n, const int64_t* cs, int32_t* fill, int32_t* perm) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  int64_t r = rank_of(g, cellc[i], cellc[n + i], cellc[2 * n + i]);
  int32_t slot = atomicAdd(&fill[r], 1);
  perm[cs[r] + slot] = (int32_t)i;
}
// per cell: sort its lattice indices ascending, then the mean velocity
__global__ void k_ic_cells(const int64_t* cs, int64_t n3c, int32_t* perm, const float* vel, int64_t n, float* vfield) {
  int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= n3c) return;
  int64_t b = cs[r], e = cs[r + 1];
  for (int64_t a = b + 1; a < e; a++) {
    int32_t v = perm[a];
    int64_t c = a - 1;
    while (c >= b && perm[c] > v) {
      perm[c + 1] = perm[c];
      c--;
    }
    perm[c + 1] = v;
  }
  for (int d = 0; d < 3; d++) {
    double s = 0;
    for (int64_t a = b; a < e; a++) s += (double)vel[d * n + perm[a]];
    vfield[r * 3 + d] = e > b ? (float)(s / (double)(e - b)) : 0.f;
  }
}
// residual velocity sums for sigma_vi and max |v_z|
__global__ void k_ic_sigma(const int32_t* perm, const int16_t* cellc, Geo g, int64_t n, const float* vel,
                           const float* vfield, double* part, uint32_t* vzmax_bits) {
  __shared__ double red[IC_NT / 64];
  __shared__ float redm[IC_NT / 64];
  int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  double s = 0;
  float m = 0.f;
  if (j < n) {
    int32_t i = perm[j];
    int64_t r = rank_of(g, cellc[i], cellc[n + i], cellc[2 * n + i]);
    for (int d = 0; d < 3; d++) {
      double vr = (double)vel[d * n + i] - (double)vfield[r * 3 + d];
      s += vr * vr;
    }
    m = fabsf(vel[2 * n + i]);
  }
  for (int o = 32; o > 0; o >>= 1) {
    s += __shfl_xor(s, o);
    m = fmaxf(m, __shfl_xor(m, o));
  }
  if ((threadIdx.x & 63) == 0) {
    red[threadIdx.x >> 6] = s;
    redm[threadIdx.x >> 6] = m;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    float mm = 0.f;
    for (int i = 0; i < IC_NT / 64; i++) {
      t += red[i];
      mm = fmaxf(mm, redm[i]);
    }
    part[blockIdx.x] = t;
    atomicMax(vzmax_bits, __float_as_uint(mm));
  }
}
__global__ void k_ic_encode(const int32_t* perm, const int16_t* cellc, const uint16_t* ucode, Geo g, int64_t n,
                            const float* vel, const float* vfield, double c, const double* venc, int32_t kmax,
                            void* xp, void* vp, int32_t* pid) {
  int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (j >= n) return;
  int32_t i = perm[j];
  int64_t r = rank_of(g, cellc[i], cellc[n + i], cellc[2 * n + i]);
  for (int d = 0; d < 3; d++) {
    store_x(xp, 3 * j + d, g.izipx, (int64_t)ucode[d * n + i]);
    double vr = (double)vel[d * n + i] - (double)vfield[r * 3 + d];
    store_v(vp, 3 * j + d, g.izipv, vencode(venc, kmax, g.izipv, c, vr));
  }
  pid[j] = i + 1;
}

extern "C" int cube_ic_zeldovich(cube_handle* h, const cube_ic_params* ip, cube_sim_header* sim_out) {
  CUBE_EXCLUSIVE;
  if (!h || !ip || !ip->lnk || !ip->lnp || ip->nk < 2 || ip->np_nc < 1) return CUBE_ERR_ARG;
  if (hipSetDevice(h->p.device) != hipSuccess) return cube_fail(h, CUBE_ERR_HIP, "hipSetDevice failed");
  const int64_t ng = h->nc * ip->np_nc;
  const int64_t n = ng * ng * ng;
  if (n > h->p.np_max) return cube_fail(h, CUBE_ERR_CAPACITY, "IC needs %lld particles > np_max", (long long)n);
  if (n >= ((int64_t)1 << 31)) return cube_fail(h, CUBE_ERR_ARG, "pid is int32: at most 2^31-1 particles");
  if (ng > 2048 || (ng & (ng - 1))) return cube_fail(h, CUBE_ERR_ARG, "lattice %lld must be a power of two <= 2048", (long long)ng);
  // scratch in the (idle) fine-mesh work buffer: cellc 6n, ucode 6n, vel 12n, perm 4n bytes
  if ((size_t)28 * n > h->fine.bytes_R()) return cube_fail(h, CUBE_ERR_CAPACITY, "IC scratch does not fit");
  char* scr = (char*)h->fine.W;
  int16_t* cellc = (int16_t*)scr;
  uint16_t* ucode = (uint16_t*)(scr + 6 * n);
  float* vel = (float*)(scr + 12 * n);
  int32_t* perm = (int32_t*)(scr + 24 * n);
  Mesh m;
  int r = mesh_alloc(h, m, ng, 4, false, false);
  if (r) return r;
  double *dlnk, *dlnp;
  HIP_TRY(h, hipMalloc(&dlnk, sizeof(double) * ip->nk));
  HIP_TRY(h, hipMalloc(&dlnp, sizeof(double) * ip->nk));
  HIP_TRY(h, hipMemcpy(dlnk, ip->lnk, sizeof(double) * ip->nk, hipMemcpyHostToDevice));
  HIP_TRY(h, hipMemcpy(dlnp, ip->lnp, sizeof(double) * ip->nk, hipMemcpyHostToDevice));
  const unsigned gn = (unsigned)((n + IC_NT - 1) / IC_NT);
  const int64_t H = ng / 2 + 1;
  k_ic_noise<<<gn, IC_NT, 0, h->stream>>>((float*)m.R, ng, splitmix64(ip->seed));
  if ((r = mesh_fft3_r2c(h, m))) return r;
  const double box = ip->box, kf = 2.0 * M_PI / box;
  const double scale = sqrt((double)n / (box * box * box)) / (double)n * ((double)h->nf / box);
  Geo g = geo_of(h);
  for (int e = 0; e < 3; e++) {
    k_ic_mult<<<(unsigned)((H * ng * ng + IC_NT - 1) / IC_NT), IC_NT, 0, h->stream>>>(
        (const float2*)m.R, (float2*)m.W, ng, e, kf, scale, dlnk, dlnp, ip->nk);
    if ((r = mesh_pass(h, m, m.W, 2, true))) return r;
    if ((r = mesh_pass(h, m, m.W, 1, true))) return r;
    if ((r = mesh_c2r(h, m, m.W, m.W, ng + 2))) return r;
    k_ic_particles<<<gn, IC_NT, 0, h->stream>>>((const float*)m.W, ng, e, ip->np_nc, g.nc, g.izipx, ip->vfac, cellc,
                                                 ucode, vel);
  }
  HIP_TRY(h, hipGetLastError());
  // binning in lattice order
  HIP_TRY(h, hipMemsetAsync(h->rhoc, 0, sizeof(int32_t) * h->n3c, h->stream));
  k_ic_count<<<gn, IC_NT, 0, h->stream>>>(g, cellc, n, h->rhoc);
  if ((r = scan_counts(h, h->rhoc, h->n3c, h->cstart))) return r;
  HIP_TRY(h, hipMemsetAsync(h->rhoc2, 0, sizeof(int32_t) * h->n3c, h->stream));
  k_ic_scatter<<<gn, IC_NT, 0, h->stream>>>(g, cellc, n, h->cstart, h->rhoc2, perm);
  k_ic_cells<<<(unsigned)((h->n3c + IC_NT - 1) / IC_NT), IC_NT, 0, h->stream>>>(h->cstart, h->n3c, perm, vel, n, h->vfield);
  HIP_TRY(h, hipMemsetAsync(h->uflags + 40, 0, sizeof(uint32_t) * 4, h->stream));
  double* part;
  HIP_TRY(h, hipMalloc(&part, sizeof(double) * gn));
  k_ic_sigma<<<gn, IC_NT, 0, h->stream>>>(perm, cellc, g, n, vel, h->vfield, part, h->uflags + 40);
  HIP_TRY(h, hipGetLastError());
  double sum;
  if ((r = reduce_doubles(h, part, gn, 1, &sum))) return r;
  uint32_t vzb;
  HIP_TRY(h, hipMemcpy(&vzb, h->uflags + 40, 4, hipMemcpyDeviceToHost));
  float vz_max;
  memcpy(&vz_max, &vzb, 4);
  const float sigma = (float)(sqrt(sum / (double)n) / sqrt(3.0));
  k_ic_encode<<<gn, IC_NT, 0, h->stream>>>(perm, cellc, ucode, g, n, vel, h->vfield, vcoef(sigma), h->venc, h->vkmax,
                                           h->xp, h->vp, h->pid);
  HIP_TRY(h, hipGetLastError());
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  (void)hipFree(part);
  (void)hipFree(dlnk);
  (void)hipFree(dlnp);
  mesh_free(m);
  h->np = n;
  h->npglobal = n;
  h->sigma_vi = sigma;
  h->sigma_vi_new = sigma;
  if (int r = sigma_new_to_device(h)) return r;
  h->vz_max = vz_max;
  // initial_conditions.f90:943, real(4): f_cdm*nf_global**3/npglobal
  const float f_cdm = ip->f_cdm > 0 ? ip->f_cdm : 1.0f;
  h->mass_p = f_cdm * (float)(h->nf * h->nf * h->nf) / (float)n;
  h->ftemp_pp = 0;
  cube_sim_header& s = h->sim;
  memset(&s, 0, sizeof s);
  s.nplocal = s.npglobal = n;
  s.izipx = h->p.izipx; s.izipv = h->p.izipv; s.izipx_nu = 2; s.izipv_nu = 2;
  s.image = 1; s.nn = 1; s.nnt = h->nnt; s.nt = h->nt; s.ncell = CUBE_NCELL; s.ncb = h->ncb;
  s.a = ip->a;
  s.dt_pp = s.dt_fine = s.dt_coarse = s.dt_vmax = s.dt_vmax_nu = 1000.0f;  // initial_conditions.f90:175-179
  s.mass_p_cdm = h->mass_p;
  s.box = ip->box;
  s.sigma_vi = sigma;
  s.z_i = 1.0f / ip->a - 1.0f;
  s.z_i_nu = 10.0f;
  s.vz_max = vz_max;
  if (sim_out) *sim_out = s;
  return CUBE_OK;
}

// ------------------------------------------------------------ fluid ICs
// Neutrino fluid initial conditions (utilities/neutrino_ic.f90:40-181): the
// CDM linear density and velocity of the same seeded field, filtered per
// fluid by the free-streaming transfer functions
//   density  tf = 1/(1 + l + l^2),
//   velocity tf = (1 + 1.5 l + 2 l^2)/(1 + l + l^2)^2,
// l = k/kfs, kfs = sqrt(1.5 omega_m/(1+z_i_nu))*100 / v_n (v_n the fluid's
// Gauss-Laguerre velocity in km/s), then block-averaged onto the fluid mesh
// (hg_set_fld, hydrogr.f90:288-308): rho = <1 + delta>, m = rho <v>,
// E = rho cs2/(g-1) + m^2/(2 rho) per component.  The CDM linear fields live
// on the particle lattice here (ng = np_nc*nc per dimension; the reference
// reads them on the fine grid), so the block factor is ng/nc_fluid.
__global__ void k_nu_ic_mult(const float2* C, float2* W, int64_t ng, int field, double kf, double scale, const double* lnk,
                             const double* lnp, int nk, float kfs, float box) {
  const int64_t H = ng / 2 + 1;
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= H * ng * ng) return;
  const int64_t kx = i % H, ky = (i / H) % ng, kz = i / (H * ng);
  const int64_t sx = kx, sy = ky < ng / 2 ? ky : ky - ng, sz = kz < ng / 2 ? kz : kz - ng;
  float2 out{0.f, 0.f};
  const int e = field - 1;
  const int64_t se = e == 0 ? sx : (e == 1 ? sy : sz);
  const bool odd_nyq = field > 0 && (se == ng / 2 || se == -ng / 2);
  if ((sx | sy | sz) != 0 && !odd_nyq) {
    const double k2 = kf * kf * (double)(sx * sx + sy * sy + sz * sz);
    const double lk = 0.5 * log(k2);
    double lp;
    if (lk <= lnk[0]) lp = lnp[0];
    else if (lk >= lnk[nk - 1]) lp = lnp[nk - 1];
    else {
      int lo = 0, hi = nk - 1;
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (lnk[mid] <= lk) lo = mid; else hi = mid;
      }
      const double t = (lk - lnk[lo]) / (lnk[hi] - lnk[lo]);
      lp = lnp[lo] + t * (lnp[hi] - lnp[lo]);
    }
    // neutrino_ic.f90:88-98 and 150-160, in real(4)
    const float kr = sqrtf((float)(sx * sx) + (float)(sy * sy) + (float)(sz * sz));
    const float kp = kr * 2.0f * PI_F / box;
    const float ll = kp / kfs;
    const float q = 1.0f + ll + ll * ll;
    const float tf = field == 0 ? 1.0f / q : (1.0f + 1.5f * ll + 2.0f * ll * ll) / (q * q);
    const float2 c = C[i];
    const double a = sqrt(exp(lp)) * scale * (double)tf;
    if (field == 0) {
      out = float2{(float)((double)c.x * a), (float)((double)c.y * a)};
    } else {
      const double s = (kf * (double)se) / k2 * a;  // i k_e/k^2
      out = float2{(float)(-(double)c.y * s), (float)((double)c.x * s)};
    }
  }
  W[i] = out;
}

// block average of the lattice field (row pitch ng+2) onto fluid field q:
// q = 0 stores <1 + delta>, q = 1..3 stores <v>; sums in float, x fastest
__global__ void k_nu_ic_block(const float* R, int64_t ng, int b, int q, float* F, int nc) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t n3 = (int64_t)nc * nc * nc;
  if (i >= n3) return;
  const int64_t x = i % nc, y = (i / nc) % nc, z = i / ((int64_t)nc * nc);
  float s = 0.f;
  for (int c = 0; c < b; c++)
    for (int bb = 0; bb < b; bb++)
      for (int a = 0; a < b; a++) {
        const float v = R[((z * b + c) * ng + (y * b + bb)) * (ng + 2) + x * b + a];
        s += q == 0 ? 1.0f + v : v;
      }
  F[(int64_t)q * n3 + i] = (float)((double)s / ((double)b * b * b));
}

// momenta and energy from rho and <v> (neutrino_ic.f90:110-113, 170-177)
__global__ void k_nu_ic_finish(float* F, int nc, float cs2, float g) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t n3 = (int64_t)nc * nc * nc;
  if (i >= n3) return;
  const float rho = F[i];
  float e = rho * cs2 / (g - 1.0f);
  for (int d = 1; d <= 3; d++) {
    const float m = rho * F[(int64_t)d * n3 + i];
    F[(int64_t)d * n3 + i] = m;
    e = e + 0.5f * (m * m) / rho;
  }
  F[4 * n3 + i] = e;
}

extern "C" int cube_nu_ic(cube_handle* h, const cube_ic_params* ip) {
  CUBE_EXCLUSIVE;
  if (!h || !ip || !ip->lnk || !ip->lnp || ip->nk < 2 || ip->np_nc < 1) return CUBE_ERR_ARG;
  if (!h->nu.on) return cube_fail(h, CUBE_ERR_ARG, "cube_nu_ic needs cube_nu_init first");
  if (hipSetDevice(h->p.device) != hipSuccess) return cube_fail(h, CUBE_ERR_HIP, "hipSetDevice failed");
  NuState& s = h->nu;
  const int64_t ng = h->nc * ip->np_nc;
  if (ng > 2048 || (ng & (ng - 1))) return cube_fail(h, CUBE_ERR_ARG, "lattice %lld must be a power of two <= 2048", (long long)ng);
  for (int n = 0; n < 3; n++)
    if (s.nc[n] > ng || ng % s.nc[n])
      return cube_fail(h, CUBE_ERR_ARG, "fluid %d of %d cells is finer than the %lld^3 particle lattice", n, s.nc[n],
                       (long long)ng);
  const cube_nu_params& p = s.par;
  Mesh m;
  int r = mesh_alloc(h, m, ng, 4, false, false);
  if (r) return r;
  double *dlnk, *dlnp;
  HIP_TRY(h, hipMalloc(&dlnk, sizeof(double) * ip->nk));
  HIP_TRY(h, hipMalloc(&dlnp, sizeof(double) * ip->nk));
  HIP_TRY(h, hipMemcpy(dlnk, ip->lnk, sizeof(double) * ip->nk, hipMemcpyHostToDevice));
  HIP_TRY(h, hipMemcpy(dlnp, ip->lnp, sizeof(double) * ip->nk, hipMemcpyHostToDevice));
  const int64_t n = ng * ng * ng;
  k_ic_noise<<<(unsigned)((n + IC_NT - 1) / IC_NT), IC_NT, 0, h->stream>>>((float*)m.R, ng, splitmix64(ip->seed));
  if ((r = mesh_fft3_r2c(h, m))) return r;
  const double box = ip->box, kf = 2.0 * M_PI / box;
  const double sdelta = sqrt((double)n / (box * box * box)) / (double)n;
  const double svel = sdelta * ((double)h->nf / box) * ip->vfac;
  // neutrino_ic.f90:12 (real(4) parameter) and :79-82
  const float kfs1 = sqrtf(1.5f * p.omega_m / (1.0f + p.z_i_nu)) * 100.0f;
  const float beta = 5966.34f;   // dfld_beta_eV (hydrodf.f90:9)
  const float sol = 2.998e5f;    // dfld_sol (hydrodf.f90:11)
  static const float gl_v[3] = {0.41577456f, 2.29428036f, 6.28994508f};
  const int64_t H = ng / 2 + 1;
  for (int fl = 0; fl < 3; fl++) {
    const float vneu = gl_v[fl] / beta * sol / p.m_ev;
    const float kfs = kfs1 / vneu;
    const int nc = s.nc[fl];
    const int b = (int)(ng / nc);
    const int64_t n3 = (int64_t)nc * nc * nc;
    for (int field = 0; field < 4; field++) {
      k_nu_ic_mult<<<(unsigned)((H * ng * ng + IC_NT - 1) / IC_NT), IC_NT, 0, h->stream>>>(
          (const float2*)m.R, (float2*)m.W, ng, field, kf, field == 0 ? sdelta : svel, dlnk, dlnp, ip->nk, kfs,
          (float)box);
      if ((r = mesh_pass(h, m, m.W, 2, true))) return r;
      if ((r = mesh_pass(h, m, m.W, 1, true))) return r;
      if ((r = mesh_c2r(h, m, m.W, m.W, ng + 2))) return r;
      k_nu_ic_block<<<(unsigned)((n3 + IC_NT - 1) / IC_NT), IC_NT, 0, h->stream>>>((const float*)m.W, ng, b, field,
                                                                                 s.F[fl], nc);
    }
    k_nu_ic_finish<<<(unsigned)((n3 + IC_NT - 1) / IC_NT), IC_NT, 0, h->stream>>>(s.F[fl], nc, s.cs2[fl], p.gamma);
  }
  HIP_TRY(h, hipGetLastError());
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  (void)hipFree(dlnk);
  (void)hipFree(dlnp);
  mesh_free(m);
  return CUBE_OK;
}
