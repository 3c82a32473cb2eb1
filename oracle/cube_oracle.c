/*
 * CPU ORACLE -- test infrastructure only.
 *
 * A plain-C restatement of CUBE's per-step gravity/update path, used by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
 * checker.  Nothing in the product (cube_nufluid_amd/) links, loads or calls
 * this file.
 *
 * It follows the reference's loop structure literally: per tile, the
 * extended-tile cell loops in the reference's processing order
 * (ilayer -> k -> j -> i -> l), tile-local float coordinates, per-tile
 * nfe^3 FFT convolution with the tabulated fine kernel, a global nc^3
 * convolution with the coarse kernel, head-insertion linked lists for PP,
 * and the reference's float/double mixing.  Ghost particles of the
 * reference's extended-tile layout (buffer_x/buffer_v, redistribute_cdm)
 * are periodic copies of physical particles; with one image they are read
 * here through periodic index wraps of the compact physical arrays, which
 * is what those routines produce.
 *
 * Reference files (CUBE_v2.0/work/main/):
 *   update_particle.f90:13-259   drift, re-bin, re-quantise, sigma_v stats
 *   pm.f90:2-451                 fine/coarse PM, kicks, dt constraints
 *   ext_pp_force.f90:9-160       PP linked lists, pair force, kick
 *   kernel_f.f90:1-43            fine kernel from wfxyzf.3.ascii
 *   kernel_c.f90:1-134           coarse kernel + LRCKCORR
 *   timestep.f90:136-180         expansion (Taylor-3 Friedmann step)
 *   parameters.f90:8-19,111-116  fixed-point codecs
 *
 * FFTs: FFTW3 single precision (legacy sfftw_* API, version unpinned in the
 * reference) is absent from this image; the transforms here are a
 * mixed-radix DFT evaluated in double precision, results stored in float
 * where the reference stores float arrays.  FFT arithmetic is therefore
 * pinned by the DFT definition only (tests compare against numpy.fft and
 * closed forms), not by FFTW's rounding.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>

/* fluid_oracle.c */
int ofl_evolve(float* P, int nc, int ncb, float g, float dt, float dx, int dir, float* dtg);
void ofl_gravity_fine(float* P, int nc, float dx, const float* F, int nf, float a_mid, float dt);
void ofl_gravity_coarse(float* P, int nc, float dx, const float* Fc, int ncc, int ncell, float a_mid, float dt);
void ofl_density_centres(float* out, int nf, int nfl, const float** rho, const int* ncf, const float* dxf, const float* w);
#endif

typedef struct { double re, im; } cpx;

typedef struct orc {
  int64_t nc, nt, nnt, ncb, ncell, izipx, izipv;
  int64_t nf, nft, nfb, nfe, nf_cutoff;
  int pp_on, pp_range;
  float rsoft;
  /* compact physical state, cell-rank (checkpoint) order */
  int64_t np;
  int16_t *xp, *vp;
  int32_t *pid;
  int32_t *rhoc;
  float *vfield;
  int64_t *cstart;
  float sigma_vi, sigma_vi_new, vz_max, mass_p;
  float ftemp_pp; /* ext_pp_force.f90:18 'real ftemp' is SAVEd, never reset */
  int64_t tile_limit; /* timing samples only: process tiles with index < tile_limit */
  double t_global;    /* seconds of the last PM call spent on box-wide coarse mesh work */
  /* neutrino fluid (fluid_oracle.c): three fluids, fields [5][nc^3] */
  int nu_on;
  float nu_fneu, nu_zi, nu_gamma, nu_w[3], nu_dx[3], nu_ai[3];
  int nu_nc[3], nu_ncb[3];
  float* nu_F[3];
  float *kern_f;  /* (nfe/2+1, nfe, nfe, 3) Fortran order */
  float *kern_c;  /* (nc/2+1, nc, nc, 3) */
  float fk_table[16 * 16 * 16 * 3];
  float ck_table[3 * 4 * 4 * 4];
  char err[256];
} orc_t;

static const float PI_F = 3.14159274101257324f; /* parameters.f90:67 pi=4*atan(1.) */

/* Planted mistakes for tests that must show they can see them (never set
   outside tests): kicks without the 1/(6 pi) of pm.f90:146, kicks with 1 in
   place of a_mid, and a drift by dt in place of dt_mid
   (update_particle.f90:34).  The kick factor macros evaluate to a_mid, 6
   and pi when no perturbation is set, so the arithmetic is unchanged. */
#define ORC_PERTURB_NO_6PI 1
#define ORC_PERTURB_NO_AMID 2
#define ORC_PERTURB_DT_NOT_MID 4
static int g_perturb = 0;
void orc_set_perturb(int flags) { g_perturb = flags; }
#define KA ((g_perturb & ORC_PERTURB_NO_AMID) ? 1.0f : a_mid)
#define K6 ((g_perturb & ORC_PERTURB_NO_6PI) ? 1.0f : 6.0f)
#define KPI ((g_perturb & ORC_PERTURB_NO_6PI) ? 1.0f : PI_F)

/* ---------------------------------------------------------------- codecs */
static inline int64_t nvbin_of(int64_t izipv) { return (int64_t)1 << (8 * izipv); }

/* int(xp+ishift,izipx)+rshift == unsigned bit pattern + 0.5 (parameters.f90:16-18) */
static inline double xoff(int16_t xp, int64_t izipx) {
  uint32_t u = izipx == 1 ? (uint32_t)(uint8_t)(int8_t)xp : (uint32_t)(uint16_t)xp;
  return ((double)u + 0.5) * (izipx == 1 ? 1.0 / 256.0 : 1.0 / 65536.0);
}
static inline int64_t xsub(int16_t xp, int64_t izipx) { /* fine sub-cell: floor(4*xoff) */
  uint32_t u = izipx == 1 ? (uint32_t)(uint8_t)(int8_t)xp : (uint32_t)(uint16_t)xp;
  return izipx == 1 ? (u >> 6) : (u >> 14);
}
static inline int16_t xwrap(int64_t v, int64_t izipx) {
  return izipx == 1 ? (int16_t)(int8_t)(uint8_t)(uint64_t)v : (int16_t)(uint16_t)(uint64_t)v;
}

/* vreal=tan((pi*real(vp))/real(nvbin-1)) / (sqrt(pi/2)/(sigma*vrel_boost))
   update_particle.f90:69; tan in single precision, divide in double. */
double orc_vdecode(int32_t vp, float sigma, int64_t izipv) {
  float nv1 = (float)(nvbin_of(izipv) - 1);
  float t = tanf((PI_F * (float)vp) / nv1);
  double c = (double)sqrtf(PI_F / 2) / ((double)sigma * 2.5);
  return (double)t / c;
}
static inline double nint_d(double y) { return y >= 0 ? floor(y + 0.5) : -floor(-y + 0.5); }
/* vp=nint(real(nvbin-1)*atan(sqrt(pi/2)/(sigma*vrel_boost)*vreal)/pi)
   update_particle.f90:119; atan in double (vrel_boost is real(8)). */
int32_t orc_vencode(double v, float sigma, int64_t izipv) {
  float nv1 = (float)(nvbin_of(izipv) - 1);
  double c = (double)sqrtf(PI_F / 2) / ((double)sigma * 2.5);
  double y = (double)nv1 * atan(c * v) / (double)PI_F;
  return (int32_t)nint_d(y);
}
/* the same encode from the atan argument x = c*v, for n values */
void orc_vencode_x(const double* x, int64_t n, int64_t izipv, int32_t* out) {
  float nv1 = (float)(nvbin_of(izipv) - 1);
  for (int64_t i = 0; i < n; i++) out[i] = (int32_t)nint_d((double)nv1 * atan(x[i]) / (double)PI_F);
}

/* ------------------------------------------------------------ geometry */
static inline int64_t wrapi(int64_t a, int64_t n) { a %= n; return a < 0 ? a + n : a; }
static inline int64_t rank_of(const orc_t* o, int64_t gx, int64_t gy, int64_t gz) {
  int64_t nt = o->nt, nnt = o->nnt;
  int64_t t = ((gz / nt) * nnt + gy / nt) * nnt + gx / nt;
  return t * nt * nt * nt + ((gz % nt) * nt + gy % nt) * nt + gx % nt;
}
/* rank of the cell at tile-local (1-based, may lie in the buffer) coords */
static inline int64_t rank_loc(const orc_t* o, int64_t tx, int64_t ty, int64_t tz,
                               int64_t i, int64_t j, int64_t k) {
  return rank_of(o, wrapi(tx * o->nt + i - 1, o->nc), wrapi(ty * o->nt + j - 1, o->nc),
                 wrapi(tz * o->nt + k - 1, o->nc));
}

/* ------------------------------------------------------------------ FFT */
static void factorize(int64_t n, int* f) {
  int nf = 0;
  int64_t p;
  for (p = 4; n % 4 == 0 && n > 4; n /= 4) f[nf++] = 4;
  for (p = 2; n > 1; ) {
    if (n % p == 0) { f[nf++] = (int)p; n /= p; }
    else p = (p == 2) ? 3 : p + 2;
  }
  f[nf] = 0;
}
/* recursive mixed-radix decimation-in-time DFT (unnormalised, sign -1 fwd) */
static void dft_rec(cpx* out, const cpx* in, int64_t n, int64_t s, const int* fac,
                    const cpx* tw, int64_t ntot, cpx* scratch) {
  int p = fac[0];
  int64_t m = n / p;
  if (m == 1) {
    for (int q = 0; q < p; q++) out[q] = in[q * s];
  } else {
    for (int q = 0; q < p; q++) dft_rec(out + q * m, in + q * s, m, s * p, fac + 1, tw, ntot, scratch);
  }
  int64_t step = ntot / n;
  for (int64_t k = 0; k < m; k++) {
    for (int q = 0; q < p; q++) {
      cpx a = out[k + q * m];
      cpx w = tw[((int64_t)q * k * step) % ntot];
      scratch[q].re = a.re * w.re - a.im * w.im;
      scratch[q].im = a.re * w.im + a.im * w.re;
    }
    for (int r = 0; r < p; r++) {
      double sr = 0, si = 0;
      for (int q = 0; q < p; q++) {
        cpx w = tw[(((int64_t)q * r) % p) * (ntot / p)];
        sr += scratch[q].re * w.re - scratch[q].im * w.im;
        si += scratch[q].re * w.im + scratch[q].im * w.re;
      }
      out[k + r * m].re = sr;
      out[k + r * m].im = si;
    }
  }
}
typedef struct { int64_t n; int fac[64]; cpx* tw; } plan1d;
static void plan_make(plan1d* pl, int64_t n, int sign) {
  pl->n = n;
  factorize(n, pl->fac);
  pl->tw = (cpx*)malloc(sizeof(cpx) * n);
  for (int64_t j = 0; j < n; j++) {
    double a = sign * 2.0 * M_PI * (double)j / (double)n;
    pl->tw[j].re = cos(a);
    pl->tw[j].im = sin(a);
  }
}
static void plan_free(plan1d* pl) { free(pl->tw); }
static void dft1(const plan1d* pl, cpx* buf, cpx* tmp, cpx* scratch) {
  dft_rec(tmp, buf, pl->n, 1, pl->fac, pl->tw, pl->n, scratch);
  memcpy(buf, tmp, sizeof(cpx) * pl->n);
}
/* 3D complex transform of a[z][y][x] (x fastest) with dims (nx,ny,nz) */
static void dft3(cpx* a, int64_t nx, int64_t ny, int64_t nz, int sign, int threads) {
  plan1d px, py, pz;
  plan_make(&px, nx, sign); plan_make(&py, ny, sign); plan_make(&pz, nz, sign);
  int64_t nmax = nx > ny ? nx : ny; if (nz > nmax) nmax = nz;
#pragma omp parallel num_threads(threads)
  {
    cpx* b = (cpx*)malloc(sizeof(cpx) * nmax);
    cpx* t = (cpx*)malloc(sizeof(cpx) * nmax);
    cpx sc[64];
#pragma omp for schedule(static)
    for (int64_t r = 0; r < ny * nz; r++) {
      memcpy(b, a + r * nx, sizeof(cpx) * nx);
      dft1(&px, b, t, sc);
      memcpy(a + r * nx, b, sizeof(cpx) * nx);
    }
#pragma omp for schedule(static)
    for (int64_t r = 0; r < nx * nz; r++) {
      int64_t x = r % nx, z = r / nx;
      for (int64_t y = 0; y < ny; y++) b[y] = a[(z * ny + y) * nx + x];
      dft1(&py, b, t, sc);
      for (int64_t y = 0; y < ny; y++) a[(z * ny + y) * nx + x] = b[y];
    }
#pragma omp for schedule(static)
    for (int64_t r = 0; r < nx * ny; r++) {
      for (int64_t z = 0; z < nz; z++) b[z] = a[z * nx * ny + r];
      dft1(&pz, b, t, sc);
      for (int64_t z = 0; z < nz; z++) a[z * nx * ny + r] = b[z];
    }
    free(b); free(t);
  }
  plan_free(&px); plan_free(&py); plan_free(&pz);
}
/* R2C of real r[n^3] -> c[(n/2+1) x n x n] (FFTW layout, unnormalised) */
static void r2c3(const float* r, float* cre, float* cim, int64_t n, int threads) {
  int64_t n3 = n * n * n, h = n / 2 + 1;
  cpx* a = (cpx*)malloc(sizeof(cpx) * n3);
  for (int64_t i = 0; i < n3; i++) { a[i].re = r[i]; a[i].im = 0; }
  dft3(a, n, n, n, -1, threads);
  for (int64_t z = 0; z < n; z++)
    for (int64_t y = 0; y < n; y++)
      for (int64_t x = 0; x < h; x++) {
        cpx v = a[(z * n + y) * n + x];
        cre[(z * n + y) * h + x] = (float)v.re;
        cim[(z * n + y) * h + x] = (float)v.im;
      }
  free(a);
}
/* C2R (unnormalised, Hermitian half-spectrum input; imaginary parts of the
   self-conjugate bins are discarded as FFTW's c2r does) */
static void c2r3(const float* cre, const float* cim, float* r, int64_t n, int threads) {
  int64_t n3 = n * n * n, h = n / 2 + 1;
  cpx* a = (cpx*)malloc(sizeof(cpx) * n3);
  for (int64_t z = 0; z < n; z++)
    for (int64_t y = 0; y < n; y++)
      for (int64_t x = 0; x < n; x++) {
        cpx v;
        if (x < h) {
          v.re = cre[(z * n + y) * h + x];
          v.im = cim[(z * n + y) * h + x];
        } else {
          int64_t xs = n - x, ys = (n - y) % n, zs = (n - z) % n;
          v.re = cre[(zs * n + ys) * h + xs];
          v.im = -cim[(zs * n + ys) * h + xs];
        }
        a[(z * n + y) * n + x] = v;
      }
  dft3(a, n, n, n, +1, threads);
  for (int64_t i = 0; i < n3; i++) r[i] = (float)a[i].re;
  free(a);
}

/* ------------------------------------------------------------- kernels */
/* kernel_f.f90:14-37: table (16^3 x 3), PP core zeroed, odd/even mirror,
   R2C on the nfe^3 tile grid, keep Im. */
static void build_kern_f(orc_t* o, int threads) {
  int64_t nfe = o->nfe, h = nfe / 2 + 1, n3 = nfe * nfe * nfe, nc16 = o->nf_cutoff;
  float* rho = (float*)calloc(n3, sizeof(float));
  float* cre = (float*)malloc(sizeof(float) * h * nfe * nfe);
  float* cim = (float*)malloc(sizeof(float) * h * nfe * nfe);
  o->kern_f = (float*)malloc(sizeof(float) * h * nfe * nfe * 3);
  for (int d = 0; d < 3; d++) {
    memset(rho, 0, sizeof(float) * n3);
#define RF(i, j, k) rho[((int64_t)(k)-1) * nfe * nfe + ((int64_t)(j)-1) * nfe + ((i)-1)]
    for (int k = 1; k <= nc16; k++)
      for (int j = 1; j <= nc16; j++)
        for (int i = 1; i <= nc16; i++)
          RF(i, j, k) = o->fk_table[((((k - 1) * 16 + (j - 1)) * 16 + (i - 1)) * 3) + d];
    if (o->pp_on)
      for (int k = 1; k <= o->pp_range + 1; k++)
        for (int j = 1; j <= o->pp_range + 1; j++)
          for (int i = 1; i <= o->pp_range + 1; i++) RF(i, j, k) = 0;
    float mf[3] = {d == 0 ? -1.f : 1.f, d == 1 ? -1.f : 1.f, d == 2 ? -1.f : 1.f};
    /* rho_f(nfe-nf_cutoff+2:nfe,:,:)=mfactor(1)*rho_f(nf_cutoff:2:-1,:,:) */
    for (int64_t k = 1; k <= nfe; k++)
      for (int64_t j = 1; j <= nfe; j++)
        for (int64_t m = 0; m < nc16 - 1; m++) RF(nfe - nc16 + 2 + m, j, k) = mf[0] * RF(nc16 - m, j, k);
    for (int64_t k = 1; k <= nfe; k++)
      for (int64_t m = 0; m < nc16 - 1; m++)
        for (int64_t i = 1; i <= nfe; i++) RF(i, nfe - nc16 + 2 + m, k) = mf[1] * RF(i, nc16 - m, k);
    for (int64_t m = 0; m < nc16 - 1; m++)
      for (int64_t j = 1; j <= nfe; j++)
        for (int64_t i = 1; i <= nfe; i++) RF(i, j, nfe - nc16 + 2 + m) = mf[2] * RF(i, j, nc16 - m);
#undef RF
    r2c3(rho, cre, cim, nfe, threads);
    memcpy(o->kern_f + (int64_t)d * h * nfe * nfe, cim, sizeof(float) * h * nfe * nfe);
  }
  free(rho); free(cre); free(cim);
}

/* kernel_c.f90:16-117 with nn=1 (every octant branch applies). */
static void build_kern_c(orc_t* o, int threads) {
  int64_t nc = o->nc, h = nc / 2 + 1, n3 = nc * nc * nc, ncell = o->ncell;
  float* ck = (float*)malloc(sizeof(float) * 3 * n3);  /* ck(3,nc,nc,nc) */
  float* ck0 = (float*)malloc(sizeof(float) * 3 * n3); /* uncorrected */
#define CK(a, d, i, j, k) a[(((int64_t)(k)-1) * nc * nc + ((int64_t)(j)-1) * nc + ((i)-1)) * 3 + (d)]
  for (int64_t k = 1; k <= nc; k++)
    for (int64_t j = 1; j <= nc; j++)
      for (int64_t i = 1; i <= nc; i++) {
        float rx[3];
        int64_t g[3] = {i, j, k};
        for (int d = 0; d < 3; d++) rx[d] = (float)(ncell * (wrapi(g[d] + nc / 2 - 1, nc) - nc / 2));
        float r = sqrtf(rx[0] * rx[0] + rx[1] * rx[1] + rx[2] * rx[2]);
        for (int d = 0; d < 3; d++) {
          float v = (r == 0) ? 0.0f : -rx[d] / (r * r * r);
          CK(ck, d, i, j, k) = v;
          CK(ck0, d, i, j, k) = v;
        }
      }
  /* ck_table(3,4,4,4) read as (3i4,3e16.8); octant copies with sign flips */
#define CT(d, i, j, k) o->ck_table[((((k)-1) * 16 + ((j)-1) * 4 + ((i)-1)) * 3) + (d)]
  /* octant assignments, kernel_c.f90:43-72 ('4:2:-1' maps nc-2..nc <- 4..2) */
  for (int k = 1; k <= 4; k++)
    for (int j = 1; j <= 4; j++)
      for (int i = 1; i <= 4; i++)
        for (int d = 0; d < 3; d++) CK(ck, d, i, j, k) = CT(d, i, j, k);
  for (int m = 0; m < 3; m++) { /* index nc-2+m <- table index 4-m */
    int t = 4 - m;
    for (int b = 1; b <= 4; b++)
      for (int c = 1; c <= 4; c++)
        for (int d = 0; d < 3; d++) {
          /* 5 x */ CK(ck, d, nc - 2 + m, b, c) = (d == 0 ? -1.f : 1.f) * CT(d, t, b, c);
          /* 3 y */ CK(ck, d, b, nc - 2 + m, c) = (d == 1 ? -1.f : 1.f) * CT(d, b, t, c);
          /* 2 z */ CK(ck, d, b, c, nc - 2 + m) = (d == 2 ? -1.f : 1.f) * CT(d, b, c, t);
        }
    for (int m2 = 0; m2 < 3; m2++) {
      int t2 = 4 - m2;
      for (int c = 1; c <= 4; c++)
        for (int d = 0; d < 3; d++) {
          /* 4 yz */ CK(ck, d, c, nc - 2 + m, nc - 2 + m2) = (d == 0 ? 1.f : -1.f) * CT(d, c, t, t2);
          /* 6 xz */ CK(ck, d, nc - 2 + m, c, nc - 2 + m2) = (d == 1 ? 1.f : -1.f) * CT(d, t, c, t2);
          /* 7 xy */ CK(ck, d, nc - 2 + m, nc - 2 + m2, c) = (d == 2 ? 1.f : -1.f) * CT(d, t, t2, c);
        }
      for (int m3 = 0; m3 < 3; m3++) {
        int t3 = 4 - m3;
        for (int d = 0; d < 3; d++) /* 8 xyz */
          CK(ck, d, nc - 2 + m, nc - 2 + m2, nc - 2 + m3) = -CT(d, t, t2, t3);
      }
    }
  }
#undef CT
  float* r3 = (float*)malloc(sizeof(float) * n3);
  float* cre = (float*)malloc(sizeof(float) * h * nc * nc);
  float* cim = (float*)malloc(sizeof(float) * h * nc * nc);
  float* cim0 = (float*)malloc(sizeof(float) * h * nc * nc);
  o->kern_c = (float*)malloc(sizeof(float) * h * nc * nc * 3);
  for (int d = 0; d < 3; d++) {
    for (int64_t q = 0; q < n3; q++) r3[q] = ck[q * 3 + d];
    r2c3(r3, cre, cim, nc, threads);
    for (int64_t q = 0; q < n3; q++) r3[q] = ck0[q * 3 + d];
    r2c3(r3, cre, cim0, nc, threads);
    float* kc = o->kern_c + (int64_t)d * h * nc * nc;
    for (int64_t k = 1; k <= nc; k++)
      for (int64_t j = 1; j <= nc; j++)
        for (int64_t i = 1; i <= h; i++) {
          int64_t q = ((k - 1) * nc + (j - 1)) * h + (i - 1);
          float kx[3] = {(float)(wrapi(i + nc / 2 - 1, nc) - nc / 2), (float)(wrapi(j + nc / 2 - 1, nc) - nc / 2),
                         (float)(wrapi(k + nc / 2 - 1, nc) - nc / 2)};
          float kr = sqrtf(kx[0] * kx[0] + kx[1] * kx[1] + kx[2] * kx[2]);
          float ks[3];
          for (int e = 0; e < 3; e++) ks[e] = 2 * sinf(PI_F * kx[e] / (float)nc);
          float v = cim[q];
          if (!(kr > 8.0f || kx[d] == 0)) {
            /* On the odd component's own Nyquist plane (kx(d) = -nc/2) both
               Im FT(ck) and Im FT(-r/r^3) vanish exactly; the reference divides
               FFTW round-off there (only reachable for nc=16, |k|=8).  The
               exact-arithmetic value 0 is used. */
            if (kx[d] == (float)(-nc / 2)) v = 0.0f;
            else {
              float s2 = ks[0] * ks[0] + ks[1] * ks[1] + ks[2] * ks[2];
              v = v * 0.25f * PI_F * ks[d] / s2 / cim0[q];
            }
          }
          kc[q] = v;
        }
  }
#undef CK
  free(ck); free(ck0); free(r3); free(cre); free(cim); free(cim0);
}

/* ------------------------------------------------------------ lifecycle */
orc_t* orc_create(const int64_t* ip, const float* fp, const float* fk_table, const float* ck_table, int threads) {
  orc_t* o = (orc_t*)calloc(1, sizeof(orc_t));
  o->nc = ip[0]; o->nnt = ip[1]; o->ncb = ip[2]; o->izipx = ip[3]; o->izipv = ip[4];
  o->pp_on = (int)ip[5]; o->pp_range = (int)ip[6];
  o->ncell = 4; o->nf_cutoff = 16;
  o->rsoft = fp[0];
  o->nt = o->nc / o->nnt;
  o->nf = o->nc * o->ncell;
  o->nft = o->nt * o->ncell;
  o->nfb = o->ncb * o->ncell;
  o->nfe = o->nft + 2 * o->nfb;
  memcpy(o->fk_table, fk_table, sizeof(o->fk_table));
  memcpy(o->ck_table, ck_table, sizeof(o->ck_table));
  build_kern_f(o, threads);
  build_kern_c(o, threads);
  return o;
}
void orc_destroy(orc_t* o) {
  if (!o) return;
  free(o->xp); free(o->vp); free(o->pid); free(o->rhoc); free(o->vfield); free(o->cstart);
  free(o->kern_f); free(o->kern_c);
  for (int n = 0; n < 3; n++) free(o->nu_F[n]);
  free(o);
}
const char* orc_error(orc_t* o) { return o->err; }
int64_t orc_np(orc_t* o) { return o->np; }
void orc_kernels(orc_t* o, float* kf, float* kc) {
  int64_t hf = o->nfe / 2 + 1, hc = o->nc / 2 + 1;
  if (kf) memcpy(kf, o->kern_f, sizeof(float) * hf * o->nfe * o->nfe * 3);
  if (kc) memcpy(kc, o->kern_c, sizeof(float) * hc * o->nc * o->nc * 3);
}
void orc_scalars(orc_t* o, float* out) {
  out[0] = o->sigma_vi; out[1] = o->sigma_vi_new; out[2] = o->vz_max; out[3] = o->mass_p; out[4] = o->ftemp_pp;
}
void orc_set_tile_limit(orc_t* o, int64_t n) { o->tile_limit = n; }
double orc_global_seconds(const orc_t* o) { return o->t_global; }

/* neutrino fluid set-up: fpar = {f_neu, z_i_nu, gamma, w[3], dx[3], a_i[3]},
   ipar = {nc[3], ncb[3]}; fields start as given by orc_nu_set */
void orc_nu_setup(orc_t* o, const float* fpar, const int* ipar) {
  o->nu_on = 1;
  o->nu_fneu = fpar[0]; o->nu_zi = fpar[1]; o->nu_gamma = fpar[2];
  for (int n = 0; n < 3; n++) {
    o->nu_w[n] = fpar[3 + n]; o->nu_dx[n] = fpar[6 + n]; o->nu_ai[n] = fpar[9 + n];
    o->nu_nc[n] = ipar[n]; o->nu_ncb[n] = ipar[3 + n];
    free(o->nu_F[n]);
    o->nu_F[n] = (float*)calloc((size_t)5 * ipar[n] * ipar[n] * ipar[n], sizeof(float));
  }
}
void orc_nu_set(orc_t* o, int n, const float* f) {
  memcpy(o->nu_F[n], f, sizeof(float) * 5 * (size_t)o->nu_nc[n] * o->nu_nc[n] * o->nu_nc[n]);
}
void orc_nu_get(const orc_t* o, int n, float* f) {
  memcpy(f, o->nu_F[n], sizeof(float) * 5 * (size_t)o->nu_nc[n] * o->nu_nc[n] * o->nu_nc[n]);
}
/* dfld_evolve (hydrodf.f90:90-110) */
int orc_nu_evolve(orc_t* o, float dt, int dir, float a, float* dt_neu) {
  for (int n = 0; n < 3; n++) {
    if (!(a >= o->nu_ai[n])) continue;
    if (ofl_evolve(o->nu_F[n], o->nu_nc[n], o->nu_ncb[n], o->nu_gamma, dt, o->nu_dx[n], dir, dt_neu)) return 1;
  }
  return 0;
}
void orc_set_scalars(orc_t* o, const float* in) {
  o->sigma_vi = in[0]; o->sigma_vi_new = in[1]; o->vz_max = in[2]; o->mass_p = in[3]; o->ftemp_pp = in[4];
}

static void build_cstart(orc_t* o) {
  int64_t n3 = o->nc * o->nc * o->nc;
  free(o->cstart);
  o->cstart = (int64_t*)malloc(sizeof(int64_t) * (n3 + 1));
  int64_t s = 0;
  for (int64_t r = 0; r < n3; r++) { o->cstart[r] = s; s += o->rhoc[r]; }
  o->cstart[n3] = s;
}

/* xp/vp are byte arrays of izipx/izipv bytes per component (checkpoint layout) */
int orc_set_state(orc_t* o, int64_t np, const void* xp, const void* vp, const int32_t* pid,
                  const int32_t* rhoc, const float* vfield, float sigma_vi, float vz_max, float mass_p) {
  int64_t n3 = o->nc * o->nc * o->nc;
  free(o->xp); free(o->vp); free(o->pid); free(o->rhoc); free(o->vfield);
  o->np = np;
  o->xp = (int16_t*)malloc(sizeof(int16_t) * 3 * (np + 1));
  o->vp = (int16_t*)malloc(sizeof(int16_t) * 3 * (np + 1));
  o->pid = (int32_t*)malloc(sizeof(int32_t) * (np + 1));
  for (int64_t q = 0; q < 3 * np; q++) {
    o->xp[q] = o->izipx == 1 ? ((const int8_t*)xp)[q] : ((const int16_t*)xp)[q];
    o->vp[q] = o->izipv == 1 ? ((const int8_t*)vp)[q] : ((const int16_t*)vp)[q];
  }
  memcpy(o->pid, pid, sizeof(int32_t) * np);
  o->rhoc = (int32_t*)malloc(sizeof(int32_t) * n3);
  o->vfield = (float*)malloc(sizeof(float) * 3 * n3);
  memcpy(o->rhoc, rhoc, sizeof(int32_t) * n3);
  memcpy(o->vfield, vfield, sizeof(float) * 3 * n3);
  o->sigma_vi = sigma_vi; o->sigma_vi_new = sigma_vi; o->vz_max = vz_max; o->mass_p = mass_p;
  build_cstart(o);
  if (o->cstart[n3] != np) { snprintf(o->err, sizeof o->err, "sum(rhoc)=%lld != np=%lld", (long long)o->cstart[n3], (long long)np); return 1; }
  return 0;
}
int orc_get_state(orc_t* o, void* xp, void* vp, int32_t* pid, int32_t* rhoc, float* vfield) {
  int64_t n3 = o->nc * o->nc * o->nc;
  for (int64_t q = 0; q < 3 * o->np; q++) {
    if (xp) { if (o->izipx == 1) ((int8_t*)xp)[q] = (int8_t)o->xp[q]; else ((int16_t*)xp)[q] = o->xp[q]; }
    if (vp) { if (o->izipv == 1) ((int8_t*)vp)[q] = (int8_t)o->vp[q]; else ((int16_t*)vp)[q] = o->vp[q]; }
  }
  if (pid) memcpy(pid, o->pid, sizeof(int32_t) * o->np);
  if (rhoc) memcpy(rhoc, o->rhoc, sizeof(int32_t) * n3);
  if (vfield) memcpy(vfield, o->vfield, sizeof(float) * 3 * n3);
  return 0;
}

/* ---------------------------------------------------------------- drift */
/* update_particle.f90:13-259.  out: [sigma_vi_new, std_vsim, std_vsim_c,
   std_vsim_res, nlayer, np, sum_c, sum_r, sum_s] */
int orc_drift(orc_t* o, float dt_old, float dt, double* out, int threads) {
  const int64_t nt = o->nt, nnt = o->nnt, ncb = o->ncb, nc = o->nc, n3 = nc * nc * nc;
  const int64_t lo = 1 - 2 * ncb, E = nt + 4 * ncb;
  const double xres = o->izipx == 1 ? 1.0 / 256.0 : 1.0 / 65536.0;
  const double wv = (double)0.1f; /* real(8),parameter :: weight_v=0.1 (a default-real literal) */
  /* update_particle.f90:34; the perturbation drifts by dt instead */
  const float dt_mid = (g_perturb & ORC_PERTURB_DT_NOT_MID) ? dt : (dt_old + dt) / 2;
  const int nlayer = 2 * (int)ceilf(dt_mid * o->vz_max / (float)o->ncell) + 1;
  const float sig = o->sigma_vi;
  int32_t* rhoc_new = (int32_t*)calloc(n3, sizeof(int32_t));
  float* vf_new = (float*)calloc(3 * n3, sizeof(float));
  int bad = 0;
  /* Tiles run in parallel when there are at least as many as threads;
     otherwise (a timing sample of few tiles) tiles run one after another and
     the planes k of one layer run in parallel, as the reference's own OpenMP
     loop does (update_particle.f90:57-63): planes of a layer lie nlayer apart
     and a particle moves less than nlayer/2 planes, so their destinations are
     disjoint and each destination still sees its arrivals in order.  A tile
     where some particle moves further is redone serially. */
  const int64_t ntiles = nnt * nnt * nnt;
  const int64_t active = o->tile_limit > 0 && o->tile_limit < ntiles ? o->tile_limit : ntiles;
  const int par_tiles = active >= threads;
  const int64_t hl = (nlayer - 1) / 2;
  /* pass 1 (update_particle.f90:57-85) per tile */
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1) collapse(3) if (par_tiles)
  for (int64_t tz = 0; tz < nnt; tz++)
    for (int64_t ty = 0; ty < nnt; ty++)
      for (int64_t tx = 0; tx < nnt; tx++) {
        if (o->tile_limit > 0 && (tz * nnt + ty) * nnt + tx >= o->tile_limit) continue;
        int32_t* rhoce = (int32_t*)malloc(sizeof(int32_t) * E * E * E);
        float* vfn = (float*)malloc(sizeof(float) * 3 * E * E * E);
#define EI(i, j, k) ((((k)-lo) * E + ((j)-lo)) * E + ((i)-lo))
        for (int attempt = par_tiles ? 1 : 0; attempt < 2; attempt++) {
          const int inner = attempt == 0;
          int far = 0;
          memset(rhoce, 0, sizeof(int32_t) * E * E * E);
          for (int64_t k = 1 - ncb; k <= nt + ncb; k++)
            for (int64_t j = 1 - ncb; j <= nt + ncb; j++)
              for (int64_t i = 1 - ncb; i <= nt + ncb; i++) {
                int64_t r = rank_loc(o, tx, ty, tz, i, j, k);
                for (int d = 0; d < 3; d++) vfn[EI(i, j, k) * 3 + d] = (float)((double)o->vfield[r * 3 + d] * wv);
              }
          for (int il = 0; il < nlayer; il++) {
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1) if (inner) reduction(| : far, bad)
            for (int64_t k = 1 - ncb + il; k <= nt + ncb; k += nlayer)
              for (int64_t j = 1 - ncb; j <= nt + ncb; j++)
                for (int64_t i = 1 - ncb; i <= nt + ncb; i++) {
                  int64_t r = rank_loc(o, tx, ty, tz, i, j, k);
                  int64_t c[3] = {i, j, k};
                  for (int64_t ip = o->cstart[r]; ip < o->cstart[r + 1]; ip++) {
                    int64_t ig[3];
                    double vreal[3];
                    for (int d = 0; d < 3; d++) {
                      double xq = ((double)c[d] - 1.0) + xoff(o->xp[ip * 3 + d], o->izipx);
                      vreal[d] = orc_vdecode(o->vp[ip * 3 + d], sig, o->izipv) + (double)o->vfield[r * 3 + d];
                      double deltax = ((double)dt_mid * vreal[d]) / (double)o->ncell;
                      ig[d] = (int64_t)ceil(xq + deltax);
                      if (ig[d] < lo || ig[d] >= lo + E) bad = 1;
                    }
                    if (ig[0] < lo || ig[0] >= lo + E || ig[1] < lo || ig[1] >= lo + E || ig[2] < lo || ig[2] >= lo + E) continue;
                    if (ig[2] - k > hl || k - ig[2] > hl) far = 1;
                    int64_t e = EI(ig[0], ig[1], ig[2]);
                    rhoce[e]++;
                    for (int d = 0; d < 3; d++) vfn[e * 3 + d] = (float)((double)vfn[e * 3 + d] + vreal[d]);
                  }
                }
          }
          if (!(inner && far)) break;
        }
        for (int64_t e = 0; e < E * E * E; e++)
          for (int d = 0; d < 3; d++) vfn[e * 3 + d] = (float)((double)vfn[e * 3 + d] / ((double)rhoce[e] + wv));
        for (int64_t k = 1; k <= nt; k++)
          for (int64_t j = 1; j <= nt; j++)
            for (int64_t i = 1; i <= nt; i++) {
              int64_t r = rank_loc(o, tx, ty, tz, i, j, k);
              rhoc_new[r] = rhoce[EI(i, j, k)];
              for (int d = 0; d < 3; d++) vf_new[r * 3 + d] = vfn[EI(i, j, k) * 3 + d];
            }
        free(rhoce); free(vfn);
      }
  if (bad) { snprintf(o->err, sizeof o->err, "particle moved beyond the 2*ncb tile buffer"); free(rhoc_new); free(vf_new); return 2; }
  int64_t* cs_new = (int64_t*)malloc(sizeof(int64_t) * (n3 + 1));
  int64_t s = 0;
  for (int64_t r = 0; r < n3; r++) { cs_new[r] = s; s += rhoc_new[r]; }
  cs_new[n3] = s;
  int64_t npn = s;
  int16_t* xpn = (int16_t*)malloc(sizeof(int16_t) * 3 * (npn + 1));
  int16_t* vpn = (int16_t*)malloc(sizeof(int16_t) * 3 * (npn + 1));
  int32_t* pidn = (int32_t*)malloc(sizeof(int32_t) * (npn + 1));
  /* pass 2 (update_particle.f90:99-128) + compaction (:134-149) */
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1) collapse(3) if (par_tiles)
  for (int64_t tz = 0; tz < nnt; tz++)
    for (int64_t ty = 0; ty < nnt; ty++)
      for (int64_t tx = 0; tx < nnt; tx++) {
        if (o->tile_limit > 0 && (tz * nnt + ty) * nnt + tx >= o->tile_limit) continue;
        int32_t* rholocal = (int32_t*)malloc(sizeof(int32_t) * nt * nt * nt);
        for (int attempt = par_tiles ? 1 : 0; attempt < 2; attempt++) {
          const int inner = attempt == 0;
          int far = 0;
          memset(rholocal, 0, sizeof(int32_t) * nt * nt * nt);
          for (int il = 0; il < nlayer; il++) {
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1) if (inner) reduction(| : far)
            for (int64_t k = 1 - ncb + il; k <= nt + ncb; k += nlayer)
              for (int64_t j = 1 - ncb; j <= nt + ncb; j++)
                for (int64_t i = 1 - ncb; i <= nt + ncb; i++) {
                  int64_t r = rank_loc(o, tx, ty, tz, i, j, k);
                  int64_t c[3] = {i, j, k};
                  for (int64_t ip = o->cstart[r]; ip < o->cstart[r + 1]; ip++) {
                    int64_t ig[3];
                    double vreal[3];
                    for (int d = 0; d < 3; d++) {
                      double xq = ((double)c[d] - 1.0) + xoff(o->xp[ip * 3 + d], o->izipx);
                      vreal[d] = orc_vdecode(o->vp[ip * 3 + d], sig, o->izipv) + (double)o->vfield[r * 3 + d];
                      double deltax = ((double)dt_mid * vreal[d]) / (double)o->ncell;
                      ig[d] = (int64_t)ceil(xq + deltax);
                    }
                    if (ig[0] < 1 || ig[0] > nt || ig[1] < 1 || ig[1] > nt || ig[2] < 1 || ig[2] > nt) continue;
                    if (ig[2] - k > hl || k - ig[2] > hl) far = 1;
                    int64_t rl = ((ig[2] - 1) * nt + (ig[1] - 1)) * nt + (ig[0] - 1);
                    int64_t rd = rank_loc(o, tx, ty, tz, ig[0], ig[1], ig[2]);
                    int64_t idx = cs_new[rd] + rholocal[rl]++;
                    for (int d = 0; d < 3; d++) {
                      double dd = (double)dt_mid * vreal[d] / (xres * (double)o->ncell);
                      xpn[idx * 3 + d] = xwrap((int64_t)o->xp[ip * 3 + d] + (int64_t)nint_d(dd), o->izipx);
                      double vr = vreal[d] - (double)vf_new[rd * 3 + d];
                      vpn[idx * 3 + d] = (int16_t)orc_vencode(vr, sig, o->izipv);
                    }
                    pidn[idx] = o->pid[ip];
                  }
                }
          }
          if (!(inner && far)) break;
        }
        free(rholocal);
      }
#undef EI
  free(o->xp); free(o->vp); free(o->pid); free(o->rhoc); free(o->vfield); free(o->cstart);
  o->xp = xpn; o->vp = vpn; o->pid = pidn; o->rhoc = rhoc_new; o->vfield = vf_new; o->cstart = cs_new;
  o->np = npn;
  /* sigma_v statistics (update_particle.f90:173-226) */
  double sum_c = 0, sum_r = 0, sum_s = 0;
  for (int64_t r = 0; r < n3; r++) {
    if (o->tile_limit > 0 && r / (nt * nt * nt) >= o->tile_limit) continue;
    float f0 = o->vfield[r * 3], f1 = o->vfield[r * 3 + 1], f2 = o->vfield[r * 3 + 2];
    sum_c += (double)(f0 * f0 + f1 * f1 + f2 * f2);
    for (int64_t ip = o->cstart[r]; ip < o->cstart[r + 1]; ip++) {
      double v[3];
      for (int d = 0; d < 3; d++) v[d] = orc_vdecode(o->vp[ip * 3 + d], sig, o->izipv);
      sum_r += v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
      for (int d = 0; d < 3; d++) v[d] += (double)o->vfield[r * 3 + d];
      sum_s += v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    }
  }
  double std_vsim = sqrt(sum_s / (double)npn);
  double std_vsim_c = sqrt(sum_c / (double)nc / (double)nc / (double)nc);
  double std_vsim_res = sqrt(sum_r / (double)npn);
  o->sigma_vi_new = (float)(std_vsim_res / (double)sqrtf(3.0f));
  if (out) {
    out[0] = o->sigma_vi_new; out[1] = std_vsim; out[2] = std_vsim_c; out[3] = std_vsim_res;
    out[4] = nlayer; out[5] = (double)npn; out[6] = sum_c; out[7] = sum_r; out[8] = sum_s;
  }
  return 0;
}

/* ------------------------------------------------------------------- PP */
/* ext_pp_force.f90:9-160.  out: [dt_pp, npairs, itest1, f2_max_pp] */
int orc_pp(orc_t* o, float a_mid, float dt, double* out, float* fpp_out, int threads) {
  const int64_t nt = o->nt, nnt = o->nnt, nft = o->nft, ncell = o->ncell;
  const int64_t H = nft + 2 * ncell; /* hoc(1-ncell:nft+ncell)^3 */
  const float sig = o->sigma_vi;
  const float GG = 1.0f / 6.0f / PI_F;
  int64_t npairs_tot = 0, itest1 = 0;
  float ftemp = o->ftemp_pp;
  const int64_t ntiles = nnt * nnt * nnt;
  const int64_t active = o->tile_limit > 0 && o->tile_limit < ntiles ? o->tile_limit : ntiles;
  const int par_tiles = active >= threads; /* else: parallel over the fine planes of one tile */
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1) collapse(3) reduction(+ : npairs_tot, itest1) reduction(max : ftemp) if (par_tiles)
  for (int64_t tz = 0; tz < nnt; tz++)
    for (int64_t ty = 0; ty < nnt; ty++)
      for (int64_t tx = 0; tx < nnt; tx++) {
        if (o->tile_limit > 0 && (tz * nnt + ty) * nnt + tx >= o->tile_limit) continue;
        int64_t nitems = 0;
        for (int64_t k = 0; k <= nt + 1; k++)
          for (int64_t j = 0; j <= nt + 1; j++)
            for (int64_t i = 0; i <= nt + 1; i++) {
              int64_t r = rank_loc(o, tx, ty, tz, i, j, k);
              nitems += o->cstart[r + 1] - o->cstart[r];
            }
        int32_t* hoc = (int32_t*)calloc(H * H * H, sizeof(int32_t));
        int32_t* ll = (int32_t*)calloc(nitems + 1, sizeof(int32_t));
        int64_t* item_ip = (int64_t*)malloc(sizeof(int64_t) * (nitems + 1));
        double* item_x = (double*)malloc(sizeof(double) * 3 * (nitems + 1));
#define HI(i, j, k) ((((k) + ncell - 1) * H + ((j) + ncell - 1)) * H + ((i) + ncell - 1))
        int32_t cnt = 0;
        for (int64_t k = 0; k <= nt + 1; k++)
          for (int64_t j = 0; j <= nt + 1; j++)
            for (int64_t i = 0; i <= nt + 1; i++) {
              int64_t r = rank_loc(o, tx, ty, tz, i, j, k);
              int64_t c[3] = {i, j, k};
              for (int64_t ip = o->cstart[r]; ip < o->cstart[r + 1]; ip++) {
                int32_t id = ++cnt;
                int64_t iv[3];
                for (int d = 0; d < 3; d++) {
                  double x = (double)(ncell * (c[d] - 1)) + (double)ncell * xoff(o->xp[ip * 3 + d], o->izipx);
                  item_x[id * 3 + d] = x;
                  iv[d] = (int64_t)floor(x) + 1;
                }
                item_ip[id] = ip;
                int64_t h = HI(iv[0], iv[1], iv[2]);
                ll[id] = hoc[h];
                hoc[h] = id;
              }
            }
        int64_t np_t = 0, it_t = 0;
        float ft = 0;
        /* each particle's force reads positions only and its kick writes its own velocity */
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1) reduction(+ : np_t, it_t) reduction(max : ft) if (!par_tiles)
        for (int64_t gz = 1; gz <= nft; gz++)
          for (int64_t gy = 1; gy <= nft; gy++)
            for (int64_t gx = 1; gx <= nft; gx++) {
              for (int32_t id1 = hoc[HI(gx, gy, gz)]; id1 != 0; id1 = ll[id1]) {
                it_t++;
                int64_t ip1 = item_ip[id1];
                const double* x1 = item_x + id1 * 3;
                double ftot[3] = {0, 0, 0};
                for (int64_t kk = gz - o->pp_range; kk <= gz + o->pp_range; kk++)
                  for (int64_t jj = gy - o->pp_range; jj <= gy + o->pp_range; jj++)
                    for (int64_t ii = gx - o->pp_range; ii <= gx + o->pp_range; ii++)
                      for (int32_t id2 = hoc[HI(ii, jj, kk)]; id2 != 0; id2 = ll[id2]) {
                        np_t++;
                        const double* x2 = item_x + id2 * 3;
                        double v21[3] = {x2[0] - x1[0], x2[1] - x1[1], x2[2] - x1[2]};
                        double rmag = sqrt(v21[0] * v21[0] + v21[1] * v21[1] + v21[2] * v21[2]);
                        if (rmag == 0) rmag = 1.0;
                        double rcut = rmag / (double)o->nf_cutoff;
                        double pcut = 1 - (1.75 * (rcut * rcut * rcut)) + (0.75 * (rcut * rcut * rcut * rcut * rcut));
                        if (rmag > (double)o->rsoft)
                          for (int d = 0; d < 3; d++)
                            ftot[d] += (double)o->mass_p * (v21[d] / (rmag * rmag * rmag)) * pcut;
                      }
                if (fpp_out)
                  for (int d = 0; d < 3; d++) fpp_out[ip1 * 3 + d] = (float)ftot[d];
                for (int d = 0; d < 3; d++) {
                  double v = orc_vdecode(o->vp[ip1 * 3 + d], sig, o->izipv);
                  v = v + ftot[d] * (double)KA * (double)dt / (double)K6 / (double)KPI;
                  o->vp[ip1 * 3 + d] = (int16_t)orc_vencode(v, sig, o->izipv);
                }
                float f2 = (float)(ftot[0] * ftot[0] + ftot[1] * ftot[1] + ftot[2] * ftot[2]);
                if (f2 > ft) ft = f2;
              }
            }
#undef HI
        npairs_tot += np_t; itest1 += it_t;
        if (ft > ftemp) ftemp = ft;
        free(hoc); free(ll); free(item_ip); free(item_x);
      }
  o->ftemp_pp = ftemp;
  float dt_pp = 1.0f * sqrtf(1.0f) / (sqrtf(ftemp) * a_mid * GG);
  if (out) { out[0] = dt_pp; out[1] = (double)npairs_tot; out[2] = (double)itest1; out[3] = ftemp; }
  if (itest1 != o->np && o->tile_limit == 0) { snprintf(o->err, sizeof o->err, "itest1 %lld != nplocal %lld", (long long)itest1, (long long)o->np); return 3; }
  return 0;
}

/* ------------------------------------------------------------------- PM */
/* pm.f90:2-451.  out: [dt_fine, dt_coarse, dt_vmax, vz_max, vmax1..3,
   testrhof, testrhoc, f2_max_fine, f2_max_coarse].  If probe != NULL the
   step does not kick; it writes per-particle interpolated forces instead:
   probe[0:3np] fine, probe[3np:6np] coarse (sum over taps of F*w1*w2*w3
   evaluated in float per tap, accumulated in double). */
static int pm_impl(orc_t* o, float a_mid, float dt, double* out, double* probe, int threads) {
  const int64_t nt = o->nt, nnt = o->nnt, ncb = o->ncb, nc = o->nc, ncell = o->ncell;
  const int64_t nfe = o->nfe, nfb = o->nfb, nft = o->nft, hf = nfe / 2 + 1;
  const int64_t n3f = nfe * nfe * nfe, nfo = nfe - 2 * nfb + 2; /* force_f(nfb:nfe-nfb+1) */
  const float GG = 1.0f / 6.0f / PI_F;
  const float sig_old = o->sigma_vi, sig_new = o->sigma_vi_new;
  (void)nft;
  double testrhof = 0;
  float f2f = 0;
  const int64_t ntiles = nnt * nnt * nnt;
  const int64_t active = o->tile_limit > 0 && o->tile_limit < ntiles ? o->tile_limit : ntiles;
  int inner = threads / (int)active;
  if (inner < 1) inner = 1;
  int outer = threads / inner;
  if (outer < 1) outer = 1;
  /* neutrino fluid: density at fine-cell centres, global fine force for the
     fluid's gravity (pm.f90:57-72, 157-171) */
  const int64_t nf = nc * ncell;
  const int nu = o->nu_on;
  const int nu_homog = nu && 1.0f / a_mid - 1.0f > o->nu_zi;
  int nu_grav = 0;
  for (int n = 0; nu && n < 3; n++) nu_grav = nu_grav || a_mid >= o->nu_ai[n];
  if (probe) nu_grav = 0;
  float* Dc = NULL;
  float* Fg = NULL;
  if (nu && !nu_homog) {
    Dc = (float*)malloc(sizeof(float) * nf * nf * nf);
    const float* rr[3] = {o->nu_F[0], o->nu_F[1], o->nu_F[2]};
    ofl_density_centres(Dc, (int)nf, 3, rr, o->nu_nc, o->nu_dx, o->nu_w);
  }
  if (nu_grav) Fg = (float*)malloc(sizeof(float) * 3 * nf * nf * nf);
  /* ---- fine PM per tile */
#pragma omp parallel for num_threads(outer) schedule(dynamic, 1) collapse(3) reduction(+ : testrhof) reduction(max : f2f)
  for (int64_t tz = 0; tz < nnt; tz++)
    for (int64_t ty = 0; ty < nnt; ty++)
      for (int64_t tx = 0; tx < nnt; tx++) {
        if (o->tile_limit > 0 && (tz * nnt + ty) * nnt + tx >= o->tile_limit) continue;
        float* rho = (float*)calloc(n3f, sizeof(float));
        float* cre = (float*)malloc(sizeof(float) * hf * nfe * nfe);
        float* cim = (float*)malloc(sizeof(float) * hf * nfe * nfe);
        float* tre = (float*)malloc(sizeof(float) * hf * nfe * nfe);
        float* tim = (float*)malloc(sizeof(float) * hf * nfe * nfe);
        float* force = (float*)malloc(sizeof(float) * 3 * nfo * nfo * nfo);
#define RF(i, j, k) rho[((int64_t)(k)-1) * nfe * nfe + ((int64_t)(j)-1) * nfe + ((i)-1)]
        if (nu) /* node i of the extended tile sits at global fine position origin+i-1-nfb */
          for (int64_t k = 1; k <= nfe; k++)
            for (int64_t j = 1; j <= nfe; j++)
              for (int64_t i = 1; i <= nfe; i++) {
                if (nu_homog) { RF(i, j, k) += o->nu_fneu; continue; }
                const int64_t gx = wrapi(tx * nt * ncell + i - 1 - nfb, nf), gy = wrapi(ty * nt * ncell + j - 1 - nfb, nf),
                              gz = wrapi(tz * nt * ncell + k - 1 - nfb, nf);
                RF(i, j, k) += Dc[(gz * nf + gy) * nf + gx] * o->nu_fneu;
              }
        for (int il = 0; il < 3; il++)
          for (int64_t k = 2 - ncb + il; k <= nt + ncb - 1; k += 3)
            for (int64_t j = 2 - ncb; j <= nt + ncb - 1; j++)
              for (int64_t i = 2 - ncb; i <= nt + ncb - 1; i++) {
                int64_t r = rank_loc(o, tx, ty, tz, i, j, k);
                int64_t c[3] = {i, j, k};
                for (int64_t ip = o->cstart[r]; ip < o->cstart[r + 1]; ip++) {
                  float tempx[3], dx1[3], dx2[3];
                  int64_t i1[3], i2[3];
                  for (int d = 0; d < 3; d++) {
                    tempx[d] = (float)((double)(ncell * (c[d] - 1)) + (double)ncell * xoff(o->xp[ip * 3 + d], o->izipx));
                    i1[d] = (int64_t)floorf(tempx[d]) + 1;
                    i2[d] = i1[d] + 1;
                    dx1[d] = (float)i1[d] - tempx[d];
                    dx2[d] = 1 - dx1[d];
                    i1[d] += nfb; i2[d] += nfb;
                  }
                  float m = o->mass_p;
                  RF(i1[0], i1[1], i1[2]) += dx1[0] * dx1[1] * dx1[2] * m;
                  RF(i2[0], i1[1], i1[2]) += dx2[0] * dx1[1] * dx1[2] * m;
                  RF(i1[0], i2[1], i1[2]) += dx1[0] * dx2[1] * dx1[2] * m;
                  RF(i1[0], i1[1], i2[2]) += dx1[0] * dx1[1] * dx2[2] * m;
                  RF(i1[0], i2[1], i2[2]) += dx1[0] * dx2[1] * dx2[2] * m;
                  RF(i2[0], i1[1], i2[2]) += dx2[0] * dx1[1] * dx2[2] * m;
                  RF(i2[0], i2[1], i1[2]) += dx2[0] * dx2[1] * dx1[2] * m;
                  RF(i2[0], i2[1], i2[2]) += dx2[0] * dx2[1] * dx2[2] * m;
                }
              }
        for (int64_t k = nfb + 1; k <= nfe - nfb; k++)
          for (int64_t j = nfb + 1; j <= nfe - nfb; j++)
            for (int64_t i = nfb + 1; i <= nfe - nfb; i++) testrhof += (double)RF(i, j, k);
        r2c3(rho, cre, cim, nfe, inner);
        for (int d = 0; d < 3; d++) {
          const float* kf = o->kern_f + (int64_t)d * hf * nfe * nfe;
          for (int64_t q = 0; q < hf * nfe * nfe; q++) {
            tre[q] = -cim[q] * kf[q];
            tim[q] = cre[q] * kf[q];
          }
          c2r3(tre, tim, rho, nfe, inner);
          for (int64_t q = 0; q < n3f; q++) rho[q] = rho[q] / (float)nfe / (float)nfe / (float)nfe;
          for (int64_t k = 0; k < nfo; k++)
            for (int64_t j = 0; j < nfo; j++)
              for (int64_t i = 0; i < nfo; i++)
                force[((k * nfo + j) * nfo + i) * 3 + d] = RF(nfb + i, nfb + j, nfb + k);
        }
#undef RF
        if (Fg) /* force index i covers global nodes origin+i-1; keep the physical ones */
          for (int64_t k = 1; k <= nt * ncell; k++)
            for (int64_t j = 1; j <= nt * ncell; j++)
              for (int64_t i = 1; i <= nt * ncell; i++)
                for (int d = 0; d < 3; d++)
                  Fg[(size_t)d * nf * nf * nf + ((tz * nt * ncell + k - 1) * nf + (ty * nt * ncell + j - 1)) * nf +
                     (tx * nt * ncell + i - 1)] = force[((k * nfo + j) * nfo + i) * 3 + d];
        for (int64_t q = 0; q < nfo * nfo * nfo; q++) {
          float f = force[q * 3] * force[q * 3] + force[q * 3 + 1] * force[q * 3 + 1] + force[q * 3 + 2] * force[q * 3 + 2];
          if (f > f2f) f2f = f;
        }
        /* fine kick, pm.f90:131-156 */
#define FF(d, i, j, k) force[((((k)-nfb) * nfo + ((j)-nfb)) * nfo + ((i)-nfb)) * 3 + (d)]
        for (int64_t k = 1; k <= nt; k++)
          for (int64_t j = 1; j <= nt; j++)
            for (int64_t i = 1; i <= nt; i++) {
              int64_t r = rank_loc(o, tx, ty, tz, i, j, k);
              int64_t c[3] = {i, j, k};
              for (int64_t ip = o->cstart[r]; ip < o->cstart[r + 1]; ip++) {
                float tempx[3], dx1[3], dx2[3];
                int64_t a1[3], a2[3];
                for (int d = 0; d < 3; d++) {
                  tempx[d] = (float)((double)(ncell * (c[d] - 1)) + (double)ncell * xoff(o->xp[ip * 3 + d], o->izipx));
                  a1[d] = (int64_t)floorf(tempx[d]) + 1;
                  a2[d] = a1[d] + 1;
                  dx1[d] = (float)a1[d] - tempx[d];
                  dx2[d] = 1 - dx1[d];
                  a1[d] += nfb; a2[d] += nfb;
                }
                for (int d = 0; d < 3; d++) {
                  const float fac = 1.0f; (void)fac;
                  if (probe) {
                    double s = 0;
                    s += (double)(FF(d, a1[0], a1[1], a1[2]) * dx1[0] * dx1[1] * dx1[2]);
                    s += (double)(FF(d, a2[0], a1[1], a1[2]) * dx2[0] * dx1[1] * dx1[2]);
                    s += (double)(FF(d, a1[0], a2[1], a1[2]) * dx1[0] * dx2[1] * dx1[2]);
                    s += (double)(FF(d, a1[0], a1[1], a2[2]) * dx1[0] * dx1[1] * dx2[2]);
                    s += (double)(FF(d, a1[0], a2[1], a2[2]) * dx1[0] * dx2[1] * dx2[2]);
                    s += (double)(FF(d, a2[0], a1[1], a2[2]) * dx2[0] * dx1[1] * dx2[2]);
                    s += (double)(FF(d, a2[0], a2[1], a1[2]) * dx2[0] * dx2[1] * dx1[2]);
                    s += (double)(FF(d, a2[0], a2[1], a2[2]) * dx2[0] * dx2[1] * dx2[2]);
                    probe[ip * 3 + d] = s;
                    continue;
                  }
                  double v = orc_vdecode(o->vp[ip * 3 + d], sig_old, o->izipv);
                  v = v + (double)(FF(d, a1[0], a1[1], a1[2]) * KA * dt / K6 / KPI * dx1[0] * dx1[1] * dx1[2]);
                  v = v + (double)(FF(d, a2[0], a1[1], a1[2]) * KA * dt / K6 / KPI * dx2[0] * dx1[1] * dx1[2]);
                  v = v + (double)(FF(d, a1[0], a2[1], a1[2]) * KA * dt / K6 / KPI * dx1[0] * dx2[1] * dx1[2]);
                  v = v + (double)(FF(d, a1[0], a1[1], a2[2]) * KA * dt / K6 / KPI * dx1[0] * dx1[1] * dx2[2]);
                  v = v + (double)(FF(d, a1[0], a2[1], a2[2]) * KA * dt / K6 / KPI * dx1[0] * dx2[1] * dx2[2]);
                  v = v + (double)(FF(d, a2[0], a1[1], a2[2]) * KA * dt / K6 / KPI * dx2[0] * dx1[1] * dx2[2]);
                  v = v + (double)(FF(d, a2[0], a2[1], a1[2]) * KA * dt / K6 / KPI * dx2[0] * dx2[1] * dx1[2]);
                  v = v + (double)(FF(d, a2[0], a2[1], a2[2]) * KA * dt / K6 / KPI * dx2[0] * dx2[1] * dx2[2]);
                  o->vp[ip * 3 + d] = (int16_t)orc_vencode(v, sig_new, o->izipv);
                }
              }
            }
#undef FF
        free(rho); free(cre); free(cim); free(tre); free(tim); free(force);
      }
  if (Fg) {
    for (int n = 0; n < 3; n++)
      if (a_mid >= o->nu_ai[n]) ofl_gravity_fine(o->nu_F[n], o->nu_nc[n], o->nu_dx[n], Fg, (int)nf, a_mid, dt);
    free(Fg);
  }
  if (!probe) o->sigma_vi = o->sigma_vi_new; /* pm.f90:180 */
  const float sig_c = probe ? sig_new : o->sigma_vi;

  /* ---- coarse PM, pm.f90:188-372.  The coarse density always covers the
     whole box (a tile limit restricts only the kicks below), so the coarse
     force of a limited sample is the full run's; its time is box-wide
     work, counted in t_global. */
  const double t_glob0 = omp_get_wtime();
  const int64_t n3c = nc * nc * nc, hc = nc / 2 + 1, T = nt + 4;
  float* r3 = (float*)calloc(n3c, sizeof(float));
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1) collapse(3)
  for (int64_t tz = 0; tz < nnt; tz++)
    for (int64_t ty = 0; ty < nnt; ty++)
      for (int64_t tx = 0; tx < nnt; tx++) {
        float* r3t = (float*)calloc(T * T * T, sizeof(float)); /* r3t(-1:nt+2)^3 */
#define RT(i, j, k) r3t[(((k) + 1) * T + ((j) + 1)) * T + ((i) + 1)]
        for (int il = 0; il < 3; il++)
          for (int64_t k = 0 + il; k <= nt + 1; k += 3)
            for (int64_t j = 0; j <= nt + 1; j++)
              for (int64_t i = 0; i <= nt + 1; i++) {
                int64_t r = rank_loc(o, tx, ty, tz, i, j, k);
                int64_t c[3] = {i, j, k};
                for (int64_t ip = o->cstart[r]; ip < o->cstart[r + 1]; ip++) {
                  float tempx[3], dx1[3], dx2[3];
                  int64_t i1[3], i2[3];
                  for (int d = 0; d < 3; d++) {
                    tempx[d] = (float)(((double)(c[d] - 1) + xoff(o->xp[ip * 3 + d], o->izipx)) - 0.5);
                    i1[d] = (int64_t)floorf(tempx[d]) + 1;
                    i2[d] = i1[d] + 1;
                    dx1[d] = (float)i1[d] - tempx[d];
                    dx2[d] = 1 - dx1[d];
                  }
                  float m = o->mass_p;
                  RT(i1[0], i1[1], i1[2]) += dx1[0] * dx1[1] * dx1[2] * m;
                  RT(i2[0], i1[1], i1[2]) += dx2[0] * dx1[1] * dx1[2] * m;
                  RT(i1[0], i2[1], i1[2]) += dx1[0] * dx2[1] * dx1[2] * m;
                  RT(i1[0], i1[1], i2[2]) += dx1[0] * dx1[1] * dx2[2] * m;
                  RT(i1[0], i2[1], i2[2]) += dx1[0] * dx2[1] * dx2[2] * m;
                  RT(i2[0], i1[1], i2[2]) += dx2[0] * dx1[1] * dx2[2] * m;
                  RT(i2[0], i2[1], i1[2]) += dx2[0] * dx2[1] * dx1[2] * m;
                  RT(i2[0], i2[1], i2[2]) += dx2[0] * dx2[1] * dx2[2] * m;
                }
              }
        for (int64_t k = 1; k <= nt; k++)
          for (int64_t j = 1; j <= nt; j++)
            for (int64_t i = 1; i <= nt; i++)
              r3[((tz * nt + k - 1) * nc + (ty * nt + j - 1)) * nc + (tx * nt + i - 1)] = RT(i, j, k);
#undef RT
        free(r3t);
      }
  if (nu) { /* pm.f90:198-242: fine-cell-centre density CIC onto coarse nodes */
    for (int64_t q = 0; nu_homog && q < n3c; q++) r3[q] += o->nu_fneu * (float)(ncell * ncell * ncell);
    if (!nu_homog)
      for (int64_t mz = 0; mz < nf; mz++)
        for (int64_t my = 0; my < nf; my++)
          for (int64_t mx = 0; mx < nf; mx++) {
            const float rnu = Dc[(mz * nf + my) * nf + mx] * o->nu_fneu;
            const int64_t m[3] = {mx, my, mz};
            float d1[3], d2[3];
            int64_t k1[3], k2[3];
            for (int d = 0; d < 3; d++) {
              const float xc = ((float)m[d] + 0.5f) / (float)ncell - 0.5f;
              const int64_t f = (int64_t)floorf(xc);
              d1[d] = (float)(f + 1) - xc;
              d2[d] = 1.0f - d1[d];
              k1[d] = wrapi(f, nc);
              k2[d] = wrapi(f + 1, nc);
            }
            for (int t = 0; t < 8; t++) {
              const int ox = t & 1, oy = (t >> 1) & 1, oz = t >> 2;
              r3[((oz ? k2[2] : k1[2]) * nc + (oy ? k2[1] : k1[1])) * nc + (ox ? k2[0] : k1[0])] +=
                  rnu * (ox ? d2[0] : d1[0]) * (oy ? d2[1] : d1[1]) * (oz ? d2[2] : d1[2]);
            }
          }
  }
  free(Dc);
  double testrhoc = 0;
  for (int64_t q = 0; q < n3c; q++) testrhoc += (double)r3[q];
  float* cre = (float*)malloc(sizeof(float) * hc * nc * nc);
  float* cim = (float*)malloc(sizeof(float) * hc * nc * nc);
  float* tre = (float*)malloc(sizeof(float) * hc * nc * nc);
  float* tim = (float*)malloc(sizeof(float) * hc * nc * nc);
  float* fc = (float*)malloc(sizeof(float) * 3 * n3c); /* force_c(d, x,y,z) physical */
  r2c3(r3, cre, cim, nc, threads);
  for (int d = 0; d < 3; d++) {
    const float* kc = o->kern_c + (int64_t)d * hc * nc * nc;
    for (int64_t q = 0; q < hc * nc * nc; q++) { tre[q] = -cim[q] * kc[q]; tim[q] = cre[q] * kc[q]; }
    c2r3(tre, tim, r3, nc, threads);
    for (int64_t q = 0; q < n3c; q++) fc[q * 3 + d] = r3[q] / (float)nc / (float)nc / (float)nc;
  }
  float f2c = 0;
  for (int64_t q = 0; q < n3c; q++) {
    float f = fc[q * 3] * fc[q * 3] + fc[q * 3 + 1] * fc[q * 3 + 1] + fc[q * 3 + 2] * fc[q * 3 + 2];
    if (f > f2c) f2c = f;
  }
  o->t_global = omp_get_wtime() - t_glob0;
  float vmax[3] = {0, 0, 0};
#define FC(d, x, y, z) fc[((wrapi((z)-1, nc) * nc + wrapi((y)-1, nc)) * nc + wrapi((x)-1, nc)) * 3 + (d)]
  float vm0 = 0, vm1 = 0, vm2 = 0;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 64) reduction(max : vm0, vm1, vm2)
  for (int64_t r = 0; r < n3c; r++) {
    int64_t t = r / (nt * nt * nt), l = r % (nt * nt * nt);
    if (o->tile_limit > 0 && t >= o->tile_limit) continue;
    int64_t txyz[3] = {t % nnt, (t / nnt) % nnt, t / (nnt * nnt)};
    int64_t c[3] = {l % nt + 1, (l / nt) % nt + 1, l / (nt * nt) + 1};
    for (int64_t ip = o->cstart[r]; ip < o->cstart[r + 1]; ip++) {
      float tempx[3], dx1[3], dx2[3];
      int64_t a1[3], a2[3];
      for (int d = 0; d < 3; d++) {
        tempx[d] = (float)(((double)(txyz[d] * nt + c[d] - 1) + xoff(o->xp[ip * 3 + d], o->izipx)) - 0.5);
        a1[d] = (int64_t)floorf(tempx[d]) + 1;
        a2[d] = a1[d] + 1;
        dx1[d] = (float)a1[d] - tempx[d];
        dx2[d] = 1 - dx1[d];
      }
      for (int d = 0; d < 3; d++) {
        if (probe) {
          double s = 0;
          s += (double)(FC(d, a1[0], a1[1], a1[2]) * dx1[0] * dx1[1] * dx1[2]);
          s += (double)(FC(d, a2[0], a1[1], a1[2]) * dx2[0] * dx1[1] * dx1[2]);
          s += (double)(FC(d, a1[0], a2[1], a1[2]) * dx1[0] * dx2[1] * dx1[2]);
          s += (double)(FC(d, a1[0], a1[1], a2[2]) * dx1[0] * dx1[1] * dx2[2]);
          s += (double)(FC(d, a1[0], a2[1], a2[2]) * dx1[0] * dx2[1] * dx2[2]);
          s += (double)(FC(d, a2[0], a1[1], a2[2]) * dx2[0] * dx1[1] * dx2[2]);
          s += (double)(FC(d, a2[0], a2[1], a1[2]) * dx2[0] * dx2[1] * dx1[2]);
          s += (double)(FC(d, a2[0], a2[1], a2[2]) * dx2[0] * dx2[1] * dx2[2]);
          probe[3 * o->np + ip * 3 + d] = s;
          continue;
        }
        double v = orc_vdecode(o->vp[ip * 3 + d], sig_c, o->izipv);
        v = v + (double)(FC(d, a1[0], a1[1], a1[2]) * KA * dt / K6 / KPI * dx1[0] * dx1[1] * dx1[2]);
        v = v + (double)(FC(d, a2[0], a1[1], a1[2]) * KA * dt / K6 / KPI * dx2[0] * dx1[1] * dx1[2]);
        v = v + (double)(FC(d, a1[0], a2[1], a1[2]) * KA * dt / K6 / KPI * dx1[0] * dx2[1] * dx1[2]);
        v = v + (double)(FC(d, a1[0], a1[1], a2[2]) * KA * dt / K6 / KPI * dx1[0] * dx1[1] * dx2[2]);
        v = v + (double)(FC(d, a1[0], a2[1], a2[2]) * KA * dt / K6 / KPI * dx1[0] * dx2[1] * dx2[2]);
        v = v + (double)(FC(d, a2[0], a1[1], a2[2]) * KA * dt / K6 / KPI * dx2[0] * dx1[1] * dx2[2]);
        v = v + (double)(FC(d, a2[0], a2[1], a1[2]) * KA * dt / K6 / KPI * dx2[0] * dx2[1] * dx1[2]);
        v = v + (double)(FC(d, a2[0], a2[1], a2[2]) * KA * dt / K6 / KPI * dx2[0] * dx2[1] * dx2[2]);
        float av = (float)fabs(v + (double)o->vfield[r * 3 + d]);
        if (d == 0 && av > vm0) vm0 = av;
        if (d == 1 && av > vm1) vm1 = av;
        if (d == 2 && av > vm2) vm2 = av;
        o->vp[ip * 3 + d] = (int16_t)orc_vencode(v, sig_c, o->izipv);
      }
    }
  }
#undef FC
  vmax[0] = vm0; vmax[1] = vm1; vmax[2] = vm2;
  if (nu_grav) { /* pm.f90:374-422 */
    float* fcs = (float*)malloc(sizeof(float) * 3 * n3c);
    for (int64_t q = 0; q < n3c; q++)
      for (int d = 0; d < 3; d++) fcs[(size_t)d * n3c + q] = fc[q * 3 + d];
    for (int n = 0; n < 3; n++)
      if (a_mid >= o->nu_ai[n]) ofl_gravity_coarse(o->nu_F[n], o->nu_nc[n], o->nu_dx[n], fcs, (int)nc, (int)ncell, a_mid, dt);
    free(fcs);
  }
  free(r3); free(cre); free(cim); free(tre); free(tim); free(fc);
  float dt_fine = sqrtf(1.0f / (sqrtf(f2f) * a_mid * GG));
  float dt_coarse = sqrtf((float)ncell / (sqrtf(f2c) * a_mid * GG));
  float vmx = vmax[0] > vmax[1] ? vmax[0] : vmax[1];
  if (vmax[2] > vmx) vmx = vmax[2];
  float dt_vmax = 0.9f * 20 / vmx;
  if (!probe) o->vz_max = vmax[2];
  if (out) {
    out[0] = dt_fine; out[1] = dt_coarse; out[2] = dt_vmax; out[3] = vmax[2];
    out[4] = vmax[0]; out[5] = vmax[1]; out[6] = vmax[2]; out[7] = testrhof; out[8] = testrhoc;
    out[9] = f2f; out[10] = f2c;
  }
  return 0;
}
int orc_pm(orc_t* o, float a_mid, float dt, double* out, int threads) { return pm_impl(o, a_mid, dt, out, NULL, threads); }
int orc_forces(orc_t* o, double* probe, double* out, int threads) { return pm_impl(o, 1.0f, 1.0f, out, probe, threads); }

/* -------------------------------------------------------------- timestep */
/* timestep.f90:136-180 (the default, non-'new_expansion' branch):
   real(4) interface, real(8) internals.  Default-real literals and
   parameters are real(4): omHsq = 4.0/9.0 is the float 0.44444445,
   1.0-omega_m-omega_l and the wde factors are float expressions, dt_x**2
   and dt_x**3 are float powers of the real(4) dt_x, and a_x = a0+da1 adds
   two real(4) values before widening.  Integer powers of the real(8) a_x
   are products ((a*a)*a); a_x**(-3*wde) is pow(). */
static void expansion_half(double a_x, float dt_x, float om, float ol, float wde, float* da) {
  double omHsq = (double)(4.0f / 9.0f);
  double a3rlm = pow(a_x, (double)(-(3.0f * wde))) * (double)ol / (double)om;
  double arkm = a_x * (double)(1.0f - om - ol) / (double)om;
  double adot = sqrt(omHsq * (a_x * a_x * a_x) * (1.0 + arkm + a3rlm));
  double addot = (a_x * a_x) * omHsq * (1.5 + 2.0 * arkm + (double)(1.5f * (1.0f - wde)) * a3rlm);
  double atdot = a_x * adot * omHsq *
                 (3.0 + 6.0 * arkm + (double)(1.5f * (2.0f - 3.0f * wde) * (1.0f - wde)) * a3rlm);
  float p2 = dt_x * dt_x, p3 = dt_x * dt_x * dt_x;
  *da = (float)(adot * (double)dt_x + (addot * (double)p2) / 2.0 + (atdot * (double)p3) / 6.0);
}

void orc_expansion(float a0, float dt0, float omega_m, float omega_l, float wde, float* da1, float* da2) {
  float dt_x = dt0 / 2;
  expansion_half((double)a0, dt_x, omega_m, omega_l, wde, da1);
  float a1 = a0 + *da1;
  expansion_half((double)a1, dt_x, omega_m, omega_l, wde, da2);
}

/* ------------------------------------------------------- test exports */
/* R2C of a real n^3 array (x fastest) into (n/2+1, n, n) re/im. */
void orc_fft_r2c(const float* r, int64_t n, float* re, float* im, int threads) { r2c3(r, re, im, n, threads); }

/* Fine-mesh CIC density of the current state assembled from the per-tile
   physical regions of pm.f90:76-108 into one nf^3 array (x fastest). */
int orc_density(orc_t* o, float* rho_out, int threads) {
  const int64_t nt = o->nt, nnt = o->nnt, ncb = o->ncb, ncell = o->ncell, nfe = o->nfe, nfb = o->nfb;
  const int64_t nf = o->nf, nft = o->nft;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1) collapse(3)
  for (int64_t tz = 0; tz < nnt; tz++)
    for (int64_t ty = 0; ty < nnt; ty++)
      for (int64_t tx = 0; tx < nnt; tx++) {
        if (o->tile_limit > 0 && (tz * nnt + ty) * nnt + tx >= o->tile_limit) continue;
        float* rho = (float*)calloc(nfe * nfe * nfe, sizeof(float));
#define RF(i, j, k) rho[((int64_t)(k)-1) * nfe * nfe + ((int64_t)(j)-1) * nfe + ((i)-1)]
        for (int il = 0; il < 3; il++)
          for (int64_t k = 2 - ncb + il; k <= nt + ncb - 1; k += 3)
            for (int64_t j = 2 - ncb; j <= nt + ncb - 1; j++)
              for (int64_t i = 2 - ncb; i <= nt + ncb - 1; i++) {
                int64_t r = rank_loc(o, tx, ty, tz, i, j, k);
                int64_t c[3] = {i, j, k};
                for (int64_t ip = o->cstart[r]; ip < o->cstart[r + 1]; ip++) {
                  float tempx[3], dx1[3], dx2[3];
                  int64_t i1[3], i2[3];
                  for (int d = 0; d < 3; d++) {
                    tempx[d] = (float)((double)(ncell * (c[d] - 1)) + (double)ncell * xoff(o->xp[ip * 3 + d], o->izipx));
                    i1[d] = (int64_t)floorf(tempx[d]) + 1;
                    i2[d] = i1[d] + 1;
                    dx1[d] = (float)i1[d] - tempx[d];
                    dx2[d] = 1 - dx1[d];
                    i1[d] += nfb; i2[d] += nfb;
                  }
                  float m = o->mass_p;
                  RF(i1[0], i1[1], i1[2]) += dx1[0] * dx1[1] * dx1[2] * m;
                  RF(i2[0], i1[1], i1[2]) += dx2[0] * dx1[1] * dx1[2] * m;
                  RF(i1[0], i2[1], i1[2]) += dx1[0] * dx2[1] * dx1[2] * m;
                  RF(i1[0], i1[1], i2[2]) += dx1[0] * dx1[1] * dx2[2] * m;
                  RF(i1[0], i2[1], i2[2]) += dx1[0] * dx2[1] * dx2[2] * m;
                  RF(i2[0], i1[1], i2[2]) += dx2[0] * dx1[1] * dx2[2] * m;
                  RF(i2[0], i2[1], i1[2]) += dx2[0] * dx2[1] * dx1[2] * m;
                  RF(i2[0], i2[1], i2[2]) += dx2[0] * dx2[1] * dx2[2] * m;
                }
              }
        for (int64_t k = 1; k <= nft; k++)
          for (int64_t j = 1; j <= nft; j++)
            for (int64_t i = 1; i <= nft; i++)
              rho_out[((tz * nft + k - 1) * nf + (ty * nft + j - 1)) * nf + (tx * nft + i - 1)] = RF(nfb + i, nfb + j, nfb + k);
#undef RF
        free(rho);
      }
  return 0;
}
