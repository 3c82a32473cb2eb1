// Neutrino fluid: the dispersive sum of three isothermal-start fluids at the
// Gauss-Laguerre velocities (hydrodf.f90), each a finite-volume fluid
// (hydrogr.f90) updated by dimensionally split MUSCL-minmod/HLL sweeps
// (hydro1d.f90:113-225, hydro3d.f90:33-108), coupled to the PM meshes
// through its density (pm.f90:55-74, 198-242) and the PM forces
// (pm.f90:157-171, 374-422).
//
// One box, so every fluid is a periodic nc^3 array with no halo cells: the
// reference's halo shells (ncb >= 2 cells) only carry periodic copies for
// the stencil, and a periodic line sweep gives the same physical cells.
// Fields are structure-of-arrays [5][nc^3] (rho, mx, my, mz, E), x fastest.
//
// Gravity: the reference adds each fine cell's impulse to the eight fluid
// cells around it one after another, the energy term using the momentum
// already kicked.  With x_k = dV_k a_k summed over all contributions k of a
// fluid cell, that sequence equals, for any order,
//   m += rho J,   E += m0.J + rho/2 (|J|^2 + Q),
//   J = sum dV_k a_k,   Q = sum dV_k (1 - dV_k) |a_k|^2,
// so each fluid cell gathers J and Q from its stencil and is updated once:
// no atomics and the same result on every run.
#include "cube_internal.hpp"
#include <math.h>
#include <string.h>
#include <vector>

#define FL_NT 256
#define FL_W 4  // lines per workgroup in a sweep
#define H_CFL 0.71f
#define H_LIM 1.5f

// ------------------------------------------------------------- 1-D pieces
__device__ __forceinline__ float sgn1(float a) { return copysignf(1.0f, a); }
__device__ __forceinline__ float minmod(float a, float b, float c) {
  const float m = fminf(fminf(fabsf(a), fabsf(b)), fabsf(c));
  return 0.25f * fabsf(sgn1(a) + sgn1(b)) * (sgn1(a) + sgn1(c)) * m;
}
__device__ __forceinline__ float kin(float r, float m1, float m2, float m3) {
  return 0.5f * ((m1 * m1 + m2 * m2) + m3 * m3) / r;
}
__device__ __forceinline__ float press(float r, float m1, float m2, float m3, float e, float g) {
  const float p = (e - kin(r, m1, m2, m3)) * (g - 1.0f);
  return p > 0.0f ? p : 0.0f;
}

// ------------------------------------------------------------------ sweep
// One sweep along axis AX over FL_W adjacent lines.  The normal momentum is
// the sweep's first momentum component and the two others follow in the
// reference's swapped order (x: x,y,z; y: y,x,z; z: z,y,x).  LDS holds the
// primitives (rho, v1, v2, v3, p) of the lines, then their interface fluxes.
template <int AX>
__global__ __launch_bounds__(FL_NT) void k_fl_sweep(float* F, int nc, float g, float dt, float dx) {
  extern __shared__ float sm[];
  const int64_t n3 = (int64_t)nc * nc * nc;
  const int nla = nc / FL_W;
  // XCD-contiguous order: the line groups adjacent in x that share cache
  // lines (strided sweeps) run on the same XCD's L2
  const uint32_t lb = xcd_remap(blockIdx.x, gridDim.x);
  const int a0 = (int)(lb % nla) * FL_W, b = (int)(lb / nla);
  const int c1 = AX == 0 ? 1 : (AX == 1 ? 2 : 3), c2 = AX == 1 ? 1 : 2, c3 = AX == 2 ? 1 : 3;
  const int L = FL_W * nc;
  auto elem = [&](int e, int& t, int& w) {
    if (AX == 0) { t = e % nc; w = e / nc; }
    else { w = e % FL_W; t = e / FL_W; }
  };
  auto cell = [&](int t, int w) -> int64_t {
    const int a = a0 + w;
    if (AX == 0) return ((int64_t)b * nc + a) * nc + t;
    if (AX == 1) return ((int64_t)b * nc + t) * nc + a;
    return ((int64_t)t * nc + b) * nc + a;
  };
  auto S = [&](int q, int w, int t) -> float& { return sm[(q * FL_W + w) * nc + t]; };
  // pass 1: primitives into LDS
  for (int e = threadIdx.x; e < L; e += FL_NT) {
    int t, w;
    elem(e, t, w);
    const int64_t c = cell(t, w);
    const float r = F[c], m1 = F[(int64_t)c1 * n3 + c], m2 = F[(int64_t)c2 * n3 + c], m3 = F[(int64_t)c3 * n3 + c],
                en = F[4 * n3 + c];
    S(0, w, t) = r;
    S(1, w, t) = m1 / r;
    S(2, w, t) = m2 / r;
    S(3, w, t) = m3 / r;
    S(4, w, t) = press(r, m1, m2, m3, en, g);
  }
  __syncthreads();
  // pass 2: HLL flux at the interface t+1/2 of each element, held in registers
  constexpr int PERMAX = 8;  // L / FL_NT for nc <= 512
  float fx[PERMAX][5];
#pragma unroll
  for (int k = 0; k < PERMAX; k++) {
    const int e = threadIdx.x + k * FL_NT;
    if (e >= L) continue;
    int t, w;
    elem(e, t, w);
    const int tm = t == 0 ? nc - 1 : t - 1, tp = t + 1 == nc ? 0 : t + 1, tq = tp + 1 == nc ? 0 : tp + 1;
    float lb[5], rb[5];
    for (int q = 0; q < 5; q++) {
      const float u0 = S(q, w, tm), u1 = S(q, w, t), u2 = S(q, w, tp), u3 = S(q, w, tq);
      lb[q] = u1 + 0.5f * minmod(H_LIM * (u1 - u0), 0.5f * (u2 - u0), H_LIM * (u2 - u1));
      rb[q] = u2 - 0.5f * minmod(H_LIM * (u2 - u1), 0.5f * (u3 - u1), H_LIM * (u3 - u2));
    }
    float hl[5], hr[5];
    hl[0] = lb[0]; hl[1] = hl[0] * lb[1]; hl[2] = hl[0] * lb[2]; hl[3] = hl[0] * lb[3];
    hl[4] = lb[4] / (g - 1.0f) + kin(hl[0], hl[1], hl[2], hl[3]);
    hr[0] = rb[0]; hr[1] = hr[0] * rb[1]; hr[2] = hr[0] * rb[2]; hr[3] = hr[0] * rb[3];
    hr[4] = rb[4] / (g - 1.0f) + kin(hr[0], hr[1], hr[2], hr[3]);
    const float pl = press(hl[0], hl[1], hl[2], hl[3], hl[4], g), vl = hl[1] / hl[0];
    const float pr = press(hr[0], hr[1], hr[2], hr[3], hr[4], g), vr = hr[1] / hr[0];
    const float fl[5] = {hl[1], hl[1] * vl + pl, hl[1] * (hl[2] / hl[0]), hl[1] * (hl[3] / hl[0]), hl[4] * vl};
    const float fr[5] = {hr[1], hr[1] * vr + pr, hr[1] * (hr[2] / hr[0]), hr[1] * (hr[3] / hr[0]), hr[4] * vr};
    const float sl = sqrtf(pl / hl[0]), sr = sqrtf(pr / hr[0]);
    const float ap = fmaxf(fmaxf(0.0f, vl + sl), vr + sr);
    const float am = fmaxf(fmaxf(0.0f, -vl + sl), -vr + sr);
    for (int q = 0; q < 5; q++) fx[k][q] = (ap * fl[q] + am * fr[q] - ap * am * (hr[q] - hl[q])) / (ap + am);
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < PERMAX; k++) {
    const int e = threadIdx.x + k * FL_NT;
    if (e >= L) continue;
    int t, w;
    elem(e, t, w);
    for (int q = 0; q < 5; q++) S(q, w, t) = fx[k][q];
  }
  __syncthreads();
  // pass 3: h += dt*(F(t-1/2) - F(t+1/2))/dx in the sweep's variable order
  for (int e = threadIdx.x; e < L; e += FL_NT) {
    int t, w;
    elem(e, t, w);
    const int tm = t == 0 ? nc - 1 : t - 1;
    const int64_t c = cell(t, w);
    const int64_t off[5] = {0, (int64_t)c1 * n3, (int64_t)c2 * n3, (int64_t)c3 * n3, 4 * n3};
    for (int q = 0; q < 5; q++) {
      const float rq = -S(q, w, t) + S(q, w, tm);
      F[off[q] + c] = F[off[q] + c] + dt * rq / dx;
    }
  }
}

// ---------------------------------------------------------------- timestep
// h3_timestep: per x-line, cfl*dx/(max|v| + max c_s); the minimum over lines
// (positive floats order as their bit patterns).
__global__ __launch_bounds__(FL_NT) void k_fl_dt(const float* F, int nc, float g, float dx, uint32_t* dtbits) {
  __shared__ float rv[FL_NT / 64], rs[FL_NT / 64];
  const int64_t n3 = (int64_t)nc * nc * nc, line = blockIdx.x;
  float v = 0.0f, s = 0.0f;
  for (int i = threadIdx.x; i < nc; i += FL_NT) {
    const int64_t c = line * nc + i;
    const float r = F[c], m1 = F[n3 + c], m2 = F[2 * n3 + c], m3 = F[3 * n3 + c], en = F[4 * n3 + c];
    v = fmaxf(v, fmaxf(fmaxf(fabsf(m1 / r), fabsf(m2 / r)), fabsf(m3 / r)));
    s = fmaxf(s, sqrtf(press(r, m1, m2, m3, en, g) / r));
  }
  for (int o = 32; o > 0; o >>= 1) {
    v = fmaxf(v, __shfl_xor(v, o));
    s = fmaxf(s, __shfl_xor(s, o));
  }
  if ((threadIdx.x & 63) == 0) { rv[threadIdx.x >> 6] = v; rs[threadIdx.x >> 6] = s; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < FL_NT / 64; i++) { v = fmaxf(v, rv[i]); s = fmaxf(s, rs[i]); }
    const float dtn = H_CFL * dx / (v + s);
    atomicMin(dtbits, __float_as_uint(dtn));
  }
}

// ---------------------------------------------------------------- density
struct FlSet {  // the three fluids of the dispersive sum
  const float* rho[3];
  int nc[3];
  float dx[3], w[3];
  int n;
};
// hg_density (CIC) of one fluid at x (fine units)
__device__ __forceinline__ float fl_density_at(const float* rho, int nc, float dx, const float x[3]) {
  float d1[3], d2[3];
  int i1[3], i2[3];
  for (int d = 0; d < 3; d++) {
    const float xx = x[d] / dx - 0.5f;
    const int f = (int)floorf(xx);
    d1[d] = (float)(f + 1) - xx;
    d2[d] = 1.0f - d1[d];
    i1[d] = f < 0 ? f + nc : f;
    i2[d] = f + 1 >= nc ? f + 1 - nc : f + 1;
  }
#define RH(a, b, c) rho[((int64_t)(c) * nc + (b)) * nc + (a)]
  return RH(i1[0], i1[1], i1[2]) * d1[0] * d1[1] * d1[2] + RH(i2[0], i1[1], i1[2]) * d2[0] * d1[1] * d1[2] +
         RH(i1[0], i2[1], i1[2]) * d1[0] * d2[1] * d1[2] + RH(i2[0], i2[1], i1[2]) * d2[0] * d2[1] * d1[2] +
         RH(i1[0], i1[1], i2[2]) * d1[0] * d1[1] * d2[2] + RH(i2[0], i1[1], i2[2]) * d2[0] * d1[1] * d2[2] +
         RH(i1[0], i2[1], i2[2]) * d1[0] * d2[1] * d2[2] + RH(i2[0], i2[1], i2[2]) * d2[0] * d2[1] * d2[2];
#undef RH
}
// dfld_density at every fine-cell centre n+0.5 into D (rows of pitch nf+2)
__global__ void k_fl_density_centres(FlSet s, int64_t nf, float* D) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nf * nf * nf) return;
  const int64_t x = i % nf, y = (i / nf) % nf, z = i / (nf * nf);
  const float p[3] = {(float)x + 0.5f, (float)y + 0.5f, (float)z + 0.5f};
  float d = 0.0f;
  for (int f = 0; f < s.n; f++) d = d + s.w[f] * fl_density_at(s.rho[f], s.nc[f], s.dx[f], p);
  D[(z * nf + y) * (nf + 2) + x] = d;
}
// fine mesh (pm.f90:57-72): node n gets f_nu * density(n + 0.5), or the
// homogeneous f_nu before the fluid starts.  Per-block sums of what was
// added go to part[blockIdx.x]: the reference's testrhof sums rho_f after
// the fluid is in (pm.f90:107).
template <typename T>
__global__ __launch_bounds__(256) void k_fl_add_fine(T* R, const float* D, int64_t nf, float fnu, int homog, double* part) {
  __shared__ double red[4];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double added = 0;
  if (i < nf * nf * nf) {
    const int64_t x = i % nf, y = (i / nf) % nf, z = i / (nf * nf), q = (z * nf + y) * (nf + 2) + x;
    const T v = (T)(homog ? fnu : D[q] * fnu);
    R[q] = R[q] + v;
    added = (double)v;
  }
  for (int o = 32; o > 0; o >>= 1) added += __shfl_xor(added, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = added;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}
// coarse mesh (pm.f90:198-242): every fine-cell centre's f_nu*density CIC
// onto coarse nodes at cell centres, gathered per node (2*ncell taps per
// dimension, the reference's dx1/dx2 weights).  Per-block sums of what was
// added go to part[blockIdx.x] (testrhoc sums r3 with the fluid in,
// pm.f90:280).
template <typename T>
__device__ __forceinline__ double fl_coarse_node(T* Rc, const float* D, int64_t nf, int ncc, float fnu, int homog, int64_t i);
template <typename T>
__global__ __launch_bounds__(256) void k_fl_add_coarse(T* Rc, const float* D, int64_t nf, int ncc, float fnu, int homog,
                                                       double* part) {
  __shared__ double red[4];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double added = i < (int64_t)ncc * ncc * ncc ? fl_coarse_node(Rc, D, nf, ncc, fnu, homog, i) : 0.0;
  for (int o = 32; o > 0; o >>= 1) added += __shfl_xor(added, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = added;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}
template <typename T>
__device__ __forceinline__ double fl_coarse_node(T* Rc, const float* D, int64_t nf, int ncc, float fnu, int homog, int64_t i) {
  const int cx = (int)(i % ncc), cy = (int)((i / ncc) % ncc), cz = (int)(i / ((int64_t)ncc * ncc));
  const int64_t q = ((int64_t)cz * ncc + cy) * (ncc + 2) + cx;
  if (homog) {
    const T v = (T)(fnu * (float)(CUBE_NCELL * CUBE_NCELL * CUBE_NCELL));
    Rc[q] = Rc[q] + v;
    return (double)v;
  }
  double acc = 0;
  const int c[3] = {cx, cy, cz};
  int m0[3];
  float wt[3][2 * CUBE_NCELL];
  for (int d = 0; d < 3; d++) {
    m0[d] = c[d] * CUBE_NCELL - CUBE_NCELL / 2;  // first fine cell whose centre reaches node c
    for (int j = 0; j < 2 * CUBE_NCELL; j++) {
      const int m = m0[d] + j;  // 0-based fine cell, centre m+0.5
      const float xc = ((float)m + 0.5f) / (float)CUBE_NCELL - 0.5f;
      const int f = (int)floorf(xc);
      const float d1 = (float)(f + 1) - xc, d2 = 1.0f - d1;
      wt[d][j] = f == c[d] ? d1 : (f + 1 == c[d] ? d2 : 0.0f);
    }
  }
  for (int jz = 0; jz < 2 * CUBE_NCELL; jz++)
    for (int jy = 0; jy < 2 * CUBE_NCELL; jy++)
      for (int jx = 0; jx < 2 * CUBE_NCELL; jx++) {
        const int64_t mx = (m0[0] + jx + nf) % nf, my = (m0[1] + jy + nf) % nf, mz = (m0[2] + jz + nf) % nf;
        const float rnu = D[(mz * nf + my) * (nf + 2) + mx] * fnu;
        acc += (double)(rnu * wt[0][jx] * wt[1][jy] * wt[2][jz]);
      }
  Rc[q] = Rc[q] + (T)acc;
  return (double)(T)acc;
}

// ---------------------------------------------------------------- gravity
// Fill A_d (row pitch p_d) with component d of the coarse force
// CIC-interpolated at every fine-cell centre, times a_mid*dt/6/pi
// (pm.f90:387-404).  Fc holds the three components [3][ncc^3].
template <typename T>
__global__ void k_fl_coarse_at_centres(const T* Fc, int ncc, int64_t nf, float a_mid, float dt, float* A0, float* A1, float* A2,
                                       int64_t p0, int64_t p1, int64_t p2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nf * nf * nf) return;
  const int64_t x = i % nf, y = (i / nf) % nf, z = i / (nf * nf);
  const float p[3] = {(float)x + 0.5f, (float)y + 0.5f, (float)z + 0.5f};
  float c1[3], c2[3];
  int k1[3], k2[3];
  for (int d = 0; d < 3; d++) {
    const float xc = p[d] / (float)CUBE_NCELL - 0.5f;
    const int f = (int)floorf(xc);
    c1[d] = (float)(f + 1) - xc;
    c2[d] = 1.0f - c1[d];
    k1[d] = f < 0 ? f + ncc : f;
    k2[d] = f + 1 >= ncc ? f + 1 - ncc : f + 1;
  }
  const float kf = a_mid * dt / 6.0f / PI_F;
  const int64_t n3c = (int64_t)ncc * ncc * ncc;
  double acc[3] = {0, 0, 0};
  for (int t = 0; t < 8; t++) {
    const int ox = t & 1, oy = (t >> 1) & 1, oz = t >> 2;
    const int64_t q = ((int64_t)(oz ? k2[2] : k1[2]) * ncc + (oy ? k2[1] : k1[1])) * ncc + (ox ? k2[0] : k1[0]);
    for (int d = 0; d < 3; d++) {
      const float fv = (float)Fc[d * n3c + q];
      acc[d] += (double)(fv * kf * (ox ? c2[0] : c1[0]) * (oy ? c2[1] : c1[1]) * (oz ? c2[2] : c1[2]));
    }
  }
  const int64_t r = z * nf + y;
  A0[r * p0 + x] = (float)acc[0];
  A1[r * p1 + x] = (float)acc[1];
  A2[r * p2 + x] = (float)acc[2];
}

// The gather of all three components at once, as separable sums.  With
// dV = w(jx) w(jy) w(jz) and a_d the acceleration on fine cell (jx,jy,jz)
// of the window,
//   J_d = sum_z w sum_y w sum_x w a_d,
//   Q   = sum dV |a|^2 - sum dV^2 |a|^2   (both separable, weights w and w^2),
// so each fine value enters three 1-D passes instead of (2 DX)^3 products.
// A workgroup owns a TXY x TXY column of fluid cells and CZ cells of it in
// z, and marches through the (CZ+1) DX fine planes those cells reach:
// per plane, the x pass (one item per window row and cell, reading the
// three components from HBM) writes five partial sums per row to LDS, and
// one thread per cell folds them along y and adds the plane, weighted, to
// the two cells whose z windows hold it.  Sums are in double in a fixed
// order; they differ from the per-tap products only by rounding.
template <typename T, bool FINE, int DX, int TXY, int CZ>
__global__ __launch_bounds__(256) void k_fl_grav_sep(const T* A0, const T* A1, const T* A2, int64_t p0, int64_t p1, int64_t p2,
                                                     int64_t nf, int nc, float a_mid, float dt, double* acc) {
  constexpr int SPAN = 2 * DX, EY = (TXY + 1) * DX, NX = EY * TXY;
  static_assert(TXY * TXY <= 256, "one thread per cell of the column");
  __shared__ double sX[5][NX];
  const int ncol = nc / TXY;
  // XCD-contiguous order: columns adjacent in x share the cache lines of
  // their window edges on one L2
  const uint32_t lb = xcd_remap(blockIdx.x, gridDim.x);
  const int bx = lb % ncol, by = (lb / ncol) % ncol, bz = lb / (ncol * ncol);
  const int c0x = bx * TXY, c0y = by * TXY, c0z = bz * CZ;
  const int s = FINE ? 1 : 0;  // fine force: node on the cell's right edge
  const int64_t m0x = (int64_t)c0x * DX - DX / 2 + s, m0y = (int64_t)c0y * DX - DX / 2 + s,
                m0z = (int64_t)c0z * DX - DX / 2 + s;
  const float kf = a_mid * dt / 6.0f / PI_F;
  const double inv3 = 1.0 / ((double)DX * (double)DX * (double)DX);
  double w[SPAN];
#pragma unroll
  for (int j = 0; j < SPAN; j++) w[j] = j < DX ? ((double)j + 0.5) / DX : 2.0 - ((double)j + 0.5) / DX;
  const int tx = threadIdx.x % TXY, ty = threadIdx.x / TXY;
  const bool owner = threadIdx.x < TXY * TXY;
  double cur[5] = {0, 0, 0, 0, 0}, prev[5] = {0, 0, 0, 0, 0};
  const int64_t n3 = (int64_t)nc * nc * nc;
  for (int pz = 0; pz < (CZ + 1) * DX; pz++) {
    const int64_t mz = (m0z + pz) & (nf - 1);  // nf is a power of two
    for (int e = threadIdx.x; e < NX; e += 256) {
      const int ry = e / TXY, cx = e % TXY;
      const int64_t my = (m0y + ry) & (nf - 1);
      const int64_t r0 = (mz * nf + my) * p0, r1 = (mz * nf + my) * p1, r2 = (mz * nf + my) * p2;
      double x0 = 0, x1 = 0, x2 = 0, s1 = 0, s2 = 0;
#pragma unroll
      for (int jx = 0; jx < SPAN; jx++) {
        const int64_t mx = (m0x + (int64_t)cx * DX + jx) & (nf - 1);
        float f0 = (float)A0[r0 + mx], f1 = (float)A1[r1 + mx], f2 = (float)A2[r2 + mx];
        if (FINE) {
          f0 = f0 * kf;
          f1 = f1 * kf;
          f2 = f2 * kf;
        }
        const double a0 = (double)f0 * inv3, a1 = (double)f1 * inv3, a2 = (double)f2 * inv3;
        const double q = a0 * a0 + a1 * a1 + a2 * a2;
        x0 += w[jx] * a0;
        x1 += w[jx] * a1;
        x2 += w[jx] * a2;
        s1 += w[jx] * q;
        s2 += w[jx] * w[jx] * q;
      }
      sX[0][e] = x0;
      sX[1][e] = x1;
      sX[2][e] = x2;
      sX[3][e] = s1;
      sX[4][e] = s2;
    }
    __syncthreads();
    if (owner) {
      double y[5] = {0, 0, 0, 0, 0};
#pragma unroll
      for (int jy = 0; jy < SPAN; jy++) {
        const int e = (ty * DX + jy) * TXY + tx;
        y[0] += w[jy] * sX[0][e];
        y[1] += w[jy] * sX[1][e];
        y[2] += w[jy] * sX[2][e];
        y[3] += w[jy] * sX[3][e];
        y[4] += w[jy] * w[jy] * sX[4][e];
      }
      // plane pz is position j of cell l = pz / DX's window and j + DX of
      // cell l - 1's
      const int j = pz % DX, l = pz / DX;
      const double wa = w[j], wb = w[j + DX];
#pragma unroll
      for (int k = 0; k < 4; k++) {
        cur[k] += wa * y[k];
        prev[k] += wb * y[k];
      }
      cur[4] += wa * wa * y[4];
      prev[4] += wb * wb * y[4];
      if (j == DX - 1) {
        if (l >= 1) {
          const int64_t i = ((int64_t)(c0z + l - 1) * nc + (c0y + ty)) * nc + (c0x + tx);
          acc[i] += prev[0];
          acc[n3 + i] += prev[1];
          acc[2 * n3 + i] += prev[2];
          acc[3 * n3 + i] += prev[3] - prev[4];
        }
#pragma unroll
        for (int k = 0; k < 5; k++) {
          prev[k] = cur[k];
          cur[k] = 0;
        }
      }
    }
    __syncthreads();
  }
}

template <typename T, bool FINE>
static bool fl_grav_sep(cube_handle* h, const T* const A[3], const int64_t pa[3], int n, float a_mid, float dt) {
  NuState& s = h->nu;
  const int nc = s.nc[n];
  const int dx = (int)s.dx[n];
  const int64_t nf = h->nf;
#define FL_SEP(DX_, TXY_, CZ_)                                                                                        \
  if (dx == DX_ && nc % TXY_ == 0 && nc % CZ_ == 0) {                                                                 \
    const unsigned g = (unsigned)((nc / TXY_) * (nc / TXY_) * (nc / CZ_));                                            \
    k_fl_grav_sep<T, FINE, DX_, TXY_, CZ_><<<g, 256, 0, h->stream>>>(A[0], A[1], A[2], pa[0], pa[1], pa[2], nf, nc, a_mid, \
                                                                      dt, s.acc[n]);                                  \
    return true;                                                                                                      \
  }
  FL_SEP(2, 16, 8)
  FL_SEP(4, 16, 8)
  FL_SEP(8, 8, 8)
#undef FL_SEP
  return false;
}

// Gather of J_d and Q over one fluid's cells from an acceleration field on
// fine-cell centres (row pitch ap).  FINE: A is the fine force component at
// nodes, the cell n taking the node n+1 on its right edge (pm.f90:163-165),
// scaled by a_mid*dt/6/pi in float; else A already holds the acceleration.
// accr = a/dx^3 (hydrogr.f90:142).  acc: [4][nc^3] doubles (J0, J1, J2, Q).
template <typename T, bool FINE>
__global__ void k_fl_grav_gather(const T* A, int64_t ap, int64_t nf, int nc, float dx, int d, float a_mid, float dt, double* acc) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n3 = (int64_t)nc * nc * nc;
  if (i >= n3) return;
  const int c[3] = {(int)(i % nc), (int)((i / nc) % nc), (int)(i / ((int64_t)nc * nc))};
  const int span = 2 * (int)dx;  // fine cells per dimension reaching this fluid cell
  const double inv3 = 1.0 / ((double)dx * (double)dx * (double)dx);
  const float kf = a_mid * dt / 6.0f / PI_F;
  double J = 0, Q = 0;
  int m0[3];
  for (int e = 0; e < 3; e++) m0[e] = (int)(c[e] * dx - dx / 2);
  for (int jz = 0; jz < span; jz++) {
    const int mz = m0[2] + jz;
    float wz;
    {
      const float xx = ((float)mz + 0.5f) / dx - 0.5f;
      const int f = (int)floorf(xx);
      const float d1 = (float)(f + 1) - xx;
      wz = f == c[2] ? d1 : (f + 1 == c[2] ? 1.0f - d1 : 0.0f);
    }
    if (wz == 0.0f) continue;
    for (int jy = 0; jy < span; jy++) {
      const int my = m0[1] + jy;
      float wy;
      {
        const float xx = ((float)my + 0.5f) / dx - 0.5f;
        const int f = (int)floorf(xx);
        const float d1 = (float)(f + 1) - xx;
        wy = f == c[1] ? d1 : (f + 1 == c[1] ? 1.0f - d1 : 0.0f);
      }
      if (wy == 0.0f) continue;
      for (int jx = 0; jx < span; jx++) {
        const int mx = m0[0] + jx;
        float wx;
        {
          const float xx = ((float)mx + 0.5f) / dx - 0.5f;
          const int f = (int)floorf(xx);
          const float d1 = (float)(f + 1) - xx;
          wx = f == c[0] ? d1 : (f + 1 == c[0] ? 1.0f - d1 : 0.0f);
        }
        if (wx == 0.0f) continue;
        const float dV = wx * wy * wz;
        const int s = FINE ? 1 : 0;  // fine force: node on the cell's right edge
        const int64_t ax = ((int64_t)mx + s + nf) % nf, ay = ((int64_t)my + s + nf) % nf, az = ((int64_t)mz + s + nf) % nf;
        const float fv = (float)A[(az * nf + ay) * ap + ax];
        const double a = (double)(FINE ? fv * kf : fv) * inv3;
        J += (double)dV * a;
        Q += (double)dV * (1.0 - (double)dV) * a * a;
      }
    }
  }
  acc[(int64_t)d * n3 + i] += J;
  acc[3 * n3 + i] += Q;
}

// m += rho J, E += m0.J + rho/2 (|J|^2 + Q); clears the accumulators
__global__ void k_fl_grav_apply(float* F, int nc, double* acc) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n3 = (int64_t)nc * nc * nc;
  if (i >= n3) return;
  const double r = F[i];
  double J[3], mj = 0, jj = 0;
  for (int d = 0; d < 3; d++) {
    J[d] = acc[(int64_t)d * n3 + i];
    mj += (double)F[(int64_t)(1 + d) * n3 + i] * J[d];
    jj += J[d] * J[d];
  }
  const double Q = acc[3 * n3 + i];
  F[4 * n3 + i] = (float)((double)F[4 * n3 + i] + mj + 0.5 * r * (jj + Q));
  for (int d = 0; d < 3; d++) {
    F[(int64_t)(1 + d) * n3 + i] = (float)((double)F[(int64_t)(1 + d) * n3 + i] + r * J[d]);
    acc[(int64_t)d * n3 + i] = 0;
  }
  acc[3 * n3 + i] = 0;
}

__global__ void k_fl_set_u32(uint32_t* p, uint32_t v) { p[threadIdx.x] = v; }

// ------------------------------------------------------------------- host
// hydrodf.f90:17-31 constants
static const float kGLv[3] = {0.41577456f, 2.29428036f, 6.28994508f};
static const float kGLw[3] = {0.71109301f, 0.27851773f, 0.01038926f};

static inline unsigned nblk(int64_t n) { return (unsigned)((n + 255) / 256); }

extern "C" int cube_nu_init(cube_handle* h, const cube_nu_params* p) {
  CUBE_EXCLUSIVE;
  if (!h || !p) return CUBE_ERR_ARG;
  if (hipSetDevice(h->p.device) != hipSuccess) return cube_fail(h, CUBE_ERR_HIP, "hipSetDevice failed");
  if (!(p->m_ev > 0) || !(p->f_neu > 0) || !(p->gamma > 1))
    return cube_fail(h, CUBE_ERR_ARG, "neutrino mass, f_neu and gamma must be positive (gamma > 1)");
  if (h->nf / 8 < FL_W) return cube_fail(h, CUBE_ERR_ARG, "fine mesh too small for the fluids");
  NuState& s = h->nu;
  s.on = true;
  s.par = *p;
  // hydrogr.f90:14 velocity unit, hydrodf.f90:9-13 start of evolution
  const float nf = (float)h->nf;
  const float dv = (2 * nf) / (3.0f * 100.0f * sqrtf(p->omega_m) * p->box);
  const float beta = 5966.34f * p->m_ev;
  const float ai_kms = (8.0f * PI_F * PI_F / 3.0f) / (p->box * p->box * (p->h0 * 100.0f) * (p->h0 * 100.0f) * p->omega_m) / (dv * dv) / 2.0f;
  const float ai_max = 1.0f / (1.0f + p->z_i_nu);
  float wsum = 0;
  for (int n = 0; n < 3; n++) {
    const float v = kGLv[n] / beta * 2.998e5f;  // km/s
    s.nc[n] = (int)(h->nf >> (n + 1));
    s.dx[n] = nf / (float)s.nc[n];
    s.cs2[n] = fmaxf(1e-6f, (v * dv) * (v * dv));
    const float x = kGLv[n];
    s.w[n] = kGLw[n] * (x * x / (expf(x) + 1.0f)) * expf(x);
    wsum += s.w[n];
    float ai = fminf(ai_kms * s.cs2[n], ai_max);
    if (p->neutrino_ic && ai < ai_max) ai = ai_max;  // NEUTRINO_IC: dfld_ai_min = dfld_ai_max
    s.ai[n] = ai;
  }
  for (int n = 0; n < 3; n++) s.w[n] /= wsum;
  for (int n = 0; n < 3; n++) {
    const int64_t n3 = (int64_t)s.nc[n] * s.nc[n] * s.nc[n];
    if (!s.F[n]) HIP_TRY(h, hipMalloc(&s.F[n], sizeof(float) * 5 * n3));
    if (!s.acc[n]) HIP_TRY(h, hipMalloc(&s.acc[n], sizeof(double) * 4 * n3));
    HIP_TRY(h, hipMemsetAsync(s.acc[n], 0, sizeof(double) * 4 * n3, h->stream));
    // hg_setup: rho = 1, m = 0, E = cs2 rho/(g-1) (q equation of state)
    std::vector<float> init(5 * n3, 0.0f);
    for (int64_t i = 0; i < n3; i++) {
      init[i] = 1.0f;
      init[4 * n3 + i] = s.cs2[n] * 1.0f / (p->gamma - 1.0f);
    }
    HIP_TRY(h, hipMemcpy(s.F[n], init.data(), sizeof(float) * 5 * n3, hipMemcpyHostToDevice));
  }
  return CUBE_OK;
}

extern "C" int cube_nu_info(cube_handle* h, int n, int32_t* nc, float* dx, float* cs2, float* weight) {
  if (!h || !h->nu.on || n < 0 || n > 2) return CUBE_ERR_ARG;
  if (nc) *nc = h->nu.nc[n];
  if (dx) *dx = h->nu.dx[n];
  if (cs2) *cs2 = h->nu.cs2[n];
  if (weight) *weight = h->nu.w[n];
  return CUBE_OK;
}

extern "C" int cube_nu_set_fluid(cube_handle* h, int n, const float* fld) {
  CUBE_EXCLUSIVE;
  if (!h || !h->nu.on || n < 0 || n > 2 || !fld) return CUBE_ERR_ARG;
  const int64_t n3 = (int64_t)h->nu.nc[n] * h->nu.nc[n] * h->nu.nc[n];
  HIP_TRY(h, hipMemcpy(h->nu.F[n], fld, sizeof(float) * 5 * n3, hipMemcpyHostToDevice));
  return CUBE_OK;
}

extern "C" int cube_nu_get_fluid(cube_handle* h, int n, float* fld) {
  CUBE_EXCLUSIVE;
  if (!h || !h->nu.on || n < 0 || n > 2 || !fld) return CUBE_ERR_ARG;
  const int64_t n3 = (int64_t)h->nu.nc[n] * h->nu.nc[n] * h->nu.nc[n];
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  HIP_TRY(h, hipMemcpy(fld, h->nu.F[n], sizeof(float) * 5 * n3, hipMemcpyDeviceToHost));
  return CUBE_OK;
}

static int fl_sweep(cube_handle* h, int n, int ax, float dt) {
  NuState& s = h->nu;
  const int nc = s.nc[n];
  if (nc * FL_W > 8 * FL_NT) return cube_fail(h, CUBE_ERR_ARG, "fluid of %d cells per dimension exceeds the sweep kernel", nc);
  const unsigned grid = (unsigned)((int64_t)(nc / FL_W) * nc);
  const size_t lds = sizeof(float) * 5 * FL_W * nc;
  const float g = s.par.gamma;
  if (ax == 0) k_fl_sweep<0><<<grid, FL_NT, lds, h->stream>>>(s.F[n], nc, g, dt, s.dx[n]);
  if (ax == 1) k_fl_sweep<1><<<grid, FL_NT, lds, h->stream>>>(s.F[n], nc, g, dt, s.dx[n]);
  if (ax == 2) k_fl_sweep<2><<<grid, FL_NT, lds, h->stream>>>(s.F[n], nc, g, dt, s.dx[n]);
  HIP_TRY(h, hipGetLastError());
  return 0;
}

// dfld_evolve (hydrodf.f90:90-110): each active fluid swept x,y,z (dir > 0)
// or z,y,x, then its timestep folded into dt_neu
extern "C" int cube_fluid_evolve(cube_handle* h, float dt, int dir, float a, float* dt_neu) {
  if (!h || !h->nu.on) return CUBE_ERR_ARG;
  if (hipSetDevice(h->p.device) != hipSuccess) return cube_fail(h, CUBE_ERR_HIP, "hipSetDevice failed");
  NuState& s = h->nu;
  int pr = prof_begin(h, PS_FLUID);
  for (int n = 0; n < 3; n++) {
    if (!(a >= s.ai[n])) continue;
    for (int k = 0; k < 3; k++) {
      const int ax = dir > 0 ? k : 2 - k;
      int r = fl_sweep(h, n, ax, dt);
      if (r) return r;
    }
    uint32_t* dtb = h->uflags + 48;
    k_fl_set_u32<<<1, 1, 0, h->stream>>>(dtb, 0x4e6e6b28u);  // 1e9f, the reference's start value
    k_fl_dt<<<(unsigned)((int64_t)s.nc[n] * s.nc[n]), FL_NT, 0, h->stream>>>(s.F[n], s.nc[n], s.par.gamma, s.dx[n], dtb);
    HIP_TRY(h, hipGetLastError());
    uint32_t hb;
    HIP_TRY(h, hipMemcpyAsync(&hb, dtb, 4, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    float dtn;
    memcpy(&dtn, &hb, 4);
    if (dtn == 1e9f) dtn = 1.0f;
    if (dt_neu) *dt_neu = fminf(*dt_neu, dtn);
  }
  prof_end(h, pr);
  return CUBE_OK;
}

// ---- coupling to the PM step (called from cube_pm)
static FlSet fl_set(const NuState& s) {
  FlSet f{};
  f.n = 3;
  for (int n = 0; n < 3; n++) {
    f.rho[n] = s.F[n];
    f.nc[n] = s.nc[n];
    f.dx[n] = s.dx[n];
    f.w[n] = s.w[n];
  }
  return f;
}

// Coarse-mesh density (before the coarse FFT); leaves the fine-cell-centre
// density in fine.W for fl_density_fine.
int fl_density_coarse(cube_handle* h, float a_mid, double* mass_added, double* mass_dev) {
  NuState& s = h->nu;
  const int homog = 1.0f / a_mid - 1.0f > s.par.z_i_nu;
  const int64_t nf = h->nf;
  float* D = (float*)h->fine.W;
  if (!homog) {
    k_fl_density_centres<<<nblk(nf * nf * nf), 256, 0, h->stream>>>(fl_set(s), nf, D);
    HIP_TRY(h, hipGetLastError());
  }
  const int64_t n3c = h->n3c;
  const int64_t nb = (int64_t)nblk(n3c);
  if (nb > h->dred_n - 16) return cube_fail(h, CUBE_ERR_ARG, "reduction buffer too small");
  if (h->prec == 4)
    k_fl_add_coarse<float><<<(unsigned)nb, 256, 0, h->stream>>>((float*)h->coarse.R, D, nf, (int)h->nc, s.par.f_neu, homog,
                                                                h->dred);
  else
    k_fl_add_coarse<double><<<(unsigned)nb, 256, 0, h->stream>>>((double*)h->coarse.R, D, nf, (int)h->nc, s.par.f_neu, homog,
                                                                 h->dred);
  HIP_TRY(h, hipGetLastError());
  if (mass_dev) return reduce_doubles_dev(h, h->dred, nb, 1, mass_dev);
  if (mass_added) return reduce_doubles(h, h->dred, nb, 1, mass_added);
  return 0;
}

int fl_density_fine(cube_handle* h, float a_mid, double* mass_added, double* mass_dev) {
  NuState& s = h->nu;
  const int homog = 1.0f / a_mid - 1.0f > s.par.z_i_nu;
  const int64_t nf = h->nf;
  const int64_t nb = (int64_t)nblk(nf * nf * nf);
  if (nb > h->dred_n - 16) return cube_fail(h, CUBE_ERR_ARG, "reduction buffer too small");
  if (h->prec == 4)
    k_fl_add_fine<float><<<(unsigned)nb, 256, 0, h->stream>>>((float*)h->fine.R, (const float*)h->fine.W, nf, s.par.f_neu, homog,
                                                              h->dred);
  else
    k_fl_add_fine<double><<<(unsigned)nb, 256, 0, h->stream>>>((double*)h->fine.R, (const float*)h->fine.W, nf, s.par.f_neu,
                                                               homog, h->dred);
  HIP_TRY(h, hipGetLastError());
  if (mass_dev) return reduce_doubles_dev(h, h->dred, nb, 1, mass_dev);
  if (mass_added) return reduce_doubles(h, h->dred, nb, 1, mass_added);
  return 0;
}

// The three fine-force components (Ff[d], row pitch pf[d]) into the
// accumulators of every active fluid.
int fl_gravity_fine(cube_handle* h, const void* const Ff[3], const int64_t pf[3], float a_mid, float dt) {
  NuState& s = h->nu;
  for (int n = 0; n < 3; n++) {
    if (!(a_mid >= s.ai[n])) continue;
    const int64_t n3 = (int64_t)s.nc[n] * s.nc[n] * s.nc[n];
    if (h->prec == 4) {
      const float* A[3] = {(const float*)Ff[0], (const float*)Ff[1], (const float*)Ff[2]};
      if (fl_grav_sep<float, true>(h, A, pf, n, a_mid, dt)) {
        HIP_TRY(h, hipGetLastError());
        continue;
      }
    } else {
      const double* A[3] = {(const double*)Ff[0], (const double*)Ff[1], (const double*)Ff[2]};
      if (fl_grav_sep<double, true>(h, A, pf, n, a_mid, dt)) {
        HIP_TRY(h, hipGetLastError());
        continue;
      }
    }
    for (int d = 0; d < 3; d++) {
      if (h->prec == 4)
        k_fl_grav_gather<float, true><<<nblk(n3), 256, 0, h->stream>>>((const float*)Ff[d], pf[d], h->nf, s.nc[n], s.dx[n], d,
                                                                      a_mid, dt, s.acc[n]);
      else
        k_fl_grav_gather<double, true><<<nblk(n3), 256, 0, h->stream>>>((const double*)Ff[d], pf[d], h->nf, s.nc[n], s.dx[n],
                                                                       d, a_mid, dt, s.acc[n]);
      HIP_TRY(h, hipGetLastError());
    }
  }
  return 0;
}

static int fl_apply(cube_handle* h, float a_mid) {
  NuState& s = h->nu;
  for (int n = 0; n < 3; n++) {
    if (!(a_mid >= s.ai[n])) continue;
    const int64_t n3 = (int64_t)s.nc[n] * s.nc[n] * s.nc[n];
    k_fl_grav_apply<<<nblk(n3), 256, 0, h->stream>>>(s.F[n], s.nc[n], s.acc[n]);
    HIP_TRY(h, hipGetLastError());
  }
  return 0;
}

// After the fine kicks: apply the fine impulses, then gather and apply the
// coarse ones (pm.f90:374-422 runs after the fine phase).  The coarse
// acceleration at fine-cell centres goes into the fine-force buffers
// Ff[d] (row pitch pf[d]), which the fine phase no longer needs.
int fl_gravity_finish(cube_handle* h, void* const Ff[3], const int64_t pf[3], float a_mid, float dt) {
  NuState& s = h->nu;
  bool any = false;
  for (int n = 0; n < 3; n++) any = any || a_mid >= s.ai[n];
  if (!any) return 0;
  int r;
  if ((r = fl_apply(h, a_mid))) return r;
  const int64_t nf = h->nf;
  float* A[3] = {(float*)Ff[0], (float*)Ff[1], (float*)Ff[2]};
  if (h->prec == 4)
    k_fl_coarse_at_centres<float><<<nblk(nf * nf * nf), 256, 0, h->stream>>>((const float*)h->fc, (int)h->nc, nf, a_mid, dt,
                                                                             A[0], A[1], A[2], pf[0], pf[1], pf[2]);
  else
    k_fl_coarse_at_centres<double><<<nblk(nf * nf * nf), 256, 0, h->stream>>>((const double*)h->fc, (int)h->nc, nf, a_mid, dt,
                                                                              A[0], A[1], A[2], pf[0], pf[1], pf[2]);
  HIP_TRY(h, hipGetLastError());
  const float* Ac[3] = {A[0], A[1], A[2]};
  for (int n = 0; n < 3; n++) {
    if (!(a_mid >= s.ai[n])) continue;
    const int64_t n3 = (int64_t)s.nc[n] * s.nc[n] * s.nc[n];
    if (!fl_grav_sep<float, false>(h, Ac, pf, n, a_mid, dt))
      for (int d = 0; d < 3; d++)
        k_fl_grav_gather<float, false><<<nblk(n3), 256, 0, h->stream>>>(A[d], pf[d], nf, s.nc[n], s.dx[n], d, a_mid, dt,
                                                                        s.acc[n]);
    HIP_TRY(h, hipGetLastError());
  }
  return fl_apply(h, a_mid);
}

void fl_free(NuState& s) {
  for (int n = 0; n < 3; n++) {
    if (s.F[n]) (void)hipFree(s.F[n]);
    if (s.acc[n]) (void)hipFree(s.acc[n]);
    s.F[n] = nullptr;
    s.acc[n] = nullptr;
  }
  s.on = false;
}
