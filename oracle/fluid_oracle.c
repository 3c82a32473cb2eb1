/* CPU restatement of CUBE's neutrino-fluid update -- test infrastructure
 * (the checker), used only by tests/.
 *
 * Follows CUBE_v2.0/work/main/hydro1d.f90:113-225 (MUSCL-minmod
 * reconstruction of density, velocities and pressure, HLL flux, q equation
 * of state: energy flux E*v, sound speed sqrt(p/rho)), hydro3d.f90:9-108
 * (x,y,z sweeps forward, z,y,x backward, momentum components swapped so the
 * sweep's normal momentum is the second variable; per-line timestep with
 * CFL 0.71), hydrogr.f90:47-81 (fields on 1-ncb:nc+ncb with
 * ncb = max(2, 1+floor(ncb_cube*ncell/dx)), uniform set-up),
 * hydrogr.f90:104-124 (buffer, evolve, buffer, timestep) and :241-284
 * (periodic halo fill).  All fluid arithmetic is real(4) in the
 * reference's expression order.  Lines run over the whole extended array
 * with cshift wrap at the line ends, exactly as the reference does; the
 * physical cells are the result.
 *
 * Gravity (hydrogr.f90:128-238) and density (:311-346) couplings use CIC
 * between fine-cell centres and fluid cells.  The reference applies each
 * fine cell's impulse to the fluid cells one after another; here the fine
 * cells are visited in z, y, x order (the reference's OpenMP loop has no
 * fixed order), and fluid cells are periodic (a single box: contributions
 * the reference CIC places in halo cells belong to the periodic images). */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define H_CFL 0.71f
#define H_LIM 1.5f

static inline float sgn1(float a) { return copysignf(1.0f, a); }
static inline float minmod(float a, float b, float c) {
  float m = fminf(fminf(fabsf(a), fabsf(b)), fabsf(c));
  return 0.25f * fabsf(sgn1(a) + sgn1(b)) * (sgn1(a) + sgn1(c)) * m;
}
static inline int cs(int i, int n) { return ((i % n) + n) % n; }

/* kinetic energy 0.5*sum(h(2:4)**2)/h(1) */
static inline float kin(float r, float m1, float m2, float m3) { return 0.5f * ((m1 * m1 + m2 * m2) + m3 * m3) / r; }
static inline float press(float r, float m1, float m2, float m3, float e, float g) {
  float p = (e - kin(r, m1, m2, m3)) * (g - 1.0f);
  return p > 0.0f ? p : 0.0f;
}

/* h_evolve on one line h[5][n] (variable-major), first order in time */
static void evolve_line(float* h, int n, float g, float dt, float dx, float* w) {
  float *r = h, *m1 = h + n, *m2 = h + 2 * n, *m3 = h + 3 * n, *e = h + 4 * n;
  float *v1 = w, *v2 = w + n, *v3 = w + 2 * n, *p = w + 3 * n, *f = w + 4 * n; /* f: 5n */
  for (int i = 0; i < n; i++) {
    v1[i] = m1[i] / r[i];
    v2[i] = m2[i] / r[i];
    v3[i] = m3[i] / r[i];
    p[i] = press(r[i], m1[i], m2[i], m3[i], e[i], g);
  }
  float* u[5] = {r, v1, v2, v3, p};
  for (int i = 0; i < n; i++) {
    float lb[5], rb[5];
    for (int q = 0; q < 5; q++) {
      const float a0 = u[q][cs(i - 1, n)], a1 = u[q][i], a2 = u[q][cs(i + 1, n)], a3 = u[q][cs(i + 2, n)];
      lb[q] = a1 + 0.5f * minmod(H_LIM * (a1 - a0), 0.5f * (a2 - a0), H_LIM * (a2 - a1));
      rb[q] = a2 - 0.5f * minmod(H_LIM * (a2 - a1), 0.5f * (a3 - a1), H_LIM * (a3 - a2));
    }
    float hl[5], hr[5];
    hl[0] = lb[0]; hl[1] = hl[0] * lb[1]; hl[2] = hl[0] * lb[2]; hl[3] = hl[0] * lb[3];
    hl[4] = lb[4] / (g - 1.0f) + kin(hl[0], hl[1], hl[2], hl[3]);
    hr[0] = rb[0]; hr[1] = hr[0] * rb[1]; hr[2] = hr[0] * rb[2]; hr[3] = hr[0] * rb[3];
    hr[4] = rb[4] / (g - 1.0f) + kin(hr[0], hr[1], hr[2], hr[3]);
    float fl[5], fr[5], sl, sr, vl, vr;
    {
      const float pl = press(hl[0], hl[1], hl[2], hl[3], hl[4], g);
      vl = hl[1] / hl[0];
      fl[0] = hl[1]; fl[1] = hl[1] * vl + pl; fl[2] = hl[1] * (hl[2] / hl[0]); fl[3] = hl[1] * (hl[3] / hl[0]);
      fl[4] = hl[4] * vl;
      sl = sqrtf(pl / hl[0]);
      const float pr = press(hr[0], hr[1], hr[2], hr[3], hr[4], g);
      vr = hr[1] / hr[0];
      fr[0] = hr[1]; fr[1] = hr[1] * vr + pr; fr[2] = hr[1] * (hr[2] / hr[0]); fr[3] = hr[1] * (hr[3] / hr[0]);
      fr[4] = hr[4] * vr;
      sr = sqrtf(pr / hr[0]);
    }
    const float ap = fmaxf(fmaxf(0.0f, vl + sl), vr + sr);
    const float am = fmaxf(fmaxf(0.0f, -vl + sl), -vr + sr);
    for (int q = 0; q < 5; q++) f[q * n + i] = (ap * fl[q] + am * fr[q] - ap * am * (hr[q] - hl[q])) / (ap + am);
  }
  for (int q = 0; q < 5; q++) {
    float* hq = h + q * n;
    const float* fq = f + q * n;
    float rq[4096];
    for (int i = 0; i < n; i++) rq[i] = -fq[i] + fq[cs(i - 1, n)];
    for (int i = 0; i < n; i++) hq[i] = hq[i] + dt * rq[i] / dx;
  }
}

/* extended array E[5][ne][ne][ne] (x fastest), ne = nc + 2 ncb */
#define EX(q, i, j, k) E[(((size_t)(q) * ne + (k)) * ne + (j)) * ne + (i)]

static void buffer(float* E, int nc, int ncb) {
  const int ne = nc + 2 * ncb;
  for (int q = 0; q < 5; q++)
    for (int k = 0; k < ne; k++)
      for (int j = 0; j < ne; j++)
        for (int i = 0; i < ncb; i++) {
          EX(q, i, j, k) = EX(q, i + nc, j, k);
          EX(q, nc + ncb + i, j, k) = EX(q, ncb + i, j, k);
        }
  for (int q = 0; q < 5; q++)
    for (int k = 0; k < ne; k++)
      for (int j = 0; j < ncb; j++)
        for (int i = 0; i < ne; i++) {
          EX(q, i, j, k) = EX(q, i, j + nc, k);
          EX(q, i, nc + ncb + j, k) = EX(q, i, ncb + j, k);
        }
  for (int q = 0; q < 5; q++)
    for (int k = 0; k < ncb; k++)
      for (int j = 0; j < ne; j++)
        for (int i = 0; i < ne; i++) {
          EX(q, i, j, k) = EX(q, i, j, k + nc);
          EX(q, i, j, nc + ncb + k) = EX(q, i, j, ncb + k);
        }
}

/* one sweep along axis ax (0 x, 1 y, 2 z): the normal momentum goes first */
static void sweep(float* E, int nc, int ncb, int ax, float g, float dt, float dx) {
  const int ne = nc + 2 * ncb;
  const int perm[3][3] = {{1, 2, 3}, {2, 1, 3}, {3, 2, 1}}; /* reference's h1d(2:4) per axis */
  /* lines of one sweep are independent: threads take whole lines */
#pragma omp parallel
  {
  float* line = (float*)malloc(sizeof(float) * 5 * ne);
  float* w = (float*)malloc(sizeof(float) * 9 * ne);
#pragma omp for collapse(2) schedule(static)
  for (int b = 0; b < ne; b++)
    for (int a = 0; a < ne; a++) {
      for (int t = 0; t < ne; t++) {
        int i, j, k;
        if (ax == 0) { i = t; j = a; k = b; } else if (ax == 1) { i = a; j = t; k = b; } else { i = a; j = b; k = t; }
        line[t] = EX(0, i, j, k);
        for (int c = 0; c < 3; c++) line[(1 + c) * ne + t] = EX(perm[ax][c], i, j, k);
        line[4 * ne + t] = EX(4, i, j, k);
      }
      evolve_line(line, ne, g, dt, dx, w);
      for (int t = 0; t < ne; t++) {
        int i, j, k;
        if (ax == 0) { i = t; j = a; k = b; } else if (ax == 1) { i = a; j = t; k = b; } else { i = a; j = b; k = t; }
        EX(0, i, j, k) = line[t];
        for (int c = 0; c < 3; c++) EX(perm[ax][c], i, j, k) = line[(1 + c) * ne + t];
        EX(4, i, j, k) = line[4 * ne + t];
      }
    }
  free(line);
  free(w);
  }
}

/* h3_timestep over x-lines of the extended array */
static float timestep(const float* E, int nc, int ncb, float g, float dx) {
  const int ne = nc + 2 * ncb;
  float dt = 1e9f;
  for (int k = 0; k < ne; k++)
    for (int j = 0; j < ne; j++) {
      float v = 0.0f, s = 0.0f;
      for (int d = 1; d <= 3; d++)
        for (int i = 0; i < ne; i++) v = fmaxf(v, fabsf(EX(d, i, j, k) / EX(0, i, j, k)));
      for (int i = 0; i < ne; i++) {
        const float p = press(EX(0, i, j, k), EX(1, i, j, k), EX(2, i, j, k), EX(3, i, j, k), EX(4, i, j, k), g);
        s = fmaxf(s, sqrtf(p / EX(0, i, j, k)));
      }
      const float dtn = H_CFL * dx / (v + s);
      dt = fminf(dt, dtn);
    }
  if (dt == 1e9f) dt = 1.0f;
  return dt;
}

/* hg_evolve on physical fields P[5][nc^3]: buffer, sweeps, buffer, timestep */
int ofl_evolve(float* P, int nc, int ncb, float g, float dt, float dx, int dir, float* dtg) {
  const int ne = nc + 2 * ncb;
  if (ne > 4096) return 1;
  float* E = (float*)malloc(sizeof(float) * 5 * (size_t)ne * ne * ne);
  for (int q = 0; q < 5; q++)
    for (int k = 0; k < nc; k++)
      for (int j = 0; j < nc; j++)
        for (int i = 0; i < nc; i++)
          EX(q, i + ncb, j + ncb, k + ncb) = P[(((size_t)q * nc + k) * nc + j) * nc + i];
  buffer(E, nc, ncb);
  const int order[2][3] = {{0, 1, 2}, {2, 1, 0}};
  for (int s = 0; s < 3; s++) {
    const int ax = order[dir > 0 ? 0 : 1][s];
    if (nc + 2 * ncb > 7) sweep(E, nc, ncb, ax, g, dt, dx);
  }
  buffer(E, nc, ncb);
  const float dtn = timestep(E, nc, ncb, g, dx);
  if (dtg) *dtg = fminf(*dtg, dtn);
  for (int q = 0; q < 5; q++)
    for (int k = 0; k < nc; k++)
      for (int j = 0; j < nc; j++)
        for (int i = 0; i < nc; i++)
          P[(((size_t)q * nc + k) * nc + j) * nc + i] = EX(q, i + ncb, j + ncb, k + ncb);
  free(E);
  return 0;
}

/* ---------------------------------------------------------------- gravity */
#define PF(q, i, j, k) P[(((size_t)(q) * nc + (k)) * nc + (j)) * nc + (i)]

/* hg_gravity with CIC (gc != 0): accr = acc/dx^3; eight cells in the
 * reference's order, each updated from the current fields */
static void gravity_at(float* P, int nc, float dx, const double acc[3], const float x[3]) {
  float accr[3], xx[3], dx1[3], dx2[3];
  int i1[3], i2[3];
  for (int d = 0; d < 3; d++) {
    accr[d] = (float)(acc[d] / (double)(dx * dx * dx));
    xx[d] = x[d] / dx - 0.5f;
    i1[d] = (int)floorf(xx[d]);  /* 0-based: 1+floor(xx) - 1 */
    i2[d] = i1[d] + 1;
    dx1[d] = (float)(i1[d] + 1) - xx[d];
    dx2[d] = 1.0f - dx1[d];
  }
  const int ord[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0}, {0, 0, 1}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
  for (int t = 0; t < 8; t++) {
    const int i = cs(ord[t][0] ? i2[0] : i1[0], nc), j = cs(ord[t][1] ? i2[1] : i1[1], nc),
              k = cs(ord[t][2] ? i2[2] : i1[2], nc);
    const float dV = (ord[t][0] ? dx2[0] : dx1[0]) * (ord[t][1] ? dx2[1] : dx1[1]) * (ord[t][2] ? dx2[2] : dx1[2]);
    const float rho = PF(0, i, j, k);
    const float dE = (PF(1, i, j, k) * accr[0] + PF(2, i, j, k) * accr[1] + PF(3, i, j, k) * accr[2]) +
                     0.5f * (accr[0] * accr[0] * rho + accr[1] * accr[1] * rho + accr[2] * accr[2] * rho);
    PF(4, i, j, k) = PF(4, i, j, k) + dV * dE;
    for (int d = 0; d < 3; d++) PF(1 + d, i, j, k) = PF(1 + d, i, j, k) + dV * (accr[d] * rho);
  }
}

/* fine phase (pm.f90:157-171): the fine force at the node on the right edge
 * of every fine cell, times a_mid*dt/6/pi, at the cell centre.  F: [3][nf^3]
 * node values, x fastest. */
void ofl_gravity_fine(float* P, int nc, float dx, const float* F, int nf, float a_mid, float dt) {
  const float pi = 3.14159274101257324f;
  for (int mz = 1; mz <= nf; mz++)
    for (int my = 1; my <= nf; my++)
      for (int mx = 1; mx <= nf; mx++) {
        const size_t node = ((size_t)(mz % nf) * nf + (my % nf)) * nf + (mx % nf);
        double acc[3];
        for (int d = 0; d < 3; d++) acc[d] = (double)(F[(size_t)d * nf * nf * nf + node] * a_mid * dt / 6 / pi);
        const float x[3] = {(float)mx - 0.5f, (float)my - 0.5f, (float)mz - 0.5f};
        gravity_at(P, nc, dx, acc, x);
      }
}

/* coarse phase (pm.f90:374-422): force_c CIC-interpolated at every fine
 * cell centre.  Fc: [3][ncc^3], x fastest, ncell fine cells per coarse cell */
void ofl_gravity_coarse(float* P, int nc, float dx, const float* Fc, int ncc, int ncell, float a_mid, float dt) {
  const float pi = 3.14159274101257324f;
  const int nf = ncc * ncell;
  for (int mz = 1; mz <= nf; mz++)
    for (int my = 1; my <= nf; my++)
      for (int mx = 1; mx <= nf; mx++) {
        const float x[3] = {(float)mx - 0.5f, (float)my - 0.5f, (float)mz - 0.5f};
        float xc[3], c1[3], c2[3];
        int k1[3], k2[3];
        for (int d = 0; d < 3; d++) {
          xc[d] = x[d] / (float)ncell - 0.5f;
          k1[d] = (int)floorf(xc[d]);
          k2[d] = k1[d] + 1;
          c1[d] = (float)(k1[d] + 1) - xc[d];
          c2[d] = 1.0f - c1[d];
        }
        const int ord[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0}, {0, 0, 1}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
        double acc[3] = {0, 0, 0};
        for (int t = 0; t < 8; t++) {
          const int i = cs(ord[t][0] ? k2[0] : k1[0], ncc), j = cs(ord[t][1] ? k2[1] : k1[1], ncc),
                    k = cs(ord[t][2] ? k2[2] : k1[2], ncc);
          const float w1 = ord[t][0] ? c2[0] : c1[0], w2 = ord[t][1] ? c2[1] : c1[1], w3 = ord[t][2] ? c2[2] : c1[2];
          for (int d = 0; d < 3; d++)
            acc[d] += (double)(Fc[(size_t)d * ncc * ncc * ncc + ((size_t)k * ncc + j) * ncc + i] * a_mid * dt / 6 / pi * w1 * w2 * w3);
        }
        gravity_at(P, nc, dx, acc, x);
      }
}

/* hg_density (CIC, gc != 0) of fluid field rho[nc^3] at x (fine units) */
static float density_at(const float* rho, int nc, float dx, const float x[3]) {
  float xx[3], d1[3], d2[3];
  int i1[3], i2[3];
  for (int d = 0; d < 3; d++) {
    xx[d] = x[d] / dx - 0.5f;
    i1[d] = (int)floorf(xx[d]);
    i2[d] = i1[d] + 1;
    d1[d] = (float)(i1[d] + 1) - xx[d];
    d2[d] = 1.0f - d1[d];
  }
#define RH(a, b, c) rho[((size_t)cs(c, nc) * nc + cs(b, nc)) * nc + cs(a, nc)]
  return RH(i1[0], i1[1], i1[2]) * d1[0] * d1[1] * d1[2] + RH(i2[0], i1[1], i1[2]) * d2[0] * d1[1] * d1[2] +
         RH(i1[0], i2[1], i1[2]) * d1[0] * d2[1] * d1[2] + RH(i2[0], i2[1], i1[2]) * d2[0] * d2[1] * d1[2] +
         RH(i1[0], i1[1], i2[2]) * d1[0] * d1[1] * d2[2] + RH(i2[0], i1[1], i2[2]) * d2[0] * d1[1] * d2[2] +
         RH(i1[0], i2[1], i2[2]) * d1[0] * d2[1] * d2[2] + RH(i2[0], i2[1], i2[2]) * d2[0] * d2[1] * d2[2];
#undef RH
}

/* dfld_density at the centres of all fine cells (pm.f90:57-72 samples it at
 * node + 0.5 for the fine mesh; :209-238 at fine-cell centres for the coarse
 * mesh): out[nf^3], weights w[nfl], fluids rho[f] of nc[f]^3 cells */
void ofl_density_centres(float* out, int nf, int nfl, const float** rho, const int* ncf, const float* dxf, const float* w) {
  for (int z = 0; z < nf; z++)
    for (int y = 0; y < nf; y++)
      for (int x = 0; x < nf; x++) {
        const float p[3] = {(float)x + 0.5f, (float)y + 0.5f, (float)z + 0.5f};
        float d = 0.0f;
        for (int f = 0; f < nfl; f++) d = d + w[f] * density_at(rho[f], ncf[f], dxf[f], p);
        out[((size_t)z * nf + y) * nf + x] = d;
      }
}
