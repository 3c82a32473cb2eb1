/*
 * CPU ORACLE -- test infrastructure only.
 *
 * Restatement of CUBE's runtime spherical-overdensity halo finder
 * (CUBE_v2.0/work/main/halofind.f90:18-598) for one image whose ghost
 * regions are periodic copies (nn = 1).  Used by tests/ as the checker of
 * the GPU halo finder; nothing in the product links or calls it.
 *
 * The reference works tile by tile on the tile's extended arrays, in which
 * ghost particles are separate copies with their own p_tag.  Restated here
 * per tile T (processed in the reference's order itz, ity, itx = nnt..1):
 *   density    fine-mesh CIC at cell centres (halofind.f90:145-177, the
 *              '-0.5' of :159); a node's value is the one its owning tile
 *              computes, summed in 2^-40 fixed point (any order gives the
 *              same sum)
 *   peaks      27-point maxima above den_peak_cutoff over T's nodes
 *              1-ncell..nft+ncell (:180-183)
 *   mesh mass  walk over the offset table sorted by radius until a
 *              completed shell's mean density drops below den_odc, cells
 *              outside T's extended array skipped (:186-195); candidates
 *              need m_mesh > mass_p*min_halo_particles/2 (:197)
 *   position   parabolic interpolation of the node maximum (:203-208)
 *   order      peaks by den_peak, highest first (:216-219, :230)
 *   particles  untagged particles within rsearch of the peak (:268-327),
 *              CIC of those within r_proxy onto a refine_ratio finer grid
 *              (:293-314), parabolic maximum of that grid (:347-356, the
 *              z denominator as written there), radius sort (:375-380),
 *              radius where the enclosed density drops below den_odc
 *              (:406-418), tags (:421)
 *   catalogue  halo_info of physical halos (:422-506), halo_pid stream
 *              (:458), hid (:460), header (:568-570)
 * Tags are kept per copy: a particle has one tag bit per tile that holds a
 * copy of it, numbered by that tile's displacement (-1..1 per axis) from
 * the particle's own tile, which is what the reference's per-tile arrays
 * of ghost copies amount to.
 *
 * Choices where the reference leaves the result to its tools (the GPU
 * path makes the same ones):
 *   - ties: the offset table is ordered by integer r^2, then by the
 *     reference's generation order (i0, j0, k0 loops); peaks of equal
 *     density by their loop index in T's extended range; particles of equal
 *     radius by gathering order (the reference's Numerical Recipes
 *     indexedsort is unstable and not restated);
 *   - norm2 is sqrt((x*x + y*y) + z*z) in float;
 *   - real**(1./3.), log10 and 10** are evaluated in double and rounded to
 *     float;
 *   - sums over a halo's particles (x_mean, var_x, v_mean, v_disp, ang_mom,
 *     inertia, q-space sums) are accumulated in double, in particle order,
 *     and rounded to float where the reference stores them; the running
 *     inertia, dx_mean and inert totals that the reference never resets
 *     between halos are carried the same way;
 *   - refined-grid sums are 2^-40 fixed point.
 * PID decode (:465-471) takes the pid lattice size nq as a parameter, with
 * spacing nf/nq fine cells (the reference: nq = nf_global).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static const float HPI_F = 3.14159274101257324f; /* parameters.f90:67 */
#define HFIX 1099511627776.0                    /* 2^40 */
#define HNCELL 4

typedef struct {
  int64_t nc, nt, nnt, ncb, izipx, izipv;
  int64_t nc_halo_max, newbox_max, max_halo_np, search_ratio, refine_ratio, min_np, nq;
  float den_odc, cutoff, mass_p, sigma_vi, a, vsim2phys, h0, box, cen[3];
} hpar;

typedef struct { int32_t dx, dy, dz, r2; } hoff;

static int cmp_off(const void* a, const void* b) {
  const hoff* x = (const hoff*)a;
  const hoff* y = (const hoff*)b;
  if (x->r2 != y->r2) return x->r2 < y->r2 ? -1 : 1;
  /* generation order: i0 (dx) outer, j0, k0 (dz) inner */
  if (x->dx != y->dx) return x->dx < y->dx ? -1 : 1;
  if (x->dy != y->dy) return x->dy < y->dy ? -1 : 1;
  return x->dz < y->dz ? -1 : (x->dz > y->dz);
}

/* offset table of halofind.f90:69-88, sorted as described above */
static hoff* offset_table(int64_t R, int64_t* n_out) {
  int64_t cap = (2 * R + 1) * (2 * R + 1) * (2 * R + 1), n = 0;
  hoff* t = (hoff*)malloc(sizeof(hoff) * cap);
  for (int64_t i = -R; i <= R; i++)
    for (int64_t j = -R; j <= R; j++)
      for (int64_t k = -R; k <= R; k++) {
        int64_t r2 = i * i + j * j + k * k;
        if (r2 > R * R) continue;
        t[n].dx = (int32_t)i; t[n].dy = (int32_t)j; t[n].dz = (int32_t)k; t[n].r2 = (int32_t)r2;
        n++;
      }
  qsort(t, n, sizeof(hoff), cmp_off);
  *n_out = n;
  return t;
}

static inline int64_t wrapi(int64_t a, int64_t n) { a %= n; return a < 0 ? a + n : a; }
static inline int64_t rank_g(const hpar* p, int64_t gx, int64_t gy, int64_t gz) {
  int64_t nt = p->nt, nnt = p->nnt;
  int64_t t = ((gz / nt) * nnt + gy / nt) * nnt + gx / nt;
  return t * nt * nt * nt + ((gz % nt) * nt + gy % nt) * nt + gx % nt;
}
static inline uint32_t xcode(const void* xp, int64_t q, int64_t izipx) {
  return izipx == 1 ? (uint32_t)((const uint8_t*)xp)[q] : (uint32_t)((const uint16_t*)xp)[q];
}
static inline int32_t vcode(const void* vp, int64_t q, int64_t izipv) {
  return izipv == 1 ? (int32_t)((const int8_t*)vp)[q] : (int32_t)((const int16_t*)vp)[q];
}
static inline double xoff(uint32_t u, int64_t izipx) {
  return ((double)u + 0.5) * (izipx == 1 ? 1.0 / 256.0 : 1.0 / 65536.0);
}
static inline float fmodulo(float a, float p) { return a - floorf(a / p) * p; }
static inline float norm3(float x, float y, float z) { return sqrtf((x * x + y * y) + z * z); }
static inline float cbrt_ref(float x) { return (float)exp(log((double)x) * (double)(1.0f / 3.0f)); }
static inline float log10_ref(float x) { return (float)log10((double)x); }
static inline float exp10_ref(float x) { return (float)pow(10.0, (double)x); }

/* decoded velocity (halofind.f90:278-279): tan in float, the divide and the
   vfield add in double (vrel_boost is real(8)), stored in float */
static inline float hvel(const hpar* p, int32_t code, float vf) {
  float nv1 = p->izipv == 1 ? 255.0f : 65535.0f;
  float t = tanf((HPI_F * (float)code) / nv1);
  double c = (double)sqrtf(HPI_F / 2) / ((double)p->sigma_vi * 2.5);
  return (float)((double)t / c + (double)vf);
}

/* cell-centre CIC density, owner-tile frame, 2^-40 fixed point */
static void density(const hpar* p, const void* xp, const int64_t* cstart, float* rho) {
  const int64_t nt = p->nt, nnt = p->nnt, ncb = p->ncb, nc = p->nc, nf = 4 * nc, nft = 4 * nt;
  int64_t* acc = (int64_t*)calloc(nft * nft * nft, sizeof(int64_t));
  for (int64_t tz = 0; tz < nnt; tz++)
    for (int64_t ty = 0; ty < nnt; ty++)
      for (int64_t tx = 0; tx < nnt; tx++) {
        memset(acc, 0, sizeof(int64_t) * nft * nft * nft);
        int64_t T[3] = {tx, ty, tz};
        for (int64_t k = 2 - ncb; k <= nt + ncb - 1; k++)
          for (int64_t j = 2 - ncb; j <= nt + ncb - 1; j++)
            for (int64_t i = 2 - ncb; i <= nt + ncb - 1; i++) {
              int64_t c[3] = {i, j, k};
              int64_t r = rank_g(p, wrapi(tx * nt + i - 1, nc), wrapi(ty * nt + j - 1, nc), wrapi(tz * nt + k - 1, nc));
              for (int64_t ip = cstart[r]; ip < cstart[r + 1]; ip++) {
                float dx1[3], dx2[3];
                int64_t i1[3];
                for (int d = 0; d < 3; d++) {
                  float pos = (float)((double)(HNCELL * (c[d] - 1)) + (double)HNCELL * xoff(xcode(xp, 3 * ip + d, p->izipx), p->izipx) - 0.5);
                  i1[d] = (int64_t)floorf(pos) + 1;
                  dx1[d] = (float)i1[d] - pos;
                  dx2[d] = 1 - dx1[d];
                }
                const float m = p->mass_p;
                const float w[8] = {dx1[0] * dx1[1] * dx1[2] * m, dx2[0] * dx1[1] * dx1[2] * m, dx1[0] * dx2[1] * dx1[2] * m,
                                    dx1[0] * dx1[1] * dx2[2] * m, dx1[0] * dx2[1] * dx2[2] * m, dx2[0] * dx1[1] * dx2[2] * m,
                                    dx2[0] * dx2[1] * dx1[2] * m, dx2[0] * dx2[1] * dx2[2] * m};
                const int o[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 1, 1}, {1, 0, 1}, {1, 1, 0}, {1, 1, 1}};
                for (int q = 0; q < 8; q++) {
                  int64_t n0 = i1[0] + o[q][0] - 1, n1 = i1[1] + o[q][1] - 1, n2 = i1[2] + o[q][2] - 1;  /* 0-based tile node */
                  if (n0 < 0 || n0 >= nft || n1 < 0 || n1 >= nft || n2 < 0 || n2 >= nft) continue;
                  acc[(n2 * nft + n1) * nft + n0] += llrint((double)w[q] * HFIX);
                }
              }
            }
        for (int64_t k = 0; k < nft; k++)
          for (int64_t j = 0; j < nft; j++)
            for (int64_t i = 0; i < nft; i++)
              rho[((T[2] * nft + k) * nf + (T[1] * nft + j)) * nf + (T[0] * nft + i)] =
                  (float)((double)acc[(k * nft + j) * nft + i] * (1.0 / HFIX));
      }
  free(acc);
}

typedef struct {
  float den, mmesh, hpos[3];
  int64_t loop;     /* loop index in the tile's extended node range */
} hpeak;
static int cmp_peak(const void* a, const void* b) {
  const hpeak* x = (const hpeak*)a;
  const hpeak* y = (const hpeak*)b;
  if (x->den != y->den) return x->den > y->den ? -1 : 1;
  return x->loop < y->loop ? -1 : (x->loop > y->loop);
}
typedef struct { float r; int64_t g; } hrank;
static int cmp_rank(const void* a, const void* b) {
  const hrank* x = (const hrank*)a;
  const hrank* y = (const hrank*)b;
  if (x->r != y->r) return x->r < y->r ? -1 : 1;
  return x->g < y->g ? -1 : (x->g > y->g);
}

/*
 * ip: nc nt nnt ncb izipx izipv nc_halo_max newbox_max max_halo_np
 *     search_ratio refine_ratio min_halo_particles nq
 * fp: den_odc den_peak_cutoff mass_p sigma_vi a vsim2phys h0 box cen[3]
 * out: halos [max_halos][42] (type_halo_info), pidw halo_pid stream, hid
 *      (optional, nhid entries, indexed by pid-1); counts[0..3] = nhalo,
 *      pid words, candidate peaks, peaks processed
 * Returns 0, or 1 (capacity / box limits, message in err).
 */
int ohf_halofind(const int64_t* ip, const float* fp, int64_t np, const void* xp, const void* vp, const int32_t* pid,
                 const int32_t* rhoc, const float* vfield, int64_t max_halos, float* halos, int64_t max_pidw,
                 int32_t* pidw, int64_t nhid, int16_t* hid, int64_t* counts, char* err) {
  hpar P;
  P.nc = ip[0]; P.nt = ip[1]; P.nnt = ip[2]; P.ncb = ip[3]; P.izipx = ip[4]; P.izipv = ip[5];
  P.nc_halo_max = ip[6]; P.newbox_max = ip[7]; P.max_halo_np = ip[8]; P.search_ratio = ip[9];
  P.refine_ratio = ip[10]; P.min_np = ip[11]; P.nq = ip[12];
  P.den_odc = fp[0]; P.cutoff = fp[1]; P.mass_p = fp[2]; P.sigma_vi = fp[3]; P.a = fp[4]; P.vsim2phys = fp[5];
  P.h0 = fp[6]; P.box = fp[7]; P.cen[0] = fp[8]; P.cen[1] = fp[9]; P.cen[2] = fp[10];
  const int64_t nc = P.nc, nt = P.nt, nnt = P.nnt, ncb = P.ncb, nf = 4 * nc, nft = 4 * nt, nfb = ncb * HNCELL;
  const int64_t n3c = nc * nc * nc;
  int rc = 0;
  int64_t* cstart = (int64_t*)malloc(sizeof(int64_t) * (n3c + 1));
  cstart[0] = 0;
  for (int64_t r = 0; r < n3c; r++) cstart[r + 1] = cstart[r] + rhoc[r];
  float* rho = (float*)malloc(sizeof(float) * nf * nf * nf);
  density(&P, xp, cstart, rho);
  int64_t ntab;
  hoff* tab = offset_table(P.nc_halo_max, &ntab);
  uint32_t* tag = (uint32_t*)calloc(np, sizeof(uint32_t));
  const int64_t E = nft + 2 * HNCELL; /* extended node range 1-ncell..nft+ncell */
  int64_t cap_pk = 1024, npk_tot = 0, nproc = 0;
  hpeak* pk = (hpeak*)malloc(sizeof(hpeak) * cap_pk);
  int64_t cap_g = 1024;
  int64_t* gl = (int64_t*)malloc(sizeof(int64_t) * cap_g);     /* gathered particle index */
  uint32_t* gc = (uint32_t*)malloc(sizeof(uint32_t) * cap_g);  /* its image bit */
  float* gx = (float*)malloc(sizeof(float) * 6 * cap_g);       /* xv_odc */
  hrank* rk = (hrank*)malloc(sizeof(hrank) * cap_g);
  int64_t* fg = NULL;
  int64_t fg_cap = 0;
  int64_t nhalo = 0, nw = 0;
  double inertia_run[6] = {0, 0, 0, 0, 0, 0}, inert_run[6] = {0, 0, 0, 0, 0, 0};
  float dx_mean[3] = {0, 0, 0};
  const float newgrid = 1.0f / (float)P.refine_ratio;
#define RHO(gx_, gy_, gz_) rho[((wrapi(gz_, nf)) * nf + wrapi(gy_, nf)) * nf + wrapi(gx_, nf)]
  for (int64_t itz = nnt - 1; itz >= 0 && !rc; itz--)
    for (int64_t ity = nnt - 1; ity >= 0 && !rc; ity--)
      for (int64_t itx = nnt - 1; itx >= 0 && !rc; itx--) {
        const int64_t T0[3] = {itx * nft, ity * nft, itz * nft};
        /* tile-local 1-based fine node i -> global 0-based node T0 + i - 1 */
        int64_t npk = 0;
        for (int64_t k0 = 1 - HNCELL; k0 <= nft + HNCELL; k0++)
          for (int64_t j0 = 1 - HNCELL; j0 <= nft + HNCELL; j0++)
            for (int64_t i0 = 1 - HNCELL; i0 <= nft + HNCELL; i0++) {
              const int64_t g0 = T0[0] + i0 - 1, g1 = T0[1] + j0 - 1, g2 = T0[2] + k0 - 1;
              float c = RHO(g0, g1, g2), denmax = c;
              for (int dz = -1; dz <= 1; dz++)
                for (int dy = -1; dy <= 1; dy++)
                  for (int dx = -1; dx <= 1; dx++) denmax = fmaxf(denmax, RHO(g0 + dx, g1 + dy, g2 + dz));
              if (!(denmax == c && denmax > P.cutoff)) continue;
              float m_mesh = 0;
              for (int64_t ii = 1; ii <= ntab; ii++) {
                const hoff* t = &tab[ii - 1];
                const int64_t ix = i0 + t->dx, iy = j0 + t->dy, iz = k0 + t->dz;
                if (ix <= 1 - nfb || ix >= nft + nfb - 1) continue;
                if (iy <= 1 - nfb || iy >= nft + nfb - 1) continue;
                if (iz <= 1 - nfb || iz >= nft + nfb - 1) continue;
                m_mesh = m_mesh + RHO(T0[0] + ix - 1, T0[1] + iy - 1, T0[2] + iz - 1);
                if (ii < ntab && tab[ii].r2 == t->r2) continue;
                if (ii >= 19 && m_mesh / (float)ii < P.den_odc) break;
              }
              if (!(m_mesh > P.mass_p * (float)P.min_np / 2.0f)) continue;
              if (npk == cap_pk) { cap_pk *= 2; pk = (hpeak*)realloc(pk, sizeof(hpeak) * cap_pk); }
              hpeak* h = &pk[npk++];
              const float xm = RHO(g0 - 1, g1, g2), xp_ = RHO(g0 + 1, g1, g2);
              const float ym = RHO(g0, g1 - 1, g2), yp = RHO(g0, g1 + 1, g2);
              const float zm = RHO(g0, g1, g2 - 1), zp = RHO(g0, g1, g2 + 1);
              h->hpos[0] = ((float)i0 - 0.5f) + 0.5f * (xp_ - xm) / (-xm + 2 * c - xp_);
              h->hpos[1] = ((float)j0 - 0.5f) + 0.5f * (yp - ym) / (-ym + 2 * c - yp);
              h->hpos[2] = ((float)k0 - 0.5f) + 0.5f * (zp - zm) / (-zm + 2 * c - zp);
              h->den = denmax;
              h->mmesh = m_mesh;
              h->loop = ((k0 - (1 - HNCELL)) * E + (j0 - (1 - HNCELL))) * E + (i0 - (1 - HNCELL));
            }
        npk_tot += npk;
        qsort(pk, npk, sizeof(hpeak), cmp_peak);
        for (int64_t ipk = 0; ipk < npk && !rc; ipk++) {
          nproc++;
          const float mass_proxy = pk[ipk].mmesh;
          float hpos[3] = {pk[ipk].hpos[0], pk[ipk].hpos[1], pk[ipk].hpos[2]};
          const float r_proxy = cbrt_ref(0.75f / HPI_F / P.den_odc * mass_proxy);
          const float rsearch = fminf((float)P.search_ratio * r_proxy, (float)(ncb - 1) * (float)HNCELL);
          int64_t cs[3][2], cr[3][2], fr[3][2], nb[3];
          for (int d = 0; d < 3; d++) {
            cs[d][0] = (int64_t)floorf((hpos[d] - rsearch) / (float)HNCELL) + 1;
            cs[d][1] = (int64_t)floorf((hpos[d] + rsearch) / (float)HNCELL) + 1;
            cr[d][0] = (int64_t)floorf((hpos[d] - r_proxy) / (float)HNCELL) + 1;
            cr[d][1] = (int64_t)floorf((hpos[d] + r_proxy) / (float)HNCELL) + 1;
            fr[d][0] = HNCELL * (cr[d][0] - 1);
            fr[d][1] = HNCELL * cr[d][1];
            nb[d] = P.refine_ratio * (fr[d][1] - fr[d][0]);
            if (cs[d][0] < 1 - ncb || cs[d][1] > nt + ncb) { snprintf(err, 256, "halo search box outside the tile buffer"); rc = 1; }
            if (nb[d] > P.newbox_max) { snprintf(err, 256, "newbox>newbox_max"); rc = 1; }
          }
          if (rc) break;
          const int64_t D0 = nb[0] + 2, D1 = nb[1] + 2, D2 = nb[2] + 2, nfgc = D0 * D1 * D2;
          if (nfgc > fg_cap) { free(fg); fg_cap = nfgc; fg = (int64_t*)malloc(sizeof(int64_t) * fg_cap); }
          memset(fg, 0, sizeof(int64_t) * nfgc);
          int64_t n_odc = 0, ng = 0;
          for (int64_t k0 = cs[2][0]; k0 <= cs[2][1]; k0++)
            for (int64_t j0 = cs[1][0]; j0 <= cs[1][1]; j0++)
              for (int64_t i0 = cs[0][0]; i0 <= cs[0][1]; i0++) {
                const int64_t u[3] = {T0[0] / HNCELL + i0 - 1, T0[1] / HNCELL + j0 - 1, T0[2] / HNCELL + k0 - 1};
                const int64_t w[3] = {wrapi(u[0], nc), wrapi(u[1], nc), wrapi(u[2], nc)};
                /* copy bit: displacement of the particle's tile from T (each tile
                   holds its own copies of its buffer particles) */
                const int64_t td[3] = {i0 >= 1 ? (i0 - 1) / nt : -1, j0 >= 1 ? (j0 - 1) / nt : -1, k0 >= 1 ? (k0 - 1) / nt : -1};
                const uint32_t bit = 1u << (uint32_t)(((td[2] + 1) * 3 + (td[1] + 1)) * 3 + (td[0] + 1));
                const int64_t r = rank_g(&P, w[0], w[1], w[2]);
                const int64_t cc[3] = {i0, j0, k0};
                for (int64_t q = cstart[r]; q < cstart[r + 1]; q++, ng++) {
                  float pos[3];
                  for (int d = 0; d < 3; d++)
                    pos[d] = (float)((double)(HNCELL * (cc[d] - 1)) + (double)HNCELL * xoff(xcode(xp, 3 * q + d, P.izipx), P.izipx));
                  const float rr = norm3(pos[0] - hpos[0], pos[1] - hpos[1], pos[2] - hpos[2]);
                  if (!(rr < rsearch) || (tag[q] & bit)) continue;
                  if (n_odc == P.max_halo_np) { snprintf(err, 256, "np_search>max_halo_np"); rc = 1; break; }
                  if (n_odc == cap_g) {
                    cap_g *= 2;
                    gl = (int64_t*)realloc(gl, sizeof(int64_t) * cap_g);
                    gc = (uint32_t*)realloc(gc, sizeof(uint32_t) * cap_g);
                    gx = (float*)realloc(gx, sizeof(float) * 6 * cap_g);
                    rk = (hrank*)realloc(rk, sizeof(hrank) * cap_g);
                  }
                  gl[n_odc] = q;
                  gc[n_odc] = bit;
                  for (int d = 0; d < 3; d++) {
                    gx[6 * n_odc + d] = pos[d];
                    gx[6 * n_odc + 3 + d] = hvel(&P, vcode(vp, 3 * q + d, P.izipv), vfield[3 * r + d]);
                  }
                  n_odc++;
                  if (rr < r_proxy) {
                    int64_t i1[3];
                    float dx1[3], dx2[3];
                    for (int d = 0; d < 3; d++) {
                      const float pb = pos[d] - 0.5f * newgrid;
                      i1[d] = (int64_t)floorf((pb - (float)fr[d][0]) / newgrid) + 1;
                      dx1[d] = (float)i1[d] - (pb - (float)fr[d][0]) / newgrid;
                      dx2[d] = 1 - dx1[d];
                    }
                    const float wq[8] = {dx1[0] * dx1[1] * dx1[2], dx2[0] * dx1[1] * dx1[2], dx1[0] * dx2[1] * dx1[2],
                                         dx1[0] * dx1[1] * dx2[2], dx1[0] * dx2[1] * dx2[2], dx2[0] * dx1[1] * dx2[2],
                                         dx2[0] * dx2[1] * dx1[2], dx2[0] * dx2[1] * dx2[2]};
                    const int o[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 1, 1}, {1, 0, 1}, {1, 1, 0}, {1, 1, 1}};
                    for (int qq = 0; qq < 8; qq++) {
                      const int64_t a0 = i1[0] + o[qq][0], a1 = i1[1] + o[qq][1], a2 = i1[2] + o[qq][2];
                      if (a0 < 0 || a0 >= D0 || a1 < 0 || a1 >= D1 || a2 < 0 || a2 >= D2) {
                        snprintf(err, 256, "refined grid index outside 0..newbox+1");
                        rc = 1;
                        continue;
                      }
                      fg[(a2 * D1 + a1) * D0 + a0] += llrint((double)wq[qq] * HFIX);
                    }
                  }
                }
                if (rc) break;
              }
          if (rc) break;
          /* maxloc over 1..newbox_max: first (x fastest) maximum; all-zero -> (1,1,1) */
          int64_t best = -1, bi = 1, bj = 1, bk = 1;
          for (int64_t k = 1; k < D2; k++)
            for (int64_t j = 1; j < D1; j++)
              for (int64_t i = 1; i < D0; i++) {
                const int64_t v = fg[(k * D1 + j) * D0 + i];
                if (v > best) { best = v; bi = i; bj = j; bk = k; }
              }
          if (best <= 0) { bi = 1; bj = 1; bk = 1; }
#define FG(i_, j_, k_) ((i_) < D0 && (j_) < D1 && (k_) < D2 ? (float)((double)fg[((k_) * D1 + (j_)) * D0 + (i_)] * (1.0 / HFIX)) : 0.0f)
          {
            const float c = FG(bi, bj, bk);
            const float xm = FG(bi - 1, bj, bk), xq = FG(bi + 1, bj, bk);
            const float ym = FG(bi, bj - 1, bk), yq = FG(bi, bj + 1, bk);
            const float zm = FG(bi, bj, bk - 1), zq = FG(bi, bj, bk + 1);
            hpos[0] = ((float)bi - 0.5f) + 0.5f * (xq - xm) / (-xm + 2 * c - xq);
            hpos[1] = ((float)bj - 0.5f) + 0.5f * (yq - ym) / (-ym + 2 * c - yq);
            hpos[2] = ((float)bk - 0.5f) + 0.5f * (zq - zm) / (-zm + 2 * c - zm);
          }
#undef FG
          for (int d = 0; d < 3; d++) hpos[d] = (float)fr[d][0] + newgrid * hpos[d];
          const int physical = hpos[0] > 0 && hpos[1] > 0 && hpos[2] > 0 && hpos[0] < (float)nft && hpos[1] < (float)nft &&
                               hpos[2] < (float)nft;
          for (int64_t g = 0; g < n_odc; g++) {
            rk[g].r = norm3(gx[6 * g] - hpos[0], gx[6 * g + 1] - hpos[1], gx[6 * g + 2] - hpos[2]);
            rk[g].g = g;
          }
          qsort(rk, n_odc, sizeof(hrank), cmp_rank);
          int64_t i_odc = 0;
          float radius = 0;
          for (int64_t ii = 2; ii <= n_odc; ii++) {
            const float r_i = rk[ii - 1].r, r_m = rk[ii - 2].r;
            const float den2 = 0.75f / HPI_F * (float)ii * P.mass_p / (r_i * r_i * r_i);
            if (den2 < P.den_odc) {
              const float den1 = 0.75f / HPI_F * (float)(ii - 1) * P.mass_p / (r_m * r_m * r_m);
              const float r2 = log10_ref(r_i), r1 = log10_ref(r_m), d2 = log10_ref(den2), d1 = log10_ref(den1);
              const float w1 = log10_ref(P.den_odc) - d2, w2 = d1 - log10_ref(P.den_odc);
              radius = exp10_ref((w1 * r1 + w2 * r2) / (d1 - d2));
              i_odc = ii - 1;
              break;
            }
          }
          if (i_odc < P.min_np) continue;
          for (int64_t m = 0; m < i_odc; m++) tag[gl[rk[m].g]] |= gc[rk[m].g];
          if (!physical) continue;
          if (nhalo == max_halos || nw + i_odc + 2 > max_pidw) { snprintf(err, 256, "halo catalogue capacity"); rc = 1; break; }
          float* H = halos + 42 * nhalo;
          nhalo++;
          const float off[3] = {(float)T0[0], (float)T0[1], (float)T0[2]};
          const float nfr = (float)nf;
          /* layout: hpos 0-2, mass_odc 3, radius_odc 4, v_disp 5, x_mean 6-8,
             v_mean 9-11, ang_mom 12-14, var_x 15-17, inertia 18-26,
             q_mean 27-29, inertia_q 30-38, s_mean 39-41 */
          for (int d = 0; d < 3; d++) H[d] = fmodulo(hpos[d] + off[d], nfr);
          H[3] = (float)i_odc;
          H[4] = radius;
          double sx[3] = {0, 0, 0}, sv[3] = {0, 0, 0};
          for (int64_t m = 0; m < i_odc; m++)
            for (int d = 0; d < 3; d++) {
              sx[d] += gx[6 * rk[m].g + d];
              sv[d] += gx[6 * rk[m].g + 3 + d];
            }
          float xm[3], vm[3];
          for (int d = 0; d < 3; d++) {
            xm[d] = (float)(sx[d] / (double)i_odc);
            vm[d] = (float)(sv[d] / (double)i_odc);
          }
          double vx[3] = {0, 0, 0}, vd = 0;
          for (int64_t m = 0; m < i_odc; m++)
            for (int d = 0; d < 3; d++) {
              const double e = (double)(gx[6 * rk[m].g + d] - xm[d]);
              const double f = (double)(gx[6 * rk[m].g + 3 + d] - vm[d]);
              vx[d] += e * e;
              vd += f * f;
            }
          float xw[3];
          for (int d = 0; d < 3; d++) {
            H[15 + d] = (float)(vx[d] / (double)(i_odc - 1));
            xw[d] = fmodulo(xm[d] + off[d], nfr);
            H[6 + d] = xw[d];
            H[9 + d] = vm[d];
          }
          H[5] = sqrtf((float)(vd / (double)(i_odc - 1)));
          {
            float los[3];
            for (int d = 0; d < 3; d++) los[d] = xw[d] - (P.cen[d] / 500.0f) * (float)nc;
            const float ln = norm3(los[0], los[1], los[2]);
            for (int d = 0; d < 3; d++) los[d] = los[d] / ln;
            const float dot = (vm[0] * los[0] + vm[1] * los[1]) + vm[2] * los[2];
            for (int d = 0; d < 3; d++) {
              float s = dot * los[d];
              s = s * P.vsim2phys / P.a / (100 * P.h0);
              s = s / (P.h0 * P.box / nfr);
              H[39 + d] = fmodulo(xw[d] + s, nfr);
            }
          }
          double am[3] = {0, 0, 0}, in[6] = {0, 0, 0, 0, 0, 0};
          for (int64_t m = 0; m < i_odc; m++) {
            const float* X = gx + 6 * rk[m].g;
            const double e0 = X[0] - (xw[0] - off[0]), e1 = X[1] - (xw[1] - off[1]), e2 = X[2] - (xw[2] - off[2]);
            const double v0 = X[3], v1 = X[4], v2 = X[5];
            am[0] += e1 * v2 - e2 * v1;
            am[1] += e2 * v0 - e0 * v2;
            am[2] += e0 * v1 - e1 * v0;
            in[0] += e0 * e0; in[1] += e1 * e1; in[2] += e2 * e2;
            in[3] += e0 * e1; in[4] += e1 * e2; in[5] += e2 * e0;
          }
          for (int d = 0; d < 3; d++) H[12 + d] = (float)am[d];
          for (int c = 0; c < 6; c++) inertia_run[c] += in[c];
          /* inertia(3,3) column-major: (1,1) (2,1) (3,1) (1,2) (2,2) (3,2) (1,3) (2,3) (3,3) */
#define SYM(dst, v)                                                         \
  do {                                                                      \
    (dst)[0] = (float)(v)[0]; (dst)[1] = (float)(v)[3]; (dst)[2] = (float)(v)[5]; \
    (dst)[3] = (float)(v)[3]; (dst)[4] = (float)(v)[1]; (dst)[5] = (float)(v)[4]; \
    (dst)[6] = (float)(v)[5]; (dst)[7] = (float)(v)[4]; (dst)[8] = (float)(v)[2]; \
  } while (0)
          SYM(H + 18, inertia_run);
          /* q-space (PID) */
          {
            const int64_t nq = P.nq;
            const float sp = (float)nf / (float)nq, half = (float)(nf / 2);
            double sdx[3] = {0, 0, 0};
            int64_t nptemp = 0;
            for (int64_t m = 0; m < i_odc; m++) {
              const int64_t qt = (int64_t)pid[gl[rk[m].g]] - 1;
              if (qt < 0) continue;
              nptemp++;
              const int64_t qi[3] = {qt % nq, (qt / nq) % nq, qt / (nq * nq)};
              for (int d = 0; d < 3; d++) {
                const float qp = (float)qi[d] * sp + 0.5f;
                const float dx = fmodulo(qp - xw[d] + half, nfr) - half;
                sdx[d] += dx;
              }
            }
            float qmean[3];
            for (int d = 0; d < 3; d++) {
              dx_mean[d] = (float)(((double)dx_mean[d] + sdx[d]) / (double)nptemp);
              qmean[d] = fmodulo(xw[d] + dx_mean[d], nfr);
              H[27 + d] = qmean[d];
            }
            double iq[6] = {0, 0, 0, 0, 0, 0};
            for (int64_t m = 0; m < i_odc; m++) {
              const int64_t qt = (int64_t)pid[gl[rk[m].g]] - 1;
              if (qt < 0) continue;
              const int64_t qi[3] = {qt % nq, (qt / nq) % nq, qt / (nq * nq)};
              double e[3];
              for (int d = 0; d < 3; d++) {
                const float qp = (float)qi[d] * sp + 0.5f;
                e[d] = fmodulo(qp + 0.5f - qmean[d] + half, nfr) - half;
              }
              iq[0] += e[0] * e[0]; iq[1] += e[1] * e[1]; iq[2] += e[2] * e[2];
              iq[3] += e[0] * e[1]; iq[4] += e[1] * e[2]; iq[5] += e[2] * e[0];
            }
            for (int c = 0; c < 6; c++) inert_run[c] += iq[c];
            SYM(H + 30, inert_run);
          }
#undef SYM
          pidw[nw++] = (int32_t)i_odc;
          for (int64_t m = 0; m < i_odc; m++) {
            const int32_t pv = pid[gl[rk[m].g]];
            pidw[nw++] = pv;
            if (hid && pv >= 1 && pv <= nhid) hid[pv - 1] = (int16_t)nhalo;
          }
          pidw[nw++] = 0;
        }
      }
#undef RHO
  counts[0] = nhalo; counts[1] = nw; counts[2] = npk_tot; counts[3] = nproc;
  free(cstart); free(rho); free(tab); free(tag); free(pk); free(gl); free(gc); free(gx); free(rk); free(fg);
  return rc;
}

/* density and peak stages alone, for the tests that pin them */
int ohf_density(const int64_t* ip, const float* fp, const void* xp, const int32_t* rhoc, float* rho) {
  hpar P;
  memset(&P, 0, sizeof(P));
  P.nc = ip[0]; P.nt = ip[1]; P.nnt = ip[2]; P.ncb = ip[3]; P.izipx = ip[4]; P.izipv = ip[5];
  P.mass_p = fp[2];
  const int64_t n3c = P.nc * P.nc * P.nc;
  int64_t* cstart = (int64_t*)malloc(sizeof(int64_t) * (n3c + 1));
  cstart[0] = 0;
  for (int64_t r = 0; r < n3c; r++) cstart[r + 1] = cstart[r] + rhoc[r];
  density(&P, xp, cstart, rho);
  free(cstart);
  return 0;
}
int64_t ohf_offset_table(int64_t R, int32_t* out, int64_t cap) {
  int64_t n;
  hoff* t = offset_table(R, &n);
  for (int64_t i = 0; i < n && i < cap; i++) {
    out[4 * i] = t[i].dx; out[4 * i + 1] = t[i].dy; out[4 * i + 2] = t[i].dz; out[4 * i + 3] = t[i].r2;
  }
  free(t);
  return n;
}
