// Internal declarations shared by the HIP translation units of libcubehip.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <mutex>
#include <shared_mutex>
#include <string>
#include <vector>
#include "cube_hip.h"

#define CUBE_NCELL 4
#define CUBE_NF_CUTOFF 16

// pi = 4*atan(1.) in single precision (parameters.f90:67)
#define PI_F 3.14159274101257324f

// ---------------------------------------------------------------- complex
template <typename T>
struct cpx { T x, y; };

template <typename T>
__host__ __device__ inline cpx<T> cmul(cpx<T> a, cpx<T> b) {
  return cpx<T>{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}

// ---------------------------------------------------------------- mesh
// One periodic N^3 mesh with its R2C/C2R work buffers and the imaginary
// part of its force kernel stored over one octant of k-space.
struct Mesh {
  int64_t N = 0;
  int prec = 4;          // sizeof(real)
  void* R = nullptr;     // real N*N*(N+2) padded rows; complex view N*N*(N/2+1)
  void* W = nullptr;     // complex work, same size as R
  void* S = nullptr;     // unpadded N^3 real mesh (fine: force component 0)
  void* X = nullptr;     // (fine) second spectrum / real work mesh, bytes_big()
  void* K = nullptr;     // kernel [3][H][H][H], layout [d][ky'][kx][kz'], scaled 1/N^3
  void* tw = nullptr;    // N twiddles exp(-2 pi i m/N), type cpx<real>
  size_t bytes_R() const { return (size_t)N * N * (N + 2) * prec; }
  // the larger of the padded real mesh and the kx-tiled spectrum
  // (ceil((N/2+1)/16)*16 complex per (ky, kz))
  size_t bytes_big() const {
    const size_t t = (size_t)((N / 2 + 1 + 15) / 16) * 16 * N * N * 2 * prec;
    return t > bytes_R() ? t : bytes_R();
  }
};

// Per-kernel timing with HIP events on the handle's stream (diagnostic).
enum ProfSlot {
  PS_DRIFT1 = 0, PS_DRIFT_SCAN, PS_DRIFT_SCATTER, PS_DRIFT_SORT, PS_DRIFT_WRITE, PS_FLUID, PS_PP, PS_CIC_C, PS_FFT_C, PS_CIC_F, PS_FFT_XF, PS_FFT_YF,
  PS_FFT_ZCONV, PS_FFT_YI, PS_FFT_XI, PS_KICK, PS_FLUID_DENS, PS_FLUID_GRAV, PS_KICK_C, PS_NSLOT
};
struct ProfRec { int slot; hipEvent_t a, b; };

// Neutrino dispersive fluid: three fluids of nf/2, nf/4, nf/8 cells per
// dimension, fields [5][nc^3] (rho, mx, my, mz, E) and gravity accumulators
// [4][nc^3] (J0, J1, J2, Q) in double.
struct NuState {
  bool on = false;
  cube_nu_params par{};
  int nc[3] = {0, 0, 0};
  float dx[3] = {0, 0, 0}, cs2[3] = {0, 0, 0}, w[3] = {0, 0, 0}, ai[3] = {0, 0, 0};
  float* F[3] = {nullptr, nullptr, nullptr};
  double* acc[3] = {nullptr, nullptr, nullptr};
};

// Device results of one step (drift, PP, PM), gathered on the device and
// read by the host with one copy: the deferred step modes launch the whole
// step without a host synchronisation and check it afterwards.
struct StepReport {
  uint32_t dflags[8];     // drift: max displacement per axis, too far / crowded, frame conflicts
  int64_t npnew;          // drift: particles after the drift
  double dsum[3];         // drift: sum_c, sum_r, sum_s
  float sigma_new;        // drift: sigma_vi_new (update_particle.f90:213), formed on the device
  int32_t pad0;
  uint32_t ppflags[12];   // PP: uflags[8..19] (dense blocks, work items, pool fallbacks, ...)
  uint64_t pp_pairs, pp_targets;
  uint32_t pp_f2;         // PP: max |F|^2 bits over the slots
  int32_t pad1;
  double mass[4];         // PM: testrhoc, its fluid share, testrhof, its fluid share
  uint32_t pmflags[8];    // PM: uflags[32..39] (f2 coarse, f2 fine, vmax[3])
};

struct cube_handle {
  cube_params p;
  int64_t nc, nt, nnt, ncb, nf, n3c;
  int64_t np = 0, npglobal = 0;
  int prec;
  hipStream_t stream = nullptr;
  // resident particle state (checkpoint order) and its double buffer
  void* xp = nullptr; void* vp = nullptr; int32_t* pid = nullptr;
  void* xp2 = nullptr; void* vp2 = nullptr; int32_t* pid2 = nullptr;
  uint16_t* disp = nullptr;   // packed cell displacement per particle (drift)
  int32_t* order = nullptr;   // 8*np_max bytes: drift arrival keys (uint64), PP overflow lists
  int32_t* dbig = nullptr;    // drift: cells with more than 1024 arrivals
  void* dscratch = nullptr;   // drift scratch, 24*np_max bytes (aliases fine.W when it fits)
  bool dscratch_own = false;
  int32_t* rhoc = nullptr; float* vfield = nullptr; int64_t* cstart = nullptr;
  int32_t* rhoc2 = nullptr; float* vfield2 = nullptr; int64_t* cstart2 = nullptr;
  // codec tables
  float* vdec = nullptr;      // tanf((pi*vp)/(nvbin-1)) for every code
  double* vdt = nullptr;      // two decoded-velocity tables (vdec/c), see vtab_for
  double vdt_c[2] = {0, 0};
  int vdt_next = 0;
  double* venc = nullptr;     // encode thresholds (see codec.cpp)
  int32_t vkmax = 0;          // largest |code| produced by the encoder
  // meshes
  Mesh fine, coarse;
  void* fc = nullptr;         // coarse force [3][nc^3] (real type of prec)
  // reductions / flags
  double* dred = nullptr;     // double partials
  int64_t dred_n = 0;
  double* dred_l1 = nullptr;  // first-level sums of large reductions
  // per-workgroup maxima, ORs and counts spread over MSLOTS slots, so that
  // millions of workgroups do not meet on one address; slots_reduce folds
  // them into the flags
  uint32_t* mslots = nullptr;
  uint32_t* uflags = nullptr; // small device scratch for maxima and flags
  int64_t* scan_tmp = nullptr;
  int64_t scan_tmp_n = 0;
  // step report (device) and its pinned host copy; the new-sigma velocity
  // coefficient and decode table, formed on the device by the drift and
  // read by the kick
  StepReport* drep = nullptr;
  StepReport* rep = nullptr;
  double* dscal = nullptr;    // [0] vcoef(sigma_vi_new)
  double* vdt_new = nullptr;  // decoded-velocity table for dscal[0]
  // step mode: 0 each call synchronises; 1 drift, PP and PM are launched
  // back to back and checked after one synchronisation; 2 as 1, the
  // launches captured into a hipGraph and replayed (kernel arguments
  // updated in place)
  int step_mode = 0;
  bool deferred = false;      // inside a deferred step: phases do not synchronise
  hipGraphExec_t gexec = nullptr;
  int64_t graph_updates = 0, graph_instantiations = 0;
  // scalars mirrored from CUBE's module state
  float sigma_vi = 0, sigma_vi_new = 0, vz_max = 0, mass_p = 0;
  float ftemp_pp = 0;         // ext_pp_force.f90:18 running max (SAVE)
  uint32_t pp_overflow_blocks = 0;  // PP blocks handled by the chunked fallback in the last call
  void* pp_meta = nullptr;    // PP dense-block lists and work items
  void* pp_red = nullptr;     // PP per-slot totals (1024 slots)
  NuState nu;
  cube_sim_header sim{};
  std::string err;
  // profiling
  bool prof_on = false;
  std::vector<ProfRec> prof_recs;
  std::vector<hipEvent_t> prof_pool;
  double prof_ms[PS_NSLOT] = {0};
  int64_t prof_n[PS_NSLOT] = {0};
};

hipEvent_t prof_event(cube_handle* h);
static inline int prof_begin(cube_handle* h, int slot) {
  if (!h->prof_on) return -1;
  ProfRec r{slot, prof_event(h), prof_event(h)};
  (void)hipEventRecord(r.a, h->stream);
  h->prof_recs.push_back(r);
  return (int)h->prof_recs.size() - 1;
}
static inline void prof_end(cube_handle* h, int id) {
  if (id >= 0) (void)hipEventRecord(h->prof_recs[id].b, h->stream);
}

// ------------------------------------------------------------- errors
int cube_fail(cube_handle* h, int code, const char* fmt, ...);
#define HIP_TRY(h, call)                                                        \
  do {                                                                           \
    hipError_t e_ = (call);                                                      \
    if (e_ != hipSuccess)                                                        \
      return cube_fail((h), CUBE_ERR_HIP, "%s failed: %s (%s:%d)", #call,      \
                       hipGetErrorString(e_), __FILE__, __LINE__);              \
  } while (0)

// ------------------------------------------------------------- codec (host)
int codec_build(cube_handle* h);  // vdec/venc tables from the host libm
// dscal[0] and vdt_new from the host's sigma_vi_new (after an upload, an IC
// or a failed drift); the drift forms them on the device otherwise
int sigma_new_to_device(cube_handle* h);
// fill n 32-bit words with v on the handle's stream (a kernel, so that a
// captured step holds no memset node)
int fill_u32(cube_handle* h, void* p, int64_t n, uint32_t v);
#define MSLOTS 1024
#define MSLOT_FIELDS 8
// out[f] (f < nf) combined with the fold of field f over the MSLOTS slots
// (slot s, field f at slots[s * MSLOT_FIELDS + f]): OR where bit f of
// or_mask is set, sum where bit f of add_mask is, max otherwise
int slots_reduce(cube_handle* h, int nf, uint32_t or_mask, uint32_t add_mask, uint32_t* out);
// While one thread captures a step into a hipGraph, no other thread may run
// a legacy-stream operation (synchronous copies and memsets) or a
// device-synchronising one (hipFree): the runtime rejects them.  Captures
// hold this lock shared; the entry points that make such calls (init,
// finalize, up/download, ICs, probes, P(k), halo finder) hold it exclusively.
std::shared_mutex& capture_mutex();
#define CUBE_EXCLUSIVE std::unique_lock<std::shared_mutex> capture_lock_(capture_mutex())

// ------------------------------------------------------------- device helpers
struct Geo {  // box geometry passed by value to kernels
  int32_t nc, nt, nnt, ncb, izipx, izipv;
  int32_t lnt, lnnt;  // log2 of nt and nnt when they are powers of two, else -1
};

// Workgroups go to the 8 XCDs round-robin by linear id and each XCD has its
// own L2.  xcd_remap gives hardware block L of T a logical id such that
// every XCD runs one contiguous range of logical ids: blocks that share
// cache lines then run on the same XCD at about the same time.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t L, uint32_t T) {
  const uint32_t q = T >> 3, r = T & 7, x = L & 7, sl = L >> 3;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + sl;
}

// Cell indices stay below 2^27 (nc <= 512), so index arithmetic is 32-bit;
// nc is a power of two and so, in the usual layouts, are nt and nnt.
__host__ __device__ __forceinline__ int32_t wrapc(int32_t a, int32_t n) { return a & (n - 1); }
__host__ __device__ __forceinline__ void divmod_nt(const Geo& g, uint32_t a, uint32_t& q, uint32_t& r) {
  if (g.lnt >= 0) { q = a >> g.lnt; r = a & ((1u << g.lnt) - 1); }
  else { q = a / (uint32_t)g.nt; r = a - q * (uint32_t)g.nt; }
}
// checkpoint-order rank of global coarse cell (0-based coords)
__host__ __device__ __forceinline__ int64_t rank_of(const Geo& g, int32_t x, int32_t y, int32_t z) {
  uint32_t tx, lx, ty, ly, tz, lz;
  divmod_nt(g, (uint32_t)x, tx, lx);
  divmod_nt(g, (uint32_t)y, ty, ly);
  divmod_nt(g, (uint32_t)z, tz, lz);
  const uint32_t nt = (uint32_t)g.nt, nnt = (uint32_t)g.nnt;
  const uint32_t t = (tz * nnt + ty) * nnt + tx;
  return (int64_t)(t * (nt * nt * nt) + (lz * nt + ly) * nt + lx);
}
__host__ __device__ __forceinline__ void cell_of_rank(const Geo& g, int64_t r64, int32_t* c) {
  const uint32_t r = (uint32_t)r64, nt = (uint32_t)g.nt, nnt = (uint32_t)g.nnt;
  uint32_t t, l, tx, ty, tz, lx, ly, lz;
  if (g.lnt >= 0 && g.lnnt >= 0) {
    const uint32_t a = g.lnt, b = g.lnnt;
    t = r >> (3 * a); l = r & ((1u << (3 * a)) - 1);
    tx = t & ((1u << b) - 1); ty = (t >> b) & ((1u << b) - 1); tz = t >> (2 * b);
    lx = l & (nt - 1); ly = (l >> a) & (nt - 1); lz = l >> (2 * a);
  } else {
    const uint32_t nt3 = nt * nt * nt;
    t = r / nt3; l = r - t * nt3;
    tx = t % nnt; ty = (t / nnt) % nnt; tz = t / (nnt * nnt);
    lx = l % nt; ly = (l / nt) % nt; lz = l / (nt * nt);
  }
  c[0] = (int32_t)(tx * nt + lx);
  c[1] = (int32_t)(ty * nt + ly);
  c[2] = (int32_t)(tz * nt + lz);
}
__device__ __forceinline__ uint32_t load_xu(const void* xp, int64_t q, int izipx) {
  return izipx == 1 ? (uint32_t)(uint8_t)((const int8_t*)xp)[q] : (uint32_t)(uint16_t)((const int16_t*)xp)[q];
}
__device__ __forceinline__ int32_t load_v(const void* vp, int64_t q, int izipv) {
  return izipv == 1 ? (int32_t)((const int8_t*)vp)[q] : (int32_t)((const int16_t*)vp)[q];
}
__device__ __forceinline__ void store_v(void* vp, int64_t q, int izipv, int32_t v) {
  if (izipv == 1) ((int8_t*)vp)[q] = (int8_t)v; else ((int16_t*)vp)[q] = (int16_t)v;
}
__device__ __forceinline__ void store_x(void* xp, int64_t q, int izipx, int64_t v) {
  if (izipx == 1) ((int8_t*)xp)[q] = (int8_t)(uint8_t)(uint64_t)v;
  else ((int16_t*)xp)[q] = (int16_t)(uint16_t)(uint64_t)v;
}
// (int(xp+ishift,izipx)+rshift)*x_resolution == (u+0.5)/2^(8*izipx)
__device__ __forceinline__ double xoff(uint32_t u, int izipx) {
  return ((double)u + 0.5) * (izipx == 1 ? (1.0 / 256.0) : (1.0 / 65536.0));
}
__device__ __forceinline__ int32_t xsub(uint32_t u, int izipx) {  // fine sub-cell 0..3
  return izipx == 1 ? (int32_t)(u >> 6) : (int32_t)(u >> 14);
}
__device__ __forceinline__ double nint_d(double y) {
  return y >= 0 ? floor(y + 0.5) : -floor(-y + 0.5);
}
// velocity decode (update_particle.f90:69, pm.f90:145): a table per sigma
// holding (double)tanf(pi*v/(nvbin-1)) / c for every code v, filled on the
// device by the reference's own double division (vtab_for)
__device__ __forceinline__ double vdecode(const double* dtab, int32_t v, int izipv) {
  return dtab[v + (izipv == 1 ? 128 : 32768)];
}
// atan(j/32) for j = 0..32, correctly rounded (gen_atan_table.py prints
// it): the breakpoints of the encoder's atan.  Kernels that encode many
// values copy it to LDS (atan_tab_to_lds).
static __constant__ double kAtanTab[33] = {
    0.0, 0.031239833430268277, 0.06241880999595735, 0.09347678115858947,
    0.12435499454676144, 0.15499674192394097, 0.18534794999569476, 0.21535769969773805,
    0.24497866312686414, 0.2741674511196588, 0.3028848683749714, 0.3310960767041321,
    0.35877067027057225, 0.38588266939807375, 0.4124104415973873, 0.43833655985795783,
    0.4636476090008061, 0.48833395105640554, 0.5123894603107377, 0.5358112379604637,
    0.5585993153435624, 0.5807563535676704, 0.6022873461349642, 0.6231993299340659,
    0.6435011087932844, 0.6632029927060933, 0.6823165548747481, 0.7008544078844502,
    0.7188299996216245, 0.7362574289814281, 0.7531512809621944, 0.7695264804056583,
    0.7853981633974483};
__device__ __forceinline__ void atan_tab_to_lds(double* s) {
  if (threadIdx.x < 33) s[threadIdx.x] = kAtanTab[threadIdx.x];
}

// velocity encode, equal to the host encoder's code for every argument.
// thr[k + kmax - 1] is the smallest x with enc(x) >= k, for k in
// [-kmax+1, kmax]; venc_fast (below) reads it only near ties.  at is
// kAtanTab or a copy of it.
__device__ __forceinline__ int32_t venc_fast(const double* thr, const double* at, int32_t kmax, int izipv, double x);
__device__ __forceinline__ int32_t vencode(const double* thr, int32_t kmax, int izipv, double c, double v) {
  return venc_fast(thr, kAtanTab, kmax, izipv, c * v);
}

// The exact table encoder, in two stages: the FP32 atan guess (within one
// code of the answer), then the correction from the guess's two thresholds
// (lo = thr[n+kmax-1], hi = thr[n+kmax], or -inf/+inf at the ends) and,
// rarely, further thresholds; the correction only ever moves in one
// direction.
__device__ __forceinline__ int32_t venc_guess(int32_t kmax, int izipv, double x) {
  const float nv1 = izipv == 1 ? 255.0f : 65535.0f;
  const float y = nv1 * atanf((float)x) * (1.0f / PI_F);
  int32_t n = (int32_t)rintf(y);
  if (n > kmax) n = kmax;
  if (n < -kmax) n = -kmax;
  return n;
}
__device__ __forceinline__ double venc_lo(const double* thr, int32_t kmax, int32_t n) {
  return n > -kmax ? thr[n + kmax - 1] : -HUGE_VAL;
}
__device__ __forceinline__ double venc_hi(const double* thr, int32_t kmax, int32_t n) {
  return n < kmax ? thr[n + kmax] : HUGE_VAL;
}
__device__ __forceinline__ int32_t venc_finish(const double* thr, int32_t kmax, double x, int32_t n, double lo, double hi) {
  if (x < lo) {
    n--;
    for (int it = 1; it < 4 && n > -kmax && x < thr[n + kmax - 1]; it++) n--;
  } else if (x >= hi) {
    n++;
    for (int it = 1; it < 4 && n < kmax && x >= thr[n + kmax]; it++) n++;
  }
  return n;
}

// atan for the encoder below.  |x| > 1 is reduced to pi/2 - atan(1/|x|),
// so u = min(|x|, 1/|x|) is in [0, 1]; with the nearest breakpoint
// c = j/32, atan(u) = atan(c) + atan(h), h = (u - c)/(1 + u*c), |h| <= 1/64,
// and atan(h) = h - h^3/3 + h^5/5 - h^7/7 leaves out less than 1e-17.
// Reciprocals come from the hardware estimate and two Newton steps.  The
// error in nv1*atan(x)/pi stays below 2e-11 (4e7 arguments checked against
// the host libm).
__device__ __forceinline__ double rcp_nr(double d) {
  double q = __builtin_amdgcn_rcp(d);
  double e = fma(-d, q, 1.0);
  q = fma(q, e, q);
  e = fma(-d, q, 1.0);
  return fma(q, e, q);
}
__device__ __forceinline__ double atan_enc(double x, const double* at) {
  const double a = fabs(x);
  const bool big = a > 1.0;
  const double u = big ? rcp_nr(fmin(a, 1e300)) : a;  // the bound keeps the Newton steps finite for |x| = inf
  int j = (int)(u * 32.0 + 0.5);
  j = j < 0 ? 0 : (j > 32 ? 32 : j);
  const double c = (double)j * 0.03125;
  const double h = (u - c) * rcp_nr(fma(u, c, 1.0));
  const double h2 = h * h;
  double t = fma(h2, -1.0 / 7.0, 1.0 / 5.0);
  t = fma(t, h2, -1.0 / 3.0);
  t = t * h2;
  double r = at[j] + fma(h, t, h);
  r = big ? 1.5707963267948966 - r : r;
  return copysign(r, x);
}

// Encode without the threshold tables for all but near-ties: the host
// encoder rounds y = nv1*atan(x)/pi (FP64 atan, pi in real(4)) to the
// nearest integer.  Here y comes from atan_enc and a product with
// nv1/pi; it differs from the host's y by less than 1e-10 (|y| < 2^15), so
// whenever y is more than 1e-8 from a half-integer both round to the same
// code, and only the rest (about 2e-8 of the arguments) settle it with the
// exact host thresholds.
__device__ __forceinline__ int32_t venc_fast(const double* thr, const double* at, int32_t kmax, int izipv, double x) {
  const double k = izipv == 1 ? 255.0 / (double)PI_F : 65535.0 / (double)PI_F;
  const double y = atan_enc(x, at) * k;
  const double r = rint(y);
  if (fabs(y - r) < 0.5 - 1e-8) {
    int32_t n = (int32_t)r;
    if (n > kmax) n = kmax;
    if (n < -kmax) n = -kmax;
    return n;
  }
  const int32_t g = venc_guess(kmax, izipv, x);
  return venc_finish(thr, kmax, x, g, venc_lo(thr, kmax, g), venc_hi(thr, kmax, g));
}

// ------------------------------------------------------------- launchers
// decoded-velocity table for coefficient c (one of two cached slots)
const double* vtab_for(cube_handle* h, double c);
// scan.hip
int scan_counts(cube_handle* h, const int32_t* cnt, int64_t n, int64_t* cstart);
int reduce_doubles(cube_handle* h, const double* part, int64_t nparts, int nfields, double* host_out);
// the same fixed-order sums into device memory, no synchronisation
int reduce_doubles_dev(cube_handle* h, const double* part, int64_t nparts, int nfields, double* dev_out);
// fft.hip
int mesh_alloc(cube_handle* h, Mesh& m, int64_t N, int prec, bool with_S, bool with_K);
int mesh_pass(cube_handle* h, Mesh& m, void* buf, int axis, bool inverse);    // strided y (1) / z (2) pass
int mesh_c2r(cube_handle* h, Mesh& m, const void* src, void* dst, int64_t opitch);
void mesh_free(Mesh& m);
int mesh_fwd_xy(cube_handle* h, Mesh& m);                  // R: X r2c then Y, in place
// dst (R or W) = Zinv(i*K_d * Zfwd(R))
int mesh_zconv(cube_handle* h, Mesh& m, int d, void* dst);
// buf (R or W): Y inverse in place, then X c2r into out (row pitch opitch);
// with f2max, also max(F0^2+F1^2+F2^2) from the other two components
int mesh_inv_yx(cube_handle* h, Mesh& m, void* buf, void* out, int64_t opitch, const void* F0, int64_t p0,
                const void* F1, int64_t p1, uint32_t* f2max);
int mesh_fft3_r2c(cube_handle* h, Mesh& m);                  // full forward on R (for kernel build)
// kx-tiled spectrum pipeline (fine mesh, per step): real rows src (pitch
// N+2) -> tiled spectrum dst with the y transform; z convolution src ->
// dst (may alias); y inverse in place then C2R into out (pitch opitch)
int mesh_t_fwd(cube_handle* h, Mesh& m, const void* src, void* dst);
int mesh_t_zconv(cube_handle* h, Mesh& m, int d, const void* src, void* dst);
// all three components of the fine convolution from one read of src
// (FP32 meshes of 32^3 and up: mesh_t_zconv3_fits)
bool mesh_t_zconv3_fits(const Mesh& m);
int mesh_t_zconv3(cube_handle* h, Mesh& m, const void* src, void* const dst[3]);
int mesh_t_inv(cube_handle* h, Mesh& m, void* buf, void* out, int64_t opitch);
int fine_kernel_build(cube_handle* h, const float* table);
int coarse_kernel_build(cube_handle* h, const float* table);
// pm.hip
// mass: host destination of the deposited mass (synchronises), or mass_dev:
// device destination (no synchronisation)
int cic_fine(cube_handle* h, double* mass, double* mass_dev = nullptr);
int pk_scratch_setup(cube_handle* h);  // device address of the P(k) scratch array on h's GPU
int cic_coarse(cube_handle* h, double* mass, double* mass_dev = nullptr);
// fine force components Ff[d] (row pitch pf[d]) and the coarse force fc
// f2max_bits (optional): max |F|^2 over all fine nodes, from the staged slabs
// new-sigma coefficient and table from h->dscal / h->vdt_new
int kick_all(cube_handle* h, const void* const Ff[3], const int64_t pf[3], float a_mid, float dt, double c_old,
             uint32_t* vmax_bits, uint32_t* f2max_bits, double* probe_f, double* probe_c);
// pp.hip
int pp_run(cube_handle* h, float a_mid, float dt, double c_old, cube_pp_out* out, double* probe);
int pp_launch_step(cube_handle* h, float a_mid, float dt, double c_old, double* probe);
int pp_finish(cube_handle* h, const StepReport* rp, cube_pp_out* out);
int pp_step_finish(cube_handle* h, float a_mid, cube_pp_out& o);  // ftemp running max, dt_pp, itest1 check
// PM launch sequence (results into h->drep) and its outputs, dt inputs and sigma update
int pm_launch_step(cube_handle* h, float a_mid, float dt);
int pm_finish_step(cube_handle* h, const StepReport* rp, float a_mid, cube_pm_out* out);
// drift.hip
int drift_run(cube_handle* h, float dt_old, float dt, cube_drift_out* out);
// the drift's launch sequence (results into h->drep, state buffers swapped)
// and its checks on a copied report
int drift_launch_step(cube_handle* h, float dt_old, float dt, int* nlayer);
int drift_finish(cube_handle* h, const StepReport* rp, int nlayer, cube_drift_out* out);
// fluid.hip: PM coupling of the neutrino fluid
int fl_density_coarse(cube_handle* h, float a_mid, double* mass_added, double* mass_dev = nullptr);
int fl_density_fine(cube_handle* h, float a_mid, double* mass_added, double* mass_dev = nullptr);
int fl_gravity_fine(cube_handle* h, const void* const Ff[3], const int64_t pf[3], float a_mid, float dt);
// reuses the fine-force buffers Ff for the coarse acceleration at fine-cell centres
int fl_gravity_finish(cube_handle* h, void* const Ff[3], const int64_t pf[3], float a_mid, float dt);
void fl_free(NuState& s);

static inline Geo geo_of(const cube_handle* h) {
  Geo g;
  g.nc = (int32_t)h->nc; g.nt = (int32_t)h->nt; g.nnt = (int32_t)h->nnt; g.ncb = (int32_t)h->ncb;
  g.izipx = (int32_t)h->p.izipx; g.izipv = (int32_t)h->p.izipv;
  auto lg = [](int64_t n) { int l = 0; while (((int64_t)1 << l) < n) l++; return ((int64_t)1 << l) == n ? l : -1; };
  g.lnt = lg(h->nt); g.lnnt = lg(h->nnt);
  return g;
}
static inline double vcoef(float sigma) {  // sqrt(pi/2)/(sigma*vrel_boost)
  return (double)sqrtf(PI_F / 2) / ((double)sigma * 2.5);
}
