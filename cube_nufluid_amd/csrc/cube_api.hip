// C ABI of libcubehip.so: handle lifecycle, state transfer and the step
// entry points that replace CUBE's update_x / ext_pp_force / particle_mesh.
#include "cube_internal.hpp"
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <algorithm>
#include <mutex>
#include <vector>

int cube_fail(cube_handle* h, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (h) h->err = buf;
  return code;
}

// ------------------------------------------------------------------ codec
// Decoding (update_particle.f90:69) evaluates tan in single precision on a
// finite set of codes, so it is tabulated here with the host libm tanf.
// Encoding (update_particle.f90:119) is nint(nv1*atan(x)/pi) with a double
// atan of x = c*v; it is a non-decreasing step function of x, so the host
// computes, for every code k, the smallest double x with enc(x) >= k by
// bisection over the ordered bit patterns.  The device evaluates its own
// atan and corrects the result against these thresholds, which makes its
// codes equal to the host libm's for every input.
static inline int32_t enc_host(double x, double nv1) {
  double y = nv1 * atan(x) / (double)PI_F;
  return (int32_t)(y >= 0 ? floor(y + 0.5) : -floor(-y + 0.5));
}
static inline uint64_t d2key(double x) {
  uint64_t u;
  memcpy(&u, &x, 8);
  return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
static inline double key2d(uint64_t k) {
  uint64_t u = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
  double x;
  memcpy(&x, &u, 8);
  return x;
}
struct CodecTables {
  std::vector<float> dec;
  std::vector<double> thr;
  int32_t kmax = 0;
};
static const CodecTables& codec_tables(int izipv) {
  static CodecTables tabs[3];
  static std::once_flag once[3];
  std::call_once(once[izipv], [izipv]() {
    CodecTables& t = tabs[izipv];
    const int64_t nvbin = (int64_t)1 << (8 * izipv);
    const float nv1f = (float)(nvbin - 1);
    const double nv1 = (double)nv1f;
    t.dec.resize(nvbin);
    for (int64_t v = -nvbin / 2; v < nvbin / 2; v++) t.dec[v + nvbin / 2] = tanf((PI_F * (float)v) / nv1f);
    t.kmax = enc_host(INFINITY, nv1);
    t.thr.resize(2 * t.kmax);
    for (int32_t k = -t.kmax + 1; k <= t.kmax; k++) {
      double x0 = tan(((double)k - 0.5) * (double)PI_F / nv1);
      double span = fabs(x0) * 1e-9 + 1e-300;
      double lo = x0 - span, hi = x0 + span;
      while (enc_host(lo, nv1) >= k) { span *= 16; lo = x0 - span; }
      while (enc_host(hi, nv1) < k) { span *= 16; hi = x0 + span; }
      uint64_t a = d2key(lo), b = d2key(hi);  // enc(a) < k <= enc(b)
      while (b - a > 1) {
        uint64_t m = a + (b - a) / 2;
        if (enc_host(key2d(m), nv1) >= k) b = m; else a = m;
      }
      t.thr[k + t.kmax - 1] = key2d(b);
    }
  });
  return tabs[izipv];
}
__global__ void k_fill_u32(uint32_t* p, int64_t n, uint32_t v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ __launch_bounds__(256) void k_slots_reduce(const uint32_t* __restrict__ slots, int nf, uint32_t or_mask,
                                                      uint32_t add_mask, uint32_t* out) {
  __shared__ uint32_t red[256];
  for (int f = 0; f < nf; f++) {
    const bool is_or = (or_mask >> f) & 1, is_add = (add_mask >> f) & 1;
    uint32_t a = 0;
    for (int s = threadIdx.x; s < MSLOTS; s += 256) {
      const uint32_t v = slots[s * MSLOT_FIELDS + f];
      a = is_or ? (a | v) : (is_add ? a + v : max(a, v));
    }
    red[threadIdx.x] = a;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) {
        const uint32_t b = red[threadIdx.x + o];
        red[threadIdx.x] = is_or ? (red[threadIdx.x] | b) : (is_add ? red[threadIdx.x] + b : max(red[threadIdx.x], b));
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      if (is_or) atomicOr(out + f, red[0]);
      else if (is_add) atomicAdd(out + f, red[0]);
      else atomicMax(out + f, red[0]);
    }
    __syncthreads();
  }
}
int slots_reduce(cube_handle* h, int nf, uint32_t or_mask, uint32_t add_mask, uint32_t* out) {
  k_slots_reduce<<<1, 256, 0, h->stream>>>(h->mslots, nf, or_mask, add_mask, out);
  HIP_TRY(h, hipGetLastError());
  return 0;
}
int fill_u32(cube_handle* h, void* p, int64_t n, uint32_t v) {
  const int64_t nb = std::min<int64_t>((n + 255) / 256, 4096);
  k_fill_u32<<<(unsigned)std::max<int64_t>(nb, 1), 256, 0, h->stream>>>((uint32_t*)p, n, v);
  HIP_TRY(h, hipGetLastError());
  return 0;
}
__global__ void k_vtab_fill(const float* __restrict__ dec, int n, double c, double* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (double)dec[i] / c;
}
const double* vtab_for(cube_handle* h, double c) {
  const int n = h->p.izipv == 1 ? 256 : 65536;
  // two slots, least recently used replaced: the kick needs the old- and
  // new-sigma tables at once
  for (int s = 0; s < 2; s++)
    if (h->vdt_c[s] == c) {
      h->vdt_next = s ^ 1;
      return h->vdt + (size_t)s * n;
    }
  const int s = h->vdt_next;
  h->vdt_next = s ^ 1;
  k_vtab_fill<<<(n + 255) / 256, 256, 0, h->stream>>>(h->vdec, n, c, h->vdt + (size_t)s * n);
  h->vdt_c[s] = c;
  return h->vdt + (size_t)s * n;
}
int codec_build(cube_handle* h) {
  const CodecTables& t = codec_tables((int)h->p.izipv);
  HIP_TRY(h, hipMalloc(&h->vdec, sizeof(float) * t.dec.size()));
  HIP_TRY(h, hipMalloc(&h->venc, sizeof(double) * t.thr.size()));
  HIP_TRY(h, hipMalloc(&h->vdt, sizeof(double) * 2 * t.dec.size()));
  HIP_TRY(h, hipMemcpy(h->vdec, t.dec.data(), sizeof(float) * t.dec.size(), hipMemcpyHostToDevice));
  HIP_TRY(h, hipMemcpy(h->venc, t.thr.data(), sizeof(double) * t.thr.size(), hipMemcpyHostToDevice));
  h->vkmax = t.kmax;
  HIP_TRY(h, hipMalloc(&h->vdt_new, sizeof(double) * t.dec.size()));
  HIP_TRY(h, hipMalloc(&h->dscal, sizeof(double) * 4));
  HIP_TRY(h, hipMalloc(&h->drep, sizeof(StepReport)));
  HIP_TRY(h, hipMalloc(&h->dred_l1, sizeof(double) * 256 * 8));  // first-level sums (scan.hip)
  HIP_TRY(h, hipMalloc(&h->mslots, sizeof(uint32_t) * MSLOTS * MSLOT_FIELDS));
  HIP_TRY(h, hipHostMalloc(&h->rep, sizeof(StepReport), hipHostMallocDefault));
  HIP_TRY(h, hipMemset(h->drep, 0, sizeof(StepReport)));
  memset(h->rep, 0, sizeof(StepReport));
  return 0;
}

// ------------------------------------------------------------ lifecycle
static bool is_pow2(int64_t n) { return n > 0 && (n & (n - 1)) == 0; }

extern "C" int cube_init(const cube_params* p, const float* kern_f_table, const float* kern_c_table, cube_handle** out) {
  CUBE_EXCLUSIVE;
  if (!p || !kern_f_table || !kern_c_table || !out) return CUBE_ERR_ARG;
  *out = nullptr;
  cube_handle* h = new cube_handle();
  h->p = *p;
  h->nc = p->nc;
  h->nnt = p->nnt;
  h->ncb = p->ncb;
  int rc = 0;
  auto fail = [&](int code) { *out = h; return code; };
  if (!is_pow2(p->nc) || p->nc < 16 || p->nc > 512)
    return fail(cube_fail(h, CUBE_ERR_ARG, "nc=%lld must be a power of two in [16,512]", (long long)p->nc));
  if (p->nnt < 1 || p->nc % p->nnt)
    return fail(cube_fail(h, CUBE_ERR_ARG, "nnt=%lld must divide nc=%lld", (long long)p->nnt, (long long)p->nc));
  h->nt = p->nc / p->nnt;
  if (h->nt % 8) return fail(cube_fail(h, CUBE_ERR_ARG, "nt=%lld must be a multiple of 8", (long long)h->nt));
  // the fine force of a node depends on the density within 15 fine cells
  // (kernel_f.f90's 16^3 table); pm.f90:79-82 deposits the particles of
  // coarse cells 2-ncb..nt+ncb-1, so a tile's density is complete that far
  // out only for 4*(ncb-1) >= 16, and only then does the one periodic fine
  // mesh used here give the reference's per-tile force
  if (p->ncb < 5 || p->ncb > h->nt)
    return fail(cube_fail(h, CUBE_ERR_ARG, "ncb=%lld outside [5, nt]: the fine force needs 4*(ncb-1) >= 16 buffer cells",
                          (long long)p->ncb));
  if ((p->izipx != 1 && p->izipx != 2) || (p->izipv != 1 && p->izipv != 2))
    return fail(cube_fail(h, CUBE_ERR_ARG, "izipx/izipv must be 1 or 2"));
  if (p->precision != 4 && p->precision != 8)
    return fail(cube_fail(h, CUBE_ERR_ARG, "precision=%d: 4 (FP32 production path) or 8 (FP64 oracle)", p->precision));
  if (p->np_max < 1 || p->np_max >= ((int64_t)1 << 31))
    return fail(cube_fail(h, CUBE_ERR_ARG, "np_max=%lld outside [1, 2^31)", (long long)p->np_max));
  h->prec = p->precision;
  h->nf = h->nc * CUBE_NCELL;
  h->n3c = h->nc * h->nc * h->nc;
  if (hipSetDevice(p->device) != hipSuccess) return fail(cube_fail(h, CUBE_ERR_HIP, "hipSetDevice(%d) failed", p->device));
  if ((rc = pk_scratch_setup(h))) return fail(rc);
#define TRY(call) do { hipError_t e_ = (call); if (e_ != hipSuccess) return fail(cube_fail(h, CUBE_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_))); } while (0)
  TRY(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  const int64_t npm = p->np_max;
  TRY(hipMalloc(&h->xp, (size_t)3 * npm * p->izipx));
  TRY(hipMalloc(&h->vp, (size_t)3 * npm * p->izipv));
  TRY(hipMalloc(&h->pid, sizeof(int32_t) * npm));
  TRY(hipMalloc(&h->xp2, (size_t)3 * npm * p->izipx));
  TRY(hipMalloc(&h->vp2, (size_t)3 * npm * p->izipv));
  TRY(hipMalloc(&h->pid2, sizeof(int32_t) * npm));
  TRY(hipMalloc(&h->disp, sizeof(uint16_t) * npm));
  TRY(hipMalloc(&h->order, sizeof(int32_t) * 2 * npm));
  TRY(hipMalloc(&h->dbig, sizeof(int32_t) * (npm / 1025 + 2)));
  {  // PP: per 2^3-cell block 2 int64 info + 2 int32 lists, plus work items
    const int64_t nb = (h->nc / 2) * (h->nc / 2) * (h->nc / 2);
    TRY(hipMalloc(&h->pp_meta, sizeof(int64_t) * (2 * nb + npm / 256 + nb + 1) + sizeof(int32_t) * 2 * nb));
    TRY(hipMalloc(&h->pp_red, 1024 * (2 * sizeof(uint64_t) + sizeof(uint32_t))));
  }
  TRY(hipMalloc(&h->rhoc, sizeof(int32_t) * h->n3c));
  TRY(hipMalloc(&h->rhoc2, sizeof(int32_t) * h->n3c));
  TRY(hipMalloc(&h->vfield, sizeof(float) * 3 * h->n3c));
  TRY(hipMalloc(&h->vfield2, sizeof(float) * 3 * h->n3c));
  TRY(hipMalloc(&h->cstart, sizeof(int64_t) * (h->n3c + 1)));
  TRY(hipMalloc(&h->cstart2, sizeof(int64_t) * (h->n3c + 1)));
  TRY(hipMalloc(&h->fc, (size_t)3 * h->n3c * h->prec));
  h->dred_n = std::max<int64_t>(h->n3c, 1 << 16) + 16;
  TRY(hipMalloc(&h->dred, sizeof(double) * h->dred_n));
  TRY(hipMalloc(&h->uflags, sizeof(uint32_t) * 64));
  TRY(hipMemset(h->uflags, 0, sizeof(uint32_t) * 64));
#undef TRY
  if ((rc = codec_build(h))) return fail(rc);
  if ((rc = mesh_alloc(h, h->fine, h->nf, h->prec, true, true))) return fail(rc);
  if ((rc = mesh_alloc(h, h->coarse, h->nc, h->prec, false, true))) return fail(rc);
  // the fine work mesh is idle during the drift and serves as its scratch
  // (crowded-cell merges, then three doubles per particle)
  if (h->fine.bytes_R() >= 3 * sizeof(double) * (size_t)npm) {
    h->dscratch = h->fine.W;
  } else {
    if (hipMalloc(&h->dscratch, 3 * sizeof(double) * (size_t)npm) != hipSuccess)
      return fail(cube_fail(h, CUBE_ERR_HIP, "hipMalloc of the drift scratch failed"));
    h->dscratch_own = true;
  }
  if ((rc = fine_kernel_build(h, kern_f_table))) return fail(rc);
  if ((rc = coarse_kernel_build(h, kern_c_table))) return fail(rc);
  *out = h;
  return CUBE_OK;
}

hipEvent_t prof_event(cube_handle* h) {
  hipEvent_t e = nullptr;
  if (!h->prof_pool.empty()) {
    e = h->prof_pool.back();
    h->prof_pool.pop_back();
  } else {
    (void)hipEventCreate(&e);
  }
  return e;
}

extern "C" int cube_profile_enable(cube_handle* h, int on) {
  if (!h) return CUBE_ERR_ARG;
  h->prof_on = on != 0;
  return CUBE_OK;
}

static const char* kSlotNames[PS_NSLOT] = {"drift_pass1", "drift_scan", "drift_scatter", "drift_sort_big", "drift_write",
                                           "fluid", "pp", "cic_coarse",
                                           "fft_coarse", "cic_fine", "fft_x_r2c", "fft_y_fwd", "fft_zconv",
                                           "fft_y_inv", "fft_x_c2r", "kick_fine", "fluid_density", "fluid_gravity",
                                           "kick_coarse"};

extern "C" int cube_profile_read(cube_handle* h, double* ms, int64_t* count, int nslot) {
  if (!h || nslot < 0) return CUBE_ERR_ARG;
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  for (ProfRec& r : h->prof_recs) {
    float t = 0;
    if (hipEventElapsedTime(&t, r.a, r.b) == hipSuccess) {
      h->prof_ms[r.slot] += t;
      h->prof_n[r.slot] += 1;
    }
    h->prof_pool.push_back(r.a);
    h->prof_pool.push_back(r.b);
  }
  h->prof_recs.clear();
  for (int i = 0; i < nslot && i < PS_NSLOT; i++) {
    if (ms) ms[i] = h->prof_ms[i];
    if (count) count[i] = h->prof_n[i];
    h->prof_ms[i] = 0;
    h->prof_n[i] = 0;
  }
  return PS_NSLOT;
}

extern "C" int cube_profile_slot_name(int slot, char* buf, size_t n) {
  if (slot < 0 || slot >= PS_NSLOT || !buf || n == 0) return CUBE_ERR_ARG;
  snprintf(buf, n, "%s", kSlotNames[slot]);
  return CUBE_OK;
}

extern "C" int cube_finalize(cube_handle* h) {
  CUBE_EXCLUSIVE;
  if (!h) return CUBE_OK;
  for (ProfRec& r : h->prof_recs) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  for (hipEvent_t e : h->prof_pool) (void)hipEventDestroy(e);
  if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
  if (h->rep) (void)hipHostFree(h->rep);
  void* bufs[] = {h->xp, h->vp, h->pid, h->xp2, h->vp2, h->pid2, h->disp, h->order, h->rhoc, h->rhoc2, h->vfield, h->vfield2,
                  h->cstart, h->cstart2, h->fc, h->dred, h->uflags, h->scan_tmp, h->vdec, h->venc, h->vdt, h->dbig,
                  h->dscratch_own ? h->dscratch : nullptr, h->pp_meta, h->pp_red, h->vdt_new, h->dscal, h->drep, h->dred_l1, h->mslots};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  mesh_free(h->fine);
  mesh_free(h->coarse);
  fl_free(h->nu);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return CUBE_OK;
}

extern "C" int cube_last_error(cube_handle* h, char* buf, size_t n) {
  if (!buf || n == 0) return CUBE_ERR_ARG;
  const char* s = h ? h->err.c_str() : "";
  snprintf(buf, n, "%s", s);
  return CUBE_OK;
}

// ------------------------------------------------------------------ state
extern "C" int cube_upload_state(cube_handle* h, const cube_sim_header* sim, const void* xp, const void* vp,
                                 const int32_t* pid, const int32_t* rhoc, const float* vfield) {
  CUBE_EXCLUSIVE;
  if (!h || !sim || !rhoc || !vfield || (sim->nplocal > 0 && (!xp || !vp || !pid))) return CUBE_ERR_ARG;
  if (sim->izipx != h->p.izipx || sim->izipv != h->p.izipv)
    return cube_fail(h, CUBE_ERR_FORMAT, "zip format incompatible: header %lld/%lld, handle %lld/%lld",
                     (long long)sim->izipx, (long long)sim->izipv, (long long)h->p.izipx, (long long)h->p.izipv);
  if (sim->nplocal < 0 || sim->nplocal > h->p.np_max)
    return cube_fail(h, CUBE_ERR_CAPACITY, "nplocal=%lld exceeds np_max=%lld", (long long)sim->nplocal, (long long)h->p.np_max);
  if (hipSetDevice(h->p.device) != hipSuccess) return cube_fail(h, CUBE_ERR_HIP, "hipSetDevice failed");
  const int64_t np = sim->nplocal;
  HIP_TRY(h, hipMemcpy(h->xp, xp, (size_t)3 * np * h->p.izipx, hipMemcpyHostToDevice));
  HIP_TRY(h, hipMemcpy(h->vp, vp, (size_t)3 * np * h->p.izipv, hipMemcpyHostToDevice));
  HIP_TRY(h, hipMemcpy(h->pid, pid, sizeof(int32_t) * np, hipMemcpyHostToDevice));
  HIP_TRY(h, hipMemcpy(h->rhoc, rhoc, sizeof(int32_t) * h->n3c, hipMemcpyHostToDevice));
  HIP_TRY(h, hipMemcpy(h->vfield, vfield, sizeof(float) * 3 * h->n3c, hipMemcpyHostToDevice));
  int r = scan_counts(h, h->rhoc, h->n3c, h->cstart);
  if (r) return r;
  int64_t tot;
  HIP_TRY(h, hipMemcpyAsync(&tot, h->cstart + h->n3c, sizeof(int64_t), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  if (tot != np) return cube_fail(h, CUBE_ERR_FORMAT, "sum(rhoc)=%lld != nplocal=%lld", (long long)tot, (long long)np);
  h->np = np;
  h->npglobal = sim->npglobal > 0 ? sim->npglobal : np;
  h->sim = *sim;
  h->sigma_vi = sim->sigma_vi;
  h->sigma_vi_new = sim->sigma_vi;
  h->vz_max = sim->vz_max;
  h->mass_p = sim->mass_p_cdm;
  return sigma_new_to_device(h);
}

extern "C" int cube_download_state(cube_handle* h, cube_sim_header* sim, void* xp, void* vp, int32_t* pid, int32_t* rhoc,
                                   float* vfield) {
  CUBE_EXCLUSIVE;
  if (!h) return CUBE_ERR_ARG;
  if (hipSetDevice(h->p.device) != hipSuccess) return cube_fail(h, CUBE_ERR_HIP, "hipSetDevice failed");
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  if (xp) HIP_TRY(h, hipMemcpy(xp, h->xp, (size_t)3 * h->np * h->p.izipx, hipMemcpyDeviceToHost));
  if (vp) HIP_TRY(h, hipMemcpy(vp, h->vp, (size_t)3 * h->np * h->p.izipv, hipMemcpyDeviceToHost));
  if (pid) HIP_TRY(h, hipMemcpy(pid, h->pid, sizeof(int32_t) * h->np, hipMemcpyDeviceToHost));
  if (rhoc) HIP_TRY(h, hipMemcpy(rhoc, h->rhoc, sizeof(int32_t) * h->n3c, hipMemcpyDeviceToHost));
  if (vfield) HIP_TRY(h, hipMemcpy(vfield, h->vfield, sizeof(float) * 3 * h->n3c, hipMemcpyDeviceToHost));
  if (sim) {
    *sim = h->sim;
    sim->nplocal = h->np;
    sim->npglobal = h->npglobal;
    sim->sigma_vi = h->sigma_vi;
    sim->vz_max = h->vz_max;
    sim->mass_p_cdm = h->mass_p;
    sim->izipx = h->p.izipx;
    sim->izipv = h->p.izipv;
  }
  return CUBE_OK;
}

// ------------------------------------------------------------------ steps
extern "C" int cube_drift(cube_handle* h, float dt_old, float dt, cube_drift_out* out) {
  if (!h) return CUBE_ERR_ARG;
  if (hipSetDevice(h->p.device) != hipSuccess) return cube_fail(h, CUBE_ERR_HIP, "hipSetDevice failed");
  return drift_run(h, dt_old, dt, out);
}

// dt_pp of ext_pp_force.f90:118-128 from the PP results
int pp_step_finish(cube_handle* h, float a_mid, cube_pp_out& o) {
  // 'ftemp' in ext_pp_force.f90:18,121 is SAVEd and never reset: dt_pp uses
  // the running maximum over every tile of every call.
  h->ftemp_pp = std::max(h->ftemp_pp, o.f2_max_pp);
  const float GG = 1.0f / 6.0f / PI_F;
  o.f2_max_pp = h->ftemp_pp;
  o.dt_pp = 1.0f * sqrtf(1.0f) / (sqrtf(h->ftemp_pp) * a_mid * GG);
  h->sim.dt_pp = o.dt_pp;
  if (o.itest1 != h->np)
    return cube_fail(h, CUBE_ERR_CHECK, "itest1=%lld /= nplocal=%lld", (long long)o.itest1, (long long)h->np);
  return CUBE_OK;
}

extern "C" int cube_pp(cube_handle* h, float a_mid, float dt, cube_pp_out* out) {
  if (!h) return CUBE_ERR_ARG;
  if (hipSetDevice(h->p.device) != hipSuccess) return cube_fail(h, CUBE_ERR_HIP, "hipSetDevice failed");
  cube_pp_out o{};
  int r = pp_run(h, a_mid, dt, vcoef(h->sigma_vi), &o, nullptr);
  if (r) return r;
  r = pp_step_finish(h, a_mid, o);
  if (out) *out = o;
  return r;
}

// particle_mesh (pm.f90:2-443) as one launch sequence; the masses, f2
// maxima and vmax go to h->drep for pm_finish
static int pm_launch(cube_handle* h, float a_mid, float dt, double* probe_f, double* probe_c) {
  int r;
  uint32_t* f2c = h->uflags + 32;
  uint32_t* f2f = h->uflags + 33;
  uint32_t* vmx = h->uflags + 34;  // 3 entries
  StepReport* dr = h->drep;
  if ((r = fill_u32(h, h->uflags + 32, 8, 0u))) return r;
  const bool nu = h->nu.on;
  // coarse mesh: force_c for all three components (pm.f90:188-326); the
  // third inverse also forms max |force_c|^2 from the two already written
  if ((r = cic_coarse(h, nullptr, &dr->mass[0]))) return r;
  int prd = nu ? prof_begin(h, PS_FLUID_DENS) : -1;
  if (nu && (r = fl_density_coarse(h, a_mid, nullptr, &dr->mass[1]))) return r;
  if (nu) prof_end(h, prd);
  if ((r = mesh_fwd_xy(h, h->coarse))) return r;
  const int64_t Nc = h->coarse.N;
  const char* fc = (const char*)h->fc;
  const size_t fcs = (size_t)h->n3c * h->prec;
  for (int d = 0; d < 3; d++) {
    if ((r = mesh_zconv(h, h->coarse, d, h->coarse.W))) return r;
    if ((r = mesh_inv_yx(h, h->coarse, h->coarse.W, (void*)(fc + d * fcs), Nc, fc, Nc, fc + fcs, Nc,
                         d == 2 ? f2c : nullptr)))
      return r;
  }
  // fine mesh, kx-tiled spectrum: the density in R goes to a tiled
  // spectrum in W.  FP32 meshes of 32^3 and up: one z pass convolves all
  // three components into X, R and W (the spectrum's last use), and each is
  // inverted into the real mesh the previous inversion left free: X -> S
  // (F0), R -> X (F1), W -> R (F2).  Otherwise components 0 and 1 are
  // convolved into X and inverted into S (F0) and R (F1), component 2 in
  // place in W and inverted into X (F2).  All three stay resident for the
  // fine kick.
  if ((r = cic_fine(h, nullptr, &dr->mass[2]))) return r;
  prd = nu ? prof_begin(h, PS_FLUID_DENS) : -1;
  if (nu && (r = fl_density_fine(h, a_mid, nullptr, &dr->mass[3]))) return r;
  if (nu) prof_end(h, prd);
  if ((r = mesh_t_fwd(h, h->fine, h->fine.R, h->fine.W))) return r;
  const int64_t Nf = h->fine.N;
  const int64_t pf[3] = {Nf, Nf + 2, Nf + 2};
  void* Fw[3];
  if (mesh_t_zconv3_fits(h->fine)) {
    void* const spec[3] = {h->fine.X, h->fine.R, h->fine.W};
    Fw[0] = h->fine.S, Fw[1] = h->fine.X, Fw[2] = h->fine.R;
    if ((r = mesh_t_zconv3(h, h->fine, h->fine.W, spec))) return r;
    for (int d = 0; d < 3; d++)
      if ((r = mesh_t_inv(h, h->fine, spec[d], Fw[d], pf[d]))) return r;
  } else {
    Fw[0] = h->fine.S, Fw[1] = h->fine.R, Fw[2] = h->fine.X;
    for (int d = 0; d < 3; d++) {
      void* spec = d == 2 ? h->fine.W : h->fine.X;
      if ((r = mesh_t_zconv(h, h->fine, d, h->fine.W, spec))) return r;
      if ((r = mesh_t_inv(h, h->fine, spec, Fw[d], pf[d]))) return r;
    }
  }
  const void* Ff[3] = {Fw[0], Fw[1], Fw[2]};
  // the kick decodes with sigma_vi and encodes with sigma_vi_new, whose
  // coefficient the drift left in h->dscal; it also forms max |F|^2
  // (dt_fine) from the force slabs it stages
  if ((r = kick_all(h, Ff, pf, a_mid, dt, vcoef(h->sigma_vi), vmx, f2f, probe_f, probe_c))) return r;
  const bool nu_kick = nu && !probe_f;  // probes leave every state unchanged
  const int prg = nu_kick ? prof_begin(h, PS_FLUID_GRAV) : -1;
  if (nu_kick && (r = fl_gravity_fine(h, Ff, pf, a_mid, dt))) return r;
  if (nu_kick && (r = fl_gravity_finish(h, Fw, pf, a_mid, dt))) return r;
  if (nu_kick) prof_end(h, prg);
  HIP_TRY(h, hipMemcpyAsync(dr->pmflags, h->uflags + 32, sizeof(dr->pmflags), hipMemcpyDeviceToDevice, h->stream));
  return 0;
}

static void pm_outputs(cube_handle* h, const StepReport* rp, cube_pm_out* out) {
  float f[5];
  memcpy(f, rp->pmflags, sizeof(f));
  out->f2_max_coarse = f[0];
  out->f2_max_fine = f[1];
  out->vmax[0] = f[2]; out->vmax[1] = f[3]; out->vmax[2] = f[4];
  // testrhoc / testrhof include the fluid's mass once it is coupled
  out->mass_c = rp->mass[0] + (h->nu.on ? rp->mass[1] : 0.0);
  out->mass_f = rp->mass[2] + (h->nu.on ? rp->mass[3] : 0.0);
}

static int pm_forces(cube_handle* h, float a_mid, float dt, cube_pm_out* out, double* probe_f, double* probe_c) {
  int r = pm_launch(h, a_mid, dt, probe_f, probe_c);
  if (r) return r;
  HIP_TRY(h, hipMemcpyAsync(h->rep, h->drep, sizeof(StepReport), hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  if (out) pm_outputs(h, h->rep, out);
  return 0;
}

// the time-step inputs of pm.f90:427-443 from the PM results; sigma_vi
// takes the new value (pm.f90:180)
static void pm_finish(cube_handle* h, float a_mid, cube_pm_out& o) {
  h->sigma_vi = h->sigma_vi_new;
  const float GG = 1.0f / 6.0f / PI_F;
  o.dt_fine = sqrtf(1.0f / (sqrtf(o.f2_max_fine) * a_mid * GG));
  o.dt_coarse = sqrtf((float)CUBE_NCELL / (sqrtf(o.f2_max_coarse) * a_mid * GG));
  float vm = std::max(o.vmax[0], std::max(o.vmax[1], o.vmax[2]));
  o.dt_vmax = 0.9f * 20 / vm;
  o.vz_max = o.vmax[2];
  h->vz_max = o.vz_max;
  h->sim.dt_fine = o.dt_fine;
  h->sim.dt_coarse = o.dt_coarse;
  h->sim.dt_vmax = o.dt_vmax;
  h->sim.vz_max = o.vz_max;
}

int pm_launch_step(cube_handle* h, float a_mid, float dt) { return pm_launch(h, a_mid, dt, nullptr, nullptr); }
int pm_finish_step(cube_handle* h, const StepReport* rp, float a_mid, cube_pm_out* out) {
  cube_pm_out o{};
  pm_outputs(h, rp, &o);
  pm_finish(h, a_mid, o);
  if (out) *out = o;
  return CUBE_OK;
}

extern "C" int cube_pm(cube_handle* h, float a_mid, float dt, cube_pm_out* out) {
  if (!h) return CUBE_ERR_ARG;
  if (hipSetDevice(h->p.device) != hipSuccess) return cube_fail(h, CUBE_ERR_HIP, "hipSetDevice failed");
  cube_pm_out o{};
  int r = pm_forces(h, a_mid, dt, &o, nullptr, nullptr);
  if (r) return r;
  pm_finish(h, a_mid, o);
  if (out) *out = o;
  return CUBE_OK;
}

// Each operation is evaluated in the precision the reference's declarations
// give it: default-real literals and parameters (4.0/9.0, 1-omega_m-omega_l,
// the wde factors), integer powers of the real(4) dt_x and the sum a0+da1
// are float operations widened afterwards; the rest is double, left to
// right; x**3 of a real(8) is (x*x)*x, x**(real) is pow().
extern "C" int cube_expansion(float a0, float dt0, float omega_m, float omega_l, float wde, float* da1, float* da2) {
  if (!da1 || !da2) return CUBE_ERR_ARG;
  const float dt_x = dt0 / 2;
  const double omHsq = (double)(4.0f / 9.0f);
  const double dt2 = (double)(dt_x * dt_x), dt3 = (double)((dt_x * dt_x) * dt_x);
  const double ekm = (double)((1.0f - omega_m) - omega_l);
  const double w1 = (double)(1.5f * (1.0f - wde));
  const double w2 = (double)((1.5f * (2.0f - 3.0f * wde)) * (1.0f - wde));
  const double ew = (double)(-(3.0f * wde));
  double a_x = a0;
  for (int half = 0; half < 2; half++) {
    double a3rlm = pow(a_x, ew) * (double)omega_l / (double)omega_m;
    double arkm = a_x * ekm / (double)omega_m;
    double adot = sqrt(omHsq * ((a_x * a_x) * a_x) * (1.0 + arkm + a3rlm));
    double addot = (a_x * a_x) * omHsq * (1.5 + 2.0 * arkm + w1 * a3rlm);
    double atdot = a_x * adot * omHsq * (3.0 + 6.0 * arkm + w2 * a3rlm);
    float da = (float)(adot * (double)dt_x + (addot * dt2) / 2.0 + (atdot * dt3) / 6.0);
    if (half == 0) {
      *da1 = da;
      a_x = (double)(a0 + da);
    } else {
      *da2 = da;
    }
  }
  return CUBE_OK;
}

// ------------------------------------------------------------- probes
extern "C" int cube_probe_pm_forces(cube_handle* h, double* f_fine, double* f_coarse) {
  CUBE_EXCLUSIVE;
  if (!h || !f_fine || !f_coarse) return CUBE_ERR_ARG;
  if (hipSetDevice(h->p.device) != hipSuccess) return cube_fail(h, CUBE_ERR_HIP, "hipSetDevice failed");
  double *pf, *pc;
  HIP_TRY(h, hipMalloc(&pf, sizeof(double) * 3 * std::max<int64_t>(h->np, 1)));
  HIP_TRY(h, hipMalloc(&pc, sizeof(double) * 3 * std::max<int64_t>(h->np, 1)));
  int r = pm_forces(h, 1.0f, 1.0f, nullptr, pf, pc);
  if (!r) {
    HIP_TRY(h, hipMemcpy(f_fine, pf, sizeof(double) * 3 * h->np, hipMemcpyDeviceToHost));
    HIP_TRY(h, hipMemcpy(f_coarse, pc, sizeof(double) * 3 * h->np, hipMemcpyDeviceToHost));
  }
  (void)hipFree(pf);
  (void)hipFree(pc);
  return r;
}

extern "C" int cube_probe_pp_forces(cube_handle* h, double* f_pp) {
  CUBE_EXCLUSIVE;
  if (!h || !f_pp) return CUBE_ERR_ARG;
  if (hipSetDevice(h->p.device) != hipSuccess) return cube_fail(h, CUBE_ERR_HIP, "hipSetDevice failed");
  double* pf;
  HIP_TRY(h, hipMalloc(&pf, sizeof(double) * 3 * std::max<int64_t>(h->np, 1)));
  cube_pp_out o{};
  int r = pp_run(h, 1.0f, 1.0f, vcoef(h->sigma_vi), &o, pf);
  if (!r) HIP_TRY(h, hipMemcpy(f_pp, pf, sizeof(double) * 3 * h->np, hipMemcpyDeviceToHost));
  (void)hipFree(pf);
  return r;
}

__global__ void k_probe_venc(const double* thr, int32_t kmax, int izipv, const double* x, int64_t n, int32_t* out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = venc_fast(thr, kAtanTab, kmax, izipv, x[i]);
}
extern "C" int cube_probe_vencode(cube_handle* h, const double* x, int64_t n, int32_t* codes) {
  CUBE_EXCLUSIVE;
  if (!h || n < 0 || (n > 0 && (!x || !codes))) return CUBE_ERR_ARG;
  if (n == 0) return CUBE_OK;
  if (hipSetDevice(h->p.device) != hipSuccess) return cube_fail(h, CUBE_ERR_HIP, "hipSetDevice failed");
  double* dx = nullptr;
  int32_t* dc = nullptr;
  hipError_t e = hipMalloc(&dx, sizeof(double) * n);
  if (e == hipSuccess) e = hipMalloc(&dc, sizeof(int32_t) * n);
  if (e == hipSuccess) e = hipMemcpy(dx, x, sizeof(double) * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    const int64_t nb = std::min<int64_t>((n + 255) / 256, 4096);
    k_probe_venc<<<(unsigned)nb, 256, 0, h->stream>>>(h->venc, h->vkmax, (int)h->p.izipv, dx, n, dc);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e == hipSuccess) e = hipMemcpy(codes, dc, sizeof(int32_t) * n, hipMemcpyDeviceToHost);
  hipFree(dx);
  hipFree(dc);
  if (e != hipSuccess) return cube_fail(h, CUBE_ERR_HIP, hipGetErrorString(e));
  return CUBE_OK;
}
extern "C" int cube_probe_density(cube_handle* h, float* rho) {
  CUBE_EXCLUSIVE;
  if (!h || !rho) return CUBE_ERR_ARG;
  if (hipSetDevice(h->p.device) != hipSuccess) return cube_fail(h, CUBE_ERR_HIP, "hipSetDevice failed");
  int r = cic_fine(h, nullptr);
  if (r) return r;
  const int64_t N = h->fine.N;
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  if (h->prec == 4) {
    HIP_TRY(h, hipMemcpy2D(rho, sizeof(float) * N, h->fine.R, sizeof(float) * (N + 2), sizeof(float) * N, N * N,
                           hipMemcpyDeviceToHost));
  } else {
    std::vector<double> row((size_t)N * N * N);
    HIP_TRY(h, hipMemcpy2D(row.data(), sizeof(double) * N, h->fine.R, sizeof(double) * (N + 2), sizeof(double) * N, N * N,
                           hipMemcpyDeviceToHost));
    for (size_t i = 0; i < row.size(); i++) rho[i] = (float)row[i];
  }
  return CUBE_OK;
}
