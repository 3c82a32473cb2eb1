// Native host loop and ensemble serving.
//
// cube_clock_init / cube_step: the body of CUBE's main loop
// (main.f90:48-97) with timestep.f90:8-95 on the host in real(4), so any C
// caller can advance a resident box without a Python or Fortran driver.
//
// cube_ens_*: a serving engine for ensembles of independent boxes of uneven
// size.  Requests are admitted as soon as their device footprint fits the
// memory budget and a slot is free (continuous batching: a request that
// arrives while others run starts at once when it fits, and a finished box
// frees its memory for the queue).  Each live box owns a handle, its HIP
// stream and a host thread that runs its steps, so the kernels of all live
// boxes overlap on the device.  Every box's arithmetic is fixed by its own
// inputs (CIC and PP sums are order-independent), so a box gives the same
// bits alone and in the mix.
#include "cube_internal.hpp"
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <string.h>
#include <thread>
#include <vector>

std::shared_mutex& capture_mutex() {
  static std::shared_mutex mu;
  return mu;
}

// ------------------------------------------------------------------ clock
extern "C" int cube_clock_init(cube_clock* c, const float* z_checkpoint, int n) {
  if (!c || !z_checkpoint || n < 1 || n > CUBE_MAX_CHECKPOINTS) return CUBE_ERR_ARG;
  memset(c, 0, sizeof *c);
  for (int i = 0; i < n; i++) c->z_checkpoint[i] = z_checkpoint[i];
  c->n_checkpoint = n;
  // initialize.f90:40-53,73, particle_initialization.f90:37-38, main.f90:42
  c->a = 1.0f / (1.0f + z_checkpoint[0]);
  c->tau = -3.0f / sqrtf(1.0f / (1.0f + z_checkpoint[0]));
  c->dt_neu = 1.0f;
  c->a_mid = c->a;
  c->cur_checkpoint = 2;
  return CUBE_OK;
}

// timestep.f90:8-95 (one image): float arithmetic in the reference's order
static void clock_timestep(cube_handle* h, cube_clock* c) {
  const float ra_max = 0.2f, dt_max = 1.0f;
  const float om = h->p.omega_m, ol = h->p.omega_l, wde = h->p.wde;
  c->dt_old = c->dt;
  h->sim.timestep += 1;
  float dt_e = dt_max, d1, d2;
  int ntry = 0;
  for (;;) {
    ntry++;
    cube_expansion(c->a, dt_e, om, ol, wde, &d1, &d2);
    const float da = d1 + d2;
    const float ra = da / (c->a + da);
    if (ra > ra_max) dt_e = dt_e * (ra_max / ra);
    else break;
    if (ntry > 10) break;
  }
  const cube_sim_header& s = h->sim;
  float dt = dt_e;
  const float cand[6] = {s.dt_fine, s.dt_coarse, s.dt_pp, s.dt_vmax, 1000.0f, c->dt_neu};
  for (float x : cand) dt = fminf(dt, x);
  cube_expansion(c->a, dt, om, ol, wde, &d1, &d2);
  float da = d1 + d2;
  c->checkpoint_step = 0;
  // z_checkpoint beyond the list reads -0.5 (initialize.f90:49)
  const float z_next = c->cur_checkpoint <= c->n_checkpoint ? c->z_checkpoint[c->cur_checkpoint - 1] : -0.5f;
  const float a_next = 1.0f / (1.0f + z_next);
  if (da >= a_next - c->a) {
    c->checkpoint_step = 1;
    if (c->cur_checkpoint == c->n_checkpoint) c->final_step = 1;
    for (int it = 0; fabsf((c->a + da) / a_next - 1.0f) >= 1e-6f && it < 100; it++) {
      dt = dt * (a_next - c->a) / da;
      cube_expansion(c->a, dt, om, ol, wde, &d1, &d2);
      da = d1 + d2;
    }
  }
  c->a_mid = c->a + da / 2;
  c->tau = c->tau + dt;
  c->t = c->t + dt;
  c->a = c->a + da;
  c->da = da;
  c->dt = dt;
  c->dt_neu = 1000.0f;
}

extern "C" int cube_set_step_mode(cube_handle* h, int mode) {
  if (!h || mode < 0 || mode > 2) return CUBE_ERR_ARG;
  h->step_mode = mode;
  return CUBE_OK;
}

extern "C" int cube_step_stats(cube_handle* h, int64_t* graph_instantiations, int64_t* graph_updates) {
  if (!h) return CUBE_ERR_ARG;
  if (graph_instantiations) *graph_instantiations = h->graph_instantiations;
  if (graph_updates) *graph_updates = h->graph_updates;
  return CUBE_OK;
}

// Drift, PP and PM of one step launched back to back (step modes 1 and 2):
// every host decision they need is known before the step (the new-sigma
// coefficient travels on the device), the error checks are flags, and the
// results come back in one copy of the step report, so the host waits once.
// Mode 2 records the launches as a hipGraph and replays it, updating the
// instantiated graph's kernel arguments (time-step values, ping-pong
// buffers) when the launch sequence matches the previous step's.
static int step_deferred(cube_handle* h, cube_clock* c) {
  if (hipSetDevice(h->p.device) != hipSuccess) return cube_fail(h, CUBE_ERR_HIP, "hipSetDevice failed");
  const bool graph = h->step_mode == 2 && !h->prof_on;
  std::shared_lock<std::shared_mutex> capture_lk(capture_mutex(), std::defer_lock);
  if (graph) {
    capture_lk.lock();
    HIP_TRY(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeRelaxed));
  }
  int nlayer = 0;
  int r = drift_launch_step(h, c->dt_old, c->dt, &nlayer);
  const bool swapped = r == 0;
  if (!r && h->p.pp_on) r = pp_launch_step(h, c->a_mid, c->dt, vcoef(h->sigma_vi), nullptr);
  if (!r) r = pm_launch_step(h, c->a_mid, c->dt);
  if (!r && hipMemcpyAsync(h->rep, h->drep, sizeof(StepReport), hipMemcpyDeviceToHost, h->stream) != hipSuccess)
    r = cube_fail(h, CUBE_ERR_HIP, "step report copy failed");
  hipGraph_t g = nullptr;
  if (graph) {
    const hipError_t e = hipStreamEndCapture(h->stream, &g);
    capture_lk.unlock();
    if (!r && e != hipSuccess) r = cube_fail(h, CUBE_ERR_HIP, "stream capture failed: %s", hipGetErrorString(e));
  }
  if (!r && graph) {
    if (h->gexec) {
      hipGraphNode_t en = nullptr;
      hipGraphExecUpdateResult res;
      if (hipGraphExecUpdate(h->gexec, g, &en, &res) == hipSuccess && res == hipGraphExecUpdateSuccess) {
        h->graph_updates++;
      } else {
        (void)hipGetLastError();
        (void)hipGraphExecDestroy(h->gexec);
        h->gexec = nullptr;
      }
    }
    if (!h->gexec) {
      if (hipGraphInstantiate(&h->gexec, g, nullptr, nullptr, 0) != hipSuccess)
        r = cube_fail(h, CUBE_ERR_HIP, "hipGraphInstantiate failed");
      else
        h->graph_instantiations++;
    }
    if (!r && hipGraphLaunch(h->gexec, h->stream) != hipSuccess) r = cube_fail(h, CUBE_ERR_HIP, "hipGraphLaunch failed");
  }
  if (g) (void)hipGraphDestroy(g);
  if (hipStreamSynchronize(h->stream) != hipSuccess && !r) r = cube_fail(h, CUBE_ERR_HIP, "hipStreamSynchronize failed");
  if (r) {
    // nothing of the step is kept: the drift's buffers are swapped back
    if (swapped) {
      std::swap(h->xp, h->xp2);
      std::swap(h->vp, h->vp2);
      std::swap(h->pid, h->pid2);
      std::swap(h->rhoc, h->rhoc2);
      std::swap(h->vfield, h->vfield2);
      std::swap(h->cstart, h->cstart2);
    }
    return r;
  }
  if ((r = drift_finish(h, h->rep, nlayer, nullptr))) return r;
  if (h->p.pp_on) {
    cube_pp_out o{};
    pp_finish(h, h->rep, &o);
    if ((r = pp_step_finish(h, c->a_mid, o))) return r;
  }
  return pm_finish_step(h, h->rep, c->a_mid, nullptr);
}

extern "C" int cube_step(cube_handle* h, cube_clock* c) {
  if (!h || !c) return CUBE_ERR_ARG;
  if (c->final_step && c->checkpoint_step) return CUBE_OK;  // already past the last checkpoint
  clock_timestep(h, c);
  c->istep += 1;
  int r;
  // main.f90:51-55, 66-70: the fluid is evolved half a step on each side
  if (h->nu.on && (r = cube_fluid_evolve(h, c->dt / 2, 1, c->a_mid, &c->dt_neu))) return r;
  if (h->step_mode > 0 && !h->nu.on) {
    if ((r = step_deferred(h, c))) return r;
  } else {
    if ((r = cube_drift(h, c->dt_old, c->dt, nullptr))) return r;
    if (h->p.pp_on && (r = cube_pp(h, c->a_mid, c->dt, nullptr))) return r;
    if ((r = cube_pm(h, c->a_mid, c->dt, nullptr))) return r;
  }
  if (h->nu.on && (r = cube_fluid_evolve(h, c->dt / 2, -1, c->a_mid, &c->dt_neu))) return r;
  if (c->checkpoint_step) {
    c->dt_old = 0;  // main.f90:74-75: the half drift closing the step
    if ((r = cube_drift(h, c->dt_old, c->dt, nullptr))) return r;
    c->checkpoint_at = c->cur_checkpoint;
    c->cur_checkpoint += 1;
    if (!c->final_step) c->dt = 0;
  } else {
    c->checkpoint_at = 0;
  }
  return CUBE_OK;
}

// ------------------------------------------------------------- footprint
extern "C" int64_t cube_footprint(const cube_params* p, int32_t ic_np_nc) {
  if (!p) return -1;
  const int64_t npm = p->np_max, nc = p->nc, n3c = nc * nc * nc, nf = 4 * nc, prec = p->precision;
  auto mesh = [&](int64_t N, bool with_S) {
    const int64_t H = N / 2 + 1;
    const int64_t big = std::max(N * N * (N + 2) * prec, (H + 15) / 16 * 16 * N * N * 2 * prec);  // Mesh::bytes_big
    return 2 * big + (with_S ? N * N * N * prec + big : 0) + 3 * H * ((H + 15) / 16 * 16) * H * prec + 2 * N * prec;
  };
  const int64_t nb2 = (nc / 2) * (nc / 2) * (nc / 2);
  int64_t b = 0;
  b += 2 * 3 * npm * (p->izipx + p->izipv) + 2 * 4 * npm;  // particles, double-buffered
  b += 2 * npm + 8 * npm + 4 * (npm / 1025 + 2);             // drift displacement, keys, lists
  b += 8 * (2 * nb2 + npm / 256 + nb2 + 1) + 8 * nb2 + 20480;  // PP lists, totals
  b += 2 * (4 * n3c + 12 * n3c + 8 * (n3c + 1)) + 3 * n3c * prec + 8 * (std::max<int64_t>(n3c, 65536) + 16);
  b += 4 * 65536 + 8 * 131072 + 16 * 65536;  // codec tables
  b += mesh(nf, true) + mesh(nc, false);
  const int64_t fine_R = nf * nf * (nf + 2) * prec;
  if (fine_R < 24 * npm) b += 24 * npm;  // drift scratch outside the fine work mesh
  if (ic_np_nc > 0) {  // Zel'dovich generator's temporary mesh
    const int64_t ng = (int64_t)ic_np_nc * nc;
    b += 2 * ng * ng * (ng + 2) * 4 + 16 * ng;
  }
  return b + b / 16 + (64 << 20);  // allocation granularity and runtime slack
}

// ---------------------------------------------------------------- ensemble
namespace {
using clk = std::chrono::steady_clock;

struct Request {
  int64_t id;
  cube_params p;
  std::vector<float> kf, kc, z;
  cube_ic_params ic;
  std::vector<double> lnk, lnp;
  int64_t bytes;
  double t_submit;
};

uint64_t fnv1a(uint64_t h, const void* data, size_t n) {
  const unsigned char* s = (const unsigned char*)data;
  for (size_t i = 0; i < n; i++) {
    h ^= s[i];
    h *= 1099511628211ull;
  }
  return h;
}
}  // namespace

struct cube_ensemble {
  cube_ens_params ep;
  clk::time_point t0;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Request> queue;
  std::vector<cube_ens_result> done;
  std::vector<std::thread> workers;
  std::vector<std::thread::id> ended;  // workers whose box is done, not yet joined
  std::thread scheduler;
  int64_t used = 0, next_id = 1;
  int live = 0;
  bool stop = false;
  double now() const { return std::chrono::duration<double>(clk::now() - t0).count(); }
};

static void run_box(cube_ensemble* e, Request rq, double t_admit) {
  cube_ens_result res{};
  res.id = rq.id;
  res.t_submit = rq.t_submit;
  res.t_admit = t_admit;
  cube_handle* h = nullptr;
  int rc = cube_init(&rq.p, rq.kf.data(), rq.kc.data(), &h);
  if (!rc) rc = cube_set_step_mode(h, e->ep.step_mode);
  if (!rc) {
    rq.ic.lnk = rq.lnk.data();
    rq.ic.lnp = rq.lnp.data();
    rc = cube_ic_zeldovich(h, &rq.ic, nullptr);
  }
  cube_clock c;
  if (!rc) rc = cube_clock_init(&c, rq.z.data(), (int)rq.z.size());
  while (!rc) {
    rc = cube_step(h, &c);
    if (rc || (c.final_step && c.checkpoint_step)) break;
    if (c.istep > 100000) { rc = CUBE_ERR_CHECK; break; }
  }
  if (!rc) {
    // fingerprint of the final state, for the alone-versus-in-the-mix check
    const int64_t n = h->np;
    std::vector<uint8_t> xp(3 * n * rq.p.izipx), vp(3 * n * rq.p.izipv);
    std::vector<int32_t> pid(n);
    cube_sim_header sim;
    rc = cube_download_state(h, &sim, xp.data(), vp.data(), pid.data(), nullptr, nullptr);
    uint64_t hv = 1469598103934665603ull;
    hv = fnv1a(hv, xp.data(), xp.size());
    hv = fnv1a(hv, vp.data(), vp.size());
    hv = fnv1a(hv, pid.data(), pid.size() * 4);
    res.state_hash = hv;
    res.nparticles = n;
    res.steps = (int32_t)c.istep;
    res.a_final = c.a;
  }
  if (rc && h) snprintf(res.message, sizeof res.message, "%s", h->err.c_str());
  else if (rc) snprintf(res.message, sizeof res.message, "box set-up failed (code %d)", rc);
  res.status = rc;
  if (h) cube_finalize(h);
  res.t_done = e->now();
  std::lock_guard<std::mutex> lk(e->mu);
  e->used -= rq.bytes;
  e->live -= 1;
  e->done.push_back(res);
  e->ended.push_back(std::this_thread::get_id());
  e->cv.notify_all();
}

static void schedule(cube_ensemble* e) {
  std::unique_lock<std::mutex> lk(e->mu);
  for (;;) {
    // join the workers whose boxes are done, with the lock released, so a
    // long-lived engine holds threads only for its live boxes; a worker
    // records itself last thing under the lock, so the join waits only for
    // its return
    if (!e->ended.empty()) {
      std::vector<std::thread> finished;
      for (std::thread::id id : e->ended) {
        auto it = std::find_if(e->workers.begin(), e->workers.end(), [id](const std::thread& t) { return t.get_id() == id; });
        if (it != e->workers.end()) {
          finished.push_back(std::move(*it));
          e->workers.erase(it);
        }
      }
      e->ended.clear();
      lk.unlock();
      for (std::thread& t : finished) t.join();
      lk.lock();
      continue;  // the queue may have changed meanwhile
    }
    // admit, in arrival order, every queued request that fits now
    bool admitted = false;
    for (auto it = e->queue.begin(); it != e->queue.end();) {
      if (e->live < e->ep.max_live && e->used + it->bytes <= e->ep.mem_budget) {
        Request rq = std::move(*it);
        it = e->queue.erase(it);
        e->used += rq.bytes;
        e->live += 1;
        admitted = true;
        const double ta = e->now();
        e->workers.emplace_back(run_box, e, std::move(rq), ta);
      } else {
        ++it;
      }
    }
    if (e->stop && e->queue.empty()) return;
    if (!admitted) e->cv.wait(lk);
  }
}

extern "C" int cube_ens_create(const cube_ens_params* p, cube_ensemble** out) {
  if (!p || !out || p->max_live < 1 || p->step_mode < 0 || p->step_mode > 2) return CUBE_ERR_ARG;
  cube_ensemble* e = new cube_ensemble();
  e->ep = *p;
  if (e->ep.mem_budget <= 0) {
    size_t fr = 0, tot = 0;
    if (hipSetDevice(p->device) != hipSuccess || hipMemGetInfo(&fr, &tot) != hipSuccess) {
      delete e;
      return CUBE_ERR_HIP;
    }
    e->ep.mem_budget = (int64_t)(fr * 0.9);
  }
  e->t0 = clk::now();
  e->scheduler = std::thread(schedule, e);
  *out = e;
  return CUBE_OK;
}

extern "C" int cube_ens_submit(cube_ensemble* e, const cube_params* p, const float* kern_f_table, const float* kern_c_table,
                               const cube_ic_params* ic, const float* z_checkpoint, int n_z, int64_t* id) {
  if (!e || !p || !kern_f_table || !kern_c_table || !ic || !z_checkpoint || n_z < 1 || n_z > CUBE_MAX_CHECKPOINTS ||
      ic->nk < 2 || !ic->lnk || !ic->lnp)
    return CUBE_ERR_ARG;
  Request rq;
  rq.p = *p;
  rq.p.device = e->ep.device;
  rq.kf.assign(kern_f_table, kern_f_table + 4096 * 3);
  rq.kc.assign(kern_c_table, kern_c_table + 64 * 3);
  rq.z.assign(z_checkpoint, z_checkpoint + n_z);
  rq.ic = *ic;
  rq.lnk.assign(ic->lnk, ic->lnk + ic->nk);
  rq.lnp.assign(ic->lnp, ic->lnp + ic->nk);
  rq.bytes = cube_footprint(p, ic->np_nc);
  if (rq.bytes > e->ep.mem_budget) return CUBE_ERR_CAPACITY;
  std::lock_guard<std::mutex> lk(e->mu);
  rq.id = e->next_id++;
  rq.t_submit = e->now();
  if (id) *id = rq.id;
  e->queue.push_back(std::move(rq));
  e->cv.notify_all();
  return CUBE_OK;
}

extern "C" int cube_ens_poll(cube_ensemble* e, cube_ens_result* out, int max_results, int* n_out) {
  if (!e || !n_out || max_results < 0 || (max_results > 0 && !out)) return CUBE_ERR_ARG;
  std::lock_guard<std::mutex> lk(e->mu);
  const int n = std::min<int>(max_results, (int)e->done.size());
  for (int i = 0; i < n; i++) out[i] = e->done[i];
  e->done.erase(e->done.begin(), e->done.begin() + n);
  *n_out = n;
  return CUBE_OK;
}

extern "C" int cube_ens_pending(cube_ensemble* e, int* queued, int* live) {
  if (!e) return CUBE_ERR_ARG;
  std::lock_guard<std::mutex> lk(e->mu);
  if (queued) *queued = (int)e->queue.size();
  if (live) *live = e->live;
  return CUBE_OK;
}

extern "C" double cube_ens_clock(cube_ensemble* e) { return e ? e->now() : 0.0; }

extern "C" int cube_ens_destroy(cube_ensemble* e) {
  if (!e) return CUBE_OK;
  {
    std::lock_guard<std::mutex> lk(e->mu);
    e->stop = true;
    e->cv.notify_all();
  }
  e->scheduler.join();
  // the scheduler has admitted everything queued; wait for the boxes
  {
    std::unique_lock<std::mutex> lk(e->mu);
    e->cv.wait(lk, [e] { return e->live == 0; });
  }
  for (auto& t : e->workers) t.join();
  delete e;
  return CUBE_OK;
}
