"""Ensemble serving over the native engine (cube_ens_* in libcubehip.so).

Requests are seeded LCDM PM+PP boxes of uneven size started from
Zel'dovich initial conditions; the engine admits each as soon as its device
footprint fits and runs all live boxes concurrently, one HIP stream and one
host thread each, each step either as synchronising phases or launched
whole (one synchronisation per step, optionally replayed as a hipGraph).
`run_ensemble` replays a list of requests with their arrival times and
returns per-request latencies, the aggregate particle-steps/s and each
box's final-state hash.
"""
import ctypes
import time

import numpy as np

from . import _lib
from .ic import power_table, velocity_factor
from .params import BoxConfig, read_kernel_tables

# Ensemble box sizes: particles per dimension -> geometry (mesh 2x particles
# per dimension: np_nc = 2 with ncell = 4), box scaled at fixed mass resolution.
C5_SIZES = (32, 64, 128, 256)


def ensemble_config(npd, pp=True):
    nc = npd // 2
    nnt = {16: 1, 32: 2, 64: 2, 128: 4}.get(nc, max(1, nc // 32))
    return BoxConfig(nc=nc, nnt=nnt, ncb=6, izipx=2, izipv=2, box=400.0 * npd / 256.0, pp_on=pp, pp_range=1)


def poisson_requests(n, rate, seed, sizes=C5_SIZES, z_i=49.0, z_f=0.0):
    """n requests with exponential inter-arrival times (rate per second),
    sizes uniform over `sizes`, seeds 1000+i."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=n)) if rate > 0 else np.zeros(n)
    t -= t[0]
    return [dict(t=float(t[i]), npd=int(rng.choice(sizes)), seed=1000 + i, z_i=z_i, z_f=z_f) for i in range(n)]


class Ensemble:
    def __init__(self, max_live=64, mem_budget=0, device=0, step_mode=0):
        self.lib = _lib.load()
        ep = _lib.EnsParams()
        ep.device, ep.max_live, ep.mem_budget, ep.step_mode = device, max_live, mem_budget, step_mode
        e = ctypes.c_void_p()
        rc = self.lib.cube_ens_create(ctypes.byref(ep), ctypes.byref(e))
        if rc:
            raise RuntimeError("cube_ens_create failed (%d)" % rc)
        self.e = e
        self.kf, self.kc = read_kernel_tables()

    def submit(self, cfg, seed, z_i, z_f):
        p = _lib.CubeParams()
        p.nc, p.nnt, p.ncb = cfg.nc, cfg.nnt, cfg.ncb
        p.izipx, p.izipv = cfg.izipx, cfg.izipv
        p.np_max = cfg.np_max
        p.pp_on, p.pp_range, p.rsoft = int(cfg.pp_on), cfg.pp_range, cfg.rsoft
        p.omega_m, p.omega_l, p.wde = cfg.omega_m, cfg.omega_l, cfg.wde
        p.precision, p.device = 4, 0
        a = 1.0 / (1.0 + z_i)
        lnk, lnp = power_table(cfg, a)
        ip = _lib.ICParams()
        ip.seed, ip.np_nc, ip.nk, ip.a, ip.box = seed, cfg.np_nc, len(lnk), a, cfg.box
        ip.vfac = velocity_factor(cfg, a)
        ip.lnk, ip.lnp = _lib.ptr(lnk), _lib.ptr(lnp)
        z = np.array([z_i, z_f], dtype=np.float32)
        rid = ctypes.c_int64()
        rc = self.lib.cube_ens_submit(self.e, ctypes.byref(p), _lib.ptr(self.kf), _lib.ptr(self.kc), ctypes.byref(ip),
                                      _lib.ptr(z), len(z), ctypes.byref(rid))
        if rc:
            raise RuntimeError("cube_ens_submit failed (%d)" % rc)
        return rid.value

    def poll(self, maxn=256):
        buf = (_lib.EnsResult * maxn)()
        n = ctypes.c_int32()
        self.lib.cube_ens_poll(self.e, buf, maxn, ctypes.byref(n))
        return [buf[i] for i in range(n.value)]

    def pending(self):
        q, l = ctypes.c_int32(), ctypes.c_int32()
        self.lib.cube_ens_pending(self.e, ctypes.byref(q), ctypes.byref(l))
        return q.value, l.value

    def clock(self):
        return self.lib.cube_ens_clock(self.e)

    def close(self):
        if self.e:
            self.lib.cube_ens_destroy(self.e)
            self.e = None


def run_ensemble(requests, max_live=64, mem_budget=0, log=None, step_mode=0):
    """Submit `requests` at their arrival times (seconds after start) and
    wait for all of them.  Returns a dict with per-request results and the
    aggregate metrics.  step_mode: cube_set_step_mode of every box (0 each
    phase synchronises, 1 one synchronisation per step, 2 steps replayed as
    hipGraphs)."""
    ens = Ensemble(max_live=max_live, mem_budget=mem_budget, step_mode=step_mode)
    try:
        t0 = time.perf_counter()
        ids = {}
        results = {}
        pending = list(requests)
        while pending or len(results) < len(requests):
            now = time.perf_counter() - t0
            while pending and pending[0]["t"] <= now:
                rq = pending.pop(0)
                ids[ens.submit(ensemble_config(rq["npd"]), rq["seed"], rq["z_i"], rq["z_f"])] = rq
            for r in ens.poll():
                rq = ids[r.id]
                if r.status:
                    raise RuntimeError("request %d (%d^3) failed: %s" % (r.id, rq["npd"], r.message.decode()))
                results[r.id] = dict(npd=rq["npd"], seed=rq["seed"], steps=r.steps, n=r.nparticles,
                                     latency=r.t_done - r.t_submit, wait=r.t_admit - r.t_submit, t_submit=r.t_submit,
                                     t_done=r.t_done, a_final=r.a_final, hash=int(r.state_hash))
                if log:
                    log("  request %d: %d^3 particles, %d steps, latency %.3f s" % (r.id, rq["npd"], r.steps,
                                                                                  results[r.id]["latency"]))
            time.sleep(0.002)
    finally:
        # runs whatever was submitted to completion, then frees the engine
        ens.close()
    res = list(results.values())
    span = max(r["t_done"] for r in res) - min(r["t_submit"] for r in res)
    lat = np.array([r["latency"] for r in res])
    work = float(sum(r["n"] * r["steps"] for r in res))
    return dict(results=res, p50=float(np.percentile(lat, 50)), p99=float(np.percentile(lat, 99)),
                particle_steps_per_s=work / span, span=span, work=work)
