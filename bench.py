"""Benchmark: CUBE PM+PP steps on one MI355X through libcubehip.so.

A step is one pass of CUBE's main loop body (main.f90:48-97): timestep,
drift/re-bin, PP force + kick, fine and coarse PM force + kick.  The timed
region holds exactly --steps steps between device synchronisations; inputs
are resident in HBM before it starts.  Prints one JSON line.

Extra fields:
  roofline     -- the HBM-bound kernel class with the largest share of step
                  time: algorithmic bytes per launch / mean launch time (HIP
                  events on the library's stream), against MI355X's 8 TB/s;
                  `traffic` is the measured HBM bytes per launch from the
                  rocprofv3 FETCH_SIZE/WRITE_SIZE summary in profiles/ when
                  that summary covers the same kernel and workload;
  pp           -- the PP force (FP32 VALU bound): pair interactions per
                  second and the FP32 rate at 30 flops per pair;
  cpu_baseline -- the CPU oracle (a C restatement of CUBE's step, OpenMP on
                  the host cores) timed on a bounded sample of the same
                  workload: the first tiles of the same box.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(slot, cfg, npart):
    """Minimal HBM bytes per launch of each profiled kernel class
    (particle arrays, meshes, offsets; re-reads of halo particles, padding
    and tile overheads are not counted)."""
    bx, bv = 3 * cfg.izipx, 3 * cfg.izipv
    mf, mc = float(cfg.nf) ** 3, float(cfg.nc) ** 3
    n = float(npart)
    table = {
        "drift_pass1": n * (bx + bv + 2) + 24 * mc,
        "drift_scan": 12 * mc,
        "drift_scatter": n * (2 + 8) + 20 * mc,
        "drift_sort_big": None,
        "drift_write": n * (8 + 2 * (bx + bv + 4)) + 36 * mc,
        "pp": n * (bx + 2 * bv) + 8 * mc,
        "cic_coarse": n * bx + 12 * mc,
        "cic_fine": n * bx + 4 * mf + 8 * mc,
        "fft_x_r2c": 8 * mf,
        "fft_y_fwd": 8 * mf,
        # one launch for all three components on FP32 meshes: the spectrum
        # read once, three convolved spectra written, three kernel octants
        "fft_zconv": 17.5 * mf,
        "fft_y_inv": 8 * mf,
        "fft_x_c2r": 8 * mf,
        # fine pass: xp, vp read + write, three fine force meshes, cell
        # offsets; coarse pass: the same particle arrays, three coarse
        # force meshes, vfield, cell offsets
        "kick_fine": n * (bx + 2 * bv) + 12 * mf + 8 * mc,
        "kick_coarse": n * (bx + 2 * bv) + 32 * mc,
        "fft_coarse": None,
    }
    return table.get(slot)


# Slots bound by FP32 VALU work rather than HBM (reported separately).
VALU_SLOTS = ("pp",)


def read_traffic(kernel_slot, workload):
    """HBM bytes per launch for `kernel_slot` from profiles/traffic.txt
    (lines 'slot workload fetch_bytes write_bytes', already scaled by the
    copy-kernel calibration), or None."""
    path = os.path.join(ROOT, "profiles", "traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 4 and not line.startswith("#") and f[0] == kernel_slot and f[1] == workload:
            return float(f[2]) + float(f[3])
    return None


def hbm_ceiling(nbytes=4 << 30, iters=5):
    """Best copy rate (read + write GB/s) of the library's streaming-copy
    probe over its four launch variants: the practical HBM ceiling."""
    from cube_nufluid_amd import _lib
    lib = _lib.load()
    out = np.zeros(4)
    rc = lib.cube_hbm_copy_probe(0, nbytes, iters, _lib.ptr(out))
    if rc:
        return None
    log("HBM copy probe (%d MiB): plain %.0f / %.0f GB/s, nontemporal %.0f / %.0f GB/s (8 / 32 workgroups per CU)"
        % (nbytes >> 20, *out))
    return {"gbs": round(float(out.max()), 1), "variants_gbs": [round(float(v), 1) for v in out],
            "kernel": "float4 grid-stride copy, 4 vectors per thread per iteration, %d MiB each way" % (nbytes >> 20)}


def copy_calibration():
    """Three plain 4 GiB device copies (torch copy kernel): under
    rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE they give the counters' reading
    for a known byte count."""
    import torch
    src = torch.zeros(1 << 30, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        torch.cuda.synchronize()
        t = time.perf_counter()
        dst.copy_(src)
        torch.cuda.synchronize()
        log("copy 4 GiB: %.3f ms" % ((time.perf_counter() - t) * 1e3))


def host_threads():
    """Host threads the CPU legs use: the cores this process may run on,
    capped by OMP_NUM_THREADS when the environment sets it (the CPU share
    given to one GPU), with the host's nproc and CPU model for the record."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS")
    threads = min(avail, int(env)) if env and env.isdigit() and int(env) > 0 else avail
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    info = {"nproc": os.cpu_count(), "affinity": avail, "omp_num_threads": env, "cpu_model": model}
    return threads, info


def make_state(cfg, seed, z):
    from cube_nufluid_amd.ic import zeldovich_ic
    return zeldovich_ic(cfg, seed=seed, z=z)


def cpu_baseline(cfg, st, steps_dt, sample_tiles, threads):
    """Time the oracle's full step (drift + PP + PM) on the first
    `sample_tiles` tiles of the same box.  The coarse mesh FFTs cover the
    whole box whatever the sample, so their time is charged pro rata
    (sample tiles / all tiles); value = particles in the sample / time."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle
    log("cpu baseline: oracle init (%d threads)" % threads)
    orc = Oracle(cfg, threads=threads)
    orc.set_state(st)
    orc.lib.orc_set_tile_limit(orc.o, sample_tiles)
    nt3 = cfg.nt ** 3
    n_sample = int(st.rhoc[: sample_tiles * nt3].sum())
    dt = float(steps_dt)
    log("cpu baseline: timing %d tile(s), %d particles" % (sample_tiles, n_sample))
    t0 = time.perf_counter()
    orc.drift(dt, dt)
    t1 = time.perf_counter()
    if cfg.pp_on:
        orc.pp(0.5, dt)
    t2 = time.perf_counter()
    orc.pm(0.5, dt)
    t3 = time.perf_counter()
    t_glob = orc.lib.orc_global_seconds(orc.o)
    orc.close()
    frac = sample_tiles / float(cfg.nnt ** 3)
    wall = (t3 - t0) - t_glob * (1.0 - frac)
    log("cpu baseline: drift %.1f s, pp %.1f s, pm %.1f s (box-wide coarse mesh %.1f s, charged %.2f of it)"
        % (t1 - t0, t2 - t1, t3 - t2, t_glob, frac))
    return {"value": n_sample / wall, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": "CPU oracle (C restatement of CUBE's step, OpenMP) one full step, drift+PP+fine/coarse PM, on "
                      "the first %d of %d tiles (%d particles) of the same box; drift %.1f s, PP %.1f s, PM %.1f s, "
                      "box-wide coarse FFTs %.1f s charged pro rata" % (sample_tiles, cfg.nnt ** 3, n_sample,
                                                                         t1 - t0, t2 - t1, t3 - t2, t_glob)}


def halofind_c3(cube, name):
    """One cube_halofind call (halofind.f90) on the state left by the timed
    steps: wall time of the synchronous call, halos and peaks found."""
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    try:
        cat = cube.halofind(max_halos=1 << 21)
    except Exception as e:  # reported, not fatal to the step measurement
        log("halofind failed: %s" % e)
        return {"error": str(e)}
    ms = (time.perf_counter() - t0) * 1e3
    members = int(cat.field("mass_odc").sum()) if cat.nhalo else 0
    log("halofind (%s): %d halos (%d particles) from %d candidate peaks in %.1f ms"
        % (name, cat.nhalo, members, cat.npeaks, ms))
    return {"workload": "%s state after the timed steps" % name, "ms": round(ms, 2), "nhalo": cat.nhalo,
            "particles_in_halos": members, "candidate_peaks": cat.npeaks, "processed_peaks": cat.nprocessed}


def halofind_evolved(config, seed, cpu):
    """The halo finder where the reference runs it: a box evolved from z=49
    to z=0 by the GPU step, halofind at the final step (z_halofind.txt =
    1.0, 0.0 as shipped, whose first entry used is z=0).  GPU wall time of
    the call; with cpu, the C restatement timed on the same state (one
    thread) and its catalogue compared."""
    from cube_nufluid_amd.params import CONFIGS
    from cube_nufluid_amd.driver import Cube
    cfg = CONFIGS[config]
    t0 = time.perf_counter()
    cube = Cube.from_ic(cfg, seed, 49.0, z_checkpoints=[49.0, 0.0], z_halofind=[1.0, 0.0])
    try:
        nsteps = cube.run()
        t_run = time.perf_counter() - t0
        if not cube.halo_catalogs:
            return {"error": "no halofind step reached"}
        z, cat = cube.halo_catalogs[-1]
        ms = cube.timings["halofind"][-1] * 1e3
        members = int(cat.field("mass_odc").sum()) if cat.nhalo else 0
        log("halofind (%s evolved, %d steps in %.1f s): %d halos (%d particles) from %d candidate peaks in %.1f ms"
            % (config, nsteps, t_run, cat.nhalo, members, cat.npeaks, ms))
        out = {"workload": "%s: %d^3 particles from z=49 to z=0 (%d steps), halofind at z=0"
                           % (config, round(cube.sim.nplocal ** (1 / 3)), nsteps),
               "ms": round(ms, 2), "nhalo": cat.nhalo, "particles_in_halos": members,
               "candidate_peaks": cat.npeaks, "processed_peaks": cat.nprocessed}
        if cpu:
            st = cube.download()
            st.sim.a = cube.sim.a
            st.sim.vsim2phys = cube.sim.vsim2phys
    finally:
        cube.close()
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib
        from cube_nufluid_amd.halo import HaloParams
        hp = HaloParams().for_box(cfg, None, st.sim)
        t1 = time.perf_counter()
        halos, words, _, counts = oracle_lib.halofind(cfg, st, hp, max_halos=1 << 20)
        cpu_s = time.perf_counter() - t1
        same = len(halos) == cat.nhalo and np.array_equal(words, cat.pid_words)
        log("halofind cpu (restatement, 1 thread): %.1f s, %d halos, catalogue membership identical: %s"
            % (cpu_s, len(halos), same))
        out["cpu"] = {"seconds": round(cpu_s, 2), "cores": 1, "kind": "port", "nhalo": len(halos),
                      "same_members_as_gpu": bool(same)}
    return out


def nfw_clumps_state(cfg, seed, frac_halo=0.5):
    """Synthetic heavy-tailed PP occupancy: a fraction frac_halo of the box's
    particles in NFW halos (concentration 4..12, virial radii from the
    halo's mass at 200 times the mean density, halo masses from
    dN/dM ~ M^-2 between 50 and 2e5 particles), the rest uniform."""
    from cube_nufluid_amd.ic import encode_particles
    rng = np.random.default_rng(seed)
    n = cfg.np_total if hasattr(cfg, "np_total") else (cfg.nc * cfg.np_nc) ** 3
    nh_target = int(frac_halo * n)
    mmin, mmax = 50.0, 2e5
    sizes = []
    tot = 0
    while tot < nh_target:
        u = rng.uniform(size=4096)
        m = mmin / (1 - u * (1 - mmin / mmax))      # dN/dM ~ M^-2 on [mmin, mmax]
        m = m.astype(np.int64)
        for x in m:
            sizes.append(int(x))
            tot += int(x)
            if tot >= nh_target:
                break
    sizes = np.array(sizes)
    nbar = n / float(cfg.nc) ** 3                       # particles per coarse cell
    rvir = (3 * sizes / (4 * np.pi * 200 * nbar)) ** (1 / 3)   # coarse-cell units
    cgrid = np.linspace(4, 12, 17)
    conc = rng.integers(0, len(cgrid), size=len(sizes))
    centres = rng.uniform(0, cfg.nc, size=(len(sizes), 3))
    # inverse CDF of the NFW enclosed mass ln(1+x) - x/(1+x), x = r/r_s,
    # tabulated per concentration
    xs = np.linspace(0, 1, 4097)
    pos = []
    for ci, cc in enumerate(cgrid):
        sel = np.nonzero(conc == ci)[0]
        if len(sel) == 0:
            continue
        m = sizes[sel]
        owner = np.repeat(np.arange(len(sel)), m)
        xg = xs * cc
        mu = np.log1p(xg) - xg / (1 + xg)
        r = np.interp(rng.uniform(0, mu[-1], size=len(owner)), mu, xg) * (rvir[sel][owner] / cc)
        d = rng.normal(size=(len(owner), 3))
        d /= np.linalg.norm(d, axis=1)[:, None]
        pos.append(centres[sel][owner] + r[:, None] * d)
    pos.append(rng.uniform(0, cfg.nc, size=(n - int(sizes.sum()), 3)))
    pos = np.concatenate(pos)[:n]
    vel = rng.normal(0, 1.0, size=pos.shape)
    st = encode_particles(cfg, pos, vel, np.arange(1, len(pos) + 1, dtype=np.int32), np.float32(1.0))
    return st, len(sizes), int(sizes.max())


def pp_unit_bench(config, seed, cpu, reps=3):
    """PP alone on a seeded NFW-clump occupancy of the config's size: GPU
    pairs/s and TFLOP/s (30 flop per pair) over `reps` passes; with cpu, the
    oracle's PP on the first tile of the same state with all host threads."""
    from cube_nufluid_amd.params import CONFIGS
    from cube_nufluid_amd.driver import Cube
    import torch
    cfg = CONFIGS[config]
    t0 = time.perf_counter()
    st, nhalo, mmax = nfw_clumps_state(cfg, seed)
    log("pp unit: %s NFW-clump state, %d particles, %d halos (largest %d) built in %.1f s"
        % (config, st.sim.nplocal, nhalo, mmax, time.perf_counter() - t0))
    cube = Cube(cfg, st)
    try:
        cube.a_mid, cube.dt = np.float32(0.9), np.float32(0.01)
        cube.ext_pp_force()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(reps):
            out = cube.ext_pp_force()
        torch.cuda.synchronize()
        sec = (time.perf_counter() - t1) / reps
    finally:
        cube.close()
    pairs = float(out.npairs)
    res = {"workload": "%s size (%d particles), %d seeded NFW halos holding half the particles, dN/dM ~ M^-2, "
                       "largest %d particles" % (config, st.sim.nplocal, nhalo, mmax),
           "ms_per_pass": round(sec * 1e3, 3), "pairs": pairs, "pairs_per_s": pairs / sec,
           "achieved_tflops": round(30 * pairs / sec / 1e12, 3), "peak_tflops": FP32_PEAK_TFLOPS,
           "dense_blocks": int(out.dense_blocks)}
    log("pp unit: %.2f ms per pass, %.3g pairs/s, %d dense blocks" % (sec * 1e3, pairs / sec, out.dense_blocks))
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle_lib import Oracle
        threads, _ = host_threads()
        orc = Oracle(cfg, threads=threads)
        orc.set_state(st)
        orc.lib.orc_set_tile_limit(orc.o, 1)
        t2 = time.perf_counter()
        r = orc.pp(0.9, 0.01)
        cs = time.perf_counter() - t2
        orc.close()
        res["cpu"] = {"pairs_per_s": r["npairs"] / cs, "cores": threads, "kind": "port",
                      "sample": "oracle PP on the first of %d tiles (%d pairs, %.1f s)" % (cfg.nnt ** 3, r["npairs"], cs)}
        log("pp unit cpu: %.3g pairs/s on one tile" % (r["npairs"] / cs))
    return res


def c4_bench(steps, warmup, seed=1004):
    """C4: 512^3 CDM particles on a 1024^3 fine mesh with the neutrino fluid
    (sum m_nu = 0.1 eV as three degenerate species; fluids 512^3, 256^3,
    128^3), seeded ICs at z = 10 with the fluids from the same linear field
    (neutrino_ic.f90 filters).  Particle-steps/s over `steps` full steps
    (both fluid half-step sweeps, drift, PP, PM with the fluid coupled), the
    fluid sweeps' time per step, and the sweep kernels against HBM: 3 sweeps
    x 2 (read + write) x 20 B per cell of each fluid's (nc+2 ncb_h)^3 mesh
    per half-step (the reference's halo-extended arrays; the periodic sweeps
    here move nc^3 cells)."""
    import torch
    from cube_nufluid_amd.params import CONFIGS
    from cube_nufluid_amd.driver import Cube
    from cube_nufluid_amd import _lib
    from cube_nufluid_amd.checkpoint import neu_ncb
    cfg = CONFIGS["C4"]
    t0 = time.perf_counter()
    cube = Cube.from_ic(cfg, seed, cfg.z_i, z_checkpoints=[cfg.z_i, -0.9])
    npart = cube.sim.nplocal
    log("C4: %d particles + fluids %s, IC in %.1f s" % (npart, [cube.nu_info(n)["nc"] for n in range(3)],
                                                         time.perf_counter() - t0))
    try:
        for _ in range(warmup):
            cube.step()
        lib = cube.lib
        lib.cube_profile_enable(cube.h, 1)
        lib.cube_profile_read(cube.h, None, None, 0)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(steps):
            cube.step()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t1
        nslot = 32
        ms = np.zeros(nslot)
        cnt = np.zeros(nslot, dtype=np.int64)
        ns = lib.cube_profile_read(cube.h, _lib.ptr(ms), _lib.ptr(cnt), nslot)
        names = []
        for s in range(ns):
            b = ctypes.create_string_buffer(64)
            lib.cube_profile_slot_name(s, b, 64)
            names.append(b.value.decode())
        per = {names[s]: round(float(ms[s]) / steps, 3) for s in range(ns) if cnt[s]}
        fluid_ms = per.get("fluid", 0.0)
        infos = [cube.nu_info(n) for n in range(3)]
        # two half-step evolves per step, each 3 sweeps of every fluid
        fbytes = sum(2 * 3 * 2 * 20.0 * (i["nc"] + 2 * neu_ncb(i["dx"], cfg.ncb)) ** 3 for i in infos)
        res = {"workload": "C4: %d^3 CDM particles, %d^3 fine mesh, fluids %s, %.0f Mpc/h, z_i = %.0f, "
                           "sum m_nu = 0.1 eV" % (round(npart ** (1 / 3)), cfg.nf, [i["nc"] for i in infos],
                                                   cfg.box, cfg.z_i),
               "particle_steps_per_s": npart * steps / wall, "ms_per_step": wall / steps * 1e3, "steps": steps,
               "z_after": round(float(1 / cube.a - 1), 4), "fluid_ms_per_step": fluid_ms,
               "fluid_algorithmic_GB_per_step": round(fbytes / 1e9, 3),
               "fluid_achieved_GBs": round(fbytes / (fluid_ms * 1e-3) / 1e9, 1) if fluid_ms else None,
               "per_kernel_ms_per_step": per}
        log("C4: %.3g particle-steps/s, %.1f ms/step, fluid %.1f ms/step" % (res["particle_steps_per_s"],
                                                                          res["ms_per_step"], fluid_ms))
        return res
    finally:
        cube.close()


def ensemble_small(n, rate):
    """The same serving comparison on boxes of {32,64,128}^3 particles only,
    boxes whose steps leave the GPU idle when run alone: one at a time with
    synchronising phases (the plain form), and concurrently with
    synchronising phases, with one synchronisation per step, and with each
    step replayed as a hipGraph."""
    from cube_nufluid_amd.serving import poisson_requests, run_ensemble
    reqs = poisson_requests(n, rate, seed=7, sizes=(32, 64, 128))
    key = lambda r: (r["npd"], r["seed"])
    plain = run_ensemble([dict(r, t=0.0) for r in reqs], max_live=1, step_mode=0)
    ref = {key(r): r["hash"] for r in plain["results"]}
    runs, same = {}, True
    for mode in (0, 1, 2):
        runs[mode] = run_ensemble(reqs, max_live=64, step_mode=mode)
        same = same and {key(r): r["hash"] for r in runs[mode]["results"]} == ref
    alone_graph = run_ensemble([dict(r, t=0.0) for r in reqs], max_live=1, step_mode=2)
    same = same and {key(r): r["hash"] for r in alone_graph["results"]} == ref
    base = plain["particle_steps_per_s"]
    log("ensemble, small boxes only: one at a time %.3g particle-steps/s; concurrent %.3g (sync phases), %.3g (one "
        "sync per step), %.3g (hipGraph steps) = %.2fx / %.2fx / %.2fx; one at a time with hipGraph steps %.3g"
        % (base, runs[0]["particle_steps_per_s"], runs[1]["particle_steps_per_s"], runs[2]["particle_steps_per_s"],
           runs[0]["particle_steps_per_s"] / base, runs[1]["particle_steps_per_s"] / base,
           runs[2]["particle_steps_per_s"] / base, alone_graph["particle_steps_per_s"]))
    return {"workload": "%d seeded LCDM PM+PP boxes, {32,64,128}^3 particles, z=49 to 0, Poisson arrivals %.1f/s"
                        % (n, rate),
            "p50_s": round(runs[2]["p50"], 4), "p99_s": round(runs[2]["p99"], 4),
            "plain_particle_steps_per_s": base,
            "concurrent_particle_steps_per_s": {"sync_phases": runs[0]["particle_steps_per_s"],
                                                "one_sync_per_step": runs[1]["particle_steps_per_s"],
                                                "hipgraph_steps": runs[2]["particle_steps_per_s"]},
            "speedup_vs_one_at_a_time": {"sync_phases": runs[0]["particle_steps_per_s"] / base,
                                         "one_sync_per_step": runs[1]["particle_steps_per_s"] / base,
                                         "hipgraph_steps": runs[2]["particle_steps_per_s"] / base},
            "one_at_a_time_hipgraph_particle_steps_per_s": alone_graph["particle_steps_per_s"],
            "same_result_alone_and_in_mix": bool(same)}


def ensemble_bench(n, rate):
    """Ensemble serving: n seeded LCDM PM+PP boxes drawn uniformly from
    {32,64,128,256}^3 particles (mesh 2x particles per dimension), Zel'dovich
    start at z=49, run to z=0, arriving as a Poisson process; served
    concurrently by the native engine with each step replayed as a hipGraph,
    then the same boxes one at a time in the plain form (synchronising
    phases).  A box takes the steps CUBE's time-step limits give it from
    z=49 to 0 (the ra <= 0.2 expansion limit alone needs about 25 steps; the
    force and velocity limits add the rest), not a fixed 50."""
    from cube_nufluid_amd.serving import poisson_requests, run_ensemble
    reqs = poisson_requests(n, rate, seed=5)
    log("ensemble: %d requests, sizes %s, Poisson rate %.1f/s" % (n, sorted(set(r["npd"] for r in reqs)), rate))
    mix = run_ensemble(reqs, max_live=64, log=log, step_mode=2)
    log("ensemble (concurrent, hipGraph steps): p50 %.3f s, p99 %.3f s, %.3g particle-steps/s"
        % (mix["p50"], mix["p99"], mix["particle_steps_per_s"]))
    plain = run_ensemble([dict(r, t=0.0) for r in reqs], max_live=1, step_mode=0)
    log("one at a time: %.3g particle-steps/s over %.2f s" % (plain["particle_steps_per_s"], plain["span"]))
    key = lambda r: (r["npd"], r["seed"])
    same = {key(r): r["hash"] for r in mix["results"]} == {key(r): r["hash"] for r in plain["results"]}
    # one-at-a-time run time per box size: the share of the largest boxes
    # bounds what overlapping can gain when one such box already fills the GPU
    by_size = {}
    for r in plain["results"]:
        by_size[r["npd"]] = by_size.get(r["npd"], 0.0) + (r["latency"] - r["wait"])
    big = max(by_size)
    log("one at a time, seconds by size: %s" % {k: round(v, 2) for k, v in sorted(by_size.items())})
    return {"workload": "C5: %d seeded LCDM PM+PP boxes, {32,64,128,256}^3 particles, z=49 to 0, Poisson arrivals "
                        "%.1f/s" % (n, rate),
            "step_mode": "hipGraph steps (mix); synchronising phases (one at a time)",
            "p50_s": round(mix["p50"], 4), "p99_s": round(mix["p99"], 4),
            "particle_steps_per_s": mix["particle_steps_per_s"],
            "plain_particle_steps_per_s": plain["particle_steps_per_s"],
            "speedup_vs_one_at_a_time": mix["particle_steps_per_s"] / plain["particle_steps_per_s"],
            "plain_seconds_by_size": {str(k): round(v, 3) for k, v in sorted(by_size.items())},
            "plain_time_over_largest_boxes_time": plain["span"] / by_size[big],
            "steps": sorted(set(r["steps"] for r in mix["results"])),
            "steps_by_size": {str(k): sorted(set(r["steps"] for r in mix["results"] if r["npd"] == k))
                              for k in sorted(set(r["npd"] for r in mix["results"]))},
            "same_result_alone_and_in_mix": bool(same)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default=os.environ.get("CUBE_BENCH_CONFIG", "C3"))
    ap.add_argument("--z", type=float, default=3.0,
                    help="redshift of the Zel'dovich start state, evolved to z=0 before the timed steps")
    ap.add_argument("--evolve-steps", type=int, default=200,
                    help="cap on the CUBE steps that evolve the start state to z=0 (0: time the Zel'dovich state)")
    ap.add_argument("--evolve-seconds", type=float, default=150.0, help="wall-clock cap on that evolution")
    ap.add_argument("--seed", type=int, default=1003)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the ensemble-serving measurement")
    ap.add_argument("--no-halofind", action="store_true", help="skip the halo-finder measurement")
    ap.add_argument("--pp-unit", action="store_true", help="also run the PP unit benchmark on NFW clumps (C2 size)")
    ap.add_argument("--halofind-evolved", action="store_true",
                    help="also time the halo finder on a C2 box evolved z=49 -> 0, with the C restatement beside it")
    ap.add_argument("--no-c4", action="store_true", help="skip the CDM + neutrino-fluid (C4) measurement")
    ap.add_argument("--c4-only", action="store_true", help="only the C4 measurement (for profiling)")
    ap.add_argument("--halofind-only", action="store_true",
                    help="only generate the state and run the halo finder on it once (for profiling)")
    ap.add_argument("--ens-requests", type=int, default=48)
    ap.add_argument("--ens-small-requests", type=int, default=24,
                    help="requests of the small-box ensemble ({32,64,128}^3 particles); 0 skips it")
    ap.add_argument("--ens-rate", type=float, default=6.0, help="Poisson arrival rate, requests per second")
    ap.add_argument("--copy-calib", action="store_true", help="only run the copy kernel used to calibrate PMC bytes")
    ap.add_argument("--cpu-tiles", type=int, default=0,
                    help="tiles in the CPU sample (0: one tile when tiles hold >4M particles, else one per thread)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs one box on one GPU; multi-GPU runs are not part of this build (asked for %d)"
              % args.gpus, file=sys.stderr)
        sys.exit(2)

    if args.copy_calib:
        copy_calibration()
        hbm_ceiling()
        return
    import torch
    from cube_nufluid_amd.params import CONFIGS
    from cube_nufluid_amd.driver import Cube
    from cube_nufluid_amd import _lib

    if args.c4_only:
        print(json.dumps({"c4": c4_bench(args.steps, args.warmup)}), flush=True)
        return
    cfg = CONFIGS[args.config]
    z0 = args.z
    t0 = time.perf_counter()
    evolve = args.evolve_steps > 0 and z0 > 0
    # evolved: the main loop from the Zel'dovich start to the z=0 checkpoint,
    # then the timed steps continue past it; otherwise the next checkpoint
    # redshift lies beyond the timed steps
    cube = Cube.from_ic(cfg, args.seed, z0, device=0, z_checkpoints=[z0, 0.0] if evolve else [z0, min(0.0, z0) - 0.9])
    npart = cube.sim.nplocal
    log("state %s: %d particles, Zel'dovich at z=%.2f generated in HBM (%.1f s incl. init)"
        % (args.config, npart, z0, time.perf_counter() - t0))
    lib = cube.lib
    state = {"start": "seeded Zel'dovich state at z=%.2f" % z0}
    if evolve:
        t1 = time.perf_counter()
        nev = 0
        done = False
        while nev < args.evolve_steps and time.perf_counter() - t1 < args.evolve_seconds and not done:
            done = cube.step()
            nev += 1
            if nev % 25 == 0 or done:
                log("evolve step %d: z=%.3f dt=%.4g (%.1f s)" % (nev, 1 / cube.a - 1, cube.dt, time.perf_counter() - t1))
        z_now = float(1 / cube.a - 1)
        state.update({"evolved_steps": nev, "evolve_seconds": round(time.perf_counter() - t1, 1),
                      "z_reached": round(z_now, 4), "reached_z0": bool(done)})
        if cfg.pp_on and getattr(cube, "last_pp", None) is not None:
            state["pairs_per_particle_at_end"] = round(cube.last_pp.npairs / npart, 2)
    state["description"] = ("%s evolved by the CUBE step on the GPU to z=%.3f in %d steps" % (
        state["start"], state["z_reached"], state["evolved_steps"]) if evolve else state["start"])
    if args.halofind_only:
        print(json.dumps({"halofind": halofind_c3(cube, args.config)}), flush=True)
        cube.close()
        return
    for i in range(args.warmup):
        cube.step()
        log("warmup step %d done" % (i + 1))
    # one profiled step for the per-kernel table (events add no host sync)
    lib.cube_profile_enable(cube.h, 1)
    lib.cube_profile_read(cube.h, None, None, 0)
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for i in range(args.steps):
        cube.step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    nslot = 32
    ms = np.zeros(nslot)
    cnt = np.zeros(nslot, dtype=np.int64)
    ns = lib.cube_profile_read(cube.h, _lib.ptr(ms), _lib.ptr(cnt), nslot)
    names = []
    for s in range(ns):
        b = ctypes.create_string_buffer(64)
        lib.cube_profile_slot_name(s, b, 64)
        names.append(b.value.decode())
    per_kernel = {names[s]: {"ms_total": round(float(ms[s]), 3), "launches": int(cnt[s])} for s in range(ns) if cnt[s]}
    log("per-kernel: " + json.dumps(per_kernel))
    if cfg.pp_on:
        lp = cube.last_pp
        log("pp: npairs %d (%.1f per particle), dense blocks %d" % (lp.npairs, lp.npairs / npart, lp.dense_blocks))
    cands = [s for s in range(ns)
             if cnt[s] and names[s] not in VALU_SLOTS and algorithmic_bytes(names[s], cfg, npart) is not None]
    dom = max(cands, key=lambda s: ms[s])
    avg_ms = ms[dom] / cnt[dom]
    abytes = algorithmic_bytes(names[dom], cfg, npart)
    achieved = abytes / (avg_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "kernel": names[dom], "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": read_traffic(names[dom], args.config),
                "algorithmic_bytes_per_launch": abytes, "mean_launch_ms": round(avg_ms, 4)}
    kick_info = None
    if all(k in names and cnt[names.index(k)] for k in ("kick_fine", "kick_coarse")):
        # the two kick passes together against the bytes one pass over the
        # particles would need (xp, vp read + write once, all six force
        # meshes, vfield, cell offsets)
        kms = [ms[names.index(k)] / cnt[names.index(k)] for k in ("kick_fine", "kick_coarse")]
        kb = npart * (3 * cfg.izipx + 6 * cfg.izipv) + 12 * float(cfg.nf) ** 3 + 32 * float(cfg.nc) ** 3
        kick_info = {"fine_ms": round(kms[0], 3), "coarse_ms": round(kms[1], 3), "ms": round(sum(kms), 3),
                     "algorithmic_bytes": kb, "achieved_gbs": round(kb / (sum(kms) * 1e-3) / 1e9, 1)}
    pp_info = None
    if cfg.pp_on and "pp" in names and cnt[names.index("pp")]:
        s_pp = names.index("pp")
        pp_s = ms[s_pp] / cnt[s_pp] * 1e-3
        pairs = float(cube.last_pp.npairs)
        pp_info = {"bound": "fp32 valu", "pairs_per_step": pairs, "mean_launch_ms": round(pp_s * 1e3, 4),
                   "pairs_per_s": pairs / pp_s, "achieved_tflops": round(30 * pairs / pp_s / 1e12, 3),
                   "peak_tflops": FP32_PEAK_TFLOPS, "flops_per_pair": 30,
                   "pairs_per_particle": round(pairs / npart, 2)}
    value = npart * args.steps / wall
    state["target_particle_steps_per_s"] = 1.5e9
    state["meets_target_on_this_state"] = bool(value >= 1.5e9)
    out = {
        "metric": "particle-steps/s (PM+PP, %d^3 on %d^3 mesh)" % (round(npart ** (1 / 3)), cfg.nf),
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: %s (EH no-wiggle LCDM)" % state["description"],
        "config": {"workload": args.config, "particles": npart, "fine_mesh": cfg.nf, "coarse_mesh": cfg.nc,
                   "box_mpc_h": cfg.box, "z_start": z0, "state": state, "izipx": cfg.izipx, "izipv": cfg.izipv,
                   "pp": bool(cfg.pp_on), "pp_range": cfg.pp_range, "parallelism": "single GPU"},
        "roofline": roofline,
        "kick": kick_info,
        "pp": pp_info,
        "per_kernel_ms": per_kernel,
    }
    st_cpu = None
    if not args.no_cpu_baseline:
        st_cpu = cube.download()
    if not args.no_halofind:
        out["halofind"] = {"state": halofind_c3(cube, args.config)}
    last_dt = float(cube.dt) or 0.05
    cube.close()
    ceil = hbm_ceiling()
    if ceil:
        roofline["measured_ceiling"] = ceil
        roofline["frac_of_measured_ceiling"] = round(achieved / ceil["gbs"], 4)
    if not args.no_halofind:
        if args.halofind_evolved:
            out["halofind"]["evolved"] = halofind_evolved("C2", 1002, cpu=not args.no_cpu_baseline)
    if args.pp_unit:
        out["pp_unit"] = pp_unit_bench("C2", 1002, cpu=not args.no_cpu_baseline)
    if not args.no_c4:
        out["c4"] = c4_bench(min(args.steps, 5), 1)
    if not args.no_ensemble:
        out["ensemble"] = ensemble_bench(args.ens_requests, args.ens_rate)
        if args.ens_small_requests > 0:
            out["ensemble"]["small_boxes_only"] = ensemble_small(args.ens_small_requests, args.ens_rate * 4)
    if st_cpu is not None:
        threads, hinfo = host_threads()
        per_tile = npart / cfg.nnt ** 3
        tiles = args.cpu_tiles or (1 if per_tile > 4e6 else min(threads, cfg.nnt ** 3))
        out["cpu_baseline"] = cpu_baseline(cfg, st_cpu, last_dt, tiles, threads)
        out["cpu_baseline"]["host"] = hinfo
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
