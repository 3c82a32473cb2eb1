"""The benchmark configurations at their own sizes, the neutrino-fluid
initial conditions, checkpoint resume and the fail-stop error returns,
through the C ABI on the GPU.

Bars as in test_gpu_parity: integer results bit-exact given identical float
inputs, per-particle forces within 1e-5 relative RMS of the oracle, kicks
within one velocity code."""
import ctypes
import os
import sys

import numpy as np
import pytest

from cube_nufluid_amd.params import BoxConfig, CONFIGS
from cube_nufluid_amd import _lib

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FORCE_RTOL = 1e-5


def rel_rms(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2)))


def _restated_fluids(cfg, seed, z):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import neutrino_ic
    return neutrino_ic.fluid_ic(cfg, seed, z, cfg.m_nu_ev, cfg.z_i_nu)


# ---------------------------------------------------------- fluid ICs
def test_gpu_neutrino_ic_matches_restatement():
    """cube_nu_ic against the numpy restatement of neutrino_ic.f90 on the
    same seeded field (fluids 64^3, 32^3, 16^3), and the CDM mass scaled by
    f_cdm: the fine-mesh mass with the fluid coupled is nf^3 (mean density
    1), the CDM share f_cdm nf^3."""
    from cube_nufluid_amd.driver import Cube
    cfg = BoxConfig(nc=32, nnt=2, ncb=6, box=500.0 * 32 / 256, pp_on=True, z_i=10.0, m_nu_ev=0.1 / 3)
    print("fluid IC nc=32", flush=True)
    cube = Cube.from_ic(cfg, seed=77, z=10.0)
    try:
        ref = _restated_fluids(cfg, 77, 10.0)
        for n in range(3):
            g = cube.nu_get(n)
            r = ref[n]
            assert g.shape == r.shape
            np.testing.assert_allclose(g[0], r[0], rtol=2e-6)
            for q in (1, 2, 3):
                scale = np.abs(r[q]).max()
                assert np.abs(g[q] - r[q]).max() <= 1e-5 * scale, (n, q)
            np.testing.assert_allclose(g[4], r[4], rtol=1e-5)
        m = float(cube.sim.mass_p_cdm)
        assert m == pytest.approx(cfg.f_cdm * cfg.nf ** 3 / cfg.npart, rel=1e-6)
        a_mid, dt = np.float32(1 / 11.0), np.float32(0.0)
        cube.a_mid, cube.dt = a_mid, dt
        out = cube.particle_mesh()
        print("mass_f", out.mass_f, "nf^3", cfg.nf ** 3, flush=True)
        assert out.mass_f == pytest.approx(cfg.nf ** 3, rel=2e-6)
    finally:
        cube.close()


# ------------------------------------------------------------- resume
def _files(opath, z):
    d = os.path.join(opath, "image1")
    from cube_nufluid_amd.checkpoint import z2str
    return sorted(f for f in os.listdir(d) if f.startswith(z2str(z) + "_"))


def test_gpu_resume_repeats_straight_run(tmp_path):
    """Run z=10 -> 6 -> 3 straight, writing both checkpoints; then resume
    from the z=6 checkpoint in a fresh handle and run to 3.  With the
    header's a, t, tau (exact=True) the final checkpoint files are byte for
    byte the straight run's.  Resuming as the reference does (a from the
    redshift list) changes a in its last bits, and the final state stays
    within a few codes of the straight run.  PP is off: the reference keeps
    the PP force maximum behind dt_pp as a running maximum over all steps
    of a run (ftemp, ext_pp_force.f90:121,128), which no checkpoint carries, so a
    resumed PP run restarts it, in the reference as here."""
    from cube_nufluid_amd.driver import Cube
    from cube_nufluid_amd.checkpoint import read_checkpoint
    cfg = BoxConfig(nc=32, nnt=2, ncb=6, box=100.0, pp_on=False)
    zs = [10.0, 6.0, 3.0]
    straight, exact, ref = (str(tmp_path / d) + "/" for d in ("straight", "exact", "ref"))
    print("resume: straight run", flush=True)
    c = Cube.from_ic(cfg, seed=4242, z=zs[0], z_checkpoints=zs)
    n_straight = c.run(max_steps=500, opath=straight)
    c.close()
    # the resumed runs read the straight run's z=6 files
    for d in (exact, ref):
        os.makedirs(d + "image1", exist_ok=True)
        for f in _files(straight, zs[1]):
            with open(straight + "image1/" + f, "rb") as a, open(d + "image1/" + f, "wb") as b:
                b.write(a.read())
    print("resume: exact", flush=True)
    c = Cube.from_checkpoint(cfg, exact, zs[1:], exact=True)
    c.run(max_steps=500, opath=exact)
    c.close()
    fa, fb = _files(straight, zs[2]), _files(exact, zs[2])
    assert fa == fb and len(fa) == 6
    for f in fa:
        with open(straight + "image1/" + f, "rb") as a, open(exact + "image1/" + f, "rb") as b:
            da, db = a.read(), b.read()
        if "_info_" in f:
            # cur_checkpoint counts entries of each run's own redshift list
            ha, hb = _lib.SimHeader.from_bytes(da), _lib.SimHeader.from_bytes(db)
            assert ha.cur_checkpoint == 3 and hb.cur_checkpoint == 2
            hb.cur_checkpoint = ha.cur_checkpoint
            db = hb.to_bytes()
        assert da == db, f
    print("resume: reference semantics", flush=True)
    c = Cube.from_checkpoint(cfg, ref, zs[1:], exact=False)
    c.run(max_steps=500, opath=ref)
    c.close()
    s, r = read_checkpoint(straight, zs[2], cfg.nc), read_checkpoint(ref, zs[2], cfg.nc)
    print("steps", n_straight, "a straight / resumed", s.sim.a, r.sim.a, flush=True)
    assert s.sim.nplocal == r.sim.nplocal
    assert abs(s.sim.a / r.sim.a - 1) < 1e-5
    # same particles; positions within a few position codes
    ps, pr = np.argsort(s.pid), np.argsort(r.pid)
    assert np.array_equal(s.pid[ps], r.pid[pr])
    assert np.mean(s.xp[ps] != r.xp[pr]) < 0.5


def test_gpu_resume_with_fluid_restores_fields(tmp_path):
    """A CDM + neutrino-fluid run writes the z-prefixed neu1 file beside the
    particle checkpoint (main.f90:75-81); resuming reads it back (hydrodf
    dfld_restart) into exactly the fields the run held, and writing it again
    gives the same bytes."""
    from cube_nufluid_amd.driver import Cube
    cfg = BoxConfig(nc=16, nnt=1, ncb=6, box=500.0 * 16 / 256, pp_on=True, z_i=10.0, m_nu_ev=0.1 / 3)
    zs = [10.0, 8.0, 7.0]
    run = str(tmp_path / "run") + "/"
    print("fluid resume run", flush=True)
    c = Cube.from_ic(cfg, seed=99, z=zs[0], z_checkpoints=zs)
    steps = 0
    while c.cur_checkpoint < 3 and steps < 200:   # until the z=8 checkpoint is written
        c.step(opath=run)
        steps += 1
    held = [c.nu_get(n) for n in range(3)]
    c.close()
    from cube_nufluid_amd.checkpoint import output_name
    fn = output_name(run, zs[1], "neu1")
    assert os.path.exists(fn)
    r = Cube.from_checkpoint(cfg, run, zs[1:], exact=True)
    try:
        for n in range(3):
            assert np.array_equal(r.nu_get(n), held[n]), n
        fn2 = str(tmp_path / "again.bin")
        r.nu_checkpoint(fn2)
        with open(fn, "rb") as a, open(fn2, "rb") as b:
            assert a.read() == b.read()
        r.run(max_steps=200, opath=run)
    finally:
        r.close()


# ----------------------------------------------------------- fail-stop
def test_gpu_error_returns():
    """The reference's stops as error returns (the Fortran host turns them
    into error stop): a header whose zip format differs from the handle's
    (particle_initialization.f90:41-45), more particles than np_max, and a
    drift that would move a particle beyond the buffer
    (update_particle.f90:90-96).  After the failed drift the handle still
    holds its state and a normal drift succeeds."""
    from cube_nufluid_amd.driver import Cube
    from cube_nufluid_amd.ic import zeldovich_ic
    cfg = BoxConfig(nc=16, nnt=1, ncb=6, box=50.0)
    st = zeldovich_ic(cfg, seed=5, z=5.0)
    print("error returns", flush=True)
    cube = Cube(cfg, st)
    try:
        lib = cube.lib
        bad = _lib.SimHeader.from_bytes(st.sim.to_bytes())
        bad.izipx = 1
        rc = lib.cube_upload_state(cube.h, ctypes.byref(bad), _lib.ptr(st.xp), _lib.ptr(st.vp), _lib.ptr(st.pid),
                                   _lib.ptr(st.rhoc), _lib.ptr(st.vfield))
        assert rc == 3
        big = _lib.SimHeader.from_bytes(st.sim.to_bytes())
        big.nplocal = cube.params.np_max + 1
        rc = lib.cube_upload_state(cube.h, ctypes.byref(big), _lib.ptr(st.xp), _lib.ptr(st.vp), _lib.ptr(st.pid),
                                   _lib.ptr(st.rhoc), _lib.ptr(st.vfield))
        assert rc == 2
        # a bulk flow of 10 coarse cells in one drift: beyond min(ncb, 7)
        fast = st.copy()
        fast.vfield[:, 0] += np.float32(40.0)
        cube.upload(fast)
        out = _lib.DriftOut()
        rc = lib.cube_drift(cube.h, ctypes.c_float(1.0), ctypes.c_float(1.0), ctypes.byref(out))
        err = ctypes.create_string_buffer(256)
        lib.cube_last_error(cube.h, err, 256)
        print("drift:", rc, err.value.decode(), flush=True)
        assert rc == 2 and b"moved more than" in err.value
        kept = cube.download()
        assert np.array_equal(kept.xp, fast.xp) and np.array_equal(kept.pid, fast.pid)
        cube.upload(st)
        cube.dt_old, cube.dt = np.float32(0.05), np.float32(0.05)
        assert cube.update_x().npcheck == st.sim.nplocal
    finally:
        cube.close()
    # the IC of a lattice larger than the capacity
    small = BoxConfig(nc=16, nnt=1, ncb=6, box=50.0, np_max_factor=0.5)
    with pytest.raises(_lib.CubeError) as e:
        Cube.from_ic(small, seed=1, z=5.0)
    assert e.value.code == 2
    # a coarse buffer whose deposited particles do not reach the fine
    # kernel's 15-cell support (4*(ncb-1) < 16)
    thin = BoxConfig(nc=16, nnt=1, ncb=4, box=50.0)
    with pytest.raises(_lib.CubeError) as e:
        Cube(thin, st)
    assert e.value.code == 1 and "ncb" in str(e.value)


# --------------------------------------------------------- sizes of C2
def test_gpu_c2_size_parity():
    """C2 at its own size (256^3 particles, 512^3 fine mesh, 400 Mpc/h,
    izipx=1, izipv=2, PP with pp_range 1): GPU IC at z=49, then the oracle
    on the same state -- drift bit-exact, fine/coarse/PP forces within
    1e-5, PP pair counts equal, kicks within two codes (one per
    re-quantisation) in under 0.1% of components."""
    from cube_nufluid_amd.driver import Cube
    from oracle_lib import Oracle
    cfg = CONFIGS["C2"]
    print("C2 at size: GPU IC", flush=True)
    cube = Cube.from_ic(cfg, seed=1002, z=49.0)
    try:
        st = cube.download()
        threads = min(16, os.cpu_count() or 1)
        orc = Oracle(cfg, threads=threads)
        orc.set_state(st)
        gf, gc = cube.probe_pm_forces()
        of, oc, _ = orc.pm_forces()
        print("C2 forces fine %.2e coarse %.2e" % (rel_rms(gf, of), rel_rms(gc, oc)), flush=True)
        assert rel_rms(gf, of) < FORCE_RTOL and rel_rms(gc, oc) < FORCE_RTOL
        gp = cube.probe_pp_forces()
        op = orc.pp(1.0, 1.0, forces=True)["f_pp"].astype(np.float64)
        print("C2 pp %.2e" % rel_rms(gp, op), flush=True)
        assert rel_rms(gp, op) < FORCE_RTOL
        orc.set_state(st)
        dt = np.float32(0.05)
        cube.dt_old, cube.dt = dt, dt
        g = cube.update_x()
        o = orc.drift(float(dt), float(dt))
        a, b = cube.download(), orc.get_state()
        assert g.npcheck == o["np"] == st.sim.nplocal
        for f in ("rhoc", "pid", "xp", "vp", "vfield"):
            assert np.array_equal(getattr(a, f), getattr(b, f)), f
        a_mid = np.float32(0.02)
        cube.a_mid, cube.dt = a_mid, dt
        gpp = cube.ext_pp_force()
        opp = orc.pp(float(a_mid), float(dt))
        assert gpp.npairs == opp["npairs"] and gpp.itest1 == opp["itest1"]
        cube.particle_mesh()
        orc.pm(float(a_mid), float(dt))
        a, b = cube.download(), orc.get_state()
        diff = np.abs(a.vp.astype(np.int64) - b.vp.astype(np.int64))
        print("C2 kick mismatch %.2e max %d" % (np.mean(diff > 0), diff.max()), flush=True)
        # each code is re-quantised after the fine and after the coarse kick,
        # and each can straddle a boundary under the two sides' force rounding
        assert diff.max() <= 2 and np.mean(diff > 0) < 1e-3
        orc.close()
    finally:
        cube.close()


# --------------------------------------------------------- sizes of C3
def test_gpu_c3_size_first_tile_and_properties():
    """C3 at its own size (1024^3 particles, 2048^3 fine and 512^3 coarse
    mesh, 1 Gpc/h, PP): GPU IC, then
      * the drift of the first of 64 tiles bit-exact against the oracle run
        on that tile (orc_set_tile_limit) with the whole box around it,
      * PP and PM kicks of the first tile's particles within two codes (one
        per re-quantisation) in under 0.1% of components,
      * npcheck = N and mass_f = N mass_p,
      * the same step repeated from the same state is bit-identical."""
    from cube_nufluid_amd.driver import Cube
    from oracle_lib import Oracle
    cfg = CONFIGS["C3"]
    print("C3 at size: GPU IC", flush=True)
    cube = Cube.from_ic(cfg, seed=1003, z=1.0)
    try:
        st0 = cube.download()
        n = st0.sim.nplocal
        assert n == 1024 ** 3
        nt3 = cfg.nt ** 3
        dt, a_mid = np.float32(0.02), np.float32(0.45)

        def full_step():
            cube.upload(st0)
            cube.dt_old, cube.dt = dt, dt
            g = cube.update_x()
            cube.a_mid, cube.dt = a_mid, dt
            cube.ext_pp_force()
            pm = cube.particle_mesh()
            return g, pm
        g1, pm1 = full_step()
        assert g1.npcheck == n
        assert pm1.mass_f == pytest.approx(n * float(st0.sim.mass_p_cdm), rel=1e-6)
        s1 = cube.download()
        g2, pm2 = full_step()
        s2 = cube.download()
        for f in ("xp", "vp", "pid", "rhoc", "vfield"):
            assert np.array_equal(getattr(s1, f), getattr(s2, f)), f
        del s1, s2
        print("C3 repeat bit-identical; oracle on the first tile", flush=True)
        threads = min(16, os.cpu_count() or 1)
        orc = Oracle(cfg, threads=threads)
        orc.lib.orc_set_tile_limit(orc.o, 1)
        # drift of tile 1
        orc.set_state(st0)
        cube.upload(st0)
        cube.dt_old, cube.dt = dt, dt
        cube.update_x()
        o = orc.drift(float(dt), float(dt))
        a, b = cube.download(), orc.get_state()
        n1 = int(b.rhoc[:nt3].sum())
        assert n1 == int(a.rhoc[:nt3].sum()) > 0
        assert np.array_equal(a.rhoc[:nt3], b.rhoc[:nt3])
        for f in ("xp", "vp", "pid"):
            assert np.array_equal(getattr(a, f)[:n1], getattr(b, f)[:n1]), f
        assert np.array_equal(a.vfield[:nt3], b.vfield[:nt3])
        del a, b
        # PP + PM kicks of tile 1 on the undrifted state
        orc.set_state(st0)
        cube.upload(st0)
        cube.a_mid, cube.dt = a_mid, dt
        gpp = cube.ext_pp_force()
        orc.pp(float(a_mid), float(dt))
        cube.particle_mesh()
        orc.pm(float(a_mid), float(dt))
        a, b = cube.download(), orc.get_state()
        m1 = int(st0.rhoc[:nt3].sum())
        diff = np.abs(a.vp[:m1].astype(np.int64) - b.vp[:m1].astype(np.int64))
        print("C3 tile-1 kick mismatch %.2e max %d; pairs/particle %.1f" % (np.mean(diff > 0), diff.max(),
                                                                           gpp.npairs / n), flush=True)
        assert diff.max() <= 2 and np.mean(diff > 0) < 1e-3
        orc.close()
    finally:
        cube.close()


# --------------------------------------------------------- sizes of C4
def test_gpu_c4_size_fluids():
    """C4 at its size: 512^3 CDM particles on the 1024^3 fine mesh with the
    three neutrino fluids on 512^3, 256^3 and 128^3 cells, generated on the
    GPU (cube_nu_ic).  Fluid means are 1 and the fine-mesh mass with the
    fluid coupled is nf^3; one hydro step of every fluid in each Strang
    order equals the restatement (oracle/fluid_oracle.c) run from the same
    downloaded fields, to rtol 1e-6, with the same CFL step."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    from cube_nufluid_amd.checkpoint import neu_ncb
    from cube_nufluid_amd.driver import Cube
    cfg = CONFIGS["C4"]
    print("C4 from_ic", flush=True)
    cube = Cube.from_ic(cfg, seed=1004, z=cfg.z_i)
    try:
        infos = [cube.nu_info(n) for n in range(3)]
        assert [i["nc"] for i in infos] == [512, 256, 128]
        cur = []
        for n in range(3):
            f = cube.nu_get(n)
            assert abs(float(f[0].mean(dtype=np.float64)) - 1) < 1e-5
            cur.append(f)
        a_mid, dt0 = np.float32(1 / 11.0), np.float32(0.0)
        cube.a_mid, cube.dt = a_mid, dt0
        out = cube.particle_mesh()
        print("C4 mass_f", out.mass_f, "nf^3", cfg.nf ** 3, flush=True)
        assert out.mass_f == pytest.approx(cfg.nf ** 3, rel=2e-6)
        for n in range(3):
            assert np.array_equal(cube.nu_get(n), cur[n])      # dt = 0 leaves the fluids
        dt = np.float32(0.5) * np.float32(cube.dt_neu)
        for direction in (1, -1):
            cube.dt_neu = np.float32(1000)
            print("C4 fluid step, direction", direction, "dt", float(dt), flush=True)
            cube.fluid_evolve(dt, direction, float(a_mid))
            dt_all = 1000.0
            for n in range(3):
                ref, dt_all = ol.fluid_evolve(cur[n], neu_ncb(infos[n]["dx"]), np.float32(5.0 / 3.0), float(dt),
                                              infos[n]["dx"], direction, dt_all)
                got = cube.nu_get(n)
                err = float(np.abs(got - ref).max())
                print("  fluid", n, "max abs diff", err, flush=True)
                np.testing.assert_allclose(got, ref, rtol=1e-6, atol=1e-7)
                assert not np.array_equal(got, cur[n])
                cur[n] = got
            assert abs(float(cube.dt_neu) / dt_all - 1) < 1e-6
    finally:
        cube.close()


# ------------------------------------------------ P(k) against the CPU path
def test_gpu_pk_z0_against_cpu_path():
    """P(k) at z = 0 of the GPU run against the CPU path (the oracle's
    restatement of CUBE's step): a 32^3-particle, 100 Mpc/h box with PP,
    Zel'dovich start at z = 49, run to z = 0 by the GPU host loop; the
    oracle replays the same time steps (drift, PP, PM, and the closing half
    drift) from the same initial state.  Both final states go through the
    same estimator (oracle/power.py); P(k) agrees within 0.1% up to k_Nyq/2
    of the fine mesh.  The oracle's wall time is printed."""
    import time
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_lib as ol
    import power
    from cube_nufluid_amd.driver import Cube
    from cube_nufluid_amd.ic import zeldovich_ic
    cfg = BoxConfig(nc=16, nnt=1, ncb=6, box=100.0, pp_on=True, pp_range=1)
    st0 = zeldovich_ic(cfg, seed=2024, z=49.0)
    cube = Cube(cfg, st0, z_checkpoints=[49.0, 0.0])
    seq = []
    step_dt = cube.timestep

    def timestep():
        step_dt()
        seq.append((float(cube.dt_old), float(cube.dt), float(cube.a_mid), bool(cube.checkpoint_step)))
    cube.timestep = timestep
    t = time.perf_counter()
    n = cube.run(max_steps=5000)
    print("GPU: %d steps to a = %.6f in %.1f s" % (n, float(cube.a), time.perf_counter() - t), flush=True)
    assert abs(float(cube.a) - 1) < 1e-5
    sg = cube.download()
    cube.close()
    o = ol.Oracle(cfg, threads=min(16, os.cpu_count() or 1))
    o.set_state(st0)
    t = time.perf_counter()
    for dt_old, dt, a_mid, ck in seq:
        o.drift(dt_old, dt)
        o.pp(a_mid, dt)
        o.pm(a_mid, dt)
        if ck:
            o.drift(0.0, dt)
    t_cpu = time.perf_counter() - t
    so = o.get_state()
    o.close()
    print("CPU path (oracle, %d threads): %d steps in %.1f s" % (o.threads, len(seq), t_cpu), flush=True)

    def pk(s):
        pos = power.positions_fine(s, cfg.nc, cfg.nt, cfg.nnt, cfg.izipx)
        return power.power_bins(power.cic_density(pos, cfg.nf))
    mg, kg, dg = pk(sg)
    mo, ko, do = pk(so)
    assert np.array_equal(mg, mo)
    nb = cfg.nf // 4
    err = np.abs(dg[:nb] / do[:nb] - 1)
    print("P(k) z=0 GPU vs CPU path, max |dP/P| up to k_Nyq/2: %.2e; per bin %s" % (err.max(), np.round(err, 5)),
          flush=True)
    assert err.max() < 1e-3
