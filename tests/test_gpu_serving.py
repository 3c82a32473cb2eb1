"""Ensemble serving through the native engine: boxes that arrive while
others run are admitted and overlap on the device; each box's final state
is bit-identical alone (one at a time) and in the mix, and equal to the
state the Python host reaches stepping the same box through the C ABI."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def fnv1a(*arrays):
    h = 1469598103934665603
    for a in arrays:
        for b in np.ascontiguousarray(a).view(np.uint8).tobytes():
            h ^= b
            h = (h * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def test_mix_equals_alone_equals_python_host():
    from cube_nufluid_amd.serving import run_ensemble, ensemble_config
    from cube_nufluid_amd.driver import Cube
    reqs = [dict(t=0.002 * i, npd=[32, 64, 32, 64, 32][i], seed=1000 + i, z_i=20.0, z_f=8.0) for i in range(5)]
    print("mix (steps replayed as hipGraphs)", flush=True)
    mix = run_ensemble(reqs, max_live=8, step_mode=2)
    print("alone (synchronising phases)", flush=True)
    alone = run_ensemble([dict(r, t=0.0) for r in reqs], max_live=1, step_mode=0)
    key = lambda r: (r["npd"], r["seed"])
    hm = {key(r): (r["hash"], r["steps"]) for r in mix["results"]}
    ha = {key(r): (r["hash"], r["steps"]) for r in alone["results"]}
    assert hm == ha
    # overlap: in the mix some two boxes were live on the device together
    iv = [(r["t_submit"] + r["wait"], r["t_done"]) for r in mix["results"]]
    assert any(a0 < b1 and b0 < a1 for i, (a0, a1) in enumerate(iv) for (b0, b1) in iv[i + 1:])
    # the native host loop (cube_step) equals the Python host loop
    cfg = ensemble_config(32)
    cube = Cube.from_ic(cfg, 1000, 20.0, z_checkpoints=[20.0, 8.0])
    cube.run(max_steps=10000)
    st = cube.download()
    cube.close()
    assert fnv1a(st.xp, st.vp, st.pid) == hm[(32, 1000)][0]


def test_c5_sizes_mix_equals_alone_z49_to_0():
    """C5 as configured: boxes of 32^3, 64^3, 128^3 and 256^3 particles
    (meshes 64^3 to 512^3), Zel'dovich start at z = 49, run to z = 0 with
    CUBE's time-step limits; submitted together, every box's final state is
    bit-identical to the same box run alone, and each takes the step count
    it takes alone."""
    from cube_nufluid_amd.serving import run_ensemble
    reqs = [dict(t=0.0, npd=n, seed=1100 + i, z_i=49.0, z_f=0.0) for i, n in enumerate((256, 128, 64, 32, 128))]
    print("C5 mix (hipGraph steps)", flush=True)
    mix = run_ensemble(reqs, max_live=8, step_mode=2)
    print("C5 alone (synchronising phases)", flush=True)
    alone = run_ensemble(reqs, max_live=1, step_mode=0)
    key = lambda r: (r["npd"], r["seed"])
    hm = {key(r): (r["hash"], r["steps"]) for r in mix["results"]}
    ha = {key(r): (r["hash"], r["steps"]) for r in alone["results"]}
    print("steps by box:", sorted((k[0], v[1]) for k, v in hm.items()), flush=True)
    assert hm == ha
    assert all(abs(r["a_final"] - 1.0) < 1e-5 for r in mix["results"])
    print("mix %.3g particle-steps/s over %.1f s, alone %.3g over %.1f s" % (
        mix["particle_steps_per_s"], mix["span"], alone["particle_steps_per_s"], alone["span"]), flush=True)


def _native_run(cfg, seed, z0, z1, mode, max_steps):
    """The box stepped by the native host loop (cube_step) in one step mode;
    returns the final-state hash, the clock and the graph counters."""
    import ctypes
    from cube_nufluid_amd import _lib
    from cube_nufluid_amd.driver import Cube
    cube = Cube.from_ic(cfg, seed, z0, z_checkpoints=[z0, z1])
    lib = cube.lib
    try:
        assert lib.cube_set_step_mode(cube.h, mode) == 0
        clk = _lib.CubeClock()
        z = np.array([z0, z1], dtype=np.float32)
        assert lib.cube_clock_init(ctypes.byref(clk), _lib.ptr(z), 2) == 0
        n = 0
        while n < max_steps and not (clk.final_step and clk.checkpoint_step):
            _lib.check(lib, cube.h, lib.cube_step(cube.h, ctypes.byref(clk)))
            n += 1
        st = cube.download()
        gi, gu = ctypes.c_int64(), ctypes.c_int64()
        lib.cube_step_stats(cube.h, ctypes.byref(gi), ctypes.byref(gu))
        return fnv1a(st.xp, st.vp, st.pid, st.rhoc, st.vfield), (clk.a, clk.dt, clk.istep), (gi.value, gu.value)
    finally:
        cube.close()


def test_step_modes_give_the_same_bits():
    """cube_step with drift, PP and PM as synchronising phases (mode 0),
    launched back to back with one synchronisation per step (mode 1), and
    replayed as hipGraphs with their arguments updated in place (mode 2):
    the same state, bit for bit, after the same steps, for a 64^3-particle
    PM+PP box (two tiles per dimension) from z = 49 to z = 5."""
    from cube_nufluid_amd.serving import ensemble_config
    cfg = ensemble_config(64)
    runs = {}
    for mode in (0, 1, 2):
        print("step mode", mode, flush=True)
        runs[mode] = _native_run(cfg, 1200, 49.0, 5.0, mode, 400)
        print("  hash %x clock %s graphs %s" % (runs[mode][0], runs[mode][1], runs[mode][2]), flush=True)
    assert runs[0][:2] == runs[1][:2] == runs[2][:2]
    gi, gu = runs[2][2]
    # every step was captured; all but the first updated the instantiated graph
    assert gi >= 1 and gu >= 1 and gi + gu == runs[2][1][2]
