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
    print("mix", flush=True)
    mix = run_ensemble(reqs, max_live=8)
    print("alone", flush=True)
    alone = run_ensemble([dict(r, t=0.0) for r in reqs], max_live=1)
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
    print("C5 mix", flush=True)
    mix = run_ensemble(reqs, max_live=8)
    print("C5 alone", flush=True)
    alone = run_ensemble(reqs, max_live=1)
    key = lambda r: (r["npd"], r["seed"])
    hm = {key(r): (r["hash"], r["steps"]) for r in mix["results"]}
    ha = {key(r): (r["hash"], r["steps"]) for r in alone["results"]}
    print("steps by box:", sorted((k[0], v[1]) for k, v in hm.items()), flush=True)
    assert hm == ha
    assert all(abs(r["a_final"] - 1.0) < 1e-5 for r in mix["results"])
    print("mix %.3g particle-steps/s over %.1f s, alone %.3g over %.1f s" % (
        mix["particle_steps_per_s"], mix["span"], alone["particle_steps_per_s"], alone["span"]), flush=True)
