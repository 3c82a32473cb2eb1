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
