"""FP32 production path against the FP64 GPU oracle (precision=8: FP64
meshes, CIC weights, FFTs, kernels, kicks and PP pair sums): per-particle
forces on the same state, and P(k) after the same steps from the same
initial conditions, within 0.1% up to half the fine-mesh Nyquist
frequency."""
import numpy as np
import pytest

from cube_nufluid_amd.params import BoxConfig
from cube_nufluid_amd.ic import zeldovich_ic

pytestmark = pytest.mark.gpu


def rel_rms(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2)))


def test_forces_fp32_vs_fp64():
    from cube_nufluid_amd.driver import Cube
    cfg = BoxConfig(nc=32, nnt=2, ncb=6, box=100.0, pp_on=True)
    st = zeldovich_ic(cfg, seed=41, z=1.0)
    out = {}
    for prec in (4, 8):
        cube = Cube(cfg, st, precision=prec)
        out[prec] = cube.probe_pm_forces() + (cube.probe_pp_forces(),)
        cube.close()
    for name, a, b in zip(("fine", "coarse", "pp"), out[4], out[8]):
        err = rel_rms(a, b)
        print("force", name, "rel RMS FP32 vs FP64", err, flush=True)
        assert err < 1e-5, name


def test_power_spectrum_fp32_vs_fp64():
    from cube_nufluid_amd.driver import Cube
    cfg = BoxConfig(nc=32, nnt=2, ncb=6, box=100.0, pp_on=True)
    z0 = 5.0
    st = zeldovich_ic(cfg, seed=42, z=z0)
    res = {}
    for prec in (4, 8):
        cube = Cube(cfg, st, precision=prec, z_checkpoints=[z0, 0.0])
        for _ in range(8):
            if cube.step():
                break
        res[prec] = cube.power_spectrum()
        print("precision", prec, "a =", float(cube.a), flush=True)
        cube.close()
    k, d32, modes = res[4]
    _, d64, _ = res[8]
    half = np.arange(1, len(k) + 1) <= cfg.nf // 4  # bins up to k_Nyq/2
    err = np.abs(d32[half] / d64[half] - 1)
    print("P(k) max rel err up to k_Nyq/2:", float(err.max()), flush=True)
    assert err.max() < 1e-3


TS_FIELDS = ("dt_old", "dt", "da", "a", "a_mid", "t", "tau", "checkpoint_step", "final_step", "halofind_step")


def run_to_z0(cfg, seed, prec, replay=None):
    """z=49 -> 0 through the host loop; records the time steps, or replays a
    recorded sequence (timestep.f90 outputs) so that both precisions take
    the same steps."""
    from cube_nufluid_amd.driver import Cube
    st = zeldovich_ic(cfg, seed=seed, z=49.0)
    cube = Cube(cfg, st, precision=prec, z_checkpoints=[49.0, 0.0])
    rec = []
    if replay is None:
        step_dt = cube.timestep

        def timestep():
            step_dt()
            rec.append(tuple(getattr(cube, f) for f in TS_FIELDS))
    else:
        seq = iter(replay)

        def timestep():
            for f, v in zip(TS_FIELDS, next(seq)):
                setattr(cube, f, v)
    cube.timestep = timestep
    n = cube.run(max_steps=5000)
    pk = cube.power_spectrum()
    a = float(cube.a)
    cube.close()
    return pk, rec, n, a


def test_power_spectrum_z0_fp32_vs_fp64():
    """A 128^3-particle box (200 Mpc/h, PP on) run from z=49 to z=0 in FP32
    and in FP64 over the same time steps: P(k) at z=0 within 0.1% up to
    k_Nyq/2."""
    cfg = BoxConfig(nc=64, nnt=2, ncb=6, box=200.0, pp_on=True)
    p32, rec, n32, a32 = run_to_z0(cfg, 1002, 4)
    p64, _, n64, a64 = run_to_z0(cfg, 1002, 8, replay=rec)
    assert n32 == n64 and a32 == a64 and abs(a32 - 1.0) < 1e-5
    k, d32, _ = p32
    _, d64, _ = p64
    half = np.arange(1, len(k) + 1) <= cfg.nf // 4
    err = np.abs(d32[half] / d64[half] - 1)
    print("z=0 after %d steps: P(k) max rel err up to k_Nyq/2 %.2e" % (n32, err.max()), flush=True)
    assert err.max() < 1e-3
