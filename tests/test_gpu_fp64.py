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
