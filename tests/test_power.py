"""The power-spectrum estimator restated on the CPU (oracle/power.py),
pinned by an exact mode-count identity and by the closed-form expectation
of a Poisson sample (shot noise with the CIC aliasing sum); the GPU
estimator through the C ABI against it."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import power  # noqa: E402


@pytest.mark.parametrize("ng", [8, 16, 32])
def test_mode_count_identity(ng):
    """Half-space modes without k=0: (ng^3 - 8)/2 conjugate pairs plus the 7
    self-conjugate modes (components in {0, ng/2}); all land in bins 1..nbin."""
    rng = np.random.default_rng(ng)
    modes, _, _ = power.power_bins(rng.random((ng, ng, ng)) + 1.0)
    assert modes.sum() == (ng ** 3 - 8) // 2 + 7
    assert len(modes) == int(np.rint(ng // 2 * np.sqrt(3)))


def test_poisson_shot_noise_closed_form():
    """Uniform random particles: <|delta_k|^2> = prod_d(1 - 2/3 sin^2(pi k_d/ng)) / Np
    (CIC window summed over aliases), so Delta^2 * sinc^4 / (4 pi k^3) per bin
    matches that average."""
    ng, npart = 32, 400000
    rng = np.random.default_rng(5)
    pos = rng.random((npart, 3)) * ng
    rho = power.cic_density(pos, ng)
    modes, ksum, d2 = power.power_bins(rho)
    # expected Delta^2 per bin from the closed form, binned like the estimator
    nyq = ng // 2
    kz, ky, kx = np.meshgrid(np.arange(ng), np.arange(ng), np.arange(nyq + 1), indexing="ij")
    keep = ~((kx == 0) & (ky == 0) & (kz == 0))
    edge = (kx == 0) | (kx == nyq)
    keep &= ~(edge & (ky > nyq)) & ~(edge & ((ky == 0) | (ky == nyq)) & (kz > nyq))
    q = [(a + nyq) % ng - nyq for a in (kx, ky, kz)]
    kr = np.sqrt((q[0] ** 2 + q[1] ** 2 + q[2] ** 2).astype(np.float32))
    ib = np.floor(kr + np.float32(0.5)).astype(np.int64)
    alias = np.ones(kr.shape)
    sinc = np.ones(kr.shape)
    for a in q:
        x = np.pi * a / ng
        alias *= 1 - 2.0 / 3.0 * np.sin(x) ** 2
        sinc *= np.where(a == 0, 1.0, np.sin(x) / np.where(a == 0, 1.0, x))
    expect = alias / sinc ** 4 / npart * 4 * np.pi * kr.astype(np.float64) ** 3
    nbin = len(modes)
    e2 = np.bincount(ib[keep], weights=expect[keep], minlength=nbin + 1)[1:]
    sel = modes >= 200
    ratio = d2[sel] / e2[sel]
    sigma = 1 / np.sqrt(modes[sel])
    assert np.all(np.abs(ratio - 1) < 5 * sigma)
    assert abs(np.average(ratio, weights=modes[sel]) - 1) < 0.01


@pytest.mark.gpu
def test_gpu_power_spectrum_matches_restatement():
    from cube_nufluid_amd.params import BoxConfig
    from cube_nufluid_amd.ic import zeldovich_ic
    from cube_nufluid_amd.driver import Cube
    cfg = BoxConfig(nc=32, nnt=2, ncb=6, box=100.0)
    st = zeldovich_ic(cfg, seed=31, z=0.0)
    cube = Cube(cfg, st)
    k, d2, modes = cube.power_spectrum()
    cube.close()
    pos = power.positions_fine(st, cfg.nc, cfg.nt, cfg.nnt, cfg.izipx)
    om, oks, od2 = power.power_bins(power.cic_density(pos, cfg.nf))
    assert np.array_equal(modes, om)
    ok = om > 0
    np.testing.assert_allclose(k[ok], oks[ok] / om[ok] * 2 * np.pi / cfg.box, rtol=1e-12)
    np.testing.assert_allclose(d2[ok], od2[ok] / om[ok], rtol=1e-4)
