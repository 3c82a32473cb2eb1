"""CPU restatement of CUBE's power-spectrum estimator -- test
infrastructure (the checker), used only by the tests.

Follows utilities/cicpower.f90:85-257 (CIC density of all particles, mass 1,
delta = rho/mean - 1) and utilities/powerspectrum.f90:27-137 (forward
transform divided by ng^3, half-space modes with the kx = 0 / ng/2 plane
duplicates skipped as at :63-66, signed frequencies mod(i+nyq-1,ng)-nyq,
linear bins nint(|k|) with |k| in real(4), sinc^4 CIC window correction,
Delta^2 = 4 pi k^3 |delta_k|^2).  The density is deposited on nodes at
fine-cell left edges; cicpower's cell-centred nodes differ by a uniform
half-cell shift, which leaves |delta_k|^2 unchanged.
"""
import numpy as np


def positions_fine(state, nc, nt, nnt, izipx):
    """Particle positions in fine-cell units (checkpoint order)."""
    from cube_nufluid_amd.ic import rank_to_cell
    n3 = nc ** 3
    r = np.repeat(np.arange(n3), state.rhoc.astype(np.int64))
    cx, cy, cz = rank_to_cell(r, nt, nnt)
    bits = 8 * izipx
    u = state.xp.astype(np.int64) & ((1 << bits) - 1)
    off = (u + 0.5) / (1 << bits)
    return 4.0 * (np.stack([cx, cy, cz], 1) + off)


def cic_density(pos, ng):
    """CIC density on an ng^3 periodic mesh, x fastest (index [z, y, x])."""
    rho = np.zeros((ng, ng, ng))
    i0 = np.floor(pos).astype(np.int64)
    w1 = (i0 + 1) - pos
    w2 = 1.0 - w1
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                w = (w1[:, 0] if dx == 0 else w2[:, 0]) * (w1[:, 1] if dy == 0 else w2[:, 1]) * \
                    (w1[:, 2] if dz == 0 else w2[:, 2])
                np.add.at(rho, ((i0[:, 2] + dz) % ng, (i0[:, 1] + dy) % ng, (i0[:, 0] + dx) % ng), w)
    return rho


def power_bins(rho):
    """Mode counts, sum |k| and sum Delta^2 per linear bin 1..nbin."""
    ng = rho.shape[0]
    nyq = ng // 2
    nbin = int(np.rint(nyq * np.sqrt(3.0)))
    delta = rho / rho.mean() - 1.0
    c = np.fft.rfftn(delta) / ng ** 3          # [kz, ky, kx], kx in 0..nyq
    kz, ky, kx = np.meshgrid(np.arange(ng), np.arange(ng), np.arange(nyq + 1), indexing="ij")
    keep = ~((kx == 0) & (ky == 0) & (kz == 0))
    edge = (kx == 0) | (kx == nyq)
    keep &= ~(edge & (ky > nyq))
    keep &= ~(edge & ((ky == 0) | (ky == nyq)) & (kz > nyq))
    q = [(a + nyq) % ng - nyq for a in (kx, ky, kz)]
    kr = np.sqrt((q[0] ** 2 + q[1] ** 2 + q[2] ** 2).astype(np.float32))
    ib = np.floor(kr + np.float32(0.5)).astype(np.int64)
    sinc = np.ones_like(kr, dtype=np.float64)
    for a in q:
        x = np.pi * a / ng
        sinc *= np.where(a == 0, 1.0, np.sin(x) / np.where(a == 0, 1.0, x))
    amp = np.abs(c) ** 2 / sinc ** 4 * 4 * np.pi * kr.astype(np.float64) ** 3
    ib, kr, amp = ib[keep], kr[keep].astype(np.float64), amp[keep]
    modes = np.bincount(ib, minlength=nbin + 1)[1:nbin + 1].astype(np.float64)
    ksum = np.bincount(ib, weights=kr, minlength=nbin + 1)[1:nbin + 1]
    d2 = np.bincount(ib, weights=amp, minlength=nbin + 1)[1:nbin + 1]
    return modes, ksum, d2
