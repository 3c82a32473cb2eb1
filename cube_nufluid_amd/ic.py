"""Seeded Zel'dovich initial conditions in CUBE's compressed format.

Follows the encoding conventions of util/initial_conditions.f90:759-954:
particles start on a lattice of np_nc per coarse cell per dimension at
q = (i-1)/np_nc + 0.5/ncell (coarse units), are displaced by the
Zel'dovich field, binned into cells, stored as the in-cell offset
floor(frac(x)/x_resolution) narrowed to izipx bytes, velocities stored
relative to the coarse-cell mean vfield and encoded with
nint((nvbin-1)*atan(sqrt(pi/2)/(sigma_vi*vrel_boost)*v)/pi).

Differences from the reference generator, which reads CAMB tables and
sigma_v tables that the target configurations do not use:
  * the linear spectrum is the Eisenstein & Hu (1998) no-wiggle fit,
    eqs. (26)-(31), normalised to sigma_8 with a top-hat of 8 Mpc/h;
  * the noise is SplitMix64 + Box-Muller per lattice site (reproducible on
    the GPU by cube_ic_zeldovich);
  * the displacement is computed spectrally on the particle lattice,
    psi_k = i k delta_k / k^2, rather than by central differences of the
    fine-grid potential;
  * the growth factor is the fitting form the reference uses (Dgrow,
    main.f90:104-117), the velocity is psi * a^2 H(a) * f(a) with
    f = Omega_m(a)^0.55;
  * sigma_vi is the measured rms residual velocity per component;
  * pid is the lattice index + 1 (the reference's fine-grid index overflows
    int32 beyond nf=1290).
"""
import numpy as np

from ._lib import SimHeader
from .checkpoint import BoxState, xdtype
from .params import NCELL, PI_F


# ----------------------------------------------------------------- cosmology
def eh_nowiggle_transfer(k_h, omega_m, omega_b, h, tcmb=2.7255):
    """Eisenstein & Hu 1998 (ApJ 496, 605) zero-baryon-oscillation fit,
    eqs. (26)-(31); k in h/Mpc."""
    k = np.asarray(k_h, dtype=np.float64) * h  # 1/Mpc
    theta = tcmb / 2.7
    om_h2, ob_h2 = omega_m * h * h, omega_b * h * h
    fb = omega_b / omega_m
    s = 44.5 * np.log(9.83 / om_h2) / np.sqrt(1.0 + 10.0 * ob_h2 ** 0.75)               # (26)
    alpha = 1.0 - 0.328 * np.log(431.0 * om_h2) * fb + 0.38 * np.log(22.3 * om_h2) * fb ** 2  # (31)
    gamma_eff = omega_m * h * (alpha + (1.0 - alpha) / (1.0 + (0.43 * k * s) ** 4))       # (30)
    q = np.asarray(k_h, dtype=np.float64) * theta ** 2 / gamma_eff                      # (28)
    l0 = np.log(2.0 * np.e + 1.8 * q)                                                    # (29)
    c0 = 14.2 + 731.0 / (1.0 + 62.5 * q)
    return l0 / (l0 + c0 * q * q)


def tophat_w(x):
    x = np.asarray(x, dtype=np.float64)
    out = np.ones_like(x)
    m = x > 1e-6
    xm = x[m]
    out[m] = 3.0 * (np.sin(xm) - xm * np.cos(xm)) / xm ** 3
    return out


def linear_power(cfg):
    """P(k) [ (Mpc/h)^3 ] at z=0 as a callable of k [h/Mpc]."""
    lk = np.linspace(np.log(1e-5), np.log(1e3), 20001)
    k = np.exp(lk)

    def shape(kk):
        return kk ** cfg.n_s * eh_nowiggle_transfer(kk, cfg.omega_m, cfg.omega_b, cfg.h0) ** 2

    integrand = k ** 3 * shape(k) / (2 * np.pi ** 2) * tophat_w(k * 8.0) ** 2
    sig2 = np.trapezoid(integrand, lk)
    amp = cfg.sigma8 ** 2 / sig2
    return lambda kk: amp * shape(np.maximum(kk, 1e-12))


def dgrow(a, om, ol):
    """main.f90:104-117 (Carroll, Press & Turner fit), normalised to 1 at a=1."""
    def g_of(a):
        hsq = om / a ** 3 + (1 - om - ol) / a ** 2 + ol
        oma = om / (a ** 3 * hsq)
        ola = ol / hsq
        return 2.5 * oma / (oma ** (4.0 / 7) - ola + (1 + oma / 2) * (1 + ola / 70))
    return a * g_of(a) / g_of(1.0)


def hubble_sim(a, om, ol):
    """H(a) in simulation time units (vfactor of initial_conditions.f90:1068-1079)."""
    lm, km = ol / om, (1 - om - ol) / om
    return 2.0 / (3.0 * np.sqrt(a ** 3)) * np.sqrt(1 + a * km + a ** 3 * lm)


# ------------------------------------------------------------------- codecs
def vcoef(sigma):
    return float(np.sqrt(PI_F / np.float32(2))) / (float(np.float32(sigma)) * 2.5)


def vencode(v, sigma, izipv):
    nv1 = float(np.float32(2 ** (8 * izipv) - 1))
    y = nv1 * np.arctan(vcoef(sigma) * np.asarray(v, dtype=np.float64)) / float(PI_F)
    return (np.sign(y) * np.floor(np.abs(y) + 0.5)).astype(np.int64)


def vdecode(vp, sigma, izipv):
    nv1 = np.float32(2 ** (8 * izipv) - 1)
    t = np.tan((PI_F * np.asarray(vp).astype(np.float32)) / nv1)
    return t.astype(np.float64) / vcoef(sigma)


# ---------------------------------------------------------------- geometry
def cell_rank(cx, cy, cz, nt, nnt):
    """Checkpoint order of 0-based global coarse cells: tile (x fastest),
    then cell inside the tile (x fastest)."""
    t = ((cz // nt) * nnt + cy // nt) * nnt + cx // nt
    return t * nt ** 3 + ((cz % nt) * nt + cy % nt) * nt + cx % nt


def rank_to_cell(r, nt, nnt):
    r = np.asarray(r, dtype=np.int64)
    t, l = r // nt ** 3, r % nt ** 3
    return (t % nnt * nt + l % nt, (t // nnt) % nnt * nt + (l // nt) % nt, t // (nnt * nnt) * nt + l // (nt * nt))


def ic_vsim2phys(cfg, a):
    """sim%vsim2phys of the initial conditions (initial_conditions.f90:198)
    in real(4), left to right."""
    f32 = np.float32
    v = f32(f32(f32(150.0) / f32(a)) * f32(cfg.box))
    return f32(f32(v * np.sqrt(f32(cfg.omega_m))) / f32(cfg.nf))


def encode_particles(cfg, pos, vel, pid, a, sigma_vi=None, z_i=None):
    """Pack continuous positions (coarse-cell units, any real) and velocities
    (simulation units) into a BoxState in checkpoint order."""
    nc, nt, nnt = cfg.nc, cfg.nt, cfg.nnt
    pos = np.mod(np.asarray(pos, dtype=np.float64), nc)
    # cell = floor(x) and u = floor(frac(x)/x_resolution): the decoded position
    # (cell + (u+0.5)*x_resolution) always lies in the cell.  The reference's
    # g=ceiling(x) with xp=floor(x/x_resolution) moves a particle sitting
    # exactly on a cell face into the cell below it.
    c0 = np.floor(pos).astype(np.int64) % nc
    nbin = 2 ** (8 * cfg.izipx)
    u = np.clip(np.floor((pos - np.floor(pos)) * nbin).astype(np.int64), 0, nbin - 1)
    xp = u.astype(np.uint16 if cfg.izipx == 2 else np.uint8).view(xdtype(cfg.izipx))
    r = cell_rank(c0[:, 0], c0[:, 1], c0[:, 2], nt, nnt)
    order = np.lexsort((pid, r))
    r, xp, vel, pid = r[order], xp[order], vel[order], pid[order]
    n3 = nc ** 3
    rhoc = np.bincount(r, minlength=n3).astype(np.int32)
    vsum = np.stack([np.bincount(r, weights=vel[:, e], minlength=n3) for e in range(3)], axis=1)
    vfield = (vsum / np.maximum(rhoc, 1)[:, None]).astype(np.float32)
    vres = vel - vfield[r].astype(np.float64)
    if sigma_vi is None:
        sigma_vi = np.float32(np.sqrt(np.mean(np.sum(vres ** 2, axis=1))) / np.sqrt(3.0))
    vp = vencode(vres, sigma_vi, cfg.izipv).astype(xdtype(cfg.izipv))
    n = pos.shape[0]
    sim = SimHeader()
    sim.nplocal = n
    sim.npglobal = n
    sim.izipx, sim.izipv = cfg.izipx, cfg.izipv
    sim.izipx_nu, sim.izipv_nu = 2, 2
    sim.image = 1
    sim.nn, sim.nnt, sim.nt, sim.ncell, sim.ncb = 1, nnt, nt, NCELL, cfg.ncb
    sim.a = a
    for f in ("dt_pp", "dt_fine", "dt_coarse", "dt_vmax", "dt_vmax_nu"):
        setattr(sim, f, 1000.0)                                 # initial_conditions.f90:175-179
    sim.mass_p_cdm = cfg.mass_p(n)
    sim.box = cfg.box
    sim.h0 = cfg.h0
    sim.omega_m, sim.omega_l, sim.s8 = cfg.omega_m, cfg.omega_l, cfg.sigma8
    sim.vsim2phys = ic_vsim2phys(cfg, a)
    sim.sigma_vi = sigma_vi
    sim.z_i = (1.0 / a - 1.0) if z_i is None else z_i
    sim.z_i_nu = 10.0
    sim.vz_max = float(np.max(np.abs(vel[:, 2]))) if n else 0.0
    return BoxState(sim, np.ascontiguousarray(xp), np.ascontiguousarray(vp), np.ascontiguousarray(pid, dtype=np.int32),
                    rhoc, np.ascontiguousarray(vfield))


def splitmix64(x):
    """SplitMix64 finaliser on uint64 arrays (wrapping arithmetic)."""
    x = np.asarray(x, dtype=np.uint64) + np.uint64(0x9E3779B97F4A7C15)
    x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return x ^ (x >> np.uint64(31))


def lattice_noise(ng, seed):
    """Unit Gaussian per lattice site (x fastest): Box-Muller of
    SplitMix64(SplitMix64(seed) + site), the same draw cube_ic_zeldovich makes."""
    seed0 = splitmix64(np.array([seed], dtype=np.uint64))[0]
    out = np.empty(ng ** 3, dtype=np.float32)
    chunk = 1 << 22
    for s0 in range(0, ng ** 3, chunk):
        idx = np.arange(s0, min(ng ** 3, s0 + chunk), dtype=np.uint64)
        h = splitmix64(idx + seed0)
        u1 = ((h >> np.uint64(32)).astype(np.float64) + 0.5) * (1.0 / 4294967296.0)
        u2 = ((h & np.uint64(0xFFFFFFFF)).astype(np.float64) + 0.5) * (1.0 / 4294967296.0)
        out[s0:s0 + len(idx)] = (np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)).astype(np.float32)
    return out.reshape(ng, ng, ng)


def power_table(cfg, a, nk=4096):
    """ln k [h/Mpc] and ln P(k) [(Mpc/h)^3] at scale factor a on a log grid."""
    lnk = np.linspace(np.log(1e-5), np.log(1e3), nk)
    p = linear_power(cfg)(np.exp(lnk)) * dgrow(a, cfg.omega_m, cfg.omega_l) ** 2
    return lnk, np.log(p)


def velocity_factor(cfg, a):
    om_a = cfg.omega_m / (cfg.omega_m + cfg.omega_l * a ** 3)
    return a * a * hubble_sim(a, cfg.omega_m, cfg.omega_l) * om_a ** 0.55


def zeldovich_ic(cfg, seed, z=None):
    """Gaussian field with the EH no-wiggle spectrum, Zel'dovich displacement
    on the particle lattice, encoded as a BoxState (host version of
    cube_ic_zeldovich: same noise, spectrum table and Nyquist treatment)."""
    z = cfg.z_i if z is None else z
    a = 1.0 / (1.0 + z)
    ng = cfg.nc * cfg.np_nc
    wk = np.fft.rfftn(lattice_noise(ng, seed).astype(np.float64))
    kf = 2 * np.pi / cfg.box
    k1 = np.fft.fftfreq(ng, d=1.0 / ng)
    kr = np.fft.rfftfreq(ng, d=1.0 / ng)
    kz, ky, kx = np.meshgrid(k1, k1, kr, indexing="ij")
    k2 = (kx ** 2 + ky ** 2 + kz ** 2) * kf * kf
    k2[0, 0, 0] = 1.0
    lnk, lnp = power_table(cfg, a)
    pk = np.exp(np.interp(0.5 * np.log(k2), lnk, lnp))
    delta_k = wk * np.sqrt(ng ** 3 * pk / cfg.box ** 3)
    delta_k[0, 0, 0] = 0
    del wk, pk
    psi = np.empty((ng, ng, ng, 3), dtype=np.float64)
    for e, ke in enumerate((kx, ky, kz)):
        fac = 1j * (ke * kf) / k2
        fac[np.abs(ke) == ng // 2] = 0  # odd derivative on its own Nyquist plane
        psi[..., e] = np.fft.irfftn(fac * delta_k, s=(ng, ng, ng), axes=(0, 1, 2))   # Mpc/h
    del delta_k, kx, ky, kz, k2
    psi *= cfg.nf / cfg.box                                                    # fine cells
    # lattice point i (0-based) at q = i/np_nc + 0.5/ncell coarse cells; array axes are (z,y,x)
    iz, iy, ix = np.meshgrid(np.arange(ng), np.arange(ng), np.arange(ng), indexing="ij")
    q = np.stack([ix, iy, iz], axis=-1).reshape(-1, 3) / cfg.np_nc + 0.5 / NCELL
    psi = psi.reshape(-1, 3)
    pid = (iz.reshape(-1).astype(np.int64) * ng + iy.reshape(-1)) * ng + ix.reshape(-1) + 1
    pos = q + psi / NCELL
    vel = psi * velocity_factor(cfg, a)
    return encode_particles(cfg, pos, vel, pid.astype(np.int32), np.float32(a), z_i=np.float32(z))
