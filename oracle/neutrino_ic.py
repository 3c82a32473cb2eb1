"""CPU restatement of the neutrino fluid initial conditions -- test
infrastructure (the checker), used only by the tests.

Follows utilities/neutrino_ic.f90:40-181 on the CDM linear field of the
seeded Zel'dovich IC: per fluid n (hydrodf.f90:17-31 Gauss-Laguerre
velocity v_n = GL_v(n)/(5966.34 m) * 2.998e5 km/s), the free-streaming scale
kfs = sqrt(1.5 omega_m/(1+z_i_nu))*100/v_n (neutrino_ic.f90:12,79-82) and
l = k/kfs with k = |k_index| 2 pi/box in real(4) (:88-96); the density
contrast is filtered by 1/(1+l+l^2) (:97) and the velocity by
(1+1.5l+2l^2)/(1+l+l^2)^2 (:159); both are block-averaged onto the fluid
mesh (hg_set_fld, hydrogr.f90:288-308) as rho = <1+delta> and <v>, then
m = rho <v> and E = rho cs2/(g-1) + m^2/(2 rho) per component (:110-113,
170-177).  cs2 = max(1e-6, (v_n hg_dv)^2), hg_dv = 2 nf/(300 sqrt(omega_m)
box) (hydrogr.f90:13,65-69).

The CDM linear fields are those of the seeded IC on the particle lattice
(noise from cube_nufluid_amd.ic.lattice_noise, spectrum table from
ic.power_table, delta_k = noise_k sqrt(ng^3 P/V), v_k = i k/k^2 delta_k
nf/box vfac with each component zeroed on its own Nyquist plane).
"""
import numpy as np

F32 = np.float32
GL_V = (0.41577456, 2.29428036, 6.28994508)
PI_F = F32(3.14159274101257324)


def fluid_setup(cfg, m_ev):
    """Cells per dimension, cell size and cs2 of the three fluids
    (hg_setup/dfld_setup)."""
    nf = cfg.nf
    dv = F32(F32(2 * nf) / (F32(3.0) * F32(100.0) * np.sqrt(F32(cfg.omega_m)) * F32(cfg.box)))
    beta = F32(F32(5966.34) * F32(m_ev))
    out = []
    for n in range(3):
        v = F32(F32(GL_V[n]) / beta * F32(2.998e5))
        out.append((nf >> (n + 1), F32(nf / (nf >> (n + 1))), max(F32(1e-6), F32(v * dv) * F32(v * dv))))
    return out


def tf_density(ll):
    ll = F32(ll)
    return F32(1.0) / (F32(1.0) + ll + ll * ll)


def tf_velocity(ll):
    ll = F32(ll)
    q = F32(1.0) + ll + ll * ll
    return (F32(1.0) + F32(1.5) * ll + F32(2.0) * ll * ll) / (q * q)


def fluid_ic(cfg, seed, z, m_ev, z_i_nu, gamma=5.0 / 3.0):
    """The three fluids as [5][nc][nc][nc] float32 arrays (rho, mx, my, mz, E)."""
    from cube_nufluid_amd.ic import lattice_noise, power_table, velocity_factor
    a = 1.0 / (1.0 + z)
    ng = cfg.nc * cfg.np_nc
    wk = np.fft.rfftn(lattice_noise(ng, seed).astype(np.float64))
    kf = 2 * np.pi / cfg.box
    k1 = np.fft.fftfreq(ng, d=1.0 / ng)
    kr1 = np.fft.rfftfreq(ng, d=1.0 / ng)
    kz, ky, kx = np.meshgrid(k1, k1, kr1, indexing="ij")
    k2 = (kx ** 2 + ky ** 2 + kz ** 2) * kf * kf
    k2[0, 0, 0] = 1.0
    lnk, lnp = power_table(cfg, a)
    delta_k = wk * np.sqrt(ng ** 3 * np.exp(np.interp(0.5 * np.log(k2), lnk, lnp)) / cfg.box ** 3)
    delta_k[0, 0, 0] = 0
    kr = np.sqrt(F32(kx) ** 2 + F32(ky) ** 2 + F32(kz) ** 2)
    kp = kr * F32(2.0) * PI_F / F32(cfg.box)
    kfs1 = np.sqrt(F32(1.5) * F32(cfg.omega_m) / (F32(1.0) + F32(z_i_nu))) * F32(100.0)
    beta = F32(5966.34)
    vfac = velocity_factor(cfg, a) * cfg.nf / cfg.box
    fluids = []
    for n, (nc, _, cs2) in enumerate(fluid_setup(cfg, m_ev)):
        vneu = F32(F32(GL_V[n]) / beta * F32(2.998e5) / F32(m_ev))
        ll = kp / F32(kfs1 / vneu)
        b = ng // nc

        def block(field):
            f = field.astype(F32).reshape(nc, b, nc, b, nc, b)
            return (f.sum(axis=(1, 3, 5), dtype=np.float64) / b ** 3).astype(F32)
        d = np.fft.irfftn(delta_k * tf_density(ll).astype(np.float64), s=(ng,) * 3, axes=(0, 1, 2))
        rho = block(F32(1.0) + d.astype(F32))
        fl = np.empty((5, nc, nc, nc), dtype=F32)
        fl[0] = rho
        e = rho * F32(cs2) / F32(gamma - 1.0)
        for c, kc in enumerate((kx, ky, kz)):
            fac = 1j * (kc * kf) / k2 * tf_velocity(ll).astype(np.float64)
            fac[np.abs(kc) == ng // 2] = 0
            v = block(np.fft.irfftn(fac * delta_k, s=(ng,) * 3, axes=(0, 1, 2)) * vfac)
            m = rho * v
            fl[1 + c] = m
            e = e + F32(0.5) * (m * m) / rho
        fl[4] = e
        fluids.append(fl)
    return fluids
