"""Host driver mirroring CUBE's main loop (main/main.f90:47-98) and
timestep (main/timestep.f90:8-95) over the C ABI of libcubehip.so.

All time-stepping scalars are real(4) in the reference; they are kept as
numpy float32 here so the dt/a sequence is computed with the same rounding.
"""
import ctypes
import os
import time

import numpy as np

from . import _lib
from .checkpoint import BoxState, output_name, write_checkpoint
from .params import NCELL, read_kernel_tables

f32 = np.float32
DT_MAX = f32(1.0)     # variables.f90:20
RA_MAX = f32(0.2)     # parameters.f90:110


class Cube:
    """One box resident on the GPU, driven like CUBE's main program."""

    def __init__(self, cfg, state, device=0, np_max=None, precision=4, z_checkpoints=None, z_halofind=None):
        self.lib = _lib.load()
        self.cfg = cfg
        p = _lib.CubeParams()
        p.nc, p.nnt, p.ncb = cfg.nc, cfg.nnt, cfg.ncb
        p.izipx, p.izipv = cfg.izipx, cfg.izipv
        p.np_max = int(np_max if np_max is not None else max(cfg.np_max, state.sim.nplocal))
        p.pp_on, p.pp_range, p.rsoft = int(cfg.pp_on), cfg.pp_range, cfg.rsoft
        p.omega_m, p.omega_l, p.wde = cfg.omega_m, cfg.omega_l, cfg.wde
        p.precision, p.device = precision, device
        self.params = p
        kf, kc = read_kernel_tables()
        h = ctypes.c_void_p()
        rc = self.lib.cube_init(ctypes.byref(p), _lib.ptr(kf), _lib.ptr(kc), ctypes.byref(h))
        self.h = h
        _lib.check(self.lib, h, rc)
        if state is None:
            raise ValueError("a state is required; use Cube.from_ic for a generated start")
        self.upload(state)
        self._init_time(state.sim.a, z_checkpoints, z_halofind)

    @classmethod
    def from_ic(cls, cfg, seed, z, device=0, z_checkpoints=None, np_max=None, z_halofind=None):
        """Generate seeded Zel'dovich initial conditions directly in HBM
        (cube_ic_zeldovich) and return the resident box."""
        from .ic import ic_vsim2phys, power_table, velocity_factor
        self = cls.__new__(cls)
        self.lib = _lib.load()
        self.cfg = cfg
        p = _lib.CubeParams()
        p.nc, p.nnt, p.ncb = cfg.nc, cfg.nnt, cfg.ncb
        p.izipx, p.izipv = cfg.izipx, cfg.izipv
        p.np_max = int(np_max if np_max is not None else cfg.np_max)
        p.pp_on, p.pp_range, p.rsoft = int(cfg.pp_on), cfg.pp_range, cfg.rsoft
        p.omega_m, p.omega_l, p.wde = cfg.omega_m, cfg.omega_l, cfg.wde
        p.precision, p.device = 4, device
        self.params = p
        kf, kc = read_kernel_tables()
        h = ctypes.c_void_p()
        rc = self.lib.cube_init(ctypes.byref(p), _lib.ptr(kf), _lib.ptr(kc), ctypes.byref(h))
        self.h = h
        _lib.check(self.lib, h, rc)
        a = 1.0 / (1.0 + z)
        lnk, lnp = power_table(cfg, a)
        ip = _lib.ICParams()
        ip.seed, ip.np_nc, ip.nk, ip.a, ip.box = seed, cfg.np_nc, len(lnk), a, cfg.box
        ip.vfac = velocity_factor(cfg, a)
        ip.lnk, ip.lnp = _lib.ptr(lnk), _lib.ptr(lnp)
        ip.f_cdm = cfg.f_cdm
        sim = _lib.SimHeader()
        _lib.check(self.lib, h, self.lib.cube_ic_zeldovich(h, ctypes.byref(ip), ctypes.byref(sim)))
        sim.h0, sim.omega_m, sim.omega_l, sim.s8 = cfg.h0, cfg.omega_m, cfg.omega_l, cfg.sigma8
        sim.vsim2phys = ic_vsim2phys(cfg, sim.a)
        sim.z_i_nu = cfg.z_i_nu
        self.sim = sim
        self._init_time(np.float32(a), z_checkpoints or [z, 0.0], z_halofind)
        if cfg.m_nu_ev > 0:
            # the neutrino IC tool's fluids from the same linear field
            # (neutrino_ic.f90), read at start as main.f90:25-32 does
            self.nu_init(m_ev=cfg.m_nu_ev, f_neu=cfg.f_neu, z_i_nu=cfg.z_i_nu)
            _lib.check(self.lib, h, self.lib.cube_nu_ic(h, ctypes.byref(ip)))
            self.nu_seed_dt()
        return self

    @classmethod
    def from_checkpoint(cls, cfg, opath, z_checkpoints, device=0, np_max=None, z_halofind=None, exact=False):
        """Resume from the checkpoint at z_checkpoints[0] (initialize.f90:
        40-61 with the list edited to start there, particle_initialization.f90:
        34-82, and for a neutrino run the fluids from the z-prefixed neu1
        file of that checkpoint, read as main.f90:25-32 reads its neu file).

        As in the reference, a = 1/(1+z) is recomputed from the list, which
        differs in its last bits from the a the run reached when it wrote
        the checkpoint.  exact=True takes a, t and tau from the header
        instead, so that a resumed CDM run repeats the uninterrupted one bit
        for bit.  The fluids' dt_neu restarts at their CFL step (the
        reference restarts it at dt_max).  z_checkpoints is a list or the
        path of a z_checkpoint.txt."""
        from .checkpoint import read_checkpoint, read_z_list
        if isinstance(z_checkpoints, (str, os.PathLike)):
            z_checkpoints = read_z_list(z_checkpoints)
        z0 = z_checkpoints[0]
        st = read_checkpoint(opath, z0, cfg.nc, izipx=cfg.izipx, izipv=cfg.izipv)
        self = cls(cfg, st, device=device, np_max=np_max, z_checkpoints=z_checkpoints, z_halofind=z_halofind)
        if exact:
            self.a, self.t, self.tau = f32(st.sim.a), f32(st.sim.t), f32(st.sim.tau)
            self.a_mid = self.a
        if cfg.m_nu_ev > 0:
            fn = output_name(opath, z0, "neu1")
            if not os.path.exists(fn):
                raise FileNotFoundError("neutrino fluid checkpoint %s missing" % fn)
            self.nu_restart(fn)
            self.nu_seed_dt()
        return self

    def _init_time(self, a0, z_checkpoints, z_halofind=None):
        # initialize.f90:40-61 and particle_initialization.f90:37-39; a path
        # names a z_checkpoint.txt
        if isinstance(z_checkpoints, (str, os.PathLike)):
            from .checkpoint import read_z_list
            z_checkpoints = read_z_list(z_checkpoints)
        self.z_checkpoints = list(z_checkpoints) if z_checkpoints else [float(1.0 / a0 - 1.0), 0.0]
        self.cur_checkpoint = 1
        self.a = f32(1.0 / (1.0 + f32(self.z_checkpoints[0])))
        self.t = f32(0)
        self.tau = f32(-3 / np.sqrt(1.0 / (1 + self.z_checkpoints[0])))
        self.dt = f32(0)
        self.dt_old = f32(0)
        self.da = f32(0)
        self.a_mid = self.a
        self.dt_neu = DT_MAX
        self.checkpoint_step = False
        self.final_step = False
        self.cur_checkpoint += 1  # main.f90:42
        # HALOFIND (initialize.f90:62-71, main.f90:43): the list as read from
        # z_halofind.txt (or that file's path); the main loop starts at entry
        # cur_checkpoint+1, and entries past the list read as 0 (the unset
        # array)
        if isinstance(z_halofind, (str, os.PathLike)):
            from .checkpoint import read_z_list
            z_halofind = read_z_list(z_halofind)
        self.z_halofind = list(z_halofind) if z_halofind else None
        self.cur_halofind = self.cur_checkpoint + 1
        self.halofind_step = False
        self.halo_params = None
        self.halo_catalogs = []
        self.timings = {}

    # ------------------------------------------------------------- state
    def upload(self, state):
        self.sim = _lib.SimHeader.from_bytes(state.sim.to_bytes())
        _lib.check(self.lib, self.h, self.lib.cube_upload_state(
            self.h, ctypes.byref(self.sim), _lib.ptr(np.ascontiguousarray(state.xp)),
            _lib.ptr(np.ascontiguousarray(state.vp)), _lib.ptr(np.ascontiguousarray(state.pid)),
            _lib.ptr(np.ascontiguousarray(state.rhoc)), _lib.ptr(np.ascontiguousarray(state.vfield))))

    def download(self):
        s = _lib.SimHeader()
        _lib.check(self.lib, self.h, self.lib.cube_download_state(self.h, ctypes.byref(s), None, None, None, None, None))
        n = s.nplocal
        xdt = np.int8 if self.cfg.izipx == 1 else np.int16
        vdt = np.int8 if self.cfg.izipv == 1 else np.int16
        xp = np.empty((n, 3), dtype=xdt)
        vp = np.empty((n, 3), dtype=vdt)
        pid = np.empty(n, dtype=np.int32)
        rhoc = np.empty(self.cfg.nc ** 3, dtype=np.int32)
        vf = np.empty((self.cfg.nc ** 3, 3), dtype=np.float32)
        _lib.check(self.lib, self.h, self.lib.cube_download_state(
            self.h, ctypes.byref(s), _lib.ptr(xp), _lib.ptr(vp), _lib.ptr(pid), _lib.ptr(rhoc), _lib.ptr(vf)))
        for f in ("a", "t", "tau"):
            setattr(s, f, float(getattr(self, f)))
        s.timestep = self.sim.timestep  # counted by this host's timestep (timestep.f90:10)
        return BoxState(s, xp, vp, pid, rhoc, vf)

    def close(self):
        if self.h:
            self.lib.cube_finalize(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ----------------------------------------------------- neutrino fluid
    def nu_init(self, m_ev=0.1, f_neu=None, z_i_nu=10.0, gamma=5.0 / 3.0, neutrino_ic=True):
        """dfld_setup for one neutrino mass; f_neu defaults to
        omega_nu/omega_m with omega_nu = 3 m/93.14/h^2 (parameters.f90:83-92)."""
        c = self.cfg
        if f_neu is None:
            f_neu = 3 * m_ev / 93.14 / c.h0 ** 2 / c.omega_m
        p = _lib.NuParams(m_ev, f_neu, z_i_nu, gamma, c.box, c.h0, c.omega_m, int(neutrino_ic))
        _lib.check(self.lib, self.h, self.lib.cube_nu_init(self.h, ctypes.byref(p)))
        self.nu_params = p
        self.nu_on = True

    def nu_info(self, n):
        nc, dx, cs2, w = ctypes.c_int32(), ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
        _lib.check(self.lib, self.h, self.lib.cube_nu_info(self.h, n, ctypes.byref(nc), ctypes.byref(dx), ctypes.byref(cs2),
                                                           ctypes.byref(w)))
        return dict(nc=nc.value, dx=dx.value, cs2=cs2.value, weight=w.value)

    def nu_get(self, n):
        nc = self.nu_info(n)["nc"]
        f = np.empty((5, nc, nc, nc), dtype=np.float32)
        _lib.check(self.lib, self.h, self.lib.cube_nu_get_fluid(self.h, n, _lib.ptr(f)))
        return f

    def nu_set(self, n, f):
        f = np.ascontiguousarray(f, dtype=np.float32)
        _lib.check(self.lib, self.h, self.lib.cube_nu_set_fluid(self.h, n, _lib.ptr(f)))

    def nu_checkpoint(self, fn):
        """dfld_checkpoint (hydrodf.f90:128-146) of the resident fluids."""
        from .checkpoint import write_neu_checkpoint
        os.makedirs(os.path.dirname(fn) or ".", exist_ok=True)
        info = [self.nu_info(n) for n in range(3)]
        beta = f32(f32(5966.34) * f32(self.nu_params.m_ev))     # d%beta, written as beta/dfld_beta_eV
        write_neu_checkpoint(fn, f32(beta / f32(5966.34)), self.nu_params.gamma, True, [i["weight"] for i in info],
                             [self.nu_get(n) for n in range(3)], [i["dx"] for i in info], ncb=self.cfg.ncb)

    def nu_restart(self, fn, f_neu=None, z_i_nu=None):
        """dfld_restart (hydrodf.f90:158-185): mass, gamma and fluids from a
        neu file; the weights are recomputed by the setup, as there."""
        from .checkpoint import read_neu_checkpoint
        with open(fn, "rb") as f:
            m_ev, gamma = np.fromfile(f, dtype=np.float32, count=2)
        self.nu_init(m_ev=float(m_ev), f_neu=f_neu if f_neu is not None else self.cfg.f_neu,
                     z_i_nu=z_i_nu if z_i_nu is not None else self.cfg.z_i_nu, gamma=float(gamma))
        info = [self.nu_info(n) for n in range(3)]
        _, _, _, _, fluids = read_neu_checkpoint(fn, [i["nc"] for i in info], [i["dx"] for i in info], ncb=self.cfg.ncb)
        for n in range(3):
            self.nu_set(n, fluids[n])

    def nu_seed_dt(self):
        """dt_neu of the fluids as loaded.  The reference starts dt_neu at
        dt_max = 1 (initialize.f90:40); with the fluids live from the first
        step (NEUTRINO_IC, fluid and CDM both starting at z_i_nu) that first
        step can exceed the fast fluid's CFL limit many times over, so the
        CFL step of the loaded fluids is taken instead (a zero-length
        evolve leaves the fields unchanged and returns it)."""
        self.dt_neu = DT_MAX
        self.fluid_evolve(f32(0.0), 1, self.a)

    def fluid_evolve(self, dt, direction, a):
        dtn = ctypes.c_float(float(self.dt_neu))
        _lib.check(self.lib, self.h, self.lib.cube_fluid_evolve(self.h, float(dt), int(direction), float(a), ctypes.byref(dtn)))
        self.dt_neu = f32(dtn.value)

    # ------------------------------------------------------- subroutines
    def expansion(self, a, dt):
        return _lib.expansion(a, dt, self.cfg.omega_m, self.cfg.omega_l, self.cfg.wde)

    def timestep(self):
        """timestep.f90:8-95 (head-image branch; one image)."""
        self.dt_old = self.dt
        self.sim.timestep += 1
        dt_e = DT_MAX
        ntry = 0
        while True:
            ntry += 1
            d1, d2 = self.expansion(self.a, dt_e)
            da = f32(d1 + d2)
            ra = f32(da / f32(self.a + da))
            if ra > RA_MAX:
                dt_e = f32(dt_e * f32(RA_MAX / ra))
            else:
                break
            if ntry > 10:
                break
        s = self.sim
        dt = min(dt_e, f32(s.dt_fine), f32(s.dt_coarse), f32(s.dt_pp), f32(s.dt_vmax), f32(1000.0), self.dt_neu)
        d1, d2 = self.expansion(self.a, dt)
        da = f32(d1 + d2)
        self.checkpoint_step = False
        self.halofind_step = False
        # initialize.f90:49 fills z_checkpoint with -0.5 beyond the list
        zc = self.z_checkpoints
        z_ck = f32(zc[self.cur_checkpoint - 1] if self.cur_checkpoint <= len(zc) else -0.5)
        z_next = z_ck
        if self.z_halofind is not None:  # timestep.f90:41-42
            zh = self.z_halofind
            z_h = f32(zh[self.cur_halofind - 1] if self.cur_halofind <= len(zh) else 0.0)
            z_next = max(z_ck, z_h)
        a_next = f32(f32(1.0) / f32(1 + z_next))
        if da >= f32(a_next - self.a):
            if z_next == z_ck:
                self.checkpoint_step = True
                if self.cur_checkpoint == len(self.z_checkpoints):
                    self.final_step = True
            if self.z_halofind is not None and z_next == z_h:
                self.halofind_step = True
            it = 0
            while abs(f32(f32(self.a + da) / a_next) - f32(1)) >= f32(1e-6) and it < 100:  # real(4) 1e-6
                dt = f32(f32(dt * f32(a_next - self.a)) / da)  # dt*(a_next-a)/da, left to right
                d1, d2 = self.expansion(self.a, dt)
                da = f32(d1 + d2)
                it += 1
        self.a_mid = f32(self.a + f32(da / 2))
        self.tau = f32(self.tau + dt)
        self.t = f32(self.t + dt)
        self.a = f32(self.a + da)
        self.da = da
        self.dt = dt
        self.dt_neu = f32(1000.0)

    def vsim2phys(self, with_h0):
        """sim%vsim2phys at the current a, in real(4) left to right: the
        PM step's (pm.f90:427, with h0) or the checkpoint's
        (checkpoint.f90:30, without)."""
        c = self.cfg
        v = f32(f32(f32(1.5) / f32(self.a)) * f32(c.box))
        if with_h0:
            v = f32(v * f32(c.h0))
        v = f32(f32(v * f32(100.0)) * np.sqrt(f32(c.omega_m)))
        return f32(v / f32(c.nf))

    def update_x(self):
        out = _lib.DriftOut()
        _lib.check(self.lib, self.h, self.lib.cube_drift(self.h, float(self.dt_old), float(self.dt), ctypes.byref(out)))
        self.last_drift = out
        return out

    def ext_pp_force(self):
        out = _lib.PPOut()
        _lib.check(self.lib, self.h, self.lib.cube_pp(self.h, float(self.a_mid), float(self.dt), ctypes.byref(out)))
        self.sim.dt_pp = out.dt_pp
        self.last_pp = out
        return out

    def particle_mesh(self):
        out = _lib.PMOut()
        _lib.check(self.lib, self.h, self.lib.cube_pm(self.h, float(self.a_mid), float(self.dt), ctypes.byref(out)))
        s = self.sim
        s.dt_fine, s.dt_coarse, s.dt_vmax, s.vz_max = out.dt_fine, out.dt_coarse, out.dt_vmax, out.vz_max
        s.vsim2phys = self.vsim2phys(with_h0=True)
        self.last_pm = out
        return out

    # --------------------------------------------------------- main loop
    def step(self, opath=None):
        """One pass of main.f90:48-97; returns True after the final checkpoint."""
        t0 = time.perf_counter()
        self.timestep()
        nu = getattr(self, "nu_on", False)
        if nu:  # main.f90:51-55
            self.fluid_evolve(f32(self.dt / 2), 1, self.a_mid)
        self.update_x()
        if self.cfg.pp_on:
            self.ext_pp_force()
        self.particle_mesh()
        if nu:  # main.f90:66-70
            self.fluid_evolve(f32(self.dt / 2), -1, self.a_mid)
        if self.checkpoint_step or self.halofind_step:  # main.f90:72-95
            self.dt_old = f32(0)
            self.update_x()
            if self.checkpoint_step:
                if opath is not None:
                    st = self.download()
                    st.sim.cur_checkpoint = self.cur_checkpoint
                    st.sim.vsim2phys = self.vsim2phys(with_h0=False)
                    zc = self.z_checkpoints[self.cur_checkpoint - 1]
                    if nu:  # main.f90:75-81, before the particle files
                        self.nu_checkpoint(output_name(opath, zc, "neu1"))
                    write_checkpoint(st, opath, zc)
                self.cur_checkpoint += 1
            if self.halofind_step:
                self._halofind_step(opath)
                self.cur_halofind += 1
            if self.final_step:
                return True
            self.dt = f32(0)
        self.timings.setdefault("step", []).append(time.perf_counter() - t0)
        return False

    def _halofind_step(self, opath):
        """halofind.f90 at z_halofind(cur_halofind): the catalogue is kept in
        halo_catalogs and, with an output path, written as the reference's
        halo, halo_pid and hid files."""
        from .halo import halo_prefix, write_halo_files
        zh = self.z_halofind
        z = float(zh[self.cur_halofind - 1]) if self.cur_halofind <= len(zh) else 0.0
        self.sim.a = self.a
        self.sim.vsim2phys = self.vsim2phys(with_h0=False)
        t0 = time.perf_counter()
        cat = self.halofind(self.halo_params, nhid=self.cfg.nf ** 3 if opath is not None else 0)
        self.timings.setdefault("halofind", []).append(time.perf_counter() - t0)
        self.halo_catalogs.append((z, cat))
        if opath is not None:
            write_halo_files(halo_prefix(opath, z), cat)

    def run(self, max_steps=2500, opath=None):
        n = 0
        while n < max_steps:
            n += 1
            if self.step(opath):
                break
        return n

    # ------------------------------------------------------ diagnostics
    def probe_pm_forces(self):
        n = self.sim.nplocal
        ff = np.empty((n, 3), dtype=np.float64)
        fc = np.empty((n, 3), dtype=np.float64)
        _lib.check(self.lib, self.h, self.lib.cube_probe_pm_forces(self.h, _lib.ptr(ff), _lib.ptr(fc)))
        return ff, fc

    def power_spectrum(self, noalloc=False):
        """(k in h/Mpc, Delta^2, modes) per linear bin 1..nbin of the fine
        mesh (cicpower.f90 / powerspectrum.f90 estimator).  noalloc selects
        the entry point that allocates no device memory."""
        nf = self.cfg.nf
        nbin = int(np.rint(nf // 2 * np.sqrt(3.0)))
        modes = np.zeros(nbin)
        ksum = np.zeros(nbin)
        d2 = np.zeros(nbin)
        fn = self.lib.cube_power_spectrum_noalloc if noalloc else self.lib.cube_power_spectrum
        _lib.check(self.lib, self.h, fn(self.h, nbin, _lib.ptr(modes), _lib.ptr(ksum), _lib.ptr(d2)))
        m = np.maximum(modes, 1)
        return ksum / m * 2 * np.pi / self.cfg.box, d2 / m, modes

    def probe_pp_forces(self):
        n = self.sim.nplocal
        fp = np.empty((n, 3), dtype=np.float64)
        _lib.check(self.lib, self.h, self.lib.cube_probe_pp_forces(self.h, _lib.ptr(fp)))
        return fp

    def halofind(self, hp=None, nq=None, nhid=0, max_halos=1 << 20, max_pid_words=None):
        """Spherical-overdensity halo catalogue of the current state
        (halofind.f90:18-598) as a HaloCatalog.  nq: pid lattice size per
        dimension (the reference numbers the fine grid, nq = nf)."""
        from .halo import HaloParams, find_halos
        hp = (hp or HaloParams()).for_box(self.cfg, nq, self.sim)
        if max_pid_words is None:
            max_pid_words = 2 * max_halos + int(self.sim.nplocal) + 16
        return find_halos(self.lib, self.h, hp, max_halos=max_halos, max_pid_words=max_pid_words, nhid=nhid)

    def probe_density(self):
        nf = self.cfg.nf
        rho = np.empty((nf, nf, nf), dtype=np.float32)
        _lib.check(self.lib, self.h, self.lib.cube_probe_density(self.h, _lib.ptr(rho)))
        return rho

    def probe_vencode(self, x):
        """device velocity codes of the atan arguments x = c*v"""
        x = np.ascontiguousarray(x, dtype=np.float64)
        codes = np.empty(len(x), dtype=np.int32)
        _lib.check(self.lib, self.h, self.lib.cube_probe_vencode(self.h, _lib.ptr(x), len(x), _lib.ptr(codes)))
        return codes
