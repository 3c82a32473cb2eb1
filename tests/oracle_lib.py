"""ctypes wrapper of the CPU oracle (oracle/liboracle.so) -- test
infrastructure, used only as the checker."""
import ctypes
import os
import subprocess

import numpy as np

from cube_nufluid_amd._lib import SimHeader
from cube_nufluid_amd.checkpoint import BoxState, xdtype
from cube_nufluid_amd.params import read_kernel_tables

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")

_lib = None
_P = ctypes.c_void_p
_i64, _i32, _f32, _f64 = ctypes.c_int64, ctypes.c_int32, ctypes.c_float, ctypes.c_double


def build_oracle():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(ORACLE_SO):
        build_oracle()
    lib = ctypes.CDLL(ORACLE_SO)
    sig = {
        "orc_create": (_P, [_P, _P, _P, _P, ctypes.c_int]),
        "orc_destroy": (None, [_P]),
        "orc_error": (ctypes.c_char_p, [_P]),
        "orc_np": (_i64, [_P]),
        "orc_kernels": (None, [_P, _P, _P]),
        "orc_scalars": (None, [_P, _P]),
        "orc_set_scalars": (None, [_P, _P]),
        "orc_set_state": (ctypes.c_int, [_P, _i64, _P, _P, _P, _P, _P, _f32, _f32, _f32]),
        "orc_get_state": (ctypes.c_int, [_P, _P, _P, _P, _P, _P]),
        "orc_drift": (ctypes.c_int, [_P, _f32, _f32, _P, ctypes.c_int]),
        "orc_pp": (ctypes.c_int, [_P, _f32, _f32, _P, _P, ctypes.c_int]),
        "orc_pm": (ctypes.c_int, [_P, _f32, _f32, _P, ctypes.c_int]),
        "orc_forces": (ctypes.c_int, [_P, _P, _P, ctypes.c_int]),
        "orc_expansion": (None, [_f32, _f32, _f32, _f32, _f32, _P, _P]),
        "orc_set_perturb": (None, [ctypes.c_int]),
        "orc_vdecode": (_f64, [_i32, _f32, _i64]),
        "orc_vencode": (_i32, [_f64, _f32, _i64]),
        "orc_vencode_x": (None, [_P, _i64, _i64, _P]),
        "orc_fft_r2c": (None, [_P, _i64, _P, _P, ctypes.c_int]),
        "orc_density": (ctypes.c_int, [_P, _P, ctypes.c_int]),
        "orc_set_tile_limit": (None, [_P, _i64]),
        "orc_global_seconds": (_f64, [_P]),
        "orc_nu_setup": (None, [_P, _P, _P]),
        "orc_nu_set": (None, [_P, ctypes.c_int, _P]),
        "orc_nu_get": (None, [_P, ctypes.c_int, _P]),
        "orc_nu_evolve": (ctypes.c_int, [_P, _f32, ctypes.c_int, _f32, _P]),
        "ofl_evolve": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, _f32, _f32, _f32, ctypes.c_int, _P]),
        "ohf_halofind": (ctypes.c_int, [_P, _P, _i64, _P, _P, _P, _P, _P, _i64, _P, _i64, _P, _i64, _P, _P, ctypes.c_char_p]),
        "ohf_density": (ctypes.c_int, [_P, _P, _P, _P, _P]),
        "ohf_offset_table": (_i64, [_i64, _P, _i64]),
    }
    for n, (r, a) in sig.items():
        f = getattr(lib, n)
        f.restype = r
        f.argtypes = a
    _lib = lib
    return lib


def p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class Oracle:
    """CPU restatement of one CUBE box (same layout as the GPU handle)."""

    def __init__(self, cfg, threads=None):
        self.lib = load()
        self.cfg = cfg
        self.threads = threads or min(8, os.cpu_count() or 1)
        kf, kc = read_kernel_tables()
        ip = np.array([cfg.nc, cfg.nnt, cfg.ncb, cfg.izipx, cfg.izipv, int(cfg.pp_on), cfg.pp_range], dtype=np.int64)
        fp = np.array([cfg.rsoft], dtype=np.float32)
        self.o = self.lib.orc_create(p(ip), p(fp), p(kf), p(kc), self.threads)

    def close(self):
        if self.o:
            self.lib.orc_destroy(self.o)
            self.o = None

    def __del__(self):
        self.close()

    def _chk(self, rc):
        if rc != 0:
            raise RuntimeError("oracle error %d: %s" % (rc, self.lib.orc_error(self.o).decode()))

    def set_state(self, st):
        s = st.sim
        self._chk(self.lib.orc_set_state(self.o, s.nplocal, p(np.ascontiguousarray(st.xp)), p(np.ascontiguousarray(st.vp)),
                                         p(np.ascontiguousarray(st.pid)), p(np.ascontiguousarray(st.rhoc)),
                                         p(np.ascontiguousarray(st.vfield)), s.sigma_vi, s.vz_max, s.mass_p_cdm))
        self.sim = SimHeader.from_bytes(s.to_bytes())

    def get_state(self):
        n = self.lib.orc_np(self.o)
        c = self.cfg
        xp = np.empty((n, 3), dtype=xdtype(c.izipx))
        vp = np.empty((n, 3), dtype=xdtype(c.izipv))
        pid = np.empty(n, dtype=np.int32)
        rhoc = np.empty(c.nc ** 3, dtype=np.int32)
        vf = np.empty((c.nc ** 3, 3), dtype=np.float32)
        self._chk(self.lib.orc_get_state(self.o, p(xp), p(vp), p(pid), p(rhoc), p(vf)))
        sim = SimHeader.from_bytes(self.sim.to_bytes())
        sim.nplocal = n
        sc = self.scalars()
        sim.sigma_vi, sim.vz_max = sc["sigma_vi"], sc["vz_max"]
        return BoxState(sim, xp, vp, pid, rhoc, vf)

    def scalars(self):
        out = np.zeros(5, dtype=np.float32)
        self.lib.orc_scalars(self.o, p(out))
        return dict(zip(["sigma_vi", "sigma_vi_new", "vz_max", "mass_p", "ftemp_pp"], out.tolist()))

    def set_scalars(self, **kw):
        sc = self.scalars()
        sc.update(kw)
        arr = np.array([sc[k] for k in ["sigma_vi", "sigma_vi_new", "vz_max", "mass_p", "ftemp_pp"]], dtype=np.float32)
        self.lib.orc_set_scalars(self.o, p(arr))

    def drift(self, dt_old, dt):
        out = np.zeros(9)
        self._chk(self.lib.orc_drift(self.o, dt_old, dt, p(out), self.threads))
        return dict(zip(["sigma_vi_new", "std_vsim", "std_vsim_c", "std_vsim_res", "nlayer", "np", "sum_c", "sum_r",
                         "sum_s"], out.tolist()))

    def pp(self, a_mid, dt, forces=False):
        out = np.zeros(4)
        n = self.lib.orc_np(self.o)
        f = np.zeros((n, 3), dtype=np.float32) if forces else None
        self._chk(self.lib.orc_pp(self.o, a_mid, dt, p(out), p(f), self.threads))
        res = dict(zip(["dt_pp", "npairs", "itest1", "f2_max_pp"], out.tolist()))
        if forces:
            res["f_pp"] = f
        return res

    def pm(self, a_mid, dt):
        out = np.zeros(11)
        self._chk(self.lib.orc_pm(self.o, a_mid, dt, p(out), self.threads))
        return dict(zip(["dt_fine", "dt_coarse", "dt_vmax", "vz_max", "vmax1", "vmax2", "vmax3", "mass_f", "mass_c",
                         "f2_max_fine", "f2_max_coarse"], out.tolist()))

    def pm_forces(self):
        n = self.lib.orc_np(self.o)
        probe = np.zeros(6 * n)
        out = np.zeros(11)
        self._chk(self.lib.orc_forces(self.o, p(probe), p(out), self.threads))
        return probe[:3 * n].reshape(n, 3), probe[3 * n:].reshape(n, 3), out

    def density(self):
        nf = self.cfg.nf
        rho = np.zeros((nf, nf, nf), dtype=np.float32)
        self._chk(self.lib.orc_density(self.o, p(rho), self.threads))
        return rho

    def kernels(self):
        c = self.cfg
        nfe = c.nt * 4 + 8 * c.ncb
        kf = np.zeros((3, nfe, nfe, nfe // 2 + 1), dtype=np.float32)
        kc = np.zeros((3, c.nc, c.nc, c.nc // 2 + 1), dtype=np.float32)
        self.lib.orc_kernels(self.o, p(kf), p(kc))
        return kf, kc

    # neutrino fluid
    def nu_setup(self, f_neu, z_i_nu, gamma, weights, dxs, ais, ncs, ncbs, fluids):
        fp = np.array([f_neu, z_i_nu, gamma, *weights, *dxs, *ais], dtype=np.float32)
        ip = np.array([*ncs, *ncbs], dtype=np.int32)
        self.lib.orc_nu_setup(self.o, p(fp), p(ip))
        self.nu_ncs = list(ncs)
        for n, f in enumerate(fluids):
            self.lib.orc_nu_set(self.o, n, p(np.ascontiguousarray(f, dtype=np.float32)))

    def nu_get(self, n):
        nc = self.nu_ncs[n]
        f = np.empty((5, nc, nc, nc), dtype=np.float32)
        self.lib.orc_nu_get(self.o, n, p(f))
        return f

    def nu_evolve(self, dt, direction, a, dt_neu=1000.0):
        d = np.array([dt_neu], dtype=np.float32)
        self._chk(self.lib.orc_nu_evolve(self.o, dt, direction, a, p(d)))
        return float(d[0])


def expansion(a, dt, om, ol, wde=-1.0):
    lib = load()
    d1, d2 = _f32(), _f32()
    lib.orc_expansion(a, dt, om, ol, wde, ctypes.byref(d1), ctypes.byref(d2))
    return np.float32(d1.value), np.float32(d2.value)


def vdecode(v, sigma, izipv):
    return load().orc_vdecode(int(v), float(sigma), int(izipv))


def vencode(v, sigma, izipv):
    return load().orc_vencode(float(v), float(sigma), int(izipv))


def vencode_x(x, izipv):
    """codes of the atan arguments x = c*v (update_particle.f90:119)"""
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty(len(x), dtype=np.int32)
    load().orc_vencode_x(p(x), len(x), int(izipv), p(out))
    return out


def fft_r2c(r):
    n = r.shape[0]
    r = np.ascontiguousarray(r, dtype=np.float32)
    re = np.zeros((n, n, n // 2 + 1), dtype=np.float32)
    im = np.zeros_like(re)
    load().orc_fft_r2c(p(r), n, p(re), p(im), 4)
    return re + 1j * im.astype(np.float64)




def fluid_evolve(fld, ncb, gamma, dt, dx, direction, dtg=1000.0):
    """ofl_evolve on one fluid's physical fields [5][nc][nc][nc] (in place copy returned)."""
    lib = load()
    f = np.ascontiguousarray(fld, dtype=np.float32).copy()
    d = np.array([dtg], dtype=np.float32)
    rc = lib.ofl_evolve(p(f), f.shape[1], ncb, gamma, dt, dx, direction, p(d))
    if rc:
        raise RuntimeError("ofl_evolve failed")
    return f, float(d[0])


def halo_arrays(cfg, st, hp):
    """Parameter arrays of ohf_halofind for state st and HaloParams hp."""
    ip = np.array([cfg.nc, cfg.nt, cfg.nnt, cfg.ncb, cfg.izipx, cfg.izipv, hp.nc_halo_max, hp.newbox_max,
                   hp.max_halo_np, hp.search_ratio, hp.refine_ratio, hp.min_halo_particles, hp.nq], dtype=np.int64)
    s = st.sim
    fp = np.array([hp.den_odc, hp.den_peak_cutoff, s.mass_p_cdm, s.sigma_vi, hp.a or s.a, hp.vsim2phys or s.vsim2phys,
                   hp.h0, hp.box, *hp.los_centre], dtype=np.float32)
    return ip, fp


def halofind(cfg, st, hp, max_halos=100000, nhid=0):
    """CPU restatement of halofind.f90 on state st: (halos [n,42] float32,
    pid stream int32, hid int16 or None, counts)."""
    lib = load()
    ip, fp = halo_arrays(cfg, st, hp)
    n = st.sim.nplocal
    halos = np.zeros((max_halos, 42), dtype=np.float32)
    pidw = np.zeros(2 * max_halos + n + 16, dtype=np.int32)
    hid = np.zeros(nhid, dtype=np.int16) if nhid else None
    counts = np.zeros(4, dtype=np.int64)
    err = ctypes.create_string_buffer(256)
    rc = lib.ohf_halofind(p(ip), p(fp), n, p(np.ascontiguousarray(st.xp)), p(np.ascontiguousarray(st.vp)),
                          p(np.ascontiguousarray(st.pid)), p(np.ascontiguousarray(st.rhoc)),
                          p(np.ascontiguousarray(st.vfield)), max_halos, p(halos), len(pidw), p(pidw), nhid, p(hid),
                          p(counts), err)
    if rc:
        raise RuntimeError("halo oracle: " + err.value.decode())
    return halos[:counts[0]].copy(), pidw[:counts[1]].copy(), hid, counts.copy()


def halo_density(cfg, st, hp):
    lib = load()
    ip, fp = halo_arrays(cfg, st, hp)
    nf = cfg.nf
    rho = np.zeros((nf, nf, nf), dtype=np.float32)
    lib.ohf_density(p(ip), p(fp), p(np.ascontiguousarray(st.xp)), p(np.ascontiguousarray(st.rhoc)), p(rho))
    return rho


def halo_offset_table(R):
    lib = load()
    cap = (2 * R + 1) ** 3
    out = np.zeros((cap, 4), dtype=np.int32)
    n = lib.ohf_offset_table(R, p(out), cap)
    return out[:n].copy()
