"""ctypes binding of libcubehip.so (declarations in include/cube_hip.h).

The library is the only compute path: if it has not been built, importing
the binding raises instead of falling back to anything else.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcubehip.so")

c_i64 = ctypes.c_int64
c_i32 = ctypes.c_int32
c_f32 = ctypes.c_float
c_f64 = ctypes.c_double


class CubeParams(ctypes.Structure):
    _fields_ = [("nc", c_i64), ("nnt", c_i64), ("ncb", c_i64), ("izipx", c_i64), ("izipv", c_i64),
                ("np_max", c_i64), ("pp_on", c_i32), ("pp_range", c_i32), ("rsoft", c_f32),
                ("omega_m", c_f32), ("omega_l", c_f32), ("wde", c_f32), ("precision", c_i32),
                ("device", c_i32)]


SIM_INT_FIELDS = ["nplocal", "npglobal", "nplocal_nu", "npglobal_nu", "izipx", "izipv", "izipx_nu",
                  "izipv_nu", "image", "nn", "nnt", "nt", "ncell", "ncb", "timestep", "cur_checkpoint",
                  "cur_halofind"]
SIM_REAL_FIELDS = ["a", "t", "tau", "dt_pp", "dt_fine", "dt_coarse", "dt_vmax", "dt_vmax_nu", "mass_p_cdm",
                   "mass_p_nu", "box", "h0", "omega_m", "omega_l", "s8", "vsim2phys", "sigma_vres",
                   "sigma_vi", "sigma_vi_nu", "z_i", "z_i_nu", "vz_max"]


class SimHeader(ctypes.Structure):
    """The 224-byte 'sim_header' of parameters.f90:129-153."""
    _fields_ = [(n, c_i64) for n in SIM_INT_FIELDS] + [(n, c_f32) for n in SIM_REAL_FIELDS]

    def to_bytes(self):
        return ctypes.string_at(ctypes.addressof(self), ctypes.sizeof(self))

    @classmethod
    def from_bytes(cls, b):
        if len(b) != ctypes.sizeof(cls):
            raise ValueError("sim header must be %d bytes, got %d" % (ctypes.sizeof(cls), len(b)))
        h = cls()
        ctypes.memmove(ctypes.addressof(h), b, len(b))
        return h

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class DriftOut(ctypes.Structure):
    _fields_ = [("sigma_vi_new", c_f32), ("nlayer", c_i32), ("std_vsim", c_f64), ("std_vsim_c", c_f64),
                ("std_vsim_res", c_f64), ("npcheck", c_i64), ("max_disp", c_i32 * 3), ("frame_conflicts", c_i32)]


class PPOut(ctypes.Structure):
    _fields_ = [("dt_pp", c_f32), ("f2_max_pp", c_f32), ("npairs", c_i64), ("itest1", c_i64), ("dense_blocks", c_i64)]


class PMOut(ctypes.Structure):
    _fields_ = [("dt_fine", c_f32), ("dt_coarse", c_f32), ("dt_vmax", c_f32), ("vz_max", c_f32),
                ("vmax", c_f32 * 3), ("f2_max_fine", c_f32), ("f2_max_coarse", c_f32),
                ("mass_f", c_f64), ("mass_c", c_f64)]


class ICParams(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_uint64), ("np_nc", c_i32), ("nk", c_i32), ("a", c_f32), ("box", c_f32),
                ("vfac", c_f64), ("lnk", ctypes.c_void_p), ("lnp", ctypes.c_void_p), ("f_cdm", c_f32),
                ("reserved", c_i32)]


assert ctypes.sizeof(SimHeader) == 224
assert ctypes.sizeof(CubeParams) == 80

# exported symbols and their signatures (include/cube_hip.h)
_P = ctypes.c_void_p
class CubeClock(ctypes.Structure):
    _fields_ = [("a", c_f32), ("t", c_f32), ("tau", c_f32), ("dt", c_f32), ("dt_old", c_f32), ("da", c_f32),
                ("a_mid", c_f32), ("dt_neu", c_f32), ("cur_checkpoint", c_i32), ("n_checkpoint", c_i32),
                ("checkpoint_step", c_i32), ("final_step", c_i32), ("checkpoint_at", c_i32), ("pad", c_i32),
                ("istep", c_i64), ("z_checkpoint", c_f32 * 64)]


class NuParams(ctypes.Structure):
    _fields_ = [("m_ev", c_f32), ("f_neu", c_f32), ("z_i_nu", c_f32), ("gamma", c_f32), ("box", c_f32), ("h0", c_f32),
                ("omega_m", c_f32), ("neutrino_ic", c_i32)]


class EnsParams(ctypes.Structure):
    _fields_ = [("device", c_i32), ("max_live", c_i32), ("mem_budget", c_i64), ("step_mode", c_i32), ("reserved", c_i32)]


class EnsResult(ctypes.Structure):
    _fields_ = [("id", c_i64), ("status", c_i32), ("steps", c_i32), ("t_submit", c_f64), ("t_admit", c_f64),
                ("t_done", c_f64), ("nparticles", c_i64), ("a_final", c_f32), ("pad", c_i32),
                ("state_hash", ctypes.c_uint64), ("message", ctypes.c_char * 128)]


SIGNATURES = {
    "cube_init": (c_i32, [ctypes.POINTER(CubeParams), _P, _P, ctypes.POINTER(_P)]),
    "cube_upload_state": (c_i32, [_P, ctypes.POINTER(SimHeader), _P, _P, _P, _P, _P]),
    "cube_download_state": (c_i32, [_P, ctypes.POINTER(SimHeader), _P, _P, _P, _P, _P]),
    "cube_drift": (c_i32, [_P, c_f32, c_f32, ctypes.POINTER(DriftOut)]),
    "cube_pp": (c_i32, [_P, c_f32, c_f32, ctypes.POINTER(PPOut)]),
    "cube_pm": (c_i32, [_P, c_f32, c_f32, ctypes.POINTER(PMOut)]),
    "cube_expansion": (c_i32, [c_f32, c_f32, c_f32, c_f32, c_f32, ctypes.POINTER(c_f32), ctypes.POINTER(c_f32)]),
    "cube_last_error": (c_i32, [_P, ctypes.c_char_p, ctypes.c_size_t]),
    "cube_finalize": (c_i32, [_P]),
    "cube_probe_pm_forces": (c_i32, [_P, _P, _P]),
    "cube_probe_pp_forces": (c_i32, [_P, _P]),
    "cube_probe_density": (c_i32, [_P, _P]),
    "cube_probe_vencode": (c_i32, [_P, _P, ctypes.c_int64, _P]),
    "cube_power_spectrum": (c_i32, [_P, c_i32, _P, _P, _P]),
    "cube_power_spectrum_noalloc": (c_i32, [_P, c_i32, _P, _P, _P]),
    "cube_halofind": (c_i32, [_P, _P, _P]),
    "cube_nu_init": (c_i32, [_P, ctypes.POINTER(NuParams)]),
    "cube_nu_info": (c_i32, [_P, c_i32, ctypes.POINTER(c_i32), ctypes.POINTER(c_f32), ctypes.POINTER(c_f32),
                             ctypes.POINTER(c_f32)]),
    "cube_nu_set_fluid": (c_i32, [_P, c_i32, _P]),
    "cube_nu_get_fluid": (c_i32, [_P, c_i32, _P]),
    "cube_fluid_evolve": (c_i32, [_P, c_f32, c_i32, c_f32, ctypes.POINTER(c_f32)]),
    "cube_clock_init": (c_i32, [ctypes.POINTER(CubeClock), _P, c_i32]),
    "cube_step": (c_i32, [_P, ctypes.POINTER(CubeClock)]),
    "cube_footprint": (c_i64, [ctypes.POINTER(CubeParams), c_i32]),
    "cube_ens_create": (c_i32, [ctypes.POINTER(EnsParams), ctypes.POINTER(_P)]),
    "cube_ens_submit": (c_i32, [_P, ctypes.POINTER(CubeParams), _P, _P, ctypes.POINTER(ICParams), _P, c_i32,
                                ctypes.POINTER(c_i64)]),
    "cube_ens_poll": (c_i32, [_P, ctypes.POINTER(EnsResult), c_i32, ctypes.POINTER(c_i32)]),
    "cube_ens_pending": (c_i32, [_P, ctypes.POINTER(c_i32), ctypes.POINTER(c_i32)]),
    "cube_ens_clock": (c_f64, [_P]),
    "cube_ens_destroy": (c_i32, [_P]),
    "cube_ic_zeldovich": (c_i32, [_P, ctypes.POINTER(ICParams), ctypes.POINTER(SimHeader)]),
    "cube_nu_ic": (c_i32, [_P, ctypes.POINTER(ICParams)]),
    "cube_profile_enable": (c_i32, [_P, c_i32]),
    "cube_profile_read": (c_i32, [_P, _P, _P, c_i32]),
    "cube_profile_slot_name": (c_i32, [c_i32, ctypes.c_char_p, ctypes.c_size_t]),
    "cube_hbm_copy_probe": (c_i32, [c_i32, c_i64, c_i32, _P]),
    "cube_set_step_mode": (c_i32, [_P, c_i32]),
    "cube_step_stats": (c_i32, [_P, _P, _P]),
}

_lib = None


def load():
    """Load libcubehip.so; raises if the HIP library has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libcubehip.so not built (%s); run __graft_entry__.build() or "
                           "python cube_nufluid_amd/csrc/build.py" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def ptr(a):
    """Raw pointer of a C-contiguous numpy array (or None)."""
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data_as(ctypes.c_void_p)


class CubeError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("cube error %d: %s" % (code, msg))
        self.code = code


def check(lib, h, rc):
    if rc != 0:
        buf = ctypes.create_string_buffer(1024)
        lib.cube_last_error(h, buf, 1024)
        raise CubeError(rc, buf.value.decode(errors="replace"))


def expansion(a, dt, omega_m, omega_l, wde=-1.0):
    """timestep.f90:136-180 through the library (float interface)."""
    lib = load()
    d1, d2 = c_f32(), c_f32()
    check(lib, None, lib.cube_expansion(a, dt, omega_m, omega_l, wde, ctypes.byref(d1), ctypes.byref(d2)))
    return np.float32(d1.value), np.float32(d2.value)
