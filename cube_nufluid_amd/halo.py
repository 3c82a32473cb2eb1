"""Runtime halo finder: parameters, the C-ABI call and CUBE's halo files.

The finder is CUBE's spherical-overdensity search (halofind.f90:18-598),
run on the GPU by `cube_halofind` over the resident state.  This module
holds its parameters (the constants at halofind.f90:27-40 plus the few
module parameters it reads), calls the entry point, and writes the three
files of halofind.f90:117-125 in the reference's formats:

    <prefix>halo_1.bin      int32 nhalo_tot, int32 nhalo, float32 den_odc,
                            then nhalo records of 42 float32 (type_halo_info,
                            halofind.f90:9-15)
    <prefix>halo_pid_1.bin  per halo: int32 i_odc, i_odc int32 pids, int32 0
    <prefix>hid_1.bin       int16 halo number per pid (nf^3 entries)

where <prefix> is opath + 'image1/' + z + '_' (basic_functions.fh:68-74).
"""
import ctypes
import os
from dataclasses import dataclass

import numpy as np

from . import _lib

# column layout of one type_halo_info record (halofind.f90:9-15)
HALO_FIELDS = {"hpos": slice(0, 3), "mass_odc": 3, "radius_odc": 4, "v_disp": 5, "x_mean": slice(6, 9),
               "v_mean": slice(9, 12), "ang_mom": slice(12, 15), "var_x": slice(15, 18), "inertia": slice(18, 27),
               "q_mean": slice(27, 30), "inertia_q": slice(30, 39), "s_mean": slice(39, 42)}
HALO_RECORD_FLOATS = 42


@dataclass
class HaloParams:
    """halofind.f90:27-40 (den_odc = 200: odc_vir is .false.) and the module
    parameters the finder reads (h0 and box of parameters.f90; h0 = 0 takes
    the box's own)."""
    den_odc: float = 200.0
    den_peak_cutoff: float = 100.0
    min_halo_particles: int = 100
    nc_halo_max: int = 128
    newbox_max: int = 300
    max_halo_np: int = 5 * 129 ** 3 // 3
    search_ratio: int = 2
    refine_ratio: int = 3
    nq: int = 0                  # pid lattice size per dimension (pids are q-lattice indices + 1)
    h0: float = 0.0
    box: float = 0.0
    los_centre: tuple = (30.0, 370.0, 370.0)   # halofind.f90:58 'cen'
    a: float = 0.0               # sim%a and sim%vsim2phys of the current state
    vsim2phys: float = 0.0

    def for_box(self, cfg, nq=None, sim=None):
        """Copy with box, pid lattice and (from header sim) a and vsim2phys
        filled in for box cfg.  The reference's pids index the fine grid
        (nq = nf); the IC writer here numbers its particle lattice (nq =
        particles per dimension)."""
        import dataclasses
        a, vs = self.a, self.vsim2phys
        if sim is not None:
            a, vs = a or float(sim.a), vs or float(sim.vsim2phys)
        return dataclasses.replace(self, box=self.box or cfg.box, h0=self.h0 or cfg.h0, nq=nq or self.nq or cfg.nf,
                                   a=a, vsim2phys=vs)


class cube_halo_params(ctypes.Structure):
    _fields_ = [("den_odc", ctypes.c_float), ("den_peak_cutoff", ctypes.c_float),
                ("min_halo_particles", ctypes.c_int32), ("nc_halo_max", ctypes.c_int32),
                ("newbox_max", ctypes.c_int32), ("search_ratio", ctypes.c_int32),
                ("refine_ratio", ctypes.c_int32), ("pad", ctypes.c_int32),
                ("max_halo_np", ctypes.c_int64), ("nq", ctypes.c_int64),
                ("h0", ctypes.c_float), ("box", ctypes.c_float), ("los_centre", ctypes.c_float * 3),
                ("a", ctypes.c_float), ("vsim2phys", ctypes.c_float), ("pad2", ctypes.c_float)]


class cube_halo_catalog(ctypes.Structure):
    _fields_ = [("max_halos", ctypes.c_int64), ("halos", ctypes.c_void_p),
                ("max_pid_words", ctypes.c_int64), ("pid_words", ctypes.c_void_p),
                ("nhid", ctypes.c_int64), ("hid", ctypes.c_void_p),
                ("nhalo", ctypes.c_int64), ("npid_words", ctypes.c_int64),
                ("npeaks", ctypes.c_int64), ("nprocessed", ctypes.c_int64)]


@dataclass
class HaloCatalog:
    den_odc: float
    halos: np.ndarray        # [nhalo, 42] float32
    pid_words: np.ndarray    # halo_pid stream
    hid: np.ndarray          # int16 per pid, or None
    npeaks: int
    nprocessed: int

    @property
    def nhalo(self):
        return len(self.halos)

    def field(self, name):
        return self.halos[:, HALO_FIELDS[name]]

    def members(self):
        """pid arrays of the halos, in catalogue order."""
        out, i = [], 0
        w = self.pid_words
        while i < len(w):
            n = int(w[i])
            out.append(w[i + 1:i + 1 + n])
            i += n + 2
        return out


def params_struct(hp):
    s = cube_halo_params()
    s.den_odc, s.den_peak_cutoff = hp.den_odc, hp.den_peak_cutoff
    s.min_halo_particles, s.nc_halo_max, s.newbox_max = hp.min_halo_particles, hp.nc_halo_max, hp.newbox_max
    s.search_ratio, s.refine_ratio = hp.search_ratio, hp.refine_ratio
    s.max_halo_np, s.nq = hp.max_halo_np, hp.nq
    s.h0, s.box, s.a, s.vsim2phys = hp.h0, hp.box, hp.a, hp.vsim2phys
    for d in range(3):
        s.los_centre[d] = hp.los_centre[d]
    return s


def find_halos(lib, handle, hp, max_halos=1 << 20, max_pid_words=None, nhid=0):
    """Run cube_halofind on the resident state of `handle`."""
    if max_pid_words is None:
        max_pid_words = 2 * max_halos + (1 << 22)
    halos = np.zeros((max_halos, HALO_RECORD_FLOATS), dtype=np.float32)
    words = np.zeros(max_pid_words, dtype=np.int32)
    hid = np.zeros(nhid, dtype=np.int16) if nhid else None
    cat = cube_halo_catalog()
    cat.max_halos, cat.halos = max_halos, halos.ctypes.data
    cat.max_pid_words, cat.pid_words = max_pid_words, words.ctypes.data
    cat.nhid, cat.hid = nhid, (hid.ctypes.data if hid is not None else None)
    ps = params_struct(hp)
    rc = lib.cube_halofind(handle, ctypes.byref(ps), ctypes.byref(cat))
    _lib.check(lib, handle, rc)
    return HaloCatalog(hp.den_odc, halos[:cat.nhalo].copy(), words[:cat.npid_words].copy(), hid,
                       int(cat.npeaks), int(cat.nprocessed))


def halo_prefix(opath, z, image=1):
    """opath//'image1/'//z_str//'_' with z written as f7.3 (basic_functions.fh:68-74),
    the same z string as the checkpoint files."""
    from .checkpoint import z2str
    return os.path.join(opath, "image%d" % image, z2str(z) + "_")


def write_halo_files(prefix, cat, image=1):
    """The halo, halo_pid and hid files of halofind.f90 (one image)."""
    d = os.path.dirname(prefix)
    if d:
        os.makedirs(d, exist_ok=True)
    suffix = "_%d.bin" % image
    with open(prefix + "halo" + suffix, "wb") as f:
        np.array([cat.nhalo, cat.nhalo], dtype=np.int32).tofile(f)
        np.array([cat.den_odc], dtype=np.float32).tofile(f)
        cat.halos.astype(np.float32).tofile(f)
    with open(prefix + "halo_pid" + suffix, "wb") as f:
        cat.pid_words.astype(np.int32).tofile(f)
    if cat.hid is not None:
        with open(prefix + "hid" + suffix, "wb") as f:
            cat.hid.astype(np.int16).tofile(f)


def read_halo_file(path):
    """(nhalo_tot, nhalo, den_odc, records) of a halo file; one whose
    records do not match its count is rejected."""
    with open(path, "rb") as f:
        hdr = np.fromfile(f, dtype=np.int32, count=2)
        den = np.fromfile(f, dtype=np.float32, count=1)
        recs = np.fromfile(f, dtype=np.float32)
    if hdr.size != 2 or den.size != 1 or recs.size != int(hdr[1]) * HALO_RECORD_FLOATS:
        raise ValueError("halo file %s: %d values after the header, %s halos of %d declared"
                         % (path, recs.size, hdr[1] if hdr.size == 2 else "?", HALO_RECORD_FLOATS))
    return int(hdr[0]), int(hdr[1]), float(den[0]), recs.reshape(-1, HALO_RECORD_FLOATS)
