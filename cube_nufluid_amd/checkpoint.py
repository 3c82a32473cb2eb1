"""CUBE checkpoint files, byte-compatible with main/checkpoint.f90:12-55 and
main/particle_initialization.f90:34-82.

Files are stream binaries named '{z:f7.3}_{content}_{image}.bin'
(basic_functions.fh:52-87) inside '{opath}/image{image}/':
  info : 224-byte sim_header
  xp   : 3*nplocal components of izipx bytes
  vp   : 3*nplocal components of izipv bytes
  np   : rhoc over physical cells, int32, tile-major (nt^3 * nnt^3)
  vc   : vfield over physical cells, 3 float32 per cell
  id   : pid, int32 per particle
"""
import dataclasses
import os
import re
from decimal import Decimal

import numpy as np

from ._lib import SimHeader

CONTENTS = ("info", "xp", "vp", "np", "vc", "id")


def z2str(z):
    """basic_functions.fh:36-42: write(str,'(f7.3)') z, then trim."""
    return ("%7.3f" % np.float32(z)).strip()


def output_name(opath, z, content, image=1):
    return os.path.join(opath, "image%d" % image, "%s_%s_%d.bin" % (z2str(z), content, image))


@dataclasses.dataclass
class BoxState:
    """Host copy of the resident state in checkpoint layout."""
    sim: SimHeader
    xp: np.ndarray      # (nplocal, 3) int8/int16
    vp: np.ndarray      # (nplocal, 3) int8/int16
    pid: np.ndarray     # (nplocal,) int32
    rhoc: np.ndarray    # (nc^3,) int32, checkpoint (tile-major) order
    vfield: np.ndarray  # (nc^3, 3) float32

    def copy(self):
        s = SimHeader.from_bytes(self.sim.to_bytes())
        return BoxState(s, self.xp.copy(), self.vp.copy(), self.pid.copy(), self.rhoc.copy(), self.vfield.copy())


def xdtype(izip):
    return np.int8 if izip == 1 else np.int16


def write_checkpoint(state, opath, z, image=1):
    """checkpoint.f90:36-47."""
    os.makedirs(os.path.join(opath, "image%d" % image), exist_ok=True)
    n = state.sim.nplocal
    with open(output_name(opath, z, "info", image), "wb") as f:
        f.write(state.sim.to_bytes())
    state.xp[:n].astype(xdtype(state.sim.izipx)).tofile(output_name(opath, z, "xp", image))
    state.vp[:n].astype(xdtype(state.sim.izipv)).tofile(output_name(opath, z, "vp", image))
    state.rhoc.astype(np.int32).tofile(output_name(opath, z, "np", image))
    state.vfield.astype(np.float32).tofile(output_name(opath, z, "vc", image))
    state.pid[:n].astype(np.int32).tofile(output_name(opath, z, "id", image))


def read_checkpoint(opath, z, nc, image=1, izipx=None, izipv=None):
    """particle_initialization.f90:34-82; a header whose izipx/izipv differ
    from the requested ones is rejected as in :41-45."""
    with open(output_name(opath, z, "info", image), "rb") as f:
        sim = SimHeader.from_bytes(f.read())
    if (izipx is not None and sim.izipx != izipx) or (izipv is not None and sim.izipv != izipv):
        raise ValueError("zip format incompatible: file %d/%d, expected %s/%s" % (sim.izipx, sim.izipv, izipx, izipv))
    n = sim.nplocal

    def load(content, dtype, count):
        fn = output_name(opath, z, content, image)
        a = np.fromfile(fn, dtype=dtype, count=count)
        if a.size != count:
            raise ValueError("truncated checkpoint: %s holds %d of %d values" % (fn, a.size, count))
        return a

    xp = load("xp", xdtype(sim.izipx), 3 * n).reshape(n, 3)
    vp = load("vp", xdtype(sim.izipv), 3 * n).reshape(n, 3)
    rhoc = load("np", np.int32, nc ** 3)
    vfield = load("vc", np.float32, 3 * nc ** 3).reshape(-1, 3)
    pid = load("id", np.int32, n)
    if int(rhoc.sum(dtype=np.int64)) != n:
        raise ValueError("sum(np) = %d != nplocal = %d" % (int(rhoc.sum()), n))
    return BoxState(sim, xp, vp, pid, rhoc, vfield)


_F84 = re.compile(r"([+-]?)(\d*)(?:\.(\d*))?(?:[eEdDqQ]([+-]?\d+)|([+-]\d+))?")


def read_f84(record):
    """The value a Fortran read with edit descriptor (f8.4) takes from one
    record, as a float32 value: the first 8 columns (a shorter record is
    padded with blanks), blanks ignored, an all-blank field reads 0, and a
    mantissa written without a decimal point has 4 implied decimal places
    ("10" reads 0.001).  Anything else in the field is an error."""
    s = record[:8].replace(" ", "")
    if not s:
        return 0.0
    m = _F84.fullmatch(s)
    if m is None or not (m.group(2) or m.group(3)):
        raise ValueError("bad real input for (f8.4): %r" % record[:8])
    sign, ip, fp, e1, e2 = m.groups()
    digits, scale = (ip, -4) if fp is None else (ip + fp, -len(fp))
    v = Decimal(sign + (digits or "0")).scaleb(scale + int(e1 or e2 or 0))
    f = float(v)
    c = np.float32(f)
    # decimal -> double -> float32 rounds twice, which can only go wrong
    # when the double lands exactly halfway between two float32 values
    lo, hi = (np.nextafter(c, np.float32(-np.inf)), c) if float(c) > f else (c, np.nextafter(c, np.float32(np.inf)))
    if (float(lo) + float(hi)) / 2 == f and Decimal(f) != v:
        c = hi if v > Decimal(f) else lo
    return float(c)


def read_z_list(path, nmax=999):
    """initialize.f90:55-61: one redshift per record, read with (f8.4)
    (read_f84) until the end of the file, at most nmax_redshift-1 = 999 of
    them.  A blank record reads 0, as in the reference; a file with no
    records is an error ('z_checkpoint.txt empty', initialize.f90:61)."""
    zs = []
    with open(path) as f:
        for line in f:
            if len(zs) == nmax:
                break
            zs.append(read_f84(line.rstrip("\n")))
    if not zs:
        raise ValueError("%s empty" % os.path.basename(path))
    return zs


# ------------------------------------------------------------ neutrino fluid
def neu_ncb(dx, ncb=6, ncell=4):
    """Halo width of a fluid of cell size dx (hydrogr.f90:53)."""
    return max(2, 1 + int(np.floor(ncb * ncell / dx)))


def write_neu_checkpoint(fn, m_ev, gamma, isothermal, weights, fluids, dxs, ncb=6):
    """dfld_checkpoint (hydrodf.f90:143-166): m/eV, gamma, isothermal
    (logical, 4 bytes), the three density weights, then every fluid's
    fld(5, 1-ncb:nc+ncb, ...) with its periodic halo, first index fastest.
    fluids: [5][nc][nc][nc] float32 arrays (z, y, x order)."""
    with open(fn, "wb") as f:
        np.array([m_ev, gamma], dtype=np.float32).tofile(f)
        np.array([1 if isothermal else 0], dtype=np.int32).tofile(f)
        np.asarray(weights, dtype=np.float32).tofile(f)
        for fl, dx in zip(fluids, dxs):
            b = neu_ncb(dx, ncb)
            ext = np.pad(np.asarray(fl, dtype=np.float32), ((0, 0), (b, b), (b, b), (b, b)), mode="wrap")
            np.ascontiguousarray(np.moveaxis(ext, 0, -1)).tofile(f)   # (k, j, i, q): q fastest


def read_neu_checkpoint(fn, ncs, dxs, ncb=6):
    """Inverse of write_neu_checkpoint: (m_ev, gamma, isothermal, weights,
    fluids as [5][nc][nc][nc] physical cells).  A file shorter than the
    fluids it should hold is rejected."""
    with open(fn, "rb") as f:
        def load(dtype, count, what):
            a = np.fromfile(f, dtype=dtype, count=count)
            if a.size != count:
                raise ValueError("truncated neutrino checkpoint: %s holds %d of %d values of %s" % (fn, a.size, count, what))
            return a

        m_ev, gamma = load(np.float32, 2, "the header")
        iso = bool(load(np.int32, 1, "the header")[0])
        w = load(np.float32, 3, "the header")
        fluids = []
        for i, (nc, dx) in enumerate(zip(ncs, dxs)):
            b = neu_ncb(dx, ncb)
            ne = nc + 2 * b
            ext = load(np.float32, 5 * ne ** 3, "fluid %d" % (i + 1)).reshape(ne, ne, ne, 5)
            fluids.append(np.ascontiguousarray(np.moveaxis(ext, -1, 0)[:, b:b + nc, b:b + nc, b:b + nc]))
    return float(m_ev), float(gamma), iso, w, fluids
