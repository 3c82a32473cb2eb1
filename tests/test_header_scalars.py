"""Scalars and strings the hosts write into CUBE's files, against the same
expressions compiled by a Fortran compiler (amdflang at -O0, default real =
real(4)).

vsim2phys, the velocity unit the hosts write into sim headers, has the
reference's real(4) bits: the initial conditions' (initial_conditions.f90:
198), the PM step's (pm.f90:427, with h0) and the checkpoint's
(checkpoint.f90:30, without h0).  Each expression is compiled here as a
Fortran program (amdflang at -O0, default real = real(4)) and evaluated
for a range of a and box parameters; the Python hosts must give the same
bits.  Evaluating in double and rounding once, or dropping the h0 factor,
fails."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from cube_nufluid_amd.driver import Cube
from cube_nufluid_amd.ic import ic_vsim2phys
from cube_nufluid_amd.params import BoxConfig

FC = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"

PROGRAM = """program v
  implicit none
  real :: a, box, h0, omega_m
  integer :: nf, i, n
  read(*,*) box, h0, omega_m, nf, n
  do i = 1, n
    read(*,*) a
    print '(3I12)', transfer((150./a)*box*sqrt(omega_m)/nf, 0), &
      transfer((1.5/a)*box*h0*100.*sqrt(omega_m)/nf, 0), &
      transfer((1.5/a)*box*100.*sqrt(omega_m)/nf, 0)
  end do
end program
"""


@pytest.fixture(scope="module")
def fortran(tmp_path_factory):
    if not os.path.exists(FC):
        pytest.skip("no Fortran compiler")
    d = tmp_path_factory.mktemp("vsim")
    (d / "v.f90").write_text(PROGRAM)
    subprocess.run([FC, "-O0", str(d / "v.f90"), "-o", str(d / "v")], check=True, capture_output=True)
    return str(d / "v")


def fortran_bits(exe, box, h0, om, nf, avals):
    inp = "%.9g %.9g %.9g %d %d\n" % (box, h0, om, nf, len(avals)) + "".join("%.9g\n" % a for a in avals)
    out = subprocess.run([exe], input=inp, capture_output=True, text=True, check=True).stdout.split()
    return np.array(out, dtype=np.int64).reshape(-1, 3).astype(np.uint32)


@pytest.mark.parametrize("nc,box,h0,om", [(16, 100.0, 0.68, 0.31), (512, 1000.0, 0.6774, 0.3089), (128, 400.0, 0.7, 0.27)])
def test_vsim2phys_matches_the_reference_expressions(fortran, nc, box, h0, om):
    cfg = BoxConfig(nc=nc, nnt=1, ncb=6, box=box)
    cfg.h0, cfg.omega_m = h0, om
    rng = np.random.default_rng(nc)
    avals = np.concatenate([np.float32(1) / (np.float32(1) + np.array([49, 10, 3, 1, 0.5, 0], np.float32)),
                            rng.uniform(0.015, 1.0, 200).astype(np.float32)])
    # the inputs as real(4) values, written with enough digits to read back exactly
    want = fortran_bits(fortran, np.float32(box), np.float32(h0), np.float32(om), cfg.nf, avals)
    host = Cube.__new__(Cube)
    host.cfg = cfg
    got = np.empty_like(want)
    for i, a in enumerate(avals):
        host.a = np.float32(a)
        got[i] = [np.float32(ic_vsim2phys(cfg, a)).view(np.uint32), host.vsim2phys(True).view(np.uint32),
                  host.vsim2phys(False).view(np.uint32)]
    for col, name in enumerate(("initial conditions", "PM step", "checkpoint")):
        bad = np.flatnonzero(got[:, col] != want[:, col])
        assert bad.size == 0, (name, avals[bad[:3]], got[bad[:3], col], want[bad[:3], col])


Z_PROGRAM = """program f
  implicit none
  real :: z
  integer :: i, n
  character(len=7) :: s
  read(*,*) n
  do i = 1, n
    read(*,*) z
    write(s, '(f7.3)') z
    print '(A)', trim(adjustl(s))
  end do
end program
"""


def test_file_name_redshift_matches_fortran_f73(tmp_path):
    """Checkpoint and halo files carry z as write(z_str,'(f7.3)') of the
    real(4) redshift, trimmed (basic_functions.fh:36-42, 52-87).  The
    redshift lists are read as (f8.4), so four-decimal values whose last
    digit is 5, and binary ties such as 0.0625, are checked too."""
    from cube_nufluid_amd.checkpoint import output_name, z2str
    from cube_nufluid_amd.halo import halo_prefix
    if not os.path.exists(FC):
        pytest.skip("no Fortran compiler")
    (tmp_path / "f.f90").write_text(Z_PROGRAM)
    subprocess.run([FC, "-O0", str(tmp_path / "f.f90"), "-o", str(tmp_path / "f")], check=True, capture_output=True)
    zs = np.concatenate([np.arange(0, 64, 1 / 16), np.round(np.arange(0, 50, 0.0005), 4),
                         [99.9995, 127.5, 999.0625]]).astype(np.float32)
    inp = "%d\n" % len(zs) + "".join("%.9g\n" % z for z in zs)
    want = subprocess.run([str(tmp_path / "f")], input=inp, capture_output=True, text=True, check=True).stdout.split()
    got = [z2str(z) for z in zs]
    bad = [(float(z), w, g) for z, w, g in zip(zs, want, got) if w != g]
    assert not bad, bad[:5]
    for z in (0.0025, 1.0005, 49.0):
        assert os.path.basename(halo_prefix("o", z)) + "halo_1.bin" == \
            os.path.basename(output_name("o", z, "halo"))


ZLIST_PROGRAM = """program r
  implicit none
  real :: z(1000)
  integer :: i, n, u
  open(newunit=u, file='z.txt', status='old')
  do i = 1, 999
    read(u, end=71, fmt='(f8.4)') z(i)
  enddo
  71 n = i - 1
  close(u)
  print '(I6)', n
  do i = 1, n
    print '(I12)', transfer(z(i), 0)
  end do
end program
"""


@pytest.mark.parametrize("nlines", [400, 1005])
def test_redshift_list_reads_as_fortran_f84(tmp_path, nlines):
    """z_checkpoint.txt as initialize.f90:55-61 reads it, (f8.4) up to 999
    records: only the first 8 columns count, blanks are ignored, a blank
    record is 0, and a number without a decimal point has 4 implied
    decimals.  A plain float() of each line gets many of these records wrong."""
    from cube_nufluid_amd.checkpoint import read_z_list
    if not os.path.exists(FC):
        pytest.skip("no Fortran compiler")
    rng = np.random.default_rng(nlines)
    lines = ["", "0", "10", "  49.0000", "49.000000", "100.00000", "   12345", "1e1", "1.5+2", "1.5d-1", "-0.5",
             "5.", " 2.25  "]
    while len(lines) < nlines:
        s = ("%.*f" % (int(rng.integers(0, 6)), rng.uniform(0, 200)) if rng.random() < 0.8
             else str(int(rng.integers(0, 99999))))
        line = " " * int(rng.integers(0, 3)) + s
        if len(line) >= 8 and rng.random() < 0.2:
            line += "  past column 8"
        lines.append(line)
    (tmp_path / "z.txt").write_text("".join(ln + "\n" for ln in lines))
    (tmp_path / "r.f90").write_text(ZLIST_PROGRAM)
    subprocess.run([FC, "-O0", "r.f90", "-o", "r"], cwd=tmp_path, check=True, capture_output=True)
    out = subprocess.run(["./r"], cwd=tmp_path, capture_output=True, text=True, check=True).stdout.split()
    n = int(out[0])
    want = np.array(out[1:], dtype=np.int64).astype(np.uint32)
    got = np.array(read_z_list(str(tmp_path / "z.txt")), dtype=np.float32).view(np.uint32)
    assert n == len(got) == min(nlines, 999)
    bad = np.flatnonzero(got != want)
    assert bad.size == 0, [(lines[i], got[i], want[i]) for i in bad[:5]]
