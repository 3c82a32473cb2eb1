"""The Fortran host (cube_nufluid_amd/fortran/cube_main, amdflang over the
iso_c_binding module) runs CUBE's main loop from checkpoint files and writes
checkpoints byte-identical to the Python host driving the same library."""
import os
import subprocess

import numpy as np
import pytest

from cube_nufluid_amd.params import BoxConfig, DATA_DIR
from cube_nufluid_amd.ic import zeldovich_ic
from cube_nufluid_amd.checkpoint import write_checkpoint, read_checkpoint

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "cube_nufluid_amd", "fortran", "cube_main")


def test_fortran_host_matches_python_host(tmp_path):
    if not os.path.exists(EXE):
        pytest.fail("Fortran host not built (make -C cube_nufluid_amd/fortran)")
    from cube_nufluid_amd.driver import Cube
    cfg = BoxConfig(nc=32, nnt=2, ncb=6, box=100.0, pp_on=True)
    zs = [10.0, 8.0, 6.0]
    st = zeldovich_ic(cfg, seed=1001, z=zs[0])
    fdir = tmp_path / "fortran"
    pdir = tmp_path / "python"
    write_checkpoint(st, str(fdir) + "/", zs[0])
    (tmp_path / "z.txt").write_text("".join("%8.4f\n" % z for z in zs))
    nml = tmp_path / "cube.nml"
    nml.write_text("&cube\n nc=32, nnt=2, ncb=6, izipx=2, izipv=2, Extended_pp_force=.true., pp_range=1,\n"
                   " omega_m=%r, omega_l=%r, box=100.0, opath='%s/', kernel_dir='%s/', zfile='%s', istep_max=200\n/\n"
                   % (float(np.float32(cfg.omega_m)), float(np.float32(cfg.omega_l)), fdir, DATA_DIR, tmp_path / "z.txt"))
    print("running", EXE, flush=True)
    r = subprocess.run([EXE, str(nml)], capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    cube = Cube(cfg, st, z_checkpoints=zs)
    cube.run(max_steps=200, opath=str(pdir) + "/")
    cube.close()
    for z in zs[1:]:
        a = read_checkpoint(str(fdir) + "/", z, cfg.nc)
        b = read_checkpoint(str(pdir) + "/", z, cfg.nc)
        for f in ("xp", "vp", "pid", "rhoc", "vfield"):
            assert np.array_equal(getattr(a, f), getattr(b, f)), (z, f)
        assert a.sim.nplocal == b.sim.nplocal and a.sim.a == pytest.approx(b.sim.a, rel=1e-6)
        # sim%timestep counts the steps taken so far (timestep.f90:10)
        assert a.sim.timestep == b.sim.timestep > 0, (z, a.sim.timestep, b.sim.timestep)
    h8 = read_checkpoint(str(pdir) + "/", zs[1], cfg.nc).sim.timestep
    h6 = read_checkpoint(str(pdir) + "/", zs[2], cfg.nc).sim.timestep
    assert 0 < h8 < h6


def test_fortran_host_halofind_matches_python_host(tmp_path):
    """HALOFIND through both hosts: z_halofind.txt as shipped (1.0, 0.0),
    whose first entry used (main.f90:43) is past the list and reads 0, so the
    finder runs at the final step; the halo_pid and hid files agree byte for
    byte and the halo records to FP32 rounding."""
    if not os.path.exists(EXE):
        pytest.fail("Fortran host not built (make -C cube_nufluid_amd/fortran)")
    from cube_nufluid_amd.driver import Cube
    from cube_nufluid_amd.halo import HaloParams, read_halo_file
    cfg = BoxConfig(nc=32, nnt=2, ncb=6, box=100.0, pp_on=True)
    zs = [10.0, 0.0]
    st = zeldovich_ic(cfg, seed=1001, z=zs[0])
    fdir = tmp_path / "fortran"
    pdir = tmp_path / "python"
    write_checkpoint(st, str(fdir) + "/", zs[0])
    (tmp_path / "z.txt").write_text("".join("%8.4f\n" % z for z in zs))
    (tmp_path / "zh.txt").write_text("%8.4f\n%8.4f\n" % (1.0, 0.0))
    nml = tmp_path / "cube.nml"
    nml.write_text("&cube\n nc=32, nnt=2, ncb=6, izipx=2, izipv=2, Extended_pp_force=.true., pp_range=1,\n"
                   " omega_m=%r, omega_l=%r, box=100.0, opath='%s/', kernel_dir='%s/', zfile='%s', istep_max=2000,\n"
                   " hfile='%s', min_halo_particles=40\n/\n"
                   % (float(np.float32(cfg.omega_m)), float(np.float32(cfg.omega_l)), fdir, DATA_DIR, tmp_path / "z.txt",
                      tmp_path / "zh.txt"))
    print("running", EXE, flush=True)
    r = subprocess.run([EXE, str(nml)], capture_output=True, text=True, timeout=300)
    print(r.stdout[-1500:], r.stderr[-1500:], flush=True)
    assert r.returncode == 0
    cube = Cube(cfg, st, z_checkpoints=zs, z_halofind=[1.0, 0.0])
    cube.halo_params = HaloParams(min_halo_particles=40)
    cube.run(max_steps=2000, opath=str(pdir) + "/")
    cube.close()
    assert len(cube.halo_catalogs) == 1 and cube.halo_catalogs[0][0] == 0.0
    for name in ("halo_pid", "hid"):
        a = np.fromfile(str(fdir / "image1" / ("0.000_%s_1.bin" % name)), dtype=np.uint8)
        b = np.fromfile(str(pdir / "image1" / ("0.000_%s_1.bin" % name)), dtype=np.uint8)
        assert np.array_equal(a, b), name
    fa = read_halo_file(str(fdir / "image1" / "0.000_halo_1.bin"))
    fb = read_halo_file(str(pdir / "image1" / "0.000_halo_1.bin"))
    print("halos:", fa[0], flush=True)
    assert fa[:3] == fb[:3] and fa[0] > 0
    np.testing.assert_allclose(fa[3], fb[3], rtol=1e-5, atol=1e-5)
