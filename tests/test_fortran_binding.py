"""The Fortran binding module (cube_nufluid_amd/fortran/cube_hip_mod.f90)
on a host without a GPU: it binds every entry point include/cube_hip.h
declares, its bind(C) types have the C structs' sizes, and calls through it
reach the library with their arguments intact.  The host-only routines
(footprint, clock set-up, expansion) must give through Fortran what they
give through ctypes; every other entry point, handed a NULL handle, engine
or an out-of-range count, must come back with CUBE_ERR_ARG."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from cube_nufluid_amd import _lib, halo
from test_lib_symbols import declared_functions

ROOT = os.path.join(os.path.dirname(__file__), "..")
MOD = os.path.join(ROOT, "cube_nufluid_amd", "fortran", "cube_hip_mod.f90")
FC = "/opt/rocm/bin/amdflang"

PROGRAM = """program chk
  use cube_hip
  implicit none
  type(cube_params) :: p
  type(cube_clock) :: c
  type(cube_sim_header) :: sim
  type(cube_nu_params) :: nup
  type(cube_ic_params) :: ip
  type(cube_ens_params) :: ep
  type(cube_ens_result) :: er(1)
  type(cube_drift_out) :: dro
  type(cube_pp_out) :: ppo
  type(cube_pm_out) :: pmo
  type(cube_halo_params) :: hp
  type(cube_halo_catalog) :: cat
  type(c_ptr) :: h, e
  real(c_float) :: z(2), da1, da2, f(4), dx, cs2, w, dtn
  real(c_double) :: d(4)
  integer(c_int32_t) :: i32(4), nc
  integer(c_int64_t) :: i64(4), id, g1, g2
  integer(c_int) :: nq, nl, nout
  character(kind=c_char) :: buf(64)
  h = c_null_ptr
  e = c_null_ptr
  print '(A,12I8)', 'sizes', c_sizeof(p), c_sizeof(sim), c_sizeof(dro), c_sizeof(ppo), c_sizeof(pmo), &
    c_sizeof(hp), c_sizeof(cat), c_sizeof(nup), c_sizeof(ip), c_sizeof(c), c_sizeof(ep), c_sizeof(er(1))
  p%nc = 64; p%nnt = 2; p%ncb = 6; p%izipx = 2; p%izipv = 1; p%np_max = 300000
  p%pp_on = 1; p%pp_range = 1; p%rsoft = 0.1; p%omega_m = 0.31; p%omega_l = 0.69; p%wde = -1
  p%precision = 4; p%device = 0
  print '(A,2I20)', 'footprint', cube_footprint(p, 0_c_int32_t), cube_footprint(p, 2_c_int32_t)
  z = [49.0, 0.0]
  print '(A,I4,2I12,2I4)', 'clock', cube_clock_init(c, z, 2), transfer(c%a, 0), transfer(c%tau, 0), &
    c%cur_checkpoint, c%n_checkpoint
  print '(A,I4,2I12)', 'expansion', cube_expansion(0.1, 0.01, 0.31, 0.69, -1.0, da1, da2), &
    transfer(da1, 0), transfer(da2, 0)
  ep%device = 0; ep%max_live = 0; ep%mem_budget = 1; ep%step_mode = 0
  print '(A,31I3)', 'null', cube_drift(h, 0., 0., dro), cube_pp(h, 0., 0., ppo), cube_pm(h, 0., 0., pmo), &
    cube_upload_state(h, sim, c_null_ptr, c_null_ptr, i32, i32, f), &
    cube_download_state(h, sim, c_null_ptr, c_null_ptr, i32, i32, f), &
    cube_halofind(h, hp, cat), cube_ic_zeldovich(h, ip, sim), cube_nu_init(h, nup), cube_nu_ic(h, ip), &
    cube_nu_info(h, 0, nc, dx, cs2, w), cube_nu_set_fluid(h, 0, f), cube_nu_get_fluid(h, 0, f), &
    cube_fluid_evolve(h, 0., 1, 1., dtn), cube_clock_init(c, z, 0), cube_step(h, c), cube_set_step_mode(h, 0), &
    cube_step_stats(h, g1, g2), cube_power_spectrum(h, 1, d, d, d), cube_power_spectrum_noalloc(h, 1, d, d, d), &
    cube_probe_density(h, f), cube_probe_pm_forces(h, d, d), cube_probe_pp_forces(h, d), &
    cube_probe_vencode(h, d, 1_c_int64_t, i32), cube_profile_enable(h, 1), cube_profile_read(h, d, i64, 1), &
    cube_profile_slot_name(-1, buf, 64_c_size_t), cube_ens_create(ep, e), cube_ens_submit(e, p, f, f, ip, z, 2, id), &
    cube_ens_poll(e, er, 1, nout), cube_ens_pending(e, nq, nl), cube_hbm_copy_probe(0, 15_c_int64_t, 1, d)
  print '(A,2I3,F8.3)', 'free', cube_finalize(h), cube_ens_destroy(e), cube_ens_clock(e)
end program
"""


def bound_names():
    with open(MOD) as f:
        return set(re.findall(r'bind\(C,\s*name="(cube_\w+)"\)', f.read()))


def test_module_binds_every_declared_entry_point():
    assert bound_names() == set(declared_functions())


@pytest.fixture(scope="module")
def run(tmp_path_factory):
    if not os.path.exists(FC):
        pytest.skip("no Fortran compiler")
    libdir = os.path.dirname(_lib.LIB_PATH)
    d = tmp_path_factory.mktemp("fbind")
    (d / "chk.f90").write_text(PROGRAM)
    subprocess.run([FC, "-O0", "-c", MOD, "-o", "mod.o"], cwd=d, check=True, capture_output=True)
    subprocess.run([FC, "-O0", "chk.f90", "mod.o", "-L" + libdir, "-lcubehip", "-Wl,-rpath," + libdir, "-o", "chk"],
                   cwd=d, check=True, capture_output=True)
    out = subprocess.run([str(d / "chk")], capture_output=True, text=True, check=True, timeout=120).stdout
    return {ln.split()[0]: ln.split()[1:] for ln in out.splitlines() if ln.strip()}


def test_type_sizes(run):
    want = [_lib.CubeParams, _lib.SimHeader, _lib.DriftOut, _lib.PPOut, _lib.PMOut, halo.cube_halo_params,
            halo.cube_halo_catalog, _lib.NuParams, _lib.ICParams, _lib.CubeClock, _lib.EnsParams, _lib.EnsResult]
    assert [int(v) for v in run["sizes"]] == [ctypes.sizeof(t) for t in want]


def test_host_routines_agree_with_ctypes(run):
    lib = _lib.load()
    p = _lib.CubeParams()
    p.nc, p.nnt, p.ncb, p.izipx, p.izipv, p.np_max = 64, 2, 6, 2, 1, 300000
    p.pp_on, p.pp_range, p.rsoft, p.omega_m, p.omega_l, p.wde = 1, 1, 0.1, 0.31, 0.69, -1.0
    p.precision, p.device = 4, 0
    assert [int(v) for v in run["footprint"]] == [lib.cube_footprint(ctypes.byref(p), 0),
                                                  lib.cube_footprint(ctypes.byref(p), 2)]
    c = _lib.CubeClock()
    z = np.array([49.0, 0.0], np.float32)
    assert lib.cube_clock_init(ctypes.byref(c), z.ctypes.data_as(ctypes.c_void_p), 2) == 0
    bits = lambda x: int(np.float32(x).view(np.int32))
    assert [int(v) for v in run["clock"]] == [0, bits(c.a), bits(c.tau), c.cur_checkpoint, c.n_checkpoint]
    d1, d2 = _lib.expansion(0.1, 0.01, 0.31, 0.69, -1.0)
    assert [int(v) for v in run["expansion"]] == [0, bits(d1), bits(d2)]


def test_null_and_out_of_range_arguments(run):
    assert [int(v) for v in run["null"]] == [1] * 31
    # free()-like on NULL; the engine clock of no engine reads 0
    assert run["free"][:2] == ["0", "0"] and float(run["free"][2]) == 0.0
