"""The C-ABI library loads on a host without a GPU and exports every entry
point declared in include/cube_hip.h (no compute calls here)."""
import ctypes
import os
import re

from cube_nufluid_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "cube_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(?:int|int64_t|double)\s+(cube_\w+)\s*\(", src)))


def test_header_declares_expected_entry_points():
    names = declared_functions()
    for n in ["cube_init", "cube_upload_state", "cube_download_state", "cube_drift", "cube_pp", "cube_pm",
              "cube_expansion", "cube_last_error", "cube_finalize"]:
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for n in declared_functions():
        assert hasattr(lib, n), n
    assert set(declared_functions()) == set(_lib.SIGNATURES)


def test_struct_layouts():
    assert ctypes.sizeof(_lib.SimHeader) == 224   # parameters.f90:129-153
    assert ctypes.sizeof(_lib.CubeParams) == 80
    assert ctypes.sizeof(_lib.DriftOut) == 56
    assert ctypes.sizeof(_lib.PPOut) == 32
    assert ctypes.sizeof(_lib.PMOut) == 56
    assert ctypes.sizeof(_lib.CubeClock) == 320
    assert ctypes.sizeof(_lib.EnsParams) == 24
    assert ctypes.sizeof(_lib.EnsResult) == 192
    from cube_nufluid_amd import halo
    assert ctypes.sizeof(halo.cube_halo_params) == 80
    assert ctypes.sizeof(halo.cube_halo_catalog) == 80
    assert ctypes.sizeof(_lib.NuParams) == 32
