"""Error behaviour of the C ABI on the host, without a GPU.

Every entry point that takes a handle, an ensemble or an output pointer
rejects NULL with CUBE_ERR_ARG before it touches the device, as the
reference's routines stop on bad input rather than run on.  The exceptions
follow C's free(): cube_finalize(NULL) and cube_ens_destroy(NULL) succeed
and do nothing.  The host-only routines (clock set-up, slot names, the last
error) are checked on real arguments too, and so is the ensemble engine's
life cycle when no box is admitted."""
import ctypes
import os
import re

import numpy as np
import pytest

from cube_nufluid_amd import _lib

CUBE_OK, CUBE_ERR_ARG = 0, 1
ROOT = os.path.join(os.path.dirname(__file__), "..")
# free()-like: NULL is a no-op; the rest return a sentinel instead of a code
NULL_IS_OK = {"cube_finalize", "cube_ens_destroy"}
SENTINEL = {"cube_footprint": -1, "cube_ens_clock": 0.0}
# needs a device to say anything (its first act is hipSetDevice)
DEVICE_ONLY = {"cube_hbm_copy_probe"}


def zero_args(argtypes):
    out = []
    for a in argtypes:
        if a in (ctypes.c_float, ctypes.c_double):
            out.append(0.0)
        elif a in (ctypes.c_int32, ctypes.c_int64, ctypes.c_size_t):
            out.append(0)
        else:
            out.append(None)
    return out


@pytest.fixture(scope="module")
def lib():
    return _lib.load()


@pytest.mark.parametrize("name", sorted(set(_lib.SIGNATURES) - DEVICE_ONLY))
def test_null_arguments_are_rejected(lib, name):
    res, args = _lib.SIGNATURES[name]
    rc = getattr(lib, name)(*zero_args(args))
    if name in NULL_IS_OK:
        assert rc == CUBE_OK
    elif name in SENTINEL:
        assert rc == SENTINEL[name]
    else:
        assert rc == CUBE_ERR_ARG


def test_expansion_needs_both_outputs(lib):
    d = ctypes.c_float()
    assert lib.cube_expansion(0.1, 0.01, 0.3, 0.7, -1.0, ctypes.byref(d), None) == CUBE_ERR_ARG
    assert lib.cube_expansion(0.1, 0.01, 0.3, 0.7, -1.0, None, ctypes.byref(d)) == CUBE_ERR_ARG


def test_last_error_without_handle(lib):
    buf = ctypes.create_string_buffer(b"x" * 16)
    assert lib.cube_last_error(None, buf, 16) == CUBE_OK
    assert buf.value == b""
    assert lib.cube_last_error(None, buf, 0) == CUBE_ERR_ARG


def test_slot_names_match_the_source_and_reject_out_of_range(lib):
    with open(os.path.join(ROOT, "cube_nufluid_amd", "csrc", "cube_api.hip")) as f:
        src = f.read()
    names = re.findall(r'"([a-z0-9_]+)"', re.search(r"kSlotNames\[PS_NSLOT\]\s*=\s*\{([^}]*)\}", src).group(1))
    buf = ctypes.create_string_buffer(64)
    for i, n in enumerate(names):
        assert lib.cube_profile_slot_name(i, buf, 64) == CUBE_OK
        assert buf.value.decode() == n
    for bad in (-1, len(names)):
        assert lib.cube_profile_slot_name(bad, buf, 64) == CUBE_ERR_ARG
    assert lib.cube_profile_slot_name(0, buf, 0) == CUBE_ERR_ARG
    # truncates into a short buffer, always NUL-terminated
    short = ctypes.create_string_buffer(b"\xff" * 4)
    assert lib.cube_profile_slot_name(0, short, 4) == CUBE_OK
    assert short.raw[3] == 0 and short.value.decode() == names[0][:3]


def test_clock_init(lib):
    """initialize.f90:40-53 and :73, particle_initialization.f90:37: the run
    starts at a = 1/(1+z_1), tau = -3/sqrt(a) in default real (float32), with
    t, dt, da zero and the next checkpoint the second entry; dt_neu starts at
    dt_max = 1.  The checkpoint count runs from 1 (the reference stops on an
    empty z_checkpoint.txt, initialize.f90:61) to the clock's capacity."""
    c = _lib.CubeClock()
    cap = len(c.z_checkpoint)
    z = np.array([49.0, 10.0, 3.0, 0.0], np.float32)
    zp = z.ctypes.data_as(ctypes.c_void_p)
    assert lib.cube_clock_init(ctypes.byref(c), zp, 0) == CUBE_ERR_ARG
    big = np.zeros(cap + 1, np.float32)
    assert lib.cube_clock_init(ctypes.byref(c), big.ctypes.data_as(ctypes.c_void_p), cap + 1) == CUBE_ERR_ARG
    assert lib.cube_clock_init(ctypes.byref(c), None, 4) == CUBE_ERR_ARG
    assert lib.cube_clock_init(ctypes.byref(c), zp, 4) == CUBE_OK
    one = np.float32(1)
    a = one / (one + z[0])
    assert np.float32(c.a) == a
    assert np.float32(c.tau) == np.float32(-3) / np.sqrt(one / (one + z[0]))
    assert (c.t, c.dt, c.dt_old, c.da, c.istep) == (0, 0, 0, 0, 0)
    assert c.n_checkpoint == 4 and c.cur_checkpoint == 2
    assert list(np.float32(c.z_checkpoint[:4])) == list(z)
    assert c.dt_neu == 1.0
    assert (c.checkpoint_step, c.final_step) == (0, 0)
    zfull = np.linspace(20, 0, cap).astype(np.float32)
    assert lib.cube_clock_init(ctypes.byref(c), zfull.ctypes.data_as(ctypes.c_void_p), cap) == CUBE_OK
    assert c.n_checkpoint == cap and np.float32(c.z_checkpoint[cap - 1]) == zfull[-1]


def test_ensemble_lifecycle_without_boxes(lib):
    """An engine given its own memory budget starts and stops on the host:
    a request whose footprint exceeds the budget is refused with
    CUBE_ERR_CAPACITY before it is queued, poll rejects a negative count,
    and destroy returns with nothing left to run."""
    from cube_nufluid_amd.params import read_kernel_tables
    ep = _lib.EnsParams()
    ep.device, ep.max_live, ep.mem_budget, ep.step_mode = 0, 2, 1 << 20, 0
    e = ctypes.c_void_p()
    assert lib.cube_ens_create(ctypes.byref(ep), ctypes.byref(e)) == CUBE_OK and e.value
    try:
        p = _lib.CubeParams()
        p.nc, p.nnt, p.ncb, p.izipx, p.izipv, p.np_max = 32, 2, 6, 2, 2, 64 ** 3
        p.omega_m, p.omega_l, p.wde, p.precision = 0.31, 0.69, -1.0, 4
        assert lib.cube_footprint(ctypes.byref(p), 2) > ep.mem_budget
        kf, kc = read_kernel_tables()
        lnk = np.log(np.array([1e-3, 1e2]))
        lnp = np.zeros(2)
        ip = _lib.ICParams()
        ip.seed, ip.np_nc, ip.nk, ip.a, ip.box = 1, 2, 2, 0.02, 100.0
        ip.lnk, ip.lnp = _lib.ptr(lnk), _lib.ptr(lnp)
        z = np.array([49.0, 0.0], np.float32)
        rid = ctypes.c_int64(-7)
        assert lib.cube_ens_submit(e, ctypes.byref(p), _lib.ptr(kf), _lib.ptr(kc), ctypes.byref(ip), _lib.ptr(z), 2,
                                   ctypes.byref(rid)) == 2  # CUBE_ERR_CAPACITY
        assert rid.value == -7
        n = ctypes.c_int32(5)
        assert lib.cube_ens_poll(e, None, -1, ctypes.byref(n)) == CUBE_ERR_ARG
        assert lib.cube_ens_poll(e, None, 0, ctypes.byref(n)) == CUBE_OK and n.value == 0
        q, l = ctypes.c_int32(-1), ctypes.c_int32(-1)
        assert lib.cube_ens_pending(e, ctypes.byref(q), ctypes.byref(l)) == CUBE_OK and (q.value, l.value) == (0, 0)
        assert lib.cube_ens_clock(e) >= 0.0
    finally:
        assert lib.cube_ens_destroy(e) == CUBE_OK
