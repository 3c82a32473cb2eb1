"""cube_power_spectrum_noalloc (partial sums in the library's static device
array, no device allocation) against cube_power_spectrum (per-call
hipMalloc) through the C ABI: two calls in a row, the second on a smaller
mesh, so bins the first call filled must not leak into the second.

Tolerance: mode counts exactly equal; the sums of |k| and of Delta^2 are
double atomic sums whose order varies between launches, so within 1e-12
relative."""
import ctypes

import numpy as np
import pytest

from cube_nufluid_amd import _lib


def _pk_pair(cube):
    k0, d0, m0 = cube.power_spectrum(noalloc=False)
    k1, d1, m1 = cube.power_spectrum(noalloc=True)
    assert np.array_equal(m0, m1)
    ok = m0 > 0
    assert ok.sum() > 10
    np.testing.assert_allclose(k1[ok], k0[ok], rtol=1e-12)
    np.testing.assert_allclose(d1[ok], d0[ok], rtol=1e-12)
    return m1


@pytest.mark.gpu
def test_noalloc_power_spectrum_matches_allocating_entry_point():
    from cube_nufluid_amd.params import BoxConfig
    from cube_nufluid_amd.ic import zeldovich_ic
    from cube_nufluid_amd.driver import Cube
    sizes = []
    for nc, seed in ((32, 71), (16, 72)):   # second call: smaller mesh, fewer bins
        cfg = BoxConfig(nc=nc, nnt=2, ncb=6, box=100.0 * nc / 32)
        print("noalloc P(k): nc=%d seed=%d" % (nc, seed), flush=True)
        cube = Cube(cfg, zeldovich_ic(cfg, seed=seed, z=0.0))
        try:
            m = _pk_pair(cube)
            sizes.append(len(m))
            # more bins than the static scratch holds: rejected before any GPU work
            buf = np.zeros(4096)
            rc = cube.lib.cube_power_spectrum_noalloc(cube.h, 4096, _lib.ptr(buf), _lib.ptr(buf), _lib.ptr(buf))
            assert rc == 1
            err = ctypes.create_string_buffer(256)
            cube.lib.cube_last_error(cube.h, err, 256)
            assert b"2048" in err.value
        finally:
            cube.close()
    assert sizes[1] < sizes[0]
