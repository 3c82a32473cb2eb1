"""The device velocity encoder's method, checked on the CPU.

The kernels encode v with venc_fast (csrc/cube_internal.hpp): y =
nv1*atan(x)/pi from a breakpoint-table atan (atan(j/32) plus a 4-term
series in h = (u - c)/(1 + u c)), rounded directly unless y lies within
1e-8 of a half-integer, where the exact host thresholds decide.  Here the
table is checked against a 40-digit atan, and the method, restated in
numpy float64, must give the oracle's code (update_particle.f90:119) for
every argument: around every code boundary and over a dense random sample.
A wrong table entry, series sign or coefficient, scale factor or margin
fails one of them.  The device's own arithmetic is checked by
test_vencode_equals_host_encoder_near_ties on the GPU."""
import os
import re

import numpy as np
import pytest

from oracle_lib import vencode_x
from vcodec_args import tie_arguments

HDR = os.path.join(os.path.dirname(__file__), "..", "cube_nufluid_amd", "csrc", "cube_internal.hpp")
PI_F = float(np.float32(np.pi))


def atan_table():
    with open(HDR) as f:
        src = f.read()
    m = re.search(r"kAtanTab\[33\]\s*=\s*\{([^}]*)\}", src)
    assert m, "kAtanTab not found"
    return np.array([float(v) for v in m.group(1).replace("\n", " ").split(",") if v.strip()])


def atan_enc(x, tab):
    a = np.abs(x)
    big = a > 1.0
    with np.errstate(divide="ignore", over="ignore"):
        u = np.where(big, 1.0 / np.minimum(a, 1e300), a)
    j = np.clip(np.floor(u * 32.0 + 0.5), 0, 32).astype(np.int64)
    c = j * 0.03125
    h = (u - c) / (u * c + 1.0)
    h2 = h * h
    t = (h2 * (-1.0 / 7.0) + 1.0 / 5.0) * h2 - 1.0 / 3.0
    r = tab[j] + (h * (t * h2) + h)
    r = np.where(big, 1.5707963267948966 - r, r)
    return np.copysign(r, x)


def venc_fast(x, izipv, tab):
    """codes, and where the thresholds decided (near-ties)"""
    nv1 = 255.0 if izipv == 1 else 65535.0
    y = atan_enc(x, tab) * (nv1 / PI_F)
    r = np.rint(y)
    near = ~(np.abs(y - r) < 0.5 - 1e-8)
    kmax = int(vencode_x(np.array([np.inf]), izipv)[0])
    code = np.clip(r, -kmax, kmax).astype(np.int32)
    code[near] = vencode_x(x[near], izipv)  # the threshold path is exact by construction
    return code, near


def test_atan_table_is_correctly_rounded():
    mp = pytest.importorskip("mpmath")
    mp.mp.dps = 40
    tab = atan_table()
    assert len(tab) == 33
    exact = [float(mp.atan(mp.mpf(j) / 32)) for j in range(33)]
    assert np.array_equal(tab, np.array(exact))


@pytest.mark.parametrize("izipv", [1, 2])
def test_method_gives_the_oracle_codes(izipv):
    tab = atan_table()
    rng = np.random.default_rng(11)
    xb, _, _ = tie_arguments(izipv)
    # dense sample over the code range (x = tan of a uniform angle), near zero
    # and far out
    xr = np.concatenate([np.tan(rng.uniform(-np.pi / 2, np.pi / 2, 2_000_000)),
                         rng.normal(0, 1e-3, 200_000), rng.normal(0, 1e6, 200_000)])
    for x in (xb, xr):
        code, near = venc_fast(x, izipv, tab)
        o = vencode_x(x, izipv)
        bad = np.flatnonzero(code != o)
        assert bad.size == 0, (bad.size, x[bad[:4]], code[bad[:4]], o[bad[:4]])
    # the direct rounding decides almost every random argument
    assert near.mean() < 1e-6
