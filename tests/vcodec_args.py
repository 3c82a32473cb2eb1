"""Velocity-codec test arguments: the atan arguments around every code
boundary of the oracle's encoder (update_particle.f90:119), shared by the
CPU and GPU encoder tests."""
import numpy as np


def _key(x):
    u = np.ascontiguousarray(x, dtype=np.float64).view(np.uint64)
    return np.where(u >> np.uint64(63), ~u, u | np.uint64(1 << 63))


def _unkey(k):
    u = np.where(k >> np.uint64(63), k & np.uint64((1 << 63) - 1), ~k)
    return u.view(np.float64)


def tie_arguments(izipv, ulps=24):
    """atan arguments +-ulps around every code boundary of the oracle's
    encoder (the smallest double with code >= k, for every k above -kmax,
    found by bisection over the ordered bit patterns), plus log-spaced
    values over the whole range and the special values."""
    from oracle_lib import vencode_x
    nv1 = float(np.float32(2 ** (8 * izipv) - 1))
    pi4 = float(np.float32(np.pi))
    kmax = int(vencode_x(np.array([np.inf]), izipv)[0])
    k = np.arange(-kmax + 1, kmax + 1)
    x0 = np.tan((k - 0.5) * pi4 / nv1)
    span = np.abs(x0) * 1e-9 + 1e-300
    lo, hi = x0 - span, x0 + span
    for _ in range(40):
        bl, bh = vencode_x(lo, izipv) >= k, vencode_x(hi, izipv) < k
        if not (bl.any() or bh.any()):
            break
        span = np.where(bl | bh, span * 16, span)
        lo, hi = x0 - span, x0 + span
    a, b = _key(lo), _key(hi)  # code(a) < k <= code(b)
    while True:
        gap = b - a
        if not (gap > 1).any():
            break
        m = a + gap // np.uint64(2)
        up = vencode_x(_unkey(m), izipv) >= k
        b = np.where((gap > 1) & up, m, b)
        a = np.where((gap > 1) & ~up, m, a)
    steps = np.arange(-ulps, ulps + 1, dtype=np.int64).astype(np.uint64)
    near = _unkey((b[None, :] + steps[:, None]).ravel())
    mags = np.logspace(-300, 300, 20001)
    special = np.array([0.0, -0.0, np.inf, -np.inf, 1e308, -1e308, 5e-324, -5e-324])
    return np.concatenate([near, mags, -mags, special]), _unkey(b), k
