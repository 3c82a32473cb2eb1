"""Independent pins of oracle functions that the oracle-vs-GPU tests cannot
check (a mistake made the same way on both sides would pass those).

Each test names the mistake it is built to catch.  CPU only.
"""
import numpy as np
import pytest

import oracle_lib as ol
from cube_nufluid_amd import _lib

F32, F64 = np.float32, np.float64


# ------------------------------------------------------------ expansion
def expansion_numpy(a0, dt0, om, ol_, wde):
    """timestep.f90:136-180 evaluated with numpy scalars of the precision
    each Fortran operation has: real(4) arguments and default-real literals
    and parameters, real(8) locals.  `x**n` with an integer n on a real(8)
    is a product chain, `a_x**(-3*wde)` is pow(), and `a0+da1` is a real(4)
    sum."""
    a0, dt0, om, ol_, wde = F32(a0), F32(dt0), F32(om), F32(ol_), F32(wde)
    dt_x = F32(dt0 / F32(2))

    def half(a_x):
        omHsq = F64(F32(4.0) / F32(9.0))
        a3rlm = F64(np.power(a_x, F64(-(F32(3) * wde)))) * F64(ol_) / F64(om)
        arkm = a_x * F64(F32(F32(1.0) - om) - ol_) / F64(om)
        adot = np.sqrt(omHsq * (a_x * a_x * a_x) * (F64(1.0) + arkm + a3rlm))
        addot = (a_x * a_x) * omHsq * (F64(1.5) + F64(2.0) * arkm + F64(F32(1.5) * (F32(1.0) - wde)) * a3rlm)
        c3 = F32(F32(F32(1.5) * (F32(2.0) - F32(3.0) * wde)) * (F32(1.0) - wde))
        atdot = a_x * adot * omHsq * (F64(3.0) + F64(6.0) * arkm + F64(c3) * a3rlm)
        p2 = F32(dt_x * dt_x)
        p3 = F32(F32(dt_x * dt_x) * dt_x)
        return F32(adot * F64(dt_x) + (addot * F64(p2)) / F64(2.0) + (atdot * F64(p3)) / F64(6.0))

    da1 = half(F64(a0))
    da2 = half(F64(F32(a0 + da1)))
    return da1, da2


def expansion_all_double(a0, dt0, om, ol_, wde):
    """The same Taylor step with every operation in double (the mistake:
    4/9, 1-om-ol, dt powers and a0+da1 not rounded to real(4))."""
    dt_x = float(F32(dt0) / F32(2))
    a_x = float(F32(a0))
    out = []
    for _ in range(2):
        a3 = a_x ** (-3 * wde) * ol_ / om
        ark = a_x * (1 - om - ol_) / om
        adot = np.sqrt(4 / 9 * a_x ** 3 * (1 + ark + a3))
        addot = a_x ** 2 * 4 / 9 * (1.5 + 2 * ark + 1.5 * (1 - wde) * a3)
        atdot = a_x * adot * 4 / 9 * (3 + 6 * ark + 1.5 * (2 - 3 * wde) * (1 - wde) * a3)
        out.append(F32(adot * dt_x + addot * dt_x ** 2 / 2 + atdot * dt_x ** 3 / 6))
        a_x = float(F32(a0)) + float(out[0])
    return tuple(out)


def _cases():
    rng = np.random.default_rng(11)
    a = np.exp(rng.uniform(np.log(0.02), 0.0, 300)).astype(np.float32)
    dt = np.exp(rng.uniform(np.log(1e-4), np.log(0.3), 300)).astype(np.float32)
    return list(zip(a.tolist(), dt.tolist()))


def test_expansion_bit_exact_real4_semantics():
    """Oracle and library expansion equal a numpy restatement with the
    reference's real(4)/real(8) mixing bit for bit, over 300 (a, dt) pairs.
    Catches: omHsq in double, 1-omega_m-omega_l or dt_x**n in double,
    a_x = a0+da1 without the real(4) rounding."""
    om, ol_ = 0.31, 1.0 - 0.31 - 0.31 / (1 + 1e9)
    differs = 0
    for a, dt in _cases():
        want = expansion_numpy(a, dt, om, ol_, -1.0)
        got_o = ol.expansion(a, dt, om, ol_, -1.0)
        got_l = _lib.expansion(a, dt, om, ol_, -1.0)
        assert got_o == want, (a, dt, got_o, want)
        assert got_l == want, (a, dt, got_l, want)
        if expansion_all_double(a, dt, om, ol_, -1.0) != want:
            differs += 1
    # the all-double mistake changes the result on a share of the steps,
    # so the comparison above can see it
    assert differs > 10


def test_expansion_general_wde_and_curvature():
    """wde != -1 and a curved universe exercise the real(4) wde factors and
    1-omega_m-omega_l; restatement and library agree bit for bit."""
    for om, ol_, wde in ((0.3, 0.6, -0.9), (0.25, 0.8, -1.1), (1.0, 0.0, -1.0)):
        for a, dt in _cases()[:60]:
            assert _lib.expansion(a, dt, om, ol_, wde) == expansion_numpy(a, dt, om, ol_, wde)
            assert ol.expansion(a, dt, om, ol_, wde) == expansion_numpy(a, dt, om, ol_, wde)
