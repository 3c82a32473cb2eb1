"""Coefficients of the velocity encoder's FP64 atan (cube_internal.hpp,
atan_enc): atan(u) = u * P(u^2) on [0, 1], P the 18-term Chebyshev fit of
atan(sqrt(w))/sqrt(w) on w in [0, 1] (mpmath.chebyfit, 40 digits), printed
highest power first as C double literals.  Run once; the output is pasted
into the header."""
import mpmath as mp


def coefficients(n=18):
    mp.mp.dps = 40
    g = lambda w: mp.atan(mp.sqrt(w)) / mp.sqrt(w) if w != 0 else mp.mpf(1)
    c, err = mp.chebyfit(g, [0, 1], n, error=True)
    return c, err


if __name__ == "__main__":
    c, err = coefficients()
    print("// max abs error of the fit on [0, 1]: %s" % mp.nstr(err, 3))
    print(",\n".join("    %s" % repr(float(x)) for x in c))
