"""Breakpoint table of the velocity encoder's FP64 atan (cube_internal.hpp,
kAtanTab): atan(j/32) for j = 0..32, correctly rounded to double (mpmath at
40 digits), printed as C double literals.  Run once; the output is pasted
into the header."""
import mpmath as mp


def table():
    mp.mp.dps = 40
    return [float(mp.atan(mp.mpf(j) / 32)) for j in range(33)]


if __name__ == "__main__":
    print(",\n".join("    %s" % repr(v) for v in table()))
