"""Generates the constants of mpmx_math.h from their mathematical definitions.

* pi/2 split for Cody-Waite argument reduction: three 33-bit-or-less parts
  P1 + P2 + P3 (P1, P2 have 33 significant bits so k*P1, k*P2 are exact for
  |k| < 2^20), and the double-double pi/2 = PIO2_HI + PIO2_LO, pi = PI_HI + PI_LO.
* Maclaurin coefficients of sin (-1)^n/(2n+1)!, cos (-1)^n/(2n)! and
  asin (2n)! / (4^n (n!)^2 (2n+1)), each rounded to the nearest double.

Run: python gen_mathconst.py  (prints C++ constant definitions)
"""
from fractions import Fraction
from math import factorial

# pi to 60 decimal digits
PI = Fraction("3.141592653589793238462643383279502884197169399375105820974944")
PIO2 = PI / 2


def to_double(fr):
    return float(fr)


def round_bits(fr, bits):
    """Round a positive Fraction to `bits` significant bits."""
    x = float(fr)
    import math
    m, e = math.frexp(x)  # x = m * 2^e, 0.5 <= m < 1
    scale = 2 ** (bits - e)
    return Fraction(round(fr * scale), scale)


def main():
    p1 = round_bits(PIO2, 33)
    p2 = round_bits(PIO2 - p1, 33)
    p3 = Fraction(to_double(PIO2 - p1 - p2))
    hi = Fraction(float(PIO2))
    lo = Fraction(float(PIO2 - hi))
    phi = Fraction(float(PI))
    plo = Fraction(float(PI - phi))
    out = []
    out.append("static constexpr double kInvPio2 = %r;" % float(1 / PIO2))
    out.append("static constexpr double kP1 = %r;" % float(p1))
    out.append("static constexpr double kP2 = %r;" % float(p2))
    out.append("static constexpr double kP3 = %r;" % float(p3))
    out.append("static constexpr double kPio2Hi = %r;" % float(hi))
    out.append("static constexpr double kPio2Lo = %r;" % float(lo))
    out.append("static constexpr double kPiHi = %r;" % float(phi))
    out.append("static constexpr double kPiLo = %r;" % float(plo))
    s = [float(Fraction((-1) ** n, factorial(2 * n + 1))) for n in range(1, 9)]
    c = [float(Fraction((-1) ** n, factorial(2 * n))) for n in range(2, 10)]
    a = [float(Fraction(factorial(2 * n), 4 ** n * factorial(n) ** 2 * (2 * n + 1)))
         for n in range(1, 28)]
    out.append("static constexpr double kSin[8] = {%s};" % ", ".join(repr(v) for v in s))
    out.append("static constexpr double kCos[8] = {%s};" % ", ".join(repr(v) for v in c))
    out.append("static constexpr double kAsin[27] = {%s};" % ", ".join(repr(v) for v in a))
    print("\n".join(out))


if __name__ == "__main__":
    main()
