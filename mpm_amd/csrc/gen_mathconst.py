"""Generates the constants of mpmx_math.h from their mathematical definitions.

* pi/2 split for Cody-Waite argument reduction: three parts P1 + P2 + P3 (P1,
  P2 with 33 significant bits, so k*P1 and k*P2 are exact for |k| < 2^20),
  and the double-double pi/2 = PIO2_HI + PIO2_LO, pi = PI_HI + PI_LO.
* Polynomial kernels, each the interpolant at the Chebyshev nodes of its
  interval (near-minimax), computed in 60-digit arithmetic and rounded to
  double:
    sin r = r + r z S(z),           z = r^2 in [0, (pi/4)^2], S of degree 5;
    cos r = 1 - z/2 + z^2 C(z),     z = r^2 in [0, (pi/4)^2], C of degree 5;
    asin y = y + y z A(z),          z = y^2 in [0, 1/4],      A of degree 12.
  Interpolation error (before rounding): S 2.0e-17, C 1.3e-18, A 1.5e-17,
  each scaled by z <= 0.62 in the result, i.e. below half an ulp.

Run: python gen_mathconst.py  (prints C++ constant definitions; needs mpmath)
"""
from fractions import Fraction

import mpmath as mp

mp.mp.dps = 60
PI = Fraction("3.141592653589793238462643383279502884197169399375105820974944")
PIO2 = PI / 2


def round_bits(fr, bits):
    """Round a positive Fraction to `bits` significant bits."""
    import math
    m, e = math.frexp(float(fr))
    scale = Fraction(2) ** (bits - e)
    return Fraction(round(fr * scale)) / scale


def cheb_interp(f, a, b, n):
    """Monomial coefficients of the degree-n interpolant of f at the n+1
    Chebyshev nodes of [a, b]."""
    nodes = [(a + b) / 2 + (b - a) / 2 * mp.cos(mp.pi * (2 * k + 1) / (2 * (n + 1)))
             for k in range(n + 1)]
    A = mp.matrix(n + 1, n + 1)
    y = mp.matrix(n + 1, 1)
    for i, x in enumerate(nodes):
        for j in range(n + 1):
            A[i, j] = x ** j
        y[i] = f(x)
    c = mp.lu_solve(A, y)
    return [float(c[j]) for j in range(n + 1)]


def S(z):
    r = mp.sqrt(z)
    return -mp.mpf(1) / 6 if z == 0 else (mp.sin(r) - r) / (r * z)


def C(z):
    r = mp.sqrt(z)
    return mp.mpf(1) / 24 if z == 0 else (mp.cos(r) - 1 + z / 2) / (z * z)


def A(z):
    y = mp.sqrt(z)
    return mp.mpf(1) / 6 if z == 0 else (mp.asin(y) - y) / (z * y)


def main():
    p1 = round_bits(PIO2, 33)
    p2 = round_bits(PIO2 - p1, 33)
    p3 = Fraction(float(PIO2 - p1 - p2))
    hi = Fraction(float(PIO2))
    lo = Fraction(float(PIO2 - hi))
    phi = Fraction(float(PI))
    plo = Fraction(float(PI - phi))
    out = []
    for name, v in (("kInvPio2", 1 / PIO2), ("kP1", p1), ("kP2", p2), ("kP3", p3),
                    ("kPio2Hi", hi), ("kPio2Lo", lo), ("kPiHi", phi), ("kPiLo", plo)):
        out.append("static constexpr double %s = %r;" % (name, float(v)))
    q = (mp.pi / 4) ** 2
    for name, f, b, n in (("kSin", S, q, 5), ("kCos", C, q, 5), ("kAsin", A, mp.mpf(1) / 4, 12)):
        c = cheb_interp(f, mp.mpf(0), b, n)
        out.append("static constexpr double %s[%d] = {%s};" % (name, n + 1,
                                                          ", ".join(repr(v) for v in c)))
    print("\n".join(out))


if __name__ == "__main__":
    main()
