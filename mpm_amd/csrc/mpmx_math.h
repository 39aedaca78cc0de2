// Deterministic FP64 elementary functions for the Mohr-Coulomb update.
//
// The reference calls the platform libm (sin, cos, tan, acos, pow) inside the
// return mapping (mohr_coulomb.tcc, material_utility.tcc). Platform libms are
// not correctly rounded and differ by an ulp between host and device, and the
// return mapping amplifies such differences over steps. These functions are
// built only from IEEE +, -, *, / and sqrt (all correctly rounded on x86-64
// and gfx950), so with -ffp-contract=off they give the same bits on the host
// and on the GPU. Accuracy is within 2 ulp of the exact value on the argument
// ranges the return mapping uses (|x| < 10 for sin/cos, [-1, 1] for acos).
//
// Methods: Cody & Waite argument reduction by pi/2 with a three-part split;
// Maclaurin series of sin/cos on [-pi/4, pi/4] (Horner, exact rational
// coefficients rounded to double); acos through the half-angle identities
// acos(v) = 2 asin(sqrt((1-v)/2)) for v > 1/2 and pi - 2 asin(sqrt((1+v)/2))
// for v < -1/2 (as used by fdlibm, Sun Microsystems 1993) with the Maclaurin
// series of asin on [-1/2, 1/2]. Constants: gen_mathconst.py.
#pragma once

#include <cmath>

#if defined(__HIPCC__)
#define MPMX_MATH_FN __host__ __device__ inline
#else
#define MPMX_MATH_FN inline
#endif

namespace mpmx {
namespace dmath {

static constexpr double kInvPio2 = 0.6366197723675814;
static constexpr double kP1 = 1.5707963267341256;
static constexpr double kP2 = 6.077100506303966e-11;
static constexpr double kP3 = 2.0222662487959506e-21;
static constexpr double kPio2Hi = 1.5707963267948966;
static constexpr double kPio2Lo = 6.123233995736766e-17;
static constexpr double kPiHi = 3.141592653589793;
static constexpr double kPiLo = 1.2246467991473532e-16;

// sin r = r + r^3 * sum_n kSin[n] r^(2n), kSin[n] = (-1)^(n+1)/(2n+3)!
MPMX_MATH_FN double sin_poly(double r, double z) {
  double p = 2.8114572543455206e-15;
  p = -7.647163731819816e-13 + z * p;
  p = 1.6059043836821613e-10 + z * p;
  p = -2.505210838544172e-08 + z * p;
  p = 2.7557319223985893e-06 + z * p;
  p = -0.0001984126984126984 + z * p;
  p = 0.008333333333333333 + z * p;
  p = -0.16666666666666666 + z * p;
  return r + (r * z) * p;
}

// cos r = 1 - z/2 + z^2 * sum_n kCos[n] z^n
MPMX_MATH_FN double cos_poly(double z) {
  double p = -1.5619206968586225e-16;
  p = 4.779477332387385e-14 + z * p;
  p = -1.1470745597729725e-11 + z * p;
  p = 2.08767569878681e-09 + z * p;
  p = -2.755731922398589e-07 + z * p;
  p = 2.48015873015873e-05 + z * p;
  p = -0.001388888888888889 + z * p;
  p = 0.041666666666666664 + z * p;
  return 1.0 - (0.5 * z - (z * z) * p);
}

// sin and cos of x (|x| < 2^19).
MPMX_MATH_FN void sincos(double x, double* s, double* c) {
  const double k = rint(x * kInvPio2);
  const double r = ((x - k * kP1) - k * kP2) - k * kP3;
  const double z = r * r;
  const double sr = sin_poly(r, z);
  const double cr = cos_poly(z);
  const int q = ((int)k) & 3;
  switch (q) {
    case 0: *s = sr; *c = cr; break;
    case 1: *s = cr; *c = -sr; break;
    case 2: *s = -sr; *c = -cr; break;
    default: *s = -cr; *c = sr; break;
  }
}

MPMX_MATH_FN double sin(double x) {
  double s, c;
  sincos(x, &s, &c);
  return s;
}
MPMX_MATH_FN double cos(double x) {
  double s, c;
  sincos(x, &s, &c);
  return c;
}
MPMX_MATH_FN double tan(double x) {
  double s, c;
  sincos(x, &s, &c);
  return s / c;
}

// asin y for |y| <= 1/2: y + y^3 * sum_n kAsin[n] y^(2n)
MPMX_MATH_FN double asin_small(double y) {
  const double z = y * y;
  double p = 0.0019650336162772837;
  p = 0.0020776610325181676 + z * p;
  p = 0.0022014739737101384 + z * p;
  p = 0.002338091892111975 + z * p;
  p = 0.0024894486782468836 + z * p;
  p = 0.00265787063820729 + z * p;
  p = 0.002846178401108942 + z * p;
  p = 0.0030578216492580306 + z * p;
  p = 0.003297059503473485 + z * p;
  p = 0.0035692053938259347 + z * p;
  p = 0.003880964558837669 + z * p;
  p = 0.004240907093679363 + z * p;
  p = 0.004660143486915096 + z * p;
  p = 0.005153309682319905 + z * p;
  p = 0.005740037670841924 + z * p;
  p = 0.006447210311889649 + z * p;
  p = 0.0073125258735988454 + z * p;
  p = 0.008390335809616815 + z * p;
  p = 0.009761609529194078 + z * p;
  p = 0.011551800896139705 + z * p;
  p = 0.01396484375 + z * p;
  p = 0.017352764423076924 + z * p;
  p = 0.022372159090909092 + z * p;
  p = 0.030381944444444444 + z * p;
  p = 0.044642857142857144 + z * p;
  p = 0.075 + z * p;
  p = 0.16666666666666666 + z * p;
  return y + (y * z) * p;
}

// acos v for v in [-1, 1] (NaN outside)
MPMX_MATH_FN double acos(double v) {
  if (v > 0.5) return 2.0 * asin_small(sqrt((1.0 - v) * 0.5));
  if (v < -0.5) return kPiHi - (2.0 * asin_small(sqrt((1.0 + v) * 0.5)) - kPiLo);
  return kPio2Hi - (asin_small(v) - kPio2Lo);
}

// x^(n/2) for the odd half-integer exponents the Mohr-Coulomb update uses
MPMX_MATH_FN double pow_1p5(double x) { return x * sqrt(x); }
MPMX_MATH_FN double pow_m1p5(double x) { return 1.0 / (x * sqrt(x)); }
MPMX_MATH_FN double pow_m2p5(double x) { return 1.0 / ((x * x) * sqrt(x)); }

}  // namespace dmath
}  // namespace mpmx
