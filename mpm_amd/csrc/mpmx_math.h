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
// on [-pi/4, pi/4], sin r = r + r z S(z) and cos r = 1 - z/2 + z^2 C(z)
// (z = r^2); acos through the half-angle identities acos(v) =
// 2 asin(sqrt((1-v)/2)) for v > 1/2 and pi - 2 asin(sqrt((1+v)/2)) for
// v < -1/2 (as used by fdlibm, Sun Microsystems 1993), with asin y =
// y + y z A(z) on [-1/2, 1/2]. S, C and A are Chebyshev interpolants
// (degrees 5, 5, 12), evaluated by Estrin's scheme (short dependency chains
// for the GPU's in-order lanes). Constants: gen_mathconst.py.
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
static constexpr double kSin[6] = {-0.16666666666666666, 0.008333333333330948, -0.00019841269836758574, 2.755731610255244e-06, -2.5051131845003624e-08, 1.5918129294866608e-10};
static constexpr double kCos[6] = {0.041666666666666664, -0.0013888888888887398, 2.480158729876569e-05, -2.7557317271729793e-07, 2.08761462684032e-09, -1.1382632425521717e-11};
static constexpr double kAsin[13] = {0.16666666666666669, 0.07499999999998433, 0.04464285714635543, 0.030381944138531247, 0.02237217294214989, 0.017352392720869973, 0.013971212973552933, 0.011479177415184906, 0.01032281435018578, 0.005457506718640358, 0.01740087944269402, -0.014851887071247204, 0.028757851367421566};

// c[0] + c[1] z + ... + c[5] z^5 by Estrin's scheme
MPMX_MATH_FN double estrin6(const double* c, double z) {
  const double z2 = z * z, z4 = z2 * z2;
  return ((c[0] + c[1] * z) + (c[2] + c[3] * z) * z2) + (c[4] + c[5] * z) * z4;
}

// sin r = r + r z S(z), z = r^2
MPMX_MATH_FN double sin_poly(double r, double z) { return r + (r * z) * estrin6(kSin, z); }

// cos r = 1 - z/2 + z^2 C(z), z = r^2
MPMX_MATH_FN double cos_poly(double z) { return 1.0 - (0.5 * z - (z * z) * estrin6(kCos, z)); }

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

// asin y = y + y z A(z) for |y| <= 1/2, z = y^2 (A by Estrin's scheme)
MPMX_MATH_FN double asin_small(double y) {
  const double* a = kAsin;
  const double z = y * y, z2 = z * z, z4 = z2 * z2, z8 = z4 * z4;
  const double q0 = (a[0] + a[1] * z) + (a[2] + a[3] * z) * z2;
  const double q1 = (a[4] + a[5] * z) + (a[6] + a[7] * z) * z2;
  const double q2 = (a[8] + a[9] * z) + (a[10] + a[11] * z) * z2;
  const double P = (q0 + q1 * z4) + (q2 + a[12] * z4) * z8;
  return y + (y * z) * P;
}

// acos v for v in [-1, 1] (NaN outside):
//   v >  1/2: 2 asin(sqrt((1-v)/2)),
//   v < -1/2: pi - 2 asin(sqrt((1+v)/2)),
//   else    : pi/2 - asin(v),
// with one asin evaluation (branch-free for the GPU's lanes).
MPMX_MATH_FN double acos(double v) {
  const bool hi = v > 0.5, lo = v < -0.5;
  const double y = hi ? sqrt((1.0 - v) * 0.5) : (lo ? sqrt((1.0 + v) * 0.5) : v);
  const double a = asin_small(y);
  if (hi) return 2.0 * a;
  if (lo) return kPiHi - (2.0 * a - kPiLo);
  return kPio2Hi - (a - kPio2Lo);
}

// x^(n/2) for the odd half-integer exponents the Mohr-Coulomb update uses
MPMX_MATH_FN double pow_1p5(double x) { return x * sqrt(x); }
MPMX_MATH_FN double pow_m1p5(double x) { return 1.0 / (x * sqrt(x)); }
MPMX_MATH_FN double pow_m2p5(double x) { return 1.0 / ((x * x) * sqrt(x)); }

}  // namespace dmath
}  // namespace mpmx
