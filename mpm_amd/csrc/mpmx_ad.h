// Forward-mode derivatives of the constitutive updates, for the adjoint of
// the explicit step.
//
// Dual<K> carries a value and K tangents; every operation propagates the
// tangents by the chain rule and computes the value with exactly the
// operation the FP64 path uses (same operands, same order, the same
// mpmx_math.h elementary functions), so the value part of a Dual evaluation
// is bit-for-bit the plain update and takes the same branches.
//
// mc_update<T> is the Mohr-Coulomb return mapping of the reference
// (mohr_coulomb.tcc:348-513 with compute_stress_invariants :124-134,
// compute_yield_state :139-191, compute_df_dp :195-344 and the
// material_utility.tcc invariants) without strain softening, written once for
// T = double and T = Dual<K>. Its operation order is that of mc_stress in
// mpmx_device.h, which the parity tests hold bitwise equal to the oracle.
// The adjoint seeds the tangents with the inputs of one update (stress,
// strain increment, Young's modulus through the elastic constants, friction
// angle) and contracts the output tangents with the stress adjoint.
#pragma once

#include <cmath>

#include "mpmx_math.h"

#if defined(__HIPCC__)
#define MPMX_AD __host__ __device__ __forceinline__
#else
#define MPMX_AD inline
#endif

namespace mpmx {
namespace ad {

template <int K>
struct Dual {
  double v;
  double d[K];
};

template <class T> struct Traits;
template <> struct Traits<double> {
  static MPMX_AD double val(double x) { return x; }
  static MPMX_AD double make(double x) { return x; }
};
template <int K> struct Traits<Dual<K>> {
  static MPMX_AD double val(const Dual<K>& x) { return x.v; }
  static MPMX_AD Dual<K> make(double x) {
    Dual<K> r;
    r.v = x;
#pragma unroll
    for (int k = 0; k < K; ++k) r.d[k] = 0.;
    return r;
  }
};
template <class T> MPMX_AD double val(const T& x) { return Traits<T>::val(x); }
template <class T> MPMX_AD T cst(double x) { return Traits<T>::make(x); }

// ------------------------------ arithmetic ----------------------------------
template <int K> MPMX_AD Dual<K> operator+(const Dual<K>& a, const Dual<K>& b) {
  Dual<K> r;
  r.v = a.v + b.v;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] + b.d[k];
  return r;
}
template <int K> MPMX_AD Dual<K> operator+(const Dual<K>& a, double b) {
  Dual<K> r = a;
  r.v = a.v + b;
  return r;
}
template <int K> MPMX_AD Dual<K> operator+(double a, const Dual<K>& b) {
  Dual<K> r = b;
  r.v = a + b.v;
  return r;
}
template <int K> MPMX_AD Dual<K> operator-(const Dual<K>& a, const Dual<K>& b) {
  Dual<K> r;
  r.v = a.v - b.v;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] - b.d[k];
  return r;
}
template <int K> MPMX_AD Dual<K> operator-(const Dual<K>& a, double b) {
  Dual<K> r = a;
  r.v = a.v - b;
  return r;
}
template <int K> MPMX_AD Dual<K> operator-(double a, const Dual<K>& b) {
  Dual<K> r;
  r.v = a - b.v;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = -b.d[k];
  return r;
}
template <int K> MPMX_AD Dual<K> operator-(const Dual<K>& a) {
  Dual<K> r;
  r.v = -a.v;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = -a.d[k];
  return r;
}
template <int K> MPMX_AD Dual<K> operator*(const Dual<K>& a, const Dual<K>& b) {
  Dual<K> r;
  r.v = a.v * b.v;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
  return r;
}
template <int K> MPMX_AD Dual<K> operator*(const Dual<K>& a, double b) {
  Dual<K> r;
  r.v = a.v * b;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] * b;
  return r;
}
template <int K> MPMX_AD Dual<K> operator*(double a, const Dual<K>& b) {
  Dual<K> r;
  r.v = a * b.v;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = a * b.d[k];
  return r;
}
template <int K> MPMX_AD Dual<K> operator/(const Dual<K>& a, const Dual<K>& b) {
  Dual<K> r;
  r.v = a.v / b.v;
  const double ib = 1.0 / b.v;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) * ib;
  return r;
}
template <int K> MPMX_AD Dual<K> operator/(const Dual<K>& a, double b) {
  Dual<K> r;
  r.v = a.v / b;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] / b;
  return r;
}
template <int K> MPMX_AD Dual<K> operator/(double a, const Dual<K>& b) {
  Dual<K> r;
  r.v = a / b.v;
  const double f = -r.v / b.v;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = f * b.d[k];
  return r;
}
template <int K, class U> MPMX_AD Dual<K>& operator+=(Dual<K>& a, const U& b) {
  a = a + b;
  return a;
}
template <int K, class U> MPMX_AD Dual<K>& operator-=(Dual<K>& a, const U& b) {
  a = a - b;
  return a;
}
template <int K, class U> MPMX_AD Dual<K>& operator*=(Dual<K>& a, const U& b) {
  a = a * b;
  return a;
}

// --------------------------- elementary functions ---------------------------
MPMX_AD double t_sqrt(double x) { return sqrt(x); }
template <int K> MPMX_AD Dual<K> t_sqrt(const Dual<K>& a) {
  Dual<K> r;
  r.v = sqrt(a.v);
  const double f = r.v > 0. ? 0.5 / r.v : 0.;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = f * a.d[k];
  return r;
}
MPMX_AD double t_fabs(double x) { return fabs(x); }
template <int K> MPMX_AD Dual<K> t_fabs(const Dual<K>& a) { return a.v < 0. ? -a : a; }

MPMX_AD void t_sincos(double x, double* s, double* c) { dmath::sincos(x, s, c); }
template <int K> MPMX_AD void t_sincos(const Dual<K>& a, Dual<K>* s, Dual<K>* c) {
  double sv, cv;
  dmath::sincos(a.v, &sv, &cv);
  s->v = sv;
  c->v = cv;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    s->d[k] = cv * a.d[k];
    c->d[k] = -sv * a.d[k];
  }
}
template <class T> MPMX_AD T t_cos(const T& x) {
  T s, c;
  t_sincos(x, &s, &c);
  return c;
}
MPMX_AD double t_acos(double x) { return dmath::acos(x); }
template <int K> MPMX_AD Dual<K> t_acos(const Dual<K>& a) {
  Dual<K> r;
  r.v = dmath::acos(a.v);
  const double q = 1.0 - a.v * a.v;
  const double f = q > 0. ? -1.0 / sqrt(q) : 0.;
#pragma unroll
  for (int k = 0; k < K; ++k) r.d[k] = f * a.d[k];
  return r;
}
// x^(n/2) as in mpmx_math.h (same expressions, so the values agree)
template <class T> MPMX_AD T t_pow_1p5(const T& x) { return x * t_sqrt(x); }
template <class T> MPMX_AD T t_pow_m1p5(const T& x) { return 1.0 / (x * t_sqrt(x)); }
template <class T> MPMX_AD T t_pow_m2p5(const T& x) { return 1.0 / ((x * x) * t_sqrt(x)); }
template <class T> MPMX_AD T t_sq(const T& x) { return x * x; }

// ------------------------- material_utility.tcc -----------------------------
static constexpr double kEps = 2.220446049250313e-16;  // DBL_EPSILON

template <class T> MPMX_AD T mu_p(const T* s) { return (1. / 3. * ((s[0] + s[1]) + s[2])); }
template <class T> MPMX_AD T mu_j2(const T* s) {
  return (((t_sq(s[0] - s[1]) + t_sq(s[1] - s[2])) + t_sq(s[0] - s[2])) / 6.0 + t_sq(s[3])) +
             t_sq(s[4]) + t_sq(s[5]);
}
template <class T> MPMX_AD void mu_dev(const T* s, T* d) {
  const T p = mu_p(s);
  for (int i = 0; i < 6; ++i) d[i] = s[i];
  for (int i = 0; i < 3; ++i) d[i] -= p;
}
template <class T> MPMX_AD T mu_j3(const T* s) {
  T d[6];
  mu_dev(s, d);
  return ((d[0] * d[1]) * d[2]) - (d[2] * t_sq(d[3])) +
         ((((2 * d[3]) * d[4]) * d[5] - (d[0] * t_sq(d[4]))) - (d[1] * t_sq(d[5])));
}
template <class T> MPMX_AD T mu_q(const T* s) { return t_sqrt(3 * mu_j2(s)); }
template <class T> MPMX_AD T mu_lode(const T* s) {
  const T j2 = mu_j2(s);
  const T j3 = mu_j3(s);
  T v = cst<T>(0.0);
  if (fabs(val(j2)) > kEps) v = (3. * sqrt(3.) / 2.) * (j3 / t_pow_1p5(j2));
  if (val(v) > 1.0) v = cst<T>(1.0);
  if (val(v) < -1.0) v = cst<T>(-1.0);
  return (1. / 3.) * t_acos(v);
}
template <class T> MPMX_AD void mu_dq_dsigma(const T* s, T* out) {
  const T q = mu_q(s);
  T d[6];
  mu_dev(s, d);
  for (int i = 0; i < 6; ++i) out[i] = cst<T>(0.);
  if (fabs(val(q)) > kEps) {
    out[0] = 3. / (2. * q) * d[0];
    out[1] = 3. / (2. * q) * d[1];
    out[2] = 3. / (2. * q) * d[2];
    out[3] = 3. / q * d[3];
    out[4] = 3. / q * d[4];
    out[5] = 3. / q * d[5];
  }
}
template <class T> MPMX_AD T dot3(const T* a, const T* b) {
  return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}
template <class T> MPMX_AD void mu_dtheta_dsigma(const T* s, T* out) {
  const double tol = kEps;
  const T j2 = mu_j2(s);
  const T j3 = mu_j3(s);
  T dj2[6];
  mu_dev(s, dj2);
  dj2[3] *= 2.0;
  dj2[4] *= 2.0;
  dj2[5] *= 2.0;
  T dv[6];
  mu_dev(s, dv);
  const T d1[3] = {dv[0], dv[3], dv[5]};
  const T d2[3] = {dv[3], dv[1], dv[4]};
  const T d3[3] = {dv[5], dv[4], dv[2]};
  T dj3[6];
  dj3[0] = dot3(d1, d1) - (2. / 3.) * j2;
  dj3[1] = dot3(d2, d2) - (2. / 3.) * j2;
  dj3[2] = dot3(d3, d3) - (2. / 3.) * j2;
  dj3[3] = 2.0 * dot3(d1, d2);
  dj3[4] = 2.0 * dot3(d2, d3);
  dj3[5] = 2.0 * dot3(d1, d3);
  T dr_dj2 = -9.0 / 4.0 * sqrt(3.0) * j3;
  T dr_dj3 = cst<T>(3.0 / 2.0 * sqrt(3.0));
  T dtheta_dr = cst<T>(-1.0 / 3.0);
  if (fabs(val(j2)) > tol) {
    const T r = j3 / 2.0 * t_pow_m1p5(j2 / 3.0);
    dr_dj2 *= t_pow_m2p5(j2);
    dr_dj3 *= t_pow_m1p5(j2);
    const T omr = 1 - r * r;
    const T factor = (fabs(val(omr)) < tol) ? cst<T>(tol) : t_fabs(omr);
    dtheta_dr = -1.0 / (3.0 * t_sqrt(factor));
  }
  for (int i = 0; i < 6; ++i) out[i] = dtheta_dr * ((dr_dj2 * dj2[i]) + (dr_dj3 * dj3[i]));
}

// ------------------------------ Mohr-Coulomb --------------------------------
enum { AD_ELASTIC = 0, AD_SHEAR = 1, AD_TENSILE = 2 };

// Elastic constants D = [[a1,a2,a2],[a2,a1,a2],[a2,a2,a1]] (+) G I3 and the
// strength state of one particle (no softening: constant over the run).
template <class T>
struct McParams {
  T a1, a2, G;
  T phi;          // friction angle (rad)
  double psi;     // dilation angle (rad)
  double cohesion;
  T tcut;         // tension cutoff state variable
  double c_peak;  // peak cohesion of the material (enters compute_df_dp)
};

template <class T>
struct Inv {
  T eps, rho, theta, st, ct, s3, c3;
};
template <class T>
struct Angles {
  T sphi, cphi, tphi;
  double tpsi;
};

template <class T> MPMX_AD void de_mul(const McParams<T>& m, const T* x, T* y) {
  y[0] = (m.a1 * x[0] + m.a2 * x[1]) + m.a2 * x[2];
  y[1] = (m.a2 * x[0] + m.a1 * x[1]) + m.a2 * x[2];
  y[2] = (m.a2 * x[0] + m.a2 * x[1]) + m.a1 * x[2];
  y[3] = m.G * x[3];
  y[4] = m.G * x[4];
  y[5] = m.G * x[5];
}
template <class T> MPMX_AD void de_tmul(const McParams<T>& m, const T* x, T* y) {
  y[0] = (x[0] * m.a1 + x[1] * m.a2) + x[2] * m.a2;
  y[1] = (x[0] * m.a2 + x[1] * m.a1) + x[2] * m.a2;
  y[2] = (x[0] * m.a2 + x[1] * m.a2) + x[2] * m.a1;
  y[3] = x[3] * m.G;
  y[4] = x[4] * m.G;
  y[5] = x[5] * m.G;
}
template <class T> MPMX_AD T dot6(const T* a, const T* b) {
  T s = a[0] * b[0];
  for (int k = 1; k < 6; ++k) s = s + a[k] * b[k];
  return s;
}

template <class T> MPMX_AD Angles<T> angles(const McParams<T>& m) {
  Angles<T> a;
  t_sincos(m.phi, &a.sphi, &a.cphi);
  a.tphi = a.sphi / a.cphi;
  a.tpsi = dmath::tan(m.psi);
  return a;
}

template <class T> MPMX_AD void invariants(const T* s, Inv<T>& I) {
  I.eps = mu_p(s) * sqrt(3.);
  I.theta = mu_lode(s);
  I.rho = t_sqrt(2. * mu_j2(s));
  t_sincos(I.theta, &I.st, &I.ct);
  t_sincos(I.theta + M_PI / 3., &I.s3, &I.c3);
}

// compute_yield_state (mohr_coulomb.tcc:139-191)
template <class T>
MPMX_AD int yield(T* yf, const Inv<T>& I, const Angles<T>& A, double cohesion, const T& tcut) {
  const double Tolerance = -1E-1;
  yf[0] = sqrt(2. / 3.) * I.ct * I.rho + I.eps / sqrt(3.) - tcut;
  yf[1] = sqrt(1.5) * I.rho * ((I.s3 / (sqrt(3.) * A.cphi)) + (I.c3 * A.tphi / 3.)) +
          (I.eps / sqrt(3.)) * A.tphi - cohesion;
  const double y0 = val(yf[0]), y1 = val(yf[1]);
  int t = AD_ELASTIC;
  if (y0 > Tolerance && y1 > Tolerance) {
    // only the sign of h is used: plain FP64 suffices
    const double sphi = val(A.sphi);
    const double n_phi = (1. + sphi) / (1. - sphi);
    const double sigma_p = val(tcut) * n_phi - 2. * cohesion * sqrt(n_phi);
    const double alpha_p = sqrt(1. + n_phi * n_phi) + n_phi;
    const double h = y0 + alpha_p * (sqrt(2. / 3.) * dmath::cos(val(I.theta) - 4. * M_PI / 3.) *
                                         val(I.rho) +
                                     val(I.eps) / sqrt(3.) - sigma_p);
    t = (h > kEps) ? AD_TENSILE : AD_SHEAR;
  }
  if (y0 < Tolerance && y1 > Tolerance) t = AD_SHEAR;
  if (y0 > Tolerance && y1 < Tolerance) t = AD_TENSILE;
  return t;
}

// compute_df_dp (mohr_coulomb.tcc:195-344) without the softening term
template <class T>
MPMX_AD void df_dp(const McParams<T>& m, int yt, const T* stress, const Inv<T>& I,
                   const Angles<T>& A, T* df, T* dp, T* dp_dq) {
  const T rho = I.rho;
  const T ct = I.ct, st = I.st;
  T df_depsilon, df_drho, df_dtheta;
  if (yt == AD_TENSILE) {
    df_depsilon = cst<T>(1. / sqrt(3.));
    df_drho = sqrt(2. / 3.) * ct;
    df_dtheta = -sqrt(2. / 3.) * rho * st;
  } else {
    df_depsilon = A.tphi / sqrt(3.);
    df_drho = sqrt(1.5) * ((I.s3 / (sqrt(3.) * A.cphi)) + (I.c3 * A.tphi / 3.));
    df_dtheta = sqrt(1.5) * rho * ((I.c3 / (sqrt(3.) * A.cphi)) - (I.s3 * A.tphi / 3.));
  }
  const double depsilon_dsigma = 1. / 3. * sqrt(3.);
  T drho_dsigma[6];
  mu_dq_dsigma(stress, drho_dsigma);
  for (int i = 0; i < 6; ++i) drho_dsigma[i] = drho_dsigma[i] * sqrt(2. / 3.);
  T dtheta_dsigma[6];
  mu_dtheta_dsigma(stress, dtheta_dsigma);
  for (int i = 0; i < 6; ++i)
    df[i] = ((df_depsilon * (i < 3 ? depsilon_dsigma : 0.0)) + (df_drho * drho_dsigma[i])) +
            (df_dtheta * dtheta_dsigma[i]);
  if (yt == AD_TENSILE) {
    const double et_value = 0.6;
    const double xit = 0.1;
    T sqpart = ((4. * (1 - et_value * et_value)) * ct) * ct + 5. * et_value * et_value -
               4. * et_value;
    if (val(sqpart) < kEps) sqpart = cst<T>(1.E-5);
    T rt_den = (2. * (1 - et_value * et_value)) * ct + (2. * et_value - 1) * t_sqrt(sqpart);
    const T rt_num = ((4. * (1 - et_value * et_value)) * ct) * ct +
                     (2. * et_value - 1) * (2. * et_value - 1);
    if (fabs(val(rt_den)) < kEps) rt_den = cst<T>(1.E-5);
    const T rt = rt_num / (3. * rt_den);
    const T den = t_sqrt(((xit * xit) * m.tcut) * m.tcut + ((1.5 * rt) * rt * rho) * rho);
    const T cl = val(den) > 1.0e-15 ? den : cst<T>(1.0e-15);
    const T dp_drt = ((1.5 * rho) * rho) * rt / cl;
    const T dp_drho = ((1.5 * rho) * rt) * rt / cl;
    const double dp_depsilon = 1. / sqrt(3.);
    const T drtden_dtheta =
        (-2. * (1 - et_value * et_value)) * st -
        (((2. * et_value - 1) * 4. * (1 - et_value * et_value)) * ct) * st /
            t_sqrt(((4. * (1 - et_value * et_value)) * ct) * ct + 5. * et_value * et_value -
                   4. * et_value);
    const T drtnum_dtheta = ((-8. * (1 - et_value * et_value)) * ct) * st;
    const T drt_dtheta =
        (drtnum_dtheta * rt_den - drtden_dtheta * rt_num) / ((3. * rt_den) * rt_den);
    for (int i = 0; i < 6; ++i)
      dp[i] = ((dp_depsilon * (i < 3 ? depsilon_dsigma : 0.0)) + (dp_drho * drho_dsigma[i])) +
              ((dp_drt * drt_dtheta) * dtheta_dsigma[i]);
    *dp_dq = dp_drho * sqrt(2. / 3.);
  } else {
    const T r_mc = (3. - A.sphi) / (6 * A.cphi);
    T e_val = (3. - A.sphi) / (3. + A.sphi);
    if (val(e_val) <= 0.5) e_val = cst<T>(0.5 + 1.E-10);
    if (val(e_val) > 1.) e_val = cst<T>(1.);
    T sqpart = ((4. * (1 - e_val * e_val)) * t_sq(ct)) + ((5 * e_val) * e_val) - (4. * e_val);
    if (val(sqpart) < kEps) sqpart = cst<T>(1.E-5);
    T mm = ((2. * (1 - e_val * e_val)) * ct) + ((2. * e_val - 1) * t_sqrt(sqpart));
    if (fabs(val(mm)) < kEps) mm = cst<T>(1.E-5);
    const T l = ((4. * (1. - e_val * e_val)) * t_sq(ct)) + t_sq(2. * e_val - 1.);
    const T r_mw = (l / mm) * r_mc;
    const double xi = 0.1;
    T omega = t_sq(xi * m.c_peak * A.tpsi) + t_sq((r_mw * sqrt(1.5)) * rho);
    if (val(omega) < kEps) omega = cst<T>(1.E-5);
    const T dl_dtheta = ((-8. * (1. - e_val * e_val)) * ct) * st;
    const T dm_dtheta = ((-2. * (1. - e_val * e_val)) * st) +
                        ((0.5 * (2. * e_val - 1.)) * dl_dtheta) / t_sqrt(sqpart);
    const T drmw_dtheta = ((mm * dl_dtheta) - (l * dm_dtheta)) / (mm * mm);
    const double dp_depsilon = A.tpsi / sqrt(3.);
    const T dp_drho = ((3. * rho) * r_mw) * r_mw / (2. * t_sqrt(omega));
    const T dp_dtheta = ((((3. * rho) * rho) * r_mw) * r_mc * drmw_dtheta) / (2. * t_sqrt(omega));
    for (int i = 0; i < 6; ++i)
      dp[i] = ((dp_depsilon * (i < 3 ? depsilon_dsigma : 0.0)) + (dp_drho * drho_dsigma[i])) +
              (dp_dtheta * dtheta_dsigma[i]);
    *dp_dq = dp_drho * sqrt(2. / 3.);
  }
}

// upd += -(lam * D * dp), entries summed over k in order
template <class T>
MPMX_AD void correct(const McParams<T>& m, const T& lam, const T* dp, T* upd) {
  const T la1 = lam * m.a1, la2 = lam * m.a2, lg = lam * m.G;
  upd[0] += -(((la1 * dp[0] + la2 * dp[1]) + la2 * dp[2]));
  upd[1] += -(((la2 * dp[0] + la1 * dp[1]) + la2 * dp[2]));
  upd[2] += -(((la2 * dp[0] + la2 * dp[1]) + la1 * dp[2]));
  upd[3] += -(lg * dp[3]);
  upd[4] += -(lg * dp[4]);
  upd[5] += -(lg * dp[5]);
}

// compute_df_dp at the values of its arguments, as constants (no tangents)
template <class T>
MPMX_AD void df_dp_frozen(const McParams<T>& m, int yt, const T* stress, T* df, T* dp, T* dp_dq) {
  const McParams<double> md{val(m.a1), val(m.a2), val(m.G), val(m.phi),
                            m.psi,      m.cohesion, val(m.tcut), m.c_peak};
  double sd[6], a[6], b[6], q;
  for (int i = 0; i < 6; ++i) sd[i] = val(stress[i]);
  const Angles<double> Ad = angles(md);
  Inv<double> Id;
  invariants(sd, Id);
  df_dp(md, yt, sd, Id, Ad, a, b, &q);
  for (int i = 0; i < 6; ++i) {
    df[i] = cst<T>(a[i]);
    dp[i] = cst<T>(b[i]);
  }
  *dp_dq = cst<T>(q);
}

// The return mapping (mohr_coulomb.tcc:388-512) for a particle that is not in
// the separated state: out = updated stress. Returns the yield type of the
// trial stress (AD_ELASTIC: out = trial). With `frozen_flow`, the yield and
// plastic-potential gradients of every correction carry no tangents (values
// unchanged).
template <class T>
MPMX_AD int mc_update(const McParams<T>& m, const T* stress, const T* dstrain, T* out,
                      bool frozen_flow = false) {
  const double Tolerance = 1E-1;
  T trial[6];
  {
    T ds[6];
    de_mul(m, dstrain, ds);
    for (int i = 0; i < 6; ++i) trial[i] = stress[i] + ds[i];
  }
  const Angles<T> A = angles(m);
  Inv<T> It;
  T yf[2];
  invariants(trial, It);
  const int ytt = yield(yf, It, A, m.cohesion, m.tcut);
  if (ytt == AD_ELASTIC) {
    for (int i = 0; i < 6; ++i) out[i] = trial[i];
    return AD_ELASTIC;
  }
  T yield_trial = (ytt == AD_TENSILE) ? yf[0] : yf[1];
  // first correction: with the current stress if it lies on the surface
  Inv<T> Ic;
  invariants(stress, Ic);
  T yc[2];
  const int ytc = yield(yc, Ic, A, m.cohesion, m.tcut);
  double ycur = 1.7976931348623157e308;
  if (ytc == AD_TENSILE) ycur = val(yc[0]);
  if (ytc == AD_SHEAR) ycur = val(yc[1]);
  const bool use_current = fabs(ycur) < Tolerance;
  // The first correction (from the trial stress, with the gradients at the
  // current or the trial stress) and the loop's corrections share one
  // evaluation of the flow rule: pass -1 is the first correction.
  T x[6];
  for (int i = 0; i < 6; ++i) {
    x[i] = use_current ? stress[i] : trial[i];
    out[i] = trial[i];
  }
  Inv<T> I = use_current ? Ic : It;
  int yt = use_current ? ytc : ytt;
  for (int itr = -1; itr < 100; ++itr) {
    if (itr >= 0) {
      invariants(out, I);
      yt = yield(yf, I, A, m.cohesion, m.tcut);
      if (val(yf[0]) < Tolerance && val(yf[1]) < Tolerance) break;
      if (yt == AD_TENSILE) yield_trial = yf[0];
      if (yt == AD_SHEAR) yield_trial = yf[1];
      for (int i = 0; i < 6; ++i) x[i] = out[i];
    }
    T df[6], dp[6], tmp[6], dp_dq;
    if (frozen_flow)
      df_dp_frozen(m, yt, x, df, dp, &dp_dq);
    else
      df_dp(m, yt, x, I, A, df, dp, &dp_dq);
    de_tmul(m, df, tmp);
    const T lam = ((itr < 0 && use_current) ? dot6(tmp, dstrain) : yield_trial) /
                  (dot6(tmp, dp) + 0.);
    correct(m, lam, dp, out);
  }
  return ytt;
}

}  // namespace ad
}  // namespace mpmx
