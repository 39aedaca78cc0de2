// CPU restatement of geomechanics/mpm's explicit MPM step — TEST INFRASTRUCTURE.
//
// This file is the oracle: only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg load it, and only as the checker. It is never part of the
// product path (mpm_amd/), which fails loudly without its HIP library.
//
// Every routine follows the reference statement by statement, in the
// reference's evaluation order, and cites the file:line it restates (paths
// relative to the reference root). Built single-threaded it sums every nodal
// quantity in particle-vector order, which is the one order that defines a
// unique answer for the reference (its OpenMP run is thread-order
// nondeterministic: include/nodes/node.tcc:56-65 spin-locks). Built with
// -DORC_OMP -fopenmp it parallelises the particle and node sweeps exactly as
// include/mesh/mesh.tcc:70-103,1156-1161 do, with atomic nodal updates
// standing in for the reference's SpinMutex; that build is the timed
// "reference CPU path".
//
// Where the reference's floating-point association lives inside Eigen
// expression templates (products, norms), this file fixes one order (the
// sequential k-order of a non-vectorised product); those functions are pinned
// by the reference's known-answer tests only to their 1e-7 tolerance. Their
// list is kept with the parity notes of this project.
//
// Build: see oracle/Makefile. Compile with -ffp-contract=off.

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

// Elementary functions. The reference calls the platform libm; built with
// -DORC_PLATFORM_LIBM this file does too. The default build uses the
// deterministic FP64 functions of mpm_amd/csrc/mpmx_math.h (the same ones the
// GPU kernels use), so that the oracle and the device agree bit for bit.
#ifdef ORC_PLATFORM_LIBM
namespace orc {
namespace mth {
inline double sin(double x) { return std::sin(x); }
inline double cos(double x) { return std::cos(x); }
inline double tan(double x) { return std::tan(x); }
inline double acos(double x) { return std::acos(x); }
inline double pow(double x, double y) { return std::pow(x, y); }
}  // namespace mth
}  // namespace orc
#else
#include "mpmx_math.h"
namespace orc {
namespace mth {
inline double sin(double x) { return mpmx::dmath::sin(x); }
inline double cos(double x) { return mpmx::dmath::cos(x); }
inline double tan(double x) { return mpmx::dmath::tan(x); }
inline double acos(double x) { return mpmx::dmath::acos(x); }
inline double pow(double x, double y) {
  if (y == 1.5) return mpmx::dmath::pow_1p5(x);
  if (y == -1.5) return mpmx::dmath::pow_m1p5(x);
  if (y == -2.5) return mpmx::dmath::pow_m2p5(x);
  return std::pow(x, y);
}
}  // namespace mth
}  // namespace orc
#endif

#ifdef ORC_OMP
#include <omp.h>
#define ORC_PRAGMA(x) _Pragma(#x)
#else
#define ORC_PRAGMA(x)
#endif

namespace orc {

typedef int64_t Index;
static const double kEps = std::numeric_limits<double>::epsilon();
static const Index kNoCell = -1;

// ---------------------------------------------------------------------------
// Materials. Parameter vector layout (shared with the product's C-ABI):
// [0] type (0 LinearElastic, 1 MohrCoulomb) [1] density [2] youngs_modulus
// [3] poisson_ratio [4] friction (deg) [5] dilation (deg) [6] cohesion
// [7] residual_friction (deg) [8] residual_dilation (deg) [9] residual_cohesion
// [10] peak_pdstrain [11] residual_pdstrain [12] tension_cutoff [13] softening
// [14] packing_fraction_minimum [15] initial packing fraction (<=0 -> 1)
// [16] MohrCoulomb stress_rate (0 none, 1 Jaumann, 2 Green-Naghdi); LinearElastic
// has its stress_rate in [4]
// ---------------------------------------------------------------------------
enum { kLE = 0, kMC = 1 };
static const int kMatParams = 17;
// MC state variables, reference order (mohr_coulomb.tcc:116-118)
enum { SV_YIELD = 0, SV_PHI, SV_PSI, SV_COH, SV_TCUT, SV_EPS, SV_RHO, SV_THETA, SV_PDS };
static const int kNsv = 9;

struct Material {
  int type = kLE;
  int rate = 0;  // objective stress rate: 0 none, 1 Jaumann, 2 Green-Naghdi
  double density = 0, E = 0, nu = 0;
  // LinearElastic (linear_elastic.tcc:20-44)
  double bulk = 0;
  double de[6][6];
  // MohrCoulomb (mohr_coulomb.tcc:9-77)
  double grain_density = 0, min_packing = 0;
  bool softening = false;
  double phi_peak = 0, psi_peak = 0, c_peak = 0;
  double phi_res = 0, psi_res = 0, c_res = 0;
  double pd_peak = 0, pd_res = 0, tcut = 0;
  double shear = 0;
};

static inline double check_low(double val) {  // mohr_coulomb.h:124-126
  return (val > 1.0e-15 ? val : 1.0e-15);
}

static void material_init(Material& m, const double* p) {
  m.type = (int)p[0];
  m.density = p[1];
  m.E = p[2];
  m.nu = p[3];
  if (m.type == kLE) {
    // linear_elastic.tcc:8-17 (stress_rate), 26-27, 53-69
    m.rate = (int)p[4];
    m.bulk = m.E / (3.0 * (1. - 2. * m.nu));
    const double G = m.E / (2.0 * (1. + m.nu));
    const double a1 = m.bulk + (4.0 / 3.0) * G;
    const double a2 = m.bulk - (2.0 / 3.0) * G;
    std::memset(m.de, 0, sizeof(m.de));
    m.de[0][0] = a1; m.de[0][1] = a2; m.de[0][2] = a2;
    m.de[1][0] = a2; m.de[1][1] = a1; m.de[1][2] = a2;
    m.de[2][0] = a2; m.de[2][1] = a2; m.de[2][2] = a1;
    m.de[3][3] = G; m.de[4][4] = G; m.de[5][5] = G;
  } else {
    // mohr_coulomb.tcc:9-77
    double initial_packing_fraction = p[15] > 0. ? p[15] : 1.;
    m.grain_density = m.density / initial_packing_fraction;
    m.min_packing = p[14];
    m.softening = p[13] != 0.;
    m.phi_peak = p[4] * M_PI / 180.;
    m.psi_peak = p[5] * M_PI / 180.;
    m.c_peak = p[6];
    m.phi_res = p[7] * M_PI / 180.;
    m.psi_res = p[8] * M_PI / 180.;
    m.c_res = p[9];
    m.pd_peak = p[10];
    m.pd_res = p[11];
    m.tcut = p[12];
    m.rate = (int)p[16];  // stress_rate, read by the LinearElastic base (linear_elastic.tcc:8-17)
    m.bulk = m.E / (3.0 * (1. - 2. * m.nu));
    m.shear = m.E / (2.0 * (1 + m.nu));
    // compute_elastic_tensor (mohr_coulomb.tcc:517-533)
    const double G = m.shear;
    const double a1 = m.bulk + (4.0 / 3.0) * G;
    const double a2 = m.bulk - (2.0 / 3.0) * G;
    std::memset(m.de, 0, sizeof(m.de));
    m.de[0][0] = a1; m.de[0][1] = a2; m.de[0][2] = a2;
    m.de[1][0] = a2; m.de[1][1] = a1; m.de[1][2] = a2;
    m.de[2][0] = a2; m.de[2][1] = a2; m.de[2][2] = a1;
    m.de[3][3] = G; m.de[4][4] = G; m.de[5][5] = G;
  }
}

// mohr_coulomb.tcc:85-111
static void mc_initialise_state(const Material& m, double* sv) {
  sv[SV_YIELD] = 0;
  sv[SV_PHI] = m.phi_peak;
  sv[SV_PSI] = m.psi_peak;
  sv[SV_COH] = m.c_peak;
  sv[SV_TCUT] = check_low((m.tcut < m.c_peak / mth::tan(m.phi_peak))
                              ? m.tcut
                              : m.c_peak / mth::tan(m.phi_peak));
  sv[SV_EPS] = 0.;
  sv[SV_RHO] = 0.;
  sv[SV_THETA] = 0.;
  sv[SV_PDS] = 0.;
}

// D * x with the sequential k-order of Eigen's coefficient-based 6x6 product
// (linear_elastic.tcc:81, infinitesimal_elasto_plastic.tcc:12).
static inline void de_mul(const double de[6][6], const double* x, double* y) {
  for (int i = 0; i < 6; ++i) {
    double s = de[i][0] * x[0];
    for (int k = 1; k < 6; ++k) s = s + de[i][k] * x[k];
    y[i] = s;
  }
}

// exp(A) of the skew matrix A with A01 = a[0], A02 = a[1], A12 = a[2], by
// Rodrigues' formula I + sin(t)/t A + (1 - cos t)/t^2 A^2 (t = |a|); below
// t = 1e-2 the two coefficients by their Taylor series to t^6 (truncation
// below 1e-21). This replaces the Pade approximant of Eigen's
// MatrixExponential used by linear_elastic.tcc:106-107; both are exact to
// rounding for the small rotations of a time step.
static void skew_exp(const double* a, double Q[3][3]) {
  const double t2 = (a[0] * a[0] + a[1] * a[1]) + a[2] * a[2];
  double c1, c2;
  if (t2 < 1e-4) {
    c1 = 1. - t2 * (1. / 6. - t2 * (1. / 120. - t2 * (1. / 5040.)));
    c2 = 0.5 - t2 * (1. / 24. - t2 * (1. / 720. - t2 * (1. / 40320.)));
  } else {
    const double t = std::sqrt(t2);
    c1 = mth::sin(t) / t;
    c2 = (1. - mth::cos(t)) / t2;
  }
  const double A[3][3] = {{0., a[0], a[1]}, {-a[0], 0., a[2]}, {-a[1], -a[2], 0.}};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double a2 = A[i][0] * A[0][j];
      for (int k = 1; k < 3; ++k) a2 = a2 + A[i][k] * A[k][j];
      Q[i][j] = ((i == j ? 1. : 0.) + c1 * A[i][j]) + c2 * a2;
    }
}

// Q X Q^T of a symmetric X, entries in Voigt order (00, 11, 22, 01, 12, 02);
// sequential k-order in both products.
static void rotate_sym(const double Q[3][3], const double X[3][3], double* y) {
  double T[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double t = Q[i][0] * X[0][j];
      for (int k = 1; k < 3; ++k) t = t + Q[i][k] * X[k][j];
      T[i][j] = t;
    }
  const int I[6] = {0, 1, 2, 0, 1, 0}, J[6] = {0, 1, 2, 1, 2, 2};
  for (int v = 0; v < 6; ++v) {
    double t = T[I[v]][0] * Q[J[v]][0];
    for (int k = 1; k < 3; ++k) t = t + T[I[v]][k] * Q[J[v]][k];
    y[v] = t;
  }
}

// LinearElastic stress with the Jaumann rate (linear_elastic.tcc:96-124):
// W = skew(dF - I); the stress rotated by exp(W), the strain increment
// (tensor shears) by exp(W/2), then the elastic increment of the rotated
// strain.
// C = A B and C = A B^T of 3x3 matrices (sequential k)
static void mat_mul(const double A[3][3], const double B[3][3], double C[3][3]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double t = A[i][0] * B[0][j];
      for (int k = 1; k < 3; ++k) t = t + A[i][k] * B[k][j];
      C[i][j] = t;
    }
}
static void mat_mul_t(const double A[3][3], const double B[3][3], double C[3][3]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double t = A[i][0] * B[j][0];
      for (int k = 1; k < 3; ++k) t = t + A[i][k] * B[j][k];
      C[i][j] = t;
    }
}

static void inverse(int dim, const double J[3][3], double Ji[3][3]);

// The rotation R of the polar decomposition F = R U (det F > 0), what
// mpm::materials::compute_rotation_tensor (material_utility.tcc) obtains
// from the eigen-decomposition of F^T F: here by Newton's iteration
// X <- (X + X^-T) / 2 from X = F (Higham 1986), until an iteration changes no
// entry by more than 1e-15 of the largest (at most 100 iterations).
static void polar_rotation(const double F[3][3], double R[3][3]) {
  double X[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) X[i][j] = F[i][j];
  for (int it = 0; it < 100; ++it) {
    double Xi[3][3], N[3][3];
    inverse(3, X, Xi);
    double diff = 0., big = 0.;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        N[i][j] = 0.5 * (X[i][j] + Xi[j][i]);
        const double d = std::fabs(N[i][j] - X[i][j]), a = std::fabs(N[i][j]);
        if (d > diff) diff = d;
        if (a > big) big = a;
      }
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) X[i][j] = N[i][j];
    if (diff <= 1.0E-15 * big) break;
  }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) R[i][j] = X[i][j];
}

// The two rotations of an objective stress rate: Lambda_delta (strain) and
// Lambda_Delta (stress). Jaumann (linear_elastic.tcc:102-107): exp(W/2) and
// its square, W the skew part of dF - I. Green-Naghdi (:134-150):
// R_new R_half^T and R_new R_current^T, R the rotations of F, (F + F_new)/2
// and F_new = dF F.
static void rate_rotations(int rate, const double dF[3][3], const double F[3][3], double L[3][3],
                           double LL[3][3]) {
  if (rate == 1) {
    const double a[3] = {0.5 * (0.5 * (dF[0][1] - dF[1][0])),
                         0.5 * (0.5 * (dF[0][2] - dF[2][0])),
                         0.5 * (0.5 * (dF[1][2] - dF[2][1]))};
    skew_exp(a, L);
    mat_mul(L, L, LL);
    return;
  }
  double Fn[3][3], Fh[3][3], Rc[3][3], Rh[3][3], Rn[3][3];
  mat_mul(dF, F, Fn);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) Fh[i][j] = 0.5 * F[i][j] + 0.5 * Fn[i][j];
  polar_rotation(F, Rc);
  polar_rotation(Fh, Rh);
  polar_rotation(Fn, Rn);
  mat_mul_t(Rn, Rh, L);
  mat_mul_t(Rn, Rc, LL);
}

// The trial stress of an objective rate: the stress rotated by Lambda_Delta,
// the strain increment (tensor shears) by Lambda_delta, then the elastic
// increment of the rotated strain (linear_elastic.tcc:109-124, 152-164).
static void rate_stress(const double de[6][6], const double* stress, const double* dstrain,
                        const double L[3][3], const double LL[3][3], double* out) {
  const double S[3][3] = {{stress[0], stress[3], stress[5]},
                          {stress[3], stress[1], stress[4]},
                          {stress[5], stress[4], stress[2]}};
  const double E[3][3] = {{dstrain[0], 0.5 * dstrain[3], 0.5 * dstrain[5]},
                          {0.5 * dstrain[3], dstrain[1], 0.5 * dstrain[4]},
                          {0.5 * dstrain[5], 0.5 * dstrain[4], dstrain[2]}};
  double rs[6], re[6], d[6];
  rotate_sym(LL, S, rs);
  rotate_sym(L, E, re);
  for (int v = 3; v < 6; ++v) re[v] = 2.0 * re[v];
  de_mul(de, re, d);
  for (int v = 0; v < 6; ++v) out[v] = rs[v] + d[v];
}

// LinearElastic with the Jaumann rate (linear_elastic.tcc:96-124)
static void jaumann_stress(const double de[6][6], const double* stress, const double* dstrain,
                           const double dF[3][3], double* out) {
  double L[3][3], LL[3][3];
  rate_rotations(1, dF, dF, L, LL);
  rate_stress(de, stress, dstrain, L, LL, out);
}

// x^T * D (row vector), same k-order.
static inline void de_tmul(const double de[6][6], const double* x, double* y) {
  for (int j = 0; j < 6; ++j) {
    double s = x[0] * de[0][j];
    for (int k = 1; k < 6; ++k) s = s + x[k] * de[k][j];
    y[j] = s;
  }
}

static inline double dot6(const double* a, const double* b) {
  double s = a[0] * b[0];
  for (int k = 1; k < 6; ++k) s = s + a[k] * b[k];
  return s;
}

// ----------------------------- material_utility.tcc -----------------------
static inline double mu_p(const double* s) {  // :2-5
  return (1. / 3. * (s[0] + s[1] + s[2]));
}
static inline void mu_dev(const double* s, double* d) {  // :8-19
  const double p = mu_p(s);
  for (int i = 0; i < 6; ++i) d[i] = s[i];
  for (int i = 0; i < 3; ++i) d[i] -= p;
}
static inline double sq(double x) { return x * x; }  // mth::pow(x, 2)
static inline double mu_j2(const double* s) {  // :22-32
  return (sq(s[0] - s[1]) + sq(s[1] - s[2]) + sq(s[0] - s[2])) / 6.0 +
         sq(s[3]) + sq(s[4]) + sq(s[5]);
}
static inline double mu_j3(const double* s) {  // :35-51
  double d[6];
  mu_dev(s, d);
  return (d[0] * d[1] * d[2]) - (d[2] * sq(d[3])) +
         ((2 * d[3] * d[4] * d[5]) - (d[0] * sq(d[4])) - (d[1] * sq(d[5])));
}
static inline double mu_q(const double* s) {  // :54-61
  const double j2 = mu_j2(s);
  return (std::sqrt(3 * j2));
}
static inline double mu_lode_angle(const double* s, double tol = kEps) {  // :64-81
  const double j2 = mu_j2(s);
  const double j3 = mu_j3(s);
  double v = 0.0;
  if (std::abs(j2) > tol) v = (3. * std::sqrt(3.) / 2.) * (j3 / mth::pow(j2, 1.5));
  if (v > 1.0) v = 1.0;
  if (v < -1.0) v = -1.0;
  return (1. / 3.) * mth::acos(v);
}
static inline void mu_dq_dsigma(const double* s, double* out) {  // :95-116
  const double q = mu_q(s);
  double d[6];
  mu_dev(s, d);
  for (int i = 0; i < 6; ++i) out[i] = 0.;
  if (std::abs(q) > kEps) {
    out[0] = 3. / (2. * q) * d[0];
    out[1] = 3. / (2. * q) * d[1];
    out[2] = 3. / (2. * q) * d[2];
    out[3] = 3. / q * d[3];
    out[4] = 3. / q * d[4];
    out[5] = 3. / q * d[5];
  }
}
static inline void mu_dtheta_dsigma(const double* s, double tol, double* out) {  // :167-204
  const double j2 = mu_j2(s);
  const double j3 = mu_j3(s);
  double dj2[6];
  mu_dev(s, dj2);  // dj2_dsigma :119-129
  dj2[3] *= 2.0;
  dj2[4] *= 2.0;
  dj2[5] *= 2.0;
  double dv[6];
  mu_dev(s, dv);  // dj3_dsigma :132-164
  const double d1[3] = {dv[0], dv[3], dv[5]};
  const double d2[3] = {dv[3], dv[1], dv[4]};
  const double d3[3] = {dv[5], dv[4], dv[2]};
  auto dot3 = [](const double* a, const double* b) {
    return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
  };
  double dj3[6];
  dj3[0] = dot3(d1, d1) - (2. / 3.) * j2;
  dj3[1] = dot3(d2, d2) - (2. / 3.) * j2;
  dj3[2] = dot3(d3, d3) - (2. / 3.) * j2;
  dj3[3] = 2.0 * dot3(d1, d2);
  dj3[4] = 2.0 * dot3(d2, d3);
  dj3[5] = 2.0 * dot3(d1, d3);
  double dr_dj2 = -9.0 / 4.0 * std::sqrt(3.0) * j3;
  double dr_dj3 = 3.0 / 2.0 * std::sqrt(3.0);
  double dtheta_dr = -1.0 / 3.0;
  if (std::abs(j2) > tol) {
    double r = j3 / 2.0 * mth::pow(j2 / 3.0, -1.5);
    dr_dj2 *= mth::pow(j2, -2.5);
    dr_dj3 *= mth::pow(j2, -1.5);
    const double factor = (std::abs(1 - r * r) < tol) ? tol : std::abs(1 - r * r);
    dtheta_dr = -1.0 / (3.0 * std::sqrt(factor));
  }
  for (int i = 0; i < 6; ++i)
    out[i] = dtheta_dr * ((dr_dj2 * dj2[i]) + (dr_dj3 * dj3[i]));
}

// ------------------------------ Mohr-Coulomb --------------------------------
enum { FS_ELASTIC = 0, FS_SHEAR = 1, FS_TENSILE = 2, FS_SEPARATED = 3 };

// mohr_coulomb.tcc:124-134
static inline void mc_invariants(const double* s, double* sv) {
  sv[SV_EPS] = mu_p(s) * std::sqrt(3.);
  sv[SV_THETA] = mu_lode_angle(s);
  sv[SV_RHO] = std::sqrt(2. * mu_j2(s));
}

// mohr_coulomb.tcc:139-191
static int mc_yield_state(double* yf, const double* sv) {
  const double Tolerance = -1E-1;
  const double epsilon = sv[SV_EPS];
  const double rho = sv[SV_RHO];
  const double theta = sv[SV_THETA];
  const double phi = sv[SV_PHI];
  const double cohesion = sv[SV_COH];
  const double tension_cutoff = sv[SV_TCUT];
  yf[0] = std::sqrt(2. / 3.) * mth::cos(theta) * rho + epsilon / std::sqrt(3.) -
          tension_cutoff;
  yf[1] = std::sqrt(1.5) * rho *
              ((mth::sin(theta + M_PI / 3.) / (std::sqrt(3.) * mth::cos(phi))) +
               (mth::cos(theta + M_PI / 3.) * mth::tan(phi) / 3.)) +
          (epsilon / std::sqrt(3.)) * mth::tan(phi) - cohesion;
  int yield_type = FS_ELASTIC;
  if (yf[0] > Tolerance && yf[1] > Tolerance) {
    const double n_phi = (1. + mth::sin(phi)) / (1. - mth::sin(phi));
    const double sigma_p = tension_cutoff * n_phi - 2. * cohesion * std::sqrt(n_phi);
    const double alpha_p = std::sqrt(1. + n_phi * n_phi) + n_phi;
    const double h =
        yf[0] + alpha_p * (std::sqrt(2. / 3.) * mth::cos(theta - 4. * M_PI / 3.) * rho +
                           epsilon / std::sqrt(3.) - sigma_p);
    if (h > kEps)
      yield_type = FS_TENSILE;
    else
      yield_type = FS_SHEAR;
  }
  if (yf[0] < Tolerance && yf[1] > Tolerance) yield_type = FS_SHEAR;
  if (yf[0] > Tolerance && yf[1] < Tolerance) yield_type = FS_TENSILE;
  return yield_type;
}

// mohr_coulomb.tcc:195-344
static void mc_df_dp(const Material& m, int yield_type, const double* sv, const double* stress,
                     double* df_dsigma, double* dp_dsigma, double* dp_dq, double* softening) {
  const double rho = sv[SV_RHO];
  const double theta = sv[SV_THETA];
  const double phi = sv[SV_PHI];
  const double psi = sv[SV_PSI];
  const double tension_cutoff = sv[SV_TCUT];
  const double pdstrain = sv[SV_PDS];
  double df_depsilon, df_drho, df_dtheta;
  if (yield_type == FS_TENSILE) {
    df_depsilon = 1. / std::sqrt(3.);
    df_drho = std::sqrt(2. / 3.) * mth::cos(theta);
    df_dtheta = -std::sqrt(2. / 3.) * rho * mth::sin(theta);
  } else {
    df_depsilon = mth::tan(phi) / std::sqrt(3.);
    df_drho = std::sqrt(1.5) * ((mth::sin(theta + M_PI / 3.) / (std::sqrt(3.) * mth::cos(phi))) +
                                (mth::cos(theta + M_PI / 3.) * mth::tan(phi) / 3.));
    df_dtheta = std::sqrt(1.5) * rho *
                ((mth::cos(theta + M_PI / 3.) / (std::sqrt(3.) * mth::cos(phi))) -
                 (mth::sin(theta + M_PI / 3.) * mth::tan(phi) / 3.));
  }
  double depsilon_dsigma[6] = {0, 0, 0, 0, 0, 0};
  for (int i = 0; i < 3; ++i) depsilon_dsigma[i] = 1. / 3. * std::sqrt(3.);
  double drho_dsigma[6];
  mu_dq_dsigma(stress, drho_dsigma);
  for (int i = 0; i < 6; ++i) drho_dsigma[i] = drho_dsigma[i] * std::sqrt(2. / 3.);
  double dtheta_dsigma[6];
  mu_dtheta_dsigma(stress, kEps, dtheta_dsigma);
  for (int i = 0; i < 6; ++i)
    df_dsigma[i] = ((df_depsilon * depsilon_dsigma[i]) + (df_drho * drho_dsigma[i])) +
                   (df_dtheta * dtheta_dsigma[i]);
  if (yield_type == FS_TENSILE) {
    const double et_value = 0.6;
    const double xit = 0.1;
    double sqpart = 4. * (1 - et_value * et_value) * mth::cos(theta) * mth::cos(theta) +
                    5. * et_value * et_value - 4. * et_value;
    if (sqpart < kEps) sqpart = 1.E-5;
    double rt_den = 2. * (1 - et_value * et_value) * mth::cos(theta) +
                    (2. * et_value - 1) * std::sqrt(sqpart);
    const double rt_num = 4. * (1 - et_value * et_value) * mth::cos(theta) * mth::cos(theta) +
                          (2. * et_value - 1) * (2. * et_value - 1);
    if (std::fabs(rt_den) < kEps) rt_den = 1.E-5;
    const double rt = rt_num / (3. * rt_den);
    const double dp_drt = 1.5 * rho * rho * rt /
                          check_low(std::sqrt(xit * xit * tension_cutoff * tension_cutoff +
                                              1.5 * rt * rt * rho * rho));
    const double dp_drho = 1.5 * rho * rt * rt /
                           check_low(std::sqrt(xit * xit * tension_cutoff * tension_cutoff +
                                               1.5 * rt * rt * rho * rho));
    const double dp_depsilon = 1. / std::sqrt(3.);
    const double drtden_dtheta =
        -2. * (1 - et_value * et_value) * mth::sin(theta) -
        (2. * et_value - 1) * 4. * (1 - et_value * et_value) * mth::cos(theta) * mth::sin(theta) /
            std::sqrt(4. * (1 - et_value * et_value) * mth::cos(theta) * mth::cos(theta) +
                      5. * et_value * et_value - 4. * et_value);
    const double drtnum_dtheta = -8. * (1 - et_value * et_value) * mth::cos(theta) * mth::sin(theta);
    const double drt_dtheta =
        (drtnum_dtheta * rt_den - drtden_dtheta * rt_num) / (3. * rt_den * rt_den);
    for (int i = 0; i < 6; ++i)
      dp_dsigma[i] = ((dp_depsilon * depsilon_dsigma[i]) + (dp_drho * drho_dsigma[i])) +
                     (dp_drt * drt_dtheta * dtheta_dsigma[i]);
    *dp_dq = dp_drho * std::sqrt(2. / 3.);
  } else {
    const double r_mc = (3. - mth::sin(phi)) / (6 * mth::cos(phi));
    double e_val = (3. - mth::sin(phi)) / (3. + mth::sin(phi));
    if (e_val <= 0.5) e_val = 0.5 + 1.E-10;
    if (e_val > 1.) e_val = 1.;
    double sqpart = (4. * (1 - e_val * e_val) * sq(mth::cos(theta))) + (5 * e_val * e_val) -
                    (4. * e_val);
    if (sqpart < kEps) sqpart = 1.E-5;
    double mm = (2. * (1 - e_val * e_val) * mth::cos(theta)) + ((2. * e_val - 1) * std::sqrt(sqpart));
    if (std::fabs(mm) < kEps) mm = 1.E-5;
    const double l = (4. * (1. - e_val * e_val) * sq(mth::cos(theta))) + sq(2. * e_val - 1.);
    const double r_mw = (l / mm) * r_mc;
    const double xi = 0.1;
    double omega = sq(xi * m.c_peak * mth::tan(psi)) + sq(r_mw * std::sqrt(1.5) * rho);
    if (omega < kEps) omega = 1.E-5;
    const double dl_dtheta = -8. * (1. - e_val * e_val) * mth::cos(theta) * mth::sin(theta);
    const double dm_dtheta = (-2. * (1. - e_val * e_val) * mth::sin(theta)) +
                             (0.5 * (2. * e_val - 1.) * dl_dtheta) / std::sqrt(sqpart);
    const double drmw_dtheta = ((mm * dl_dtheta) - (l * dm_dtheta)) / (mm * mm);
    const double dp_depsilon = mth::tan(psi) / std::sqrt(3.);
    const double dp_drho = 3. * rho * r_mw * r_mw / (2. * std::sqrt(omega));
    const double dp_dtheta = (3. * rho * rho * r_mw * r_mc * drmw_dtheta) / (2. * std::sqrt(omega));
    for (int i = 0; i < 6; ++i)
      dp_dsigma[i] = ((dp_depsilon * depsilon_dsigma[i]) + (dp_drho * drho_dsigma[i])) +
                     (dp_dtheta * dtheta_dsigma[i]);
    *dp_dq = dp_drho * std::sqrt(2. / 3.);
  }
  double dphi_dpstrain = 0.;
  double dc_dpstrain = 0.;
  *softening = 0.;
  if (m.softening && pdstrain > m.pd_peak && pdstrain < m.pd_res) {
    dphi_dpstrain = (m.phi_res - m.phi_peak) / (m.pd_res - m.pd_peak);
    dc_dpstrain = (m.c_res - m.c_peak) / (m.pd_res - m.pd_peak);
    double df_dphi = std::sqrt(1.5) * rho *
                         ((mth::sin(phi) * mth::sin(theta + M_PI / 3.) /
                           (std::sqrt(3.) * mth::cos(phi) * mth::cos(phi))) +
                          (mth::cos(theta + M_PI / 3.) / (3. * mth::cos(phi) * mth::cos(phi)))) +
                     (mu_p(stress) / (mth::cos(phi) * mth::cos(phi)));
    double df_dc = -1.;
    *softening = (-1.) * ((df_dphi * dphi_dpstrain) + (df_dc * dc_dpstrain)) * (*dp_dq);
  }
}

// mohr_coulomb.tcc:348-513. `density` is the particle's current mass density.
// trial_in: the trial stress of an objective stress rate
// (InfinitesimalElastoPlastic::compute_trial_stress,
// infinitesimal_elasto_plastic.tcc:1-25); null: stress + D dstrain.
static void mc_compute_stress(const Material& m, const double* stress, const double* dstrain,
                              double density, double* sv, double* out,
                              const double* trial_in = nullptr) {
  const double current_packing_density = density;
  const double critical_density = m.min_packing * m.grain_density;
  const double pdstrain = sv[SV_PDS];
  if (m.softening && pdstrain > m.pd_peak) {
    if (pdstrain < m.pd_res) {
      sv[SV_PHI] = m.phi_res + ((m.phi_peak - m.phi_res) * (pdstrain - m.pd_res) /
                                (m.pd_peak - m.pd_res));
      sv[SV_PSI] = m.psi_res + ((m.psi_peak - m.psi_res) * (pdstrain - m.pd_res) /
                                (m.pd_peak - m.pd_res));
      sv[SV_COH] = m.c_res + ((m.c_peak - m.c_res) * (pdstrain - m.pd_res) /
                              (m.pd_peak - m.pd_res));
    } else {
      sv[SV_PHI] = m.phi_res;
      sv[SV_PSI] = m.psi_res;
      sv[SV_COH] = m.c_res;
    }
    const double apex = sv[SV_COH] / mth::tan(sv[SV_PHI]);
    if (sv[SV_TCUT] > apex) sv[SV_TCUT] = check_low(apex);
  }
  sv[SV_YIELD] = 0;
  double trial[6];
  if (trial_in) {
    for (int i = 0; i < 6; ++i) trial[i] = trial_in[i];
  } else {
    double ds[6];
    de_mul(m.de, dstrain, ds);
    for (int i = 0; i < 6; ++i) trial[i] = stress[i] + ds[i];
  }
  if (current_packing_density <= critical_density) {
    sv[SV_PDS] += mu_q(trial) / 3.0 / m.shear;
    sv[SV_YIELD] = 3;
    for (int i = 0; i < 6; ++i) out[i] = 0.;
    return;
  }
  mc_invariants(trial, sv);
  double yft[2];
  int ytt = mc_yield_state(yft, sv);
  if (ytt == FS_ELASTIC) {
    sv[SV_YIELD] = 0;
    for (int i = 0; i < 6; ++i) out[i] = trial[i];
    return;
  }
  const double Tolerance = 1E-1;
  double softening_trial = 0., dp_dq_trial = 0.;
  double df_t[6] = {0}, dp_t[6] = {0};
  mc_df_dp(m, ytt, sv, trial, df_t, dp_t, &dp_dq_trial, &softening_trial);
  double yield_trial = 0.;
  if (ytt == FS_TENSILE) {
    sv[SV_YIELD] = 2;
    yield_trial = yft[0];
  }
  if (ytt == FS_SHEAR) {
    sv[SV_YIELD] = 1;
    yield_trial = yft[1];
  }
  double tmp[6];
  de_tmul(m.de, df_t, tmp);
  double lambda_trial = yield_trial / (dot6(tmp, dp_t) + softening_trial);
  mc_invariants(stress, sv);
  double yf[2];
  int yt = mc_yield_state(yf, sv);
  double yield = std::numeric_limits<double>::max();
  if (yt == FS_TENSILE) yield = yf[0];
  if (yt == FS_SHEAR) yield = yf[1];
  double softening = 0., dp_dq = 0.;
  double df[6] = {0}, dp[6] = {0};
  mc_df_dp(m, yt, sv, stress, df, dp, &dp_dq, &softening);
  de_tmul(m.de, df, tmp);
  const double lambda = dot6(tmp, dstrain) / (dot6(tmp, dp) + softening);
  double upd[6];
  for (int i = 0; i < 6; ++i) upd[i] = trial[i];
  double dpdstrain = 0.;
  // -(lambda * de * dp): Eigen folds the scalar into the product's left
  // operand; each entry is (lambda*D_ik) summed over k in order.
  auto corr = [&](double lam, const double* dpv) {
    for (int i = 0; i < 6; ++i) {
      double s = (lam * m.de[i][0]) * dpv[0];
      for (int k = 1; k < 6; ++k) s = s + (lam * m.de[i][k]) * dpv[k];
      upd[i] += -s;
    }
  };
  if (std::fabs(yield) < Tolerance) {
    corr(lambda, dp);
    dpdstrain = lambda * dp_dq;
  } else {
    corr(lambda_trial, dp_t);
    dpdstrain = lambda_trial * dp_dq_trial;
  }
  const int itr_max = 100;
  for (int itr = 0; itr < itr_max; ++itr) {
    mc_invariants(upd, sv);
    ytt = mc_yield_state(yft, sv);
    if (yft[0] < Tolerance && yft[1] < Tolerance) break;
    mc_df_dp(m, ytt, sv, upd, df_t, dp_t, &dp_dq_trial, &softening_trial);
    if (ytt == FS_TENSILE) yield_trial = yft[0];
    if (ytt == FS_SHEAR) yield_trial = yft[1];
    de_tmul(m.de, df_t, tmp);
    lambda_trial = yield_trial / (dot6(tmp, dp_t) + softening_trial);
    corr(lambda_trial, dp_t);
    dpdstrain += lambda_trial * dp_dq_trial;
  }
  mc_invariants(upd, sv);
  sv[SV_PDS] += dpdstrain;
  for (int i = 0; i < 6; ++i) out[i] = upd[i];
}

// ------------------------------ Elements -----------------------------------
// quadrilateral_element.tcc:11-43 ; hexahedron_element.tcc:19-73
static void shapefn(int dim, const double* xi, double* N) {
  if (dim == 2) {
    N[0] = 0.25 * (1 - xi[0]) * (1 - xi[1]);
    N[1] = 0.25 * (1 + xi[0]) * (1 - xi[1]);
    N[2] = 0.25 * (1 + xi[0]) * (1 + xi[1]);
    N[3] = 0.25 * (1 - xi[0]) * (1 + xi[1]);
  } else {
    N[0] = 0.125 * (1 - xi[0]) * (1 - xi[1]) * (1 - xi[2]);
    N[1] = 0.125 * (1 + xi[0]) * (1 - xi[1]) * (1 - xi[2]);
    N[2] = 0.125 * (1 + xi[0]) * (1 + xi[1]) * (1 - xi[2]);
    N[3] = 0.125 * (1 - xi[0]) * (1 + xi[1]) * (1 - xi[2]);
    N[4] = 0.125 * (1 - xi[0]) * (1 - xi[1]) * (1 + xi[2]);
    N[5] = 0.125 * (1 + xi[0]) * (1 - xi[1]) * (1 + xi[2]);
    N[6] = 0.125 * (1 + xi[0]) * (1 + xi[1]) * (1 + xi[2]);
    N[7] = 0.125 * (1 - xi[0]) * (1 + xi[1]) * (1 + xi[2]);
  }
}

// grad[i*3 + a] = dN_i / dxi_a
static void grad_shapefn(int dim, const double* xi, double* g) {
  if (dim == 2) {
    g[0 * 3 + 0] = -0.25 * (1 - xi[1]);
    g[1 * 3 + 0] = 0.25 * (1 - xi[1]);
    g[2 * 3 + 0] = 0.25 * (1 + xi[1]);
    g[3 * 3 + 0] = -0.25 * (1 + xi[1]);
    g[0 * 3 + 1] = -0.25 * (1 - xi[0]);
    g[1 * 3 + 1] = -0.25 * (1 + xi[0]);
    g[2 * 3 + 1] = 0.25 * (1 + xi[0]);
    g[3 * 3 + 1] = 0.25 * (1 - xi[0]);
  } else {
    g[0 * 3 + 0] = -0.125 * (1 - xi[1]) * (1 - xi[2]);
    g[1 * 3 + 0] = 0.125 * (1 - xi[1]) * (1 - xi[2]);
    g[2 * 3 + 0] = 0.125 * (1 + xi[1]) * (1 - xi[2]);
    g[3 * 3 + 0] = -0.125 * (1 + xi[1]) * (1 - xi[2]);
    g[4 * 3 + 0] = -0.125 * (1 - xi[1]) * (1 + xi[2]);
    g[5 * 3 + 0] = 0.125 * (1 - xi[1]) * (1 + xi[2]);
    g[6 * 3 + 0] = 0.125 * (1 + xi[1]) * (1 + xi[2]);
    g[7 * 3 + 0] = -0.125 * (1 + xi[1]) * (1 + xi[2]);

    g[0 * 3 + 1] = -0.125 * (1 - xi[0]) * (1 - xi[2]);
    g[1 * 3 + 1] = -0.125 * (1 + xi[0]) * (1 - xi[2]);
    g[2 * 3 + 1] = 0.125 * (1 + xi[0]) * (1 - xi[2]);
    g[3 * 3 + 1] = 0.125 * (1 - xi[0]) * (1 - xi[2]);
    g[4 * 3 + 1] = -0.125 * (1 - xi[0]) * (1 + xi[2]);
    g[5 * 3 + 1] = -0.125 * (1 + xi[0]) * (1 + xi[2]);
    g[6 * 3 + 1] = 0.125 * (1 + xi[0]) * (1 + xi[2]);
    g[7 * 3 + 1] = 0.125 * (1 - xi[0]) * (1 + xi[2]);

    g[0 * 3 + 2] = -0.125 * (1 - xi[0]) * (1 - xi[1]);
    g[1 * 3 + 2] = -0.125 * (1 + xi[0]) * (1 - xi[1]);
    g[2 * 3 + 2] = -0.125 * (1 + xi[0]) * (1 + xi[1]);
    g[3 * 3 + 2] = -0.125 * (1 - xi[0]) * (1 + xi[1]);
    g[4 * 3 + 2] = 0.125 * (1 - xi[0]) * (1 - xi[1]);
    g[5 * 3 + 2] = 0.125 * (1 + xi[0]) * (1 - xi[1]);
    g[6 * 3 + 2] = 0.125 * (1 + xi[0]) * (1 + xi[1]);
    g[7 * 3 + 2] = 0.125 * (1 - xi[0]) * (1 + xi[1]);
  }
}

// Inverse of a 2x2 / 3x3 matrix by the adjugate over the determinant
// (cofactor expansion along the first column).
static void inverse(int dim, const double J[3][3], double Ji[3][3]) {
  if (dim == 2) {
    const double det = J[0][0] * J[1][1] - J[1][0] * J[0][1];
    const double invdet = 1.0 / det;
    Ji[0][0] = J[1][1] * invdet;
    Ji[1][0] = -J[1][0] * invdet;
    Ji[0][1] = -J[0][1] * invdet;
    Ji[1][1] = J[0][0] * invdet;
  } else {
    // cofactors c_ij of J
    double c[3][3];
    c[0][0] = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    c[1][0] = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    c[2][0] = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    c[0][1] = J[2][1] * J[0][2] - J[2][2] * J[0][1];
    c[1][1] = J[2][2] * J[0][0] - J[2][0] * J[0][2];
    c[2][1] = J[2][0] * J[0][1] - J[2][1] * J[0][0];
    c[0][2] = J[0][1] * J[1][2] - J[0][2] * J[1][1];
    c[1][2] = J[0][2] * J[1][0] - J[0][0] * J[1][2];
    c[2][2] = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    // c[i][j] is here the (j,i) cofactor, i.e. the adjugate entry adj(i,j).
    const double det = (J[0][0] * c[0][0] + J[1][0] * c[0][1]) + J[2][0] * c[0][2];
    const double invdet = 1.0 / det;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) Ji[i][j] = c[i][j] * invdet;
  }
}

// mpm::geometry::rotation_matrix (src/geometry.cc:6-56): the rotation of
// Euler angles (alpha, beta) in 2D, (alpha, beta, gamma) in 3D.
static void rotation_matrix(int dim, const double* ang, double R[3][3]) {
  const double ca = mth::cos(ang[0]), sa = mth::sin(ang[0]);
  const double cb = mth::cos(ang[1]), sb = mth::sin(ang[1]);
  const double cos_alpha_cos_beta = ca * cb;
  const double cos_alpha_sin_beta = ca * sb;
  const double sin_alpha_cos_beta = sa * cb;
  const double sin_alpha_sin_beta = sa * sb;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) R[i][j] = i == j ? 1. : 0.;
  if (dim == 2) {
    R[0][0] = cos_alpha_cos_beta - sin_alpha_sin_beta;
    R[0][1] = -cos_alpha_sin_beta - sin_alpha_cos_beta;
    R[1][0] = sin_alpha_cos_beta + cos_alpha_sin_beta;
    R[1][1] = -sin_alpha_sin_beta + cos_alpha_cos_beta;
    return;
  }
  const double sg = mth::sin(ang[2]), cg = mth::cos(ang[2]);
  const double cos_beta_sin_gamma = cb * sg;
  const double sin_beta_sin_gamma = sb * sg;
  const double cos_alpha_sin_gamma = ca * sg;
  const double sin_alpha_sin_gamma = sa * sg;
  R[0][0] = cos_alpha_cos_beta - sin_alpha_sin_beta * cg;
  R[0][1] = -cos_alpha_sin_beta - sin_alpha_cos_beta * cg;
  R[0][2] = sin_alpha_sin_gamma;
  R[1][0] = sin_alpha_cos_beta + cos_alpha_sin_beta * cg;
  R[1][1] = -sin_alpha_sin_beta + cos_alpha_cos_beta * cg;
  R[1][2] = -cos_alpha_sin_gamma;
  R[2][0] = sin_beta_sin_gamma;
  R[2][1] = cos_beta_sin_gamma;
  R[2][2] = cg;
}

// dn_dx (quadrilateral_element.tcc:295-311, hexahedron_element.tcc:307-321):
// J = grad^T X (sequential over nodes), dN/dx = grad * (J^-1)^T.
static void dn_dx(int dim, const double* xi, const double X[8][3], double out[8][3]) {
  const int nn = dim == 2 ? 4 : 8;
  double g[24];
  grad_shapefn(dim, xi, g);
  double J[3][3] = {{0}};
  for (int a = 0; a < dim; ++a)
    for (int b = 0; b < dim; ++b) {
      double s = g[0 * 3 + a] * X[0][b];
      for (int i = 1; i < nn; ++i) s = s + g[i * 3 + a] * X[i][b];
      J[a][b] = s;
    }
  double Ji[3][3];
  inverse(dim, J, Ji);
  for (int i = 0; i < nn; ++i)
    for (int a = 0; a < dim; ++a) {
      double s = g[i * 3 + 0] * Ji[a][0];
      for (int b = 1; b < dim; ++b) s = s + g[i * 3 + b] * Ji[a][b];
      out[i][a] = s;
    }
}

static inline double norm_diff(int dim, const double* a, const double* b) {
  double s = (a[0] - b[0]) * (a[0] - b[0]);
  for (int k = 1; k < dim; ++k) s = s + (a[k] - b[k]) * (a[k] - b[k]);
  return std::sqrt(s);
}

// ------------------ isoparametric local coordinates -----------------------
// Cell::transform_real_to_unit_cell (cell.tcc:617-828) for quad4 / hex8.
// Vertices of the unit cell (QuadrilateralElement / HexahedronElement
// unit_cell_coordinates): quad counter-clockwise from (-1,-1); hex the same
// square at zeta = -1, then at zeta = +1.
static void unit_cell(int dim, double U[8][3]) {
  static const double sq[4][2] = {{-1., -1.}, {1., -1.}, {1., 1.}, {-1., 1.}};
  const int nn = dim == 2 ? 4 : 8;
  for (int i = 0; i < nn; ++i) {
    U[i][0] = sq[i % 4][0];
    U[i][1] = sq[i % 4][1];
    U[i][2] = i < 4 ? -1. : 1.;
  }
}

// r = X^T N(xi) - p (nodal_coords * shapefn - point), sequential over nodes.
static void iso_residual(int dim, const double X[8][3], const double* xi, const double* p,
                         double* r) {
  const int nn = dim == 2 ? 4 : 8;
  double N[8];
  shapefn(dim, xi, N);
  for (int k = 0; k < dim; ++k) {
    double s = X[0][k] * N[0];
    for (int i = 1; i < nn; ++i) s = s + X[i][k] * N[i];
    r[k] = s - p[k];
  }
}

static inline double sqnorm(int dim, const double* v) {
  double s = v[0] * v[0];
  for (int k = 1; k < dim; ++k) s = s + v[k] * v[k];
  return s;
}

// The bilinear map of a quad4 inverted in closed form
// (QuadrilateralElement<2,4>::natural_coordinates_analytical,
// quadrilateral_element.tcc:523-754, after Zhao et al. 1999). With
// 4x = a1 + a2 xi + a3 eta + a4 xi eta (and b for y), the point solves
// a2 xi + a3 eta + a4 xi eta = c1, b2 xi + b3 eta + b4 xi eta = c2. A
// parallelogram (a4 = b4 = 0) is a linear system; otherwise eliminating the
// product gives the line p xi + q eta = k, and substituting it a quadratic
// in xi (q != 0) or xi itself (q = 0). The root with both coordinates in
// [-1, 1] is returned (the one with the smaller xi first); none: DBL_MAX.
// This is a derivation of the same closed form, not the reference's Sympy
// expression: the two round differently in the last bits.
static void q4_natural_coordinates(const double X[8][3], const double* pt, double* xi) {
  const double big = std::numeric_limits<double>::max();
  xi[0] = xi[1] = big;
  const double x1 = X[0][0], y1 = X[0][1], x2 = X[1][0], y2 = X[1][1];
  const double x3 = X[2][0], y3 = X[2][1], x4 = X[3][0], y4 = X[3][1];
  const double a1 = x1 + x2 + x3 + x4, a2 = -x1 + x2 + x3 - x4;
  const double a3 = -x1 - x2 + x3 + x4, a4 = x1 - x2 + x3 - x4;
  const double b1 = y1 + y2 + y3 + y4, b2 = -y1 + y2 + y3 - y4;
  const double b3 = -y1 - y2 + y3 + y4, b4 = y1 - y2 + y3 - y4;
  const double c1 = 4. * pt[0] - a1, c2 = 4. * pt[1] - b1;
  auto inside = [](double u, double v) { return u >= -1. && u <= 1. && v >= -1. && v <= 1.; };
  if (a4 == 0. && b4 == 0.) {
    const double det = a2 * b3 - a3 * b2;
    xi[0] = (c1 * b3 - a3 * c2) / det;
    xi[1] = (a2 * c2 - b2 * c1) / det;
    return;
  }
  const double pl = a2 * b4 - a4 * b2, ql = a3 * b4 - a4 * b3, kl = c1 * b4 - c2 * a4;
  if (ql == 0.) {
    const double u = kl / pl;
    const double den = a3 + a4 * u;
    const double v = den != 0. ? (c1 - a2 * u) / den : (c2 - b2 * u) / (b3 + b4 * u);
    if (inside(u, v)) {
      xi[0] = u;
      xi[1] = v;
    }
    return;
  }
  // A xi^2 + B xi + C = 0
  const double A = -a4 * pl, B = a2 * ql + a4 * kl - a3 * pl, C = a3 * kl - c1 * ql;
  double r[2];
  int nr = 0;
  if (A == 0.) {
    r[nr++] = -C / B;
  } else {
    const double disc = B * B - 4. * A * C;
    if (!(disc >= 0.)) return;
    const double sd = std::sqrt(disc);
    const double qq = -0.5 * (B + (B >= 0. ? sd : -sd));
    const double ra = qq / A;
    const double rb = qq != 0. ? C / qq : ra;
    r[nr++] = ra < rb ? ra : rb;
    r[nr++] = ra < rb ? rb : ra;
  }
  for (int j = 0; j < nr; ++j) {
    const double u = r[j];
    const double v = (kl - pl * u) / ql;
    if (inside(u, v)) {
      xi[0] = u;
      xi[1] = v;
      return;
    }
  }
}

// transform_real_to_unit_cell (cell.tcc:617-828), isoparametric branch:
// the analytical quad4 inverse (2D) with its early exit, the affine guess
// A^-1 (p - b) of the least-squares affine map (A = X^T KA, b = X^T Kb,
// KA = unit vertex / nn, Kb = 1 / nn for the symmetric unit cell) with its
// early exit, then Newton-Raphson from the better of the two with a
// halving line search. The iteration matrix is the reference's
// jacobian_local evaluated on the unit cell's own vertices (cell.tcc:772).
static void transform_real_to_unit_cell(int dim, const double X[8][3], double mean_length,
                                        const double* p, double* xi) {
  const int nn = dim == 2 ? 4 : 8;
  const double big = std::numeric_limits<double>::max();
  for (int k = 0; k < dim; ++k) xi[k] = big;
  const double zero[3] = {0., 0., 0.};
  double an_xi[3] = {big, big, big}, an_res[3] = {big, big, big};
  const double tol2 = 1.0E-16 * mean_length * mean_length;
  if (dim == 2) {
    q4_natural_coordinates(X, p, an_xi);
    bool in_cell = true;
    for (int k = 0; k < dim; ++k)
      if (an_xi[k] < -1. || an_xi[k] > 1. || std::isnan(an_xi[k])) in_cell = false;
    iso_residual(dim, X, an_xi, p, an_res);
    if (sqnorm(dim, an_res) < tol2 && in_cell) {
      for (int k = 0; k < dim; ++k) xi[k] = an_xi[k];
      return;
    }
  }
  (void)zero;
  // affine guess
  double U[8][3];
  unit_cell(dim, U);
  const double inv_nn = 1.0 / nn;
  double A[3][3] = {{0}}, Ai[3][3], b[3];
  for (int k = 0; k < dim; ++k) {
    for (int a = 0; a < dim; ++a) {
      double s = X[0][k] * (U[0][a] * inv_nn);
      for (int i = 1; i < nn; ++i) s = s + X[i][k] * (U[i][a] * inv_nn);
      A[k][a] = s;
    }
    double s = X[0][k] * inv_nn;
    for (int i = 1; i < nn; ++i) s = s + X[i][k] * inv_nn;
    b[k] = p[k] - s;
  }
  inverse(dim, A, Ai);
  double af_xi[3] = {0., 0., 0.}, af_res[3] = {big, big, big};
  bool af_nan = false;
  for (int k = 0; k < dim; ++k) {
    double s = Ai[k][0] * b[0];
    for (int a = 1; a < dim; ++a) s = s + Ai[k][a] * b[a];
    af_xi[k] = s;
    if (std::isnan(s)) af_nan = true;
  }
  if (!af_nan) {
    iso_residual(dim, X, af_xi, p, af_res);
    if (sqnorm(dim, af_res) < tol2) {
      for (int k = 0; k < dim; ++k) xi[k] = af_xi[k];
      return;
    }
  }
  // the better of the two as the starting point
  double g_xi[3] = {0., 0., 0.}, g_res[3] = {big, big, big};
  if (std::sqrt(sqnorm(dim, an_res)) < std::sqrt(sqnorm(dim, af_res))) {
    for (int k = 0; k < dim; ++k) {
      g_xi[k] = an_xi[k];
      g_res[k] = an_res[k];
    }
  } else if (!af_nan) {
    for (int k = 0; k < dim; ++k) {
      g_xi[k] = af_xi[k];
      g_res[k] = af_res[k];
    }
  }
  double nr[3];
  for (int k = 0; k < dim; ++k) {
    nr[k] = g_xi[k];
    if (nr[k] < -1. && nr[k] > -1.001)
      nr[k] = -1. + 1.E-12;
    else if (nr[k] > 1. && nr[k] < 1.001)
      nr[k] = 1. - 1.E-12;
  }
  double nr_res[3] = {big, big, big};
  for (unsigned iter = 0; iter < 10000; ++iter) {
    double g[24], J[3][3] = {{0}}, Ji[3][3];
    grad_shapefn(dim, nr, g);
    for (int a = 0; a < dim; ++a)
      for (int c = 0; c < dim; ++c) {
        double s = g[0 * 3 + a] * U[0][c];
        for (int i = 1; i < nn; ++i) s = s + g[i * 3 + a] * U[i][c];
        J[a][c] = s;
      }
    double det;
    if (dim == 2)
      det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    else
      det = J[0][0] * (J[1][1] * J[2][2] - J[2][1] * J[1][2]) -
            J[1][0] * (J[0][1] * J[2][2] - J[2][1] * J[0][2]) +
            J[2][0] * (J[0][1] * J[1][2] - J[1][1] * J[0][2]);
    if (std::fabs(det) < 1.0E-10)
      for (int k = 0; k < dim; ++k) nr[k] = 0.;
    iso_residual(dim, X, nr, p, nr_res);
    inverse(dim, J, Ji);
    double delta[3];
    for (int k = 0; k < dim; ++k) {
      double s = Ji[k][0] * nr_res[0];
      for (int a = 1; a < dim; ++a) s = s + Ji[k][a] * nr_res[a];
      delta[k] = s;
    }
    double step = 1.;
    const double rn = std::sqrt(sqnorm(dim, nr_res));
    for (int t = 0; t < 10; ++t) {
      double trial[3], tres[3];
      for (int k = 0; k < dim; ++k) trial[k] = nr[k] - step * delta[k];
      iso_residual(dim, X, trial, p, tres);
      if (std::sqrt(sqnorm(dim, tres)) < rn) {
        for (int k = 0; k < dim; ++k) {
          nr[k] = trial[k];
          nr_res[k] = tres[k];
        }
        break;
      } else if (step > 0.05) {
        step /= 2.;
      } else {
        break;
      }
    }
    double sdel[3] = {0., 0., 0.};
    for (int k = 0; k < dim; ++k) sdel[k] = step * delta[k];
    if (std::sqrt(sqnorm(dim, sdel)) < 1.0E-10) break;
    if (std::isnan(nr[0]) || std::isnan(nr[1]))
      for (int k = 0; k < dim; ++k) nr[k] = 0.;
  }
  const bool geo = std::sqrt(sqnorm(dim, g_res)) < std::sqrt(sqnorm(dim, nr_res));
  for (int k = 0; k < dim; ++k) xi[k] = geo ? g_xi[k] : nr[k];
  if (std::isnan(xi[0]) || std::isnan(xi[1]))
    throw std::runtime_error("Local coordinates of xi is NAN");
}

// ---------------------------------------------------------------------------
// Math functions of time-dependent boundary conditions.
// LinearFunction::value (src/functions/linear_function.cc:19-41) and
// SinFunction::value (src/functions/sin_function.cc:17-20).
enum { FN_LINEAR = 0, FN_SIN = 1 };
struct MathFn {
  int type = FN_LINEAR;
  std::vector<double> x, fx;  // Linear: the (x, f(x)) table in its given order
  double x0 = 0., a = 0.;     // Sin: sin(a (x - x0)) on [xmin, xmax], else 0
  double xmin = std::numeric_limits<double>::min(), xmax = std::numeric_limits<double>::max();
  double value(double xin) const {
    if (type == FN_SIN) {
      if ((xin < xmin) || (xin > xmax)) return 0.0;
      return mth::sin(a * (xin - x0));
    }
    double v = std::numeric_limits<double>::quiet_NaN();
    if (x.empty()) throw std::runtime_error("Cannot find the f(x); no linear function is defined");
    if (xin < x.front()) return fx.front();
    if (xin > x.back()) return fx.back();
    for (size_t i = 0; i + 1 < x.size(); ++i) {
      const double f = (xin - x[i]) / (x[i + 1] - x[i]);
      if (f > (-std::numeric_limits<double>::epsilon()) &&
          f < (1.0 + std::numeric_limits<double>::epsilon()))
        return (fx[i] + f * (fx[i + 1] - fx[i]));
    }
    return v;
  }
};

struct Node {
  double x[3] = {0, 0, 0};
  double mass = 0, momentum[3] = {0, 0, 0}, velocity[3] = {0, 0, 0}, acceleration[3] = {0, 0, 0};
  double fext[3] = {0, 0, 0}, fint[3] = {0, 0, 0};
  bool status = false;
  std::map<unsigned, double> vel_c, acc_c;
  std::map<unsigned, int> acc_fn;  // math function of an acceleration constraint
  double acc_now[3] = {0, 0, 0};   // the constraint values at the current time
  bool has_cf = false;             // concentrated force (node.tcc:82-111)
  double cforce[3] = {0, 0, 0};
  int cfn = -1;
  bool friction = false;
  unsigned f_dir = 0;
  int f_sign = 0;
  double f_mu = 0;
  // general (inclined) boundary: constraints act in the frame of R
  // (Node::assign_rotation_matrix, node.h:281-285)
  bool generic = false;
  double R[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
};

struct Cell {
  Index id = 0;
  Index nodes[8];
  double X[8][3];
  double centroid[3], mean_length = 0, volume = 0;
  double dndx_c[8][3];
  std::vector<Index> neighbours;  // ascending (std::set order)
  Index nparticles = 0;
};

struct Particle {
  Index id = 0;
  double x[3] = {0, 0, 0}, u[3] = {0, 0, 0}, v[3] = {0, 0, 0}, xi[3] = {0, 0, 0};
  Index cell = kNoCell;
  Index hint = kNoCell;  // cell id held without a cell pointer (mesh.tcc:1119-1121)
  double N[8], dndx[8][3];
  double mass = 0, volume = 0, density = 0;
  double stress[6] = {0}, strain[6] = {0}, dstrain[6] = {0}, strain_rate[6] = {0};
  double dvol = 0;
  double F[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  double dF[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  int mat = 0;
  double sv[kNsv] = {0};
  double size = 0;                  // pow(volume, 1/dim) at creation (particle.tcc:563-568)
  double B[3][3] = {{0}};           // mapping_matrix_ of the affine velocity updates
  double traction[3] = {0, 0, 0};   // particle.tcc:962-979
  bool has_traction = false;
};

// A particle traction (Traction, traction.h): set (-1: every particle), dir,
// value, math function (-1: none).
struct TractionLoad {
  Index particle;
  int dir;
  double value;
  int fn;
};

enum { SCHEME_USF = 0, SCHEME_USL = 1, SCHEME_MUSL = 2 };
// mpm::VelocityUpdate (data_structures/data_types.h:50), same order
enum { VU_FLIP = 0, VU_PIC = 1, VU_ASFLIP = 2, VU_TPIC = 3, VU_APIC = 4 };

struct Solver {
  int dim = 2, nn = 4;
  int scheme = SCHEME_USF;
  int vu = VU_FLIP;
  double blending = 1.0;
  std::vector<Node> nodes;
  std::vector<Cell> cells;
  std::vector<Particle> particles;
  std::vector<Material> materials;
  double gravity[3] = {0, 0, 0};
  double ramp = 0;
  Index step = 0;
  std::vector<MathFn> functions;
  std::vector<TractionLoad> tractions;
  std::vector<Index> unlocatable;
  bool locate_all = true;  // the analysis key locate_particles (mpm_base.tcc:53-54)
  // the mesh's "isoparametric" key (mpm_base.tcc:768-783): local coordinates
  // by transform_real_to_unit_cell instead of the Cartesian formula
  bool iso = false;
  std::string error;

  // ---------------- mesh -----------------
  // cell.tcc:33-67 initialise, :254-277 centroid/mean length, elements' volume
  void init_cell(Cell& c) {
    const int d = dim;
    for (int k = 0; k < d; ++k) c.centroid[k] = 0.;
    for (int i = 0; i < nn; ++i)
      for (int k = 0; k < d; ++k) c.centroid[k] += c.X[i][k];
    for (int k = 0; k < d; ++k) c.centroid[k] /= nn;
    static const int q_sides[4][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}};
    static const int h_sides[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                                       {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
    const int ns = d == 2 ? 4 : 12;
    c.mean_length = 0.;
    for (int s = 0; s < ns; ++s) {
      const int a = d == 2 ? q_sides[s][0] : h_sides[s][0];
      const int b = d == 2 ? q_sides[s][1] : h_sides[s][1];
      c.mean_length += norm_diff(d, c.X[a], c.X[b]);
    }
    c.mean_length /= ns;
    c.volume = element_volume(c.X);
    const double zero[3] = {0, 0, 0};
    dn_dx(d, zero, c.X, c.dndx_c);
  }

  double element_volume(const double X[8][3]) const {
    if (dim == 2) {  // quadrilateral_element.tcc:473-497
      const double a = norm_diff(2, X[0], X[3]);
      const double b = norm_diff(2, X[2], X[3]);
      const double c = norm_diff(2, X[1], X[2]);
      const double d = norm_diff(2, X[0], X[1]);
      const double p = norm_diff(2, X[0], X[2]);
      const double q = norm_diff(2, X[1], X[3]);
      return 0.25 * std::sqrt(4 * p * p * q * q - sq((a * a + c * c - b * b - d * d)));
    }
    // hexahedron_element.tcc:559-597
    auto sub = [](const double* u, const double* w, double* r) {
      for (int k = 0; k < 3; ++k) r[k] = u[k] - w[k];
    };
    auto cross = [](const double* u, const double* w, double* r) {
      r[0] = u[1] * w[2] - u[2] * w[1];
      r[1] = u[2] * w[0] - u[0] * w[2];
      r[2] = u[0] * w[1] - u[1] * w[0];
    };
    auto dot = [](const double* u, const double* w) {
      return (u[0] * w[0] + u[1] * w[1]) + u[2] * w[2];
    };
    const double *a = X[7], *b = X[6], *c = X[2], *d = X[3], *e = X[4], *f = X[5], *g = X[1],
                 *h = X[0];
    double t1[3], t2[3], t3[3], t4[3], cr1[3], cr2[3], cr3[3], cr4[3], cr5[3], cr6[3], s[3];
    double bd[3], ca[3], eb[3], fa[3], de[3], ha[3], cg[3], cf[3], fg[3], hf[3], hg[3], hc[3];
    sub(b, d, bd); sub(c, a, ca); sub(e, b, eb); sub(f, a, fa); sub(d, e, de); sub(h, a, ha);
    sub(c, g, cg); sub(c, f, cf); sub(f, g, fg); sub(h, f, hf); sub(h, g, hg); sub(h, c, hc);
    cross(bd, ca, cr1);  // (b-d)x(c-a)
    cross(eb, fa, cr2);  // (e-b)x(f-a)
    cross(de, ha, cr3);  // (d-e)x(h-a)
    cross(cg, cf, cr4);  // (c-g)x(c-f)
    cross(fg, hf, cr5);  // (f-g)x(h-f)
    cross(hg, hc, cr6);  // (h-g)x(h-c)
    sub(a, g, t1); sub(b, g, t2); sub(e, g, t3); sub(d, g, t4);
    for (int k = 0; k < 3; ++k) s[k] = (cr1[k] + cr2[k]) + cr3[k];
    const double v1 = (1.0 / 12) * dot(t1, s);
    for (int k = 0; k < 3; ++k) s[k] = cr1[k] + cr4[k];
    const double v2 = (1.0 / 12) * dot(t2, s);
    for (int k = 0; k < 3; ++k) s[k] = cr2[k] + cr5[k];
    const double v3 = (1.0 / 12) * dot(t3, s);
    for (int k = 0; k < 3; ++k) s[k] = cr3[k] + cr6[k];
    const double v4 = (1.0 / 12) * dot(t4, s);
    return ((v1 + v2) + v3) + v4;
  }

  void set_mesh(Index nnodes, const double* coords, Index ncells, const Index* cn) {
    nodes.assign(nnodes, Node());
    for (Index i = 0; i < nnodes; ++i)
      for (int k = 0; k < dim; ++k) nodes[i].x[k] = coords[i * dim + k];
    cells.assign(ncells, Cell());
    for (Index c = 0; c < ncells; ++c) {
      Cell& cl = cells[c];
      cl.id = c;
      for (int i = 0; i < nn; ++i) {
        const Index n = cn[c * nn + i];
        if (n < 0 || n >= nnodes) throw std::runtime_error("cell node id out of range");
        cl.nodes[i] = n;
        for (int k = 0; k < dim; ++k) cl.X[i][k] = nodes[n].x[k];
        for (int k = dim; k < 3; ++k) cl.X[i][k] = 0.;
      }
      init_cell(cl);
      if (cl.volume <= 0) throw std::runtime_error("Negative or zero volume cell");
    }
    // mesh.tcc:287-307 find_cell_neighbours
    std::vector<std::vector<Index>> node_cells(nnodes);
    for (Index c = 0; c < ncells; ++c)
      for (int i = 0; i < nn; ++i) node_cells[cells[c].nodes[i]].push_back(c);
    for (Index c = 0; c < ncells; ++c) {
      std::set<Index> nb;
      for (int i = 0; i < nn; ++i)
        for (Index o : node_cells[cells[c].nodes[i]])
          if (o != c) nb.insert(o);
      cells[c].neighbours.assign(nb.begin(), nb.end());
    }
  }

  // ---------------- locate (cell.tcc:324-380, 423-606) -----------------
  bool is_point_in_cell(const Cell& c, const double* p, double* xi) const {
    for (int k = 0; k < dim; ++k) xi[k] = std::numeric_limits<double>::max();
    // approx_point_in_cell
    if (!(norm_diff(dim, p, c.centroid) < (c.mean_length * 2.))) return false;
    if (iso) {
      transform_real_to_unit_cell(dim, c.X, c.mean_length, p, xi);
      return clamp_unit(xi);
    }
    // local_coordinates_point (Cartesian)
    double len[3];
    len[0] = norm_diff(dim, c.X[0], c.X[1]);
    len[1] = norm_diff(dim, c.X[1], c.X[2]);
    if (dim == 3) len[2] = norm_diff(dim, c.X[1], c.X[5]);
    double centre[3];
    for (int k = 0; k < dim; ++k) {
      double s = c.X[0][k] + c.X[1][k];
      for (int i = 2; i < nn; ++i) s = s + c.X[i][k];
      centre[k] = s / (dim == 2 ? 4.0 : (double)nn);
    }
    for (int k = 0; k < dim; ++k) xi[k] = 2. * (p[k] - centre[k]) / len[k];
    return clamp_unit(xi);
  }

  // The unit-cell test of is_point_in_cell (cell.tcc:361-377): inside
  // [-1, 1] (and not NaN), then kept DBL_EPSILON inside the bounds.
  bool clamp_unit(double* xi) const {
    bool status = true;
    const double tol = kEps;
    for (int k = 0; k < dim; ++k) {
      if (xi[k] < -1. || xi[k] > 1. || std::isnan(xi[k]))
        status = false;
      else {
        if (xi[k] < -1. + tol) xi[k] = -1. + tol;
        if (xi[k] > 1. - tol) xi[k] = 1. - tol;
      }
    }
    return status;
  }

  void assign_cell_xi(Particle& p, Index c, const double* xi) {
    if (p.cell != kNoCell) cells[p.cell].nparticles -= 1;
    p.cell = c;
    p.hint = kNoCell;
    for (int k = 0; k < dim; ++k) p.xi[k] = xi[k];
    cells[c].nparticles += 1;
  }

  // mesh.tcc:1116-1153
  bool locate_particle(Particle& p) {
    double xi[3];
    if (p.cell != kNoCell || p.hint != kNoCell) {
      if (p.cell != kNoCell && is_point_in_cell(cells[p.cell], p.x, xi)) {
        for (int k = 0; k < dim; ++k) p.xi[k] = xi[k];
        return true;
      }
      const Index cid = p.cell != kNoCell ? p.cell : p.hint;
      for (Index nb : cells[cid].neighbours) {
        if (is_point_in_cell(cells[nb], p.x, xi)) {
          assign_cell_xi(p, nb, xi);
          return true;
        }
      }
    }
    for (Index c = 0; c < (Index)cells.size(); ++c) {
      if (is_point_in_cell(cells[c], p.x, xi)) {
        assign_cell_xi(p, c, xi);
        return true;
      }
    }
    return false;
  }

  // mpm_scheme.tcc:245-261 -> mesh.tcc:1097-1112 (sequential sweep)
  void locate_particles() {
    unlocatable.clear();
    std::vector<Index> lost;  // vector positions
    for (Index i = 0; i < (Index)particles.size(); ++i)
      if (!locate_particle(particles[i])) {
        unlocatable.push_back(particles[i].id);
        lost.push_back(i);
      }
    if (!unlocatable.empty() && locate_all) {
      std::string msg = "Particle(s) outside the mesh domain: ";
      for (Index id : unlocatable) msg += std::to_string(id) + " ";
      throw std::runtime_error(msg);
    }
    if (!lost.empty()) remove_particles(lost);
  }

  // Mesh::remove_particle (mesh.tcc:658-665): the particle leaves its cell
  // (Particle::remove_cell) and the particle vector (std::remove keeps the
  // order of the others, vector.tcc:23-30). Loads given by vector position
  // follow their particles; those of removed particles are dropped.
  void remove_particles(const std::vector<Index>& lost) {
    const Index np = (Index)particles.size();
    std::vector<Index> remap(np, 0);
    std::vector<char> gone(np, 0);
    for (Index i : lost) gone[i] = 1;
    std::vector<Particle> kept;
    kept.reserve(np - (Index)lost.size());
    for (Index i = 0; i < np; ++i) {
      if (gone[i]) {
        if (particles[i].cell != kNoCell) cells[particles[i].cell].nparticles -= 1;
        remap[i] = -1;
        continue;
      }
      remap[i] = (Index)kept.size();
      kept.push_back(particles[i]);
    }
    particles.swap(kept);
    std::vector<PVel> pv;
    for (const auto& c : pvel)
      if (c.particle < 0 || (c.particle < np && remap[c.particle] >= 0))
        pv.push_back({c.particle < 0 ? c.particle : remap[c.particle], c.dir, c.value});
    pvel.swap(pv);
    std::vector<TractionLoad> tl;
    for (const auto& t : tractions)
      if (t.particle < 0 || (t.particle < np && remap[t.particle] >= 0))
        tl.push_back({t.particle < 0 ? t.particle : remap[t.particle], t.dir, t.value, t.fn});
    tractions.swap(tl);
  }

  // ---------------- particles init -----------------
  void add_particle_state(Particle& p) {
    if (p.mat < 0 || p.mat >= (int)materials.size()) throw std::runtime_error("bad material id");
  }

  // mesh.tcc:508-576 generate_material_points + cell.tcc:91-143 generate_points
  // + particle.tcc:589-626 compute_volume / compute_mass.
  // Gauss points of the reference's quadratures in their order
  // (quadrilateral_quadrature.tcc, hexahedron_quadrature.tcc):
  //  1: centre; 2: the element's node order; 3 (quad): corners, edge
  //  midpoints (-y, +x, +y, -x), centre; 3 (hex): xi0 slowest, xi2 fastest;
  //  4 (quad): xi0 slowest; 4 (hex): xi2 slowest, then xi0, then xi1.
  int gauss_points(int nppd, double (*qp)[3]) const {
    for (int i = 0; i < 64; ++i) qp[i][0] = qp[i][1] = qp[i][2] = 0.;
    if (nppd == 1) return 1;
    if (nppd == 2) {
      const double g = 1. / std::sqrt(3.);
      const int sg[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                            {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
      const int n = dim == 2 ? 4 : 8;
      for (int i = 0; i < n; ++i)
        for (int k = 0; k < dim; ++k) qp[i][k] = sg[i][k] * g;
      return n;
    }
    if (nppd == 3) {
      const double a = std::sqrt(3. / 5.);
      const double c[3] = {-a, 0., a};
      if (dim == 2) {
        const int idx[9][2] = {{0, 0}, {2, 0}, {2, 2}, {0, 2}, {1, 0},
                               {2, 1}, {1, 2}, {0, 1}, {1, 1}};
        for (int i = 0; i < 9; ++i) {
          qp[i][0] = c[idx[i][0]];
          qp[i][1] = c[idx[i][1]];
        }
        return 9;
      }
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          for (int k = 0; k < 3; ++k) {
            const int q = 9 * i + 3 * j + k;
            qp[q][0] = c[i];
            qp[q][1] = c[j];
            qp[q][2] = c[k];
          }
      return 27;
    }
    if (nppd == 4) {
      const double s65 = std::sqrt(6. / 5.);
      const double a = std::sqrt(3. / 7. - 2. / 7. * s65);
      const double b = std::sqrt(3. / 7. + 2. / 7. * s65);
      const double g[4] = {-b, -a, a, b};
      if (dim == 2) {
        for (int i = 0; i < 4; ++i)
          for (int j = 0; j < 4; ++j) {
            qp[4 * i + j][0] = g[i];
            qp[4 * i + j][1] = g[j];
          }
        return 16;
      }
      for (int k = 0; k < 4; ++k)
        for (int i = 0; i < 4; ++i)
          for (int j = 0; j < 4; ++j) {
            const int q = 16 * k + 4 * i + j;
            qp[q][0] = g[i];
            qp[q][1] = g[j];
            qp[q][2] = g[k];
          }
      return 64;
    }
    throw std::runtime_error("gauss generator supports nparticles_per_dir 1 to 4");
  }

  void generate_gauss(int nppd, Index nset, const Index* cset, int mat) {
    if (mat < 0 || mat >= (int)materials.size()) throw std::runtime_error("bad material id");
    const Index before = (Index)particles.size();
    double qp[64][3];
    const int nq = gauss_points(nppd, qp);
    const Index nc = nset < 0 ? (Index)cells.size() : nset;
    for (Index ic = 0; ic < nc; ++ic) {
      const Index c = nset < 0 ? ic : cset[ic];
      Cell& cl = cells[c];
      for (int q = 0; q < nq; ++q) {
        double N[8];
        shapefn(dim, qp[q], N);
        Particle p;
        p.id = (Index)particles.size();
        for (int k = 0; k < dim; ++k) {
          double s = cl.X[0][k] * N[0];
          for (int i = 1; i < nn; ++i) s = s + cl.X[i][k] * N[i];
          p.x[k] = s;
        }
        double xi[3];
        if (!is_point_in_cell(cl, p.x, xi)) continue;  // generate_points status
        p.mat = mat;
        if (materials[mat].type == kMC) mc_initialise_state(materials[mat], p.sv);
        particles.push_back(p);
        assign_cell_xi(particles.back(), c, xi);
      }
    }
    if ((Index)particles.size() == before) throw std::runtime_error("No particles were generated!");
    // locate_particles_mesh (all particles already hold a cell)
    locate_particles();
    // compute_volume then compute_mass
    for (Index i = before; i < (Index)particles.size(); ++i) {
      Particle& p = particles[i];
      const Cell& cl = cells[p.cell];
      p.volume = cl.volume / cl.nparticles;
      p.size = std::pow(p.volume, static_cast<double>(1. / dim));
      p.density = materials[p.mat].density;
      p.mass = p.volume * p.density;
    }
  }

  // Every particle: Particle::compute_volume (cell volume / particles in the
  // cell, particle.tcc:589-595) then compute_mass (volume * the material's
  // density, :618-626), as MPMBase::initialise_particles and
  // MPMExplicit::solve run them over all particles (mpm_base.tcc:413-415,
  // mpm_explicit.tcc:92-95).
  void compute_volume_mass() {
    for (auto& p : particles) {
      if (p.cell == kNoCell) throw std::runtime_error("particle without a cell");
      const Cell& cl = cells[p.cell];
      p.volume = cl.volume / cl.nparticles;
      p.size = std::pow(p.volume, static_cast<double>(1. / dim));
      p.density = materials[p.mat].density;
      p.mass = p.volume * p.density;
    }
  }

  // ---------------- step phases -----------------
  // particle.tcc:537-554 compute_shapefn
  void compute_shapefn(Particle& p) {
    const Cell& c = cells[p.cell];
    shapefn(dim, p.xi, p.N);
    dn_dx(dim, p.xi, c.X, p.dndx);
  }

  // node.tcc:56-65, 149-162 (mass_ * factor + mass with factor 1.0)
  inline void add(double& dst, double v) {
    ORC_PRAGMA(omp atomic)
    dst += v;
  }

  // particle.tcc:630-658
  // particle.tcc:630-658; APIC and ASFLIP map with the affine velocity
  // v + B D^-1 (x_i - x_p), D = sum_i N_i (x_i - x_p)(x_i - x_p)^T
  // (map_mass_momentum_to_nodes_affine, :661-696), TPIC with v + B (x_i - x_p)
  // (map_mass_momentum_to_nodes_taylor, :699-724)
  void map_mass_momentum(const Particle& p) {
    const Cell& c = cells[p.cell];
    const bool affine = vu == VU_APIC || vu == VU_ASFLIP;
    double C[3][3] = {{0}};
    if (affine) {
      double D[3][3] = {{0}}, Di[3][3];
      for (int i = 0; i < nn; ++i) {
        const Node& n = nodes[c.nodes[i]];
        double r[3];
        for (int k = 0; k < dim; ++k) r[k] = n.x[k] - p.x[k];
        for (int a = 0; a < dim; ++a)
          for (int b = 0; b < dim; ++b) D[a][b] = D[a][b] + (p.N[i] * r[a]) * r[b];
      }
      inverse(dim, D, Di);
      for (int a = 0; a < dim; ++a)
        for (int b = 0; b < dim; ++b) {
          double t = p.B[a][0] * Di[0][b];
          for (int k = 1; k < dim; ++k) t = t + p.B[a][k] * Di[k][b];
          C[a][b] = t;
        }
    } else if (vu == VU_TPIC) {
      for (int a = 0; a < dim; ++a)
        for (int b = 0; b < dim; ++b) C[a][b] = p.B[a][b];
    }
    for (int i = 0; i < nn; ++i) {
      Node& n = nodes[c.nodes[i]];
      add(n.mass, p.mass * p.N[i]);
      if (!affine && vu != VU_TPIC) {
        for (int k = 0; k < dim; ++k) add(n.momentum[k], p.mass * p.N[i] * p.v[k]);
        continue;
      }
      double r[3], mv[3];
      for (int k = 0; k < dim; ++k) r[k] = n.x[k] - p.x[k];
      for (int a = 0; a < dim; ++a) {
        double t = C[a][0] * r[0];
        for (int k = 1; k < dim; ++k) t = t + C[a][k] * r[k];
        mv[a] = p.v[a] + t;
      }
      for (int k = 0; k < dim; ++k) add(n.momentum[k], p.mass * p.N[i] * mv[k]);
    }
  }

  // The mapping matrices of the affine updates, |B_ij| < 1e-15 -> 0:
  // compute_affine_mapping_matrix (particle.tcc:1687-1708),
  // B = sum_i N_i v_i (x_i - x_p)^T, and compute_velocity_gradient
  // (:1662-1684), B = sum_i v_i dN_i/dx^T.
  void mapping_matrix(Particle& p, bool gradient) {
    const Cell& c = cells[p.cell];
    double B[3][3] = {{0}};
    for (int i = 0; i < nn; ++i) {
      const Node& n = nodes[c.nodes[i]];
      for (int a = 0; a < dim; ++a)
        for (int b = 0; b < dim; ++b) {
          const double t = gradient ? n.velocity[a] * p.dndx[i][b]
                                    : (p.N[i] * n.velocity[a]) * (n.x[b] - p.x[b]);
          B[a][b] = B[a][b] + t;
        }
    }
    for (int a = 0; a < dim; ++a)
      for (int b = 0; b < dim; ++b) p.B[a][b] = std::fabs(B[a][b]) < 1.E-15 ? 0. : B[a][b];
  }

  // compute_asflip_beta (particle.tcc:1711-1730): 0 next to a node with a
  // velocity constraint or where the deformation gradient increment of the
  // current nodal velocities compresses (det < 1), else 1
  double asflip_beta(const Particle& p, double dt) {
    const Cell& c = cells[p.cell];
    for (int i = 0; i < nn; ++i)
      if (!nodes[c.nodes[i]].vel_c.empty()) return 0.0;
    double R[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int i = 0; i < nn; ++i) {
      const Node& n = nodes[c.nodes[i]];
      for (int a = 0; a < dim; ++a)
        for (int b = 0; b < dim; ++b) R[a][b] = R[a][b] + (n.velocity[a] * p.dndx[i][b]) * dt;
    }
    for (int a = 0; a < dim; ++a)
      for (int b = 0; b < dim; ++b) {
        if (a != b && std::fabs(R[a][b]) < 1.E-15) R[a][b] = 0.;
        if (a == b && std::fabs(R[a][b] - 1.) < 1.E-15) R[a][b] = 1.;
      }
    const double det = R[0][0] * (R[1][1] * R[2][2] - R[2][1] * R[1][2]) -
                       R[1][0] * (R[0][1] * R[2][2] - R[2][1] * R[0][2]) +
                       R[2][0] * (R[0][1] * R[1][2] - R[1][1] * R[0][2]);
    return det < 1.0 ? 0.0 : 1.0;
  }

  // y = A x for a dim x dim matrix (sequential over k)
  void matvec(const double A[3][3], const double* x, double* y) const {
    for (int i = 0; i < dim; ++i) {
      double s = A[i][0] * x[0];
      for (int k = 1; k < dim; ++k) s = s + A[i][k] * x[k];
      y[i] = s;
    }
  }

  // node.tcc:370-404: per constraint in direction order; on a general
  // boundary velocity and acceleration go to the local frame (R^-1), take
  // the constraint there and come back (R)
  void apply_velocity_constraints(Node& n) {
    for (const auto& c : n.vel_c) {
      const unsigned direction = c.first % dim;
      if (!n.generic) {
        n.velocity[direction] = c.second;
        n.acceleration[direction] = 0.;
      } else {
        double Ri[3][3], lv[3], la[3];
        inverse(dim, n.R, Ri);
        matvec(Ri, n.velocity, lv);
        matvec(Ri, n.acceleration, la);
        lv[direction] = c.second;
        la[direction] = 0.;
        matvec(n.R, lv, n.velocity);
        matvec(n.R, la, n.acceleration);
      }
    }
  }

  // node.tcc:236-252
  void compute_velocity(Node& n) {
    const double tolerance = 1.E-16;
    if (n.mass > tolerance) {
      for (int k = 0; k < dim; ++k) n.velocity[k] = n.momentum[k] / n.mass;
      for (int k = 0; k < dim; ++k)
        if (std::abs(n.velocity[k]) < 1.E-15) n.velocity[k] = 0.;
    }
    apply_velocity_constraints(n);
  }

  // particle.tcc:1605-1651 (2D block / 3D full), then :1643-1651
  void update_dF(Particle& p, double dt) {
    const Cell& c = cells[p.cell];
    double R[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int i = 0; i < nn; ++i) {
      const Node& n = nodes[c.nodes[i]];
      for (int a = 0; a < dim; ++a)
        for (int b = 0; b < dim; ++b) R[a][b] = R[a][b] + (n.velocity[a] * p.dndx[i][b]) * dt;
    }
    for (int a = 0; a < dim; ++a)
      for (int b = 0; b < dim; ++b) {
        if (a != b && std::fabs(R[a][b]) < 1.E-15) R[a][b] = 0.;
        if (a == b && std::fabs(R[a][b] - 1.) < 1.E-15) R[a][b] = 1.;
      }
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) p.dF[i][j] = R[i][j];
  }

  // particle.tcc:819-858
  void strain_rate(const Particle& p, const double dndx[8][3], double* sr) {
    const Cell& c = cells[p.cell];
    for (int k = 0; k < 6; ++k) sr[k] = 0.;
    for (int i = 0; i < nn; ++i) {
      const double* v = nodes[c.nodes[i]].velocity;
      if (dim == 2) {
        sr[0] += dndx[i][0] * v[0];
        sr[1] += dndx[i][1] * v[1];
        sr[3] += dndx[i][1] * v[0] + dndx[i][0] * v[1];
      } else {
        sr[0] += dndx[i][0] * v[0];
        sr[1] += dndx[i][1] * v[1];
        sr[2] += dndx[i][2] * v[2];
        sr[3] += dndx[i][1] * v[0] + dndx[i][0] * v[1];
        sr[4] += dndx[i][2] * v[1] + dndx[i][1] * v[2];
        sr[5] += dndx[i][2] * v[0] + dndx[i][0] * v[2];
      }
    }
    if (dim == 2) {
      if (std::fabs(sr[0]) < 1.E-15) sr[0] = 0.;
      if (std::fabs(sr[1]) < 1.E-15) sr[1] = 0.;
      if (std::fabs(sr[3]) < 1.E-15) sr[3] = 0.;
    } else {
      for (int k = 0; k < 6; ++k)
        if (std::fabs(sr[k]) < 1.E-15) sr[k] = 0.;
    }
  }

  // particle.tcc:861-877
  void compute_strain(Particle& p, double dt) {
    strain_rate(p, p.dndx, p.strain_rate);
    for (int k = 0; k < 6; ++k) p.dstrain[k] = p.strain_rate[k] * dt;
    for (int k = 0; k < 6; ++k) p.strain[k] += p.dstrain[k];
    double src[6];
    strain_rate(p, cells[p.cell].dndx_c, src);
    double s = src[0];
    for (int k = 1; k < dim; ++k) s = s + src[k];
    p.dvol = dt * s;
  }

  // particle.tcc:598-606
  void update_volume(Particle& p) {
    p.volume *= (1. + p.dvol);
    p.density = p.density / (1. + p.dvol);
  }

  // particle.tcc:880-890 -> material
  void compute_stress(Particle& p) {
    const Material& m = materials[p.mat];
    double out[6];
    double rtrial[6];
    if (m.rate != 0) {  // objective stress rate (F is the step-start one)
      double L[3][3], LL[3][3];
      rate_rotations(m.rate, p.dF, p.F, L, LL);
      rate_stress(m.de, p.stress, p.dstrain, L, LL, rtrial);
    }
    if (m.type == kLE && m.rate != 0) {
      for (int k = 0; k < 6; ++k) out[k] = rtrial[k];
    } else if (m.type == kLE) {
      double ds[6];
      de_mul(m.de, p.dstrain, ds);
      for (int k = 0; k < 6; ++k) out[k] = p.stress[k] + ds[k];
    } else {
      mc_compute_stress(m, p.stress, p.dstrain, p.density, p.sv, out,
                        m.rate != 0 ? rtrial : nullptr);
    }
    std::memcpy(p.stress, out, sizeof(out));
  }

  // particle.tcc:1654-1659 F = dF * F (sequential k)
  void update_F(Particle& p) {
    double R[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double s = p.dF[i][0] * p.F[0][j];
        for (int k = 1; k < 3; ++k) s = s + p.dF[i][k] * p.F[k][j];
        R[i][j] = s;
      }
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) p.F[i][j] = R[i][j];
  }

  // mpm_scheme.tcc:75-103
  void compute_stress_strain(double dt) {
    const Index np = (Index)particles.size();
    ORC_PRAGMA(omp parallel for schedule(static))
    for (Index i = 0; i < np; ++i) update_dF(particles[i], dt);
    ORC_PRAGMA(omp parallel for schedule(static))
    for (Index i = 0; i < np; ++i) compute_strain(particles[i], dt);
    ORC_PRAGMA(omp parallel for schedule(static))
    for (Index i = 0; i < np; ++i) update_volume(particles[i]);
    ORC_PRAGMA(omp parallel for schedule(dynamic, 256))
    for (Index i = 0; i < np; ++i) compute_stress(particles[i]);
    ORC_PRAGMA(omp parallel for schedule(static))
    for (Index i = 0; i < np; ++i) update_F(particles[i]);
  }

  // particle.tcc:893-898
  void map_body_force(const Particle& p, const double* g) {
    const Cell& c = cells[p.cell];
    for (int i = 0; i < nn; ++i) {
      Node& n = nodes[c.nodes[i]];
      for (int k = 0; k < dim; ++k) add(n.fext[k], g[k] * p.mass * p.N[i]);
    }
  }

  // particle.tcc:914-950
  void map_internal_force(const Particle& p) {
    const Cell& c = cells[p.cell];
    const double* s = p.stress;
    for (int i = 0; i < nn; ++i) {
      Node& n = nodes[c.nodes[i]];
      const double* d = p.dndx[i];
      double f[3];
      if (dim == 2) {
        f[0] = d[0] * s[0] + d[1] * s[3];
        f[1] = d[1] * s[1] + d[0] * s[3];
      } else {
        f[0] = d[0] * s[0] + d[1] * s[3] + d[2] * s[5];
        f[1] = d[1] * s[1] + d[0] * s[3] + d[2] * s[4];
        f[2] = d[2] * s[2] + d[1] * s[4] + d[0] * s[5];
      }
      const double mv = -1. * p.volume;
      for (int k = 0; k < dim; ++k) add(n.fint[k], f[k] * mv);
    }
  }

  // node.tcc:510-657; on a general boundary in the frame of R
  void apply_friction(Node& n, double dt) {
    if (!n.friction) return;
    auto sign = [](double value) { return (value > 0.) ? 1. : -1.; };
    const unsigned dir_n = n.f_dir;
    const double sign_n = sign(n.f_sign);
    const double mu = n.f_mu;
    double Ri[3][3];
    if (n.generic) inverse(dim, n.R, Ri);
    if (dim == 2) {
      const unsigned dir_t = (dim - 1) - dir_n;
      double acc_n, acc_t, vel_t;
      if (!n.generic) {
        acc_n = n.acceleration[dir_n];
        acc_t = n.acceleration[dir_t];
        vel_t = n.velocity[dir_t];
      } else {
        double la[3], lv[3];
        matvec(Ri, n.acceleration, la);
        matvec(Ri, n.velocity, lv);
        acc_n = la[dir_n];
        acc_t = la[dir_t];
        vel_t = lv[dir_t];
      }
      if ((acc_n * sign_n) > 0.0) {
        if (vel_t != 0.0) {
          const double vel_net = dt * acc_t + vel_t;
          const double vel_frictional = dt * mu * std::abs(acc_n);
          if (std::abs(vel_net) <= vel_frictional)
            acc_t = -vel_t / dt;
          else
            acc_t -= sign(vel_net) * mu * std::abs(acc_n);
        } else {
          if (std::abs(acc_t) <= mu * std::abs(acc_n))
            acc_t = 0.0;
          else
            acc_t -= sign(acc_t) * mu * std::abs(acc_n);
        }
        if (!n.generic) {
          n.acceleration[dir_t] = acc_t;
        } else {
          double la[3] = {0., 0., 0.};
          la[dir_t] = acc_t;
          la[dir_n] = acc_n;
          matvec(n.R, la, n.acceleration);
        }
      }
    } else {
      static const int tdir[3][2] = {{1, 2}, {0, 2}, {0, 1}};
      const unsigned t0 = tdir[dir_n][0], t1 = tdir[dir_n][1];
      double acc[3], vel[3];
      if (!n.generic) {
        for (int k = 0; k < 3; ++k) { acc[k] = n.acceleration[k]; vel[k] = n.velocity[k]; }
      } else {
        matvec(Ri, n.acceleration, acc);
        matvec(Ri, n.velocity, vel);
      }
      const double acc_n = acc[dir_n];
      double acc_t = std::sqrt(acc[t0] * acc[t0] + acc[t1] * acc[t1]);
      const double vel_t = std::sqrt(vel[t0] * vel[t0] + vel[t1] * vel[t1]);
      if (acc_n * sign_n > 0.0) {
        if (vel_t != 0.0) {
          double vn0 = vel[t0] + acc[t0] * dt;
          double vn1 = vel[t1] + acc[t1] * dt;
          const double vel_net_t = std::sqrt(vn0 * vn0 + vn1 * vn1);
          const double vel_fricion = mu * std::abs(acc_n) * dt;
          if (vel_net_t <= vel_fricion) {
            acc[t0] = -vel[t0] / dt;
            acc[t1] = -vel[t1] / dt;
          } else {
            acc[t0] -= mu * std::abs(acc_n) * (vn0 / vel_net_t);
            acc[t1] -= mu * std::abs(acc_n) * (vn1 / vel_net_t);
          }
        } else {
          if (acc_t <= mu * std::abs(acc_n)) {
            acc[t0] = 0;
            acc[t1] = 0;
          } else {
            acc_t -= mu * std::abs(acc_n);
            acc[t0] -= mu * std::abs(acc_n) * (acc[t0] / acc_t);
            acc[t1] -= mu * std::abs(acc_n) * (acc[t1] / acc_t);
          }
        }
        if (!n.generic)
          for (int k = 0; k < 3; ++k) n.acceleration[k] = acc[k];
        else
          matvec(n.R, acc, n.acceleration);
      }
    }
  }

  // node.tcc:456-485, with the values of the current time; on a general
  // boundary per constraint in the frame of R
  void apply_acceleration_constraints(Node& n) {
    for (const auto& c : n.acc_c) {
      const unsigned direction = c.first % dim;
      if (!n.generic) {
        n.acceleration[direction] = n.acc_now[direction];
      } else {
        double Ri[3][3], la[3];
        inverse(dim, n.R, Ri);
        matvec(Ri, n.acceleration, la);
        la[direction] = n.acc_now[direction];
        matvec(n.R, la, n.acceleration);
      }
    }
  }

  // Mesh::apply_traction_on_particles (mesh.tcc:1371-1386) at current_time:
  // each load in order assigns its value to its particles (last wins), then
  // every particle with a traction maps N_i * traction to its nodes in
  // particle order (particle.tcc:983-990)
  void apply_tractions(double current_time) {
    if (tractions.empty()) return;
    const Index np = (Index)particles.size();
    for (const auto& t : tractions) {
      const double scalar = t.fn >= 0 ? functions[t.fn].value(current_time) : 1.0;
      const double tv = t.value * scalar;  // Traction::traction
      const Index b = t.particle < 0 ? 0 : t.particle;
      const Index e = t.particle < 0 ? np : std::min<Index>(t.particle + 1, np);
      for (Index i = b; i < e; ++i) {
        Particle& p = particles[i];
        p.traction[t.dir] = tv * p.volume / p.size;
        p.has_traction = true;
      }
    }
    for (auto& p : particles) {
      if (!p.has_traction || p.cell == kNoCell) continue;
      const Cell& cl = cells[p.cell];
      for (int i = 0; i < nn; ++i) {
        Node& n = nodes[cl.nodes[i]];
        for (int k = 0; k < dim; ++k) n.fext[k] = n.fext[k] + p.N[i] * p.traction[k];
      }
    }
  }

  // Node::apply_concentrated_force (node.tcc:103-111) over all nodes at
  // current_time = step * dt (mpm_scheme.tcc:156-160): external force =
  // external force * 1 + f(t) * concentrated force
  void apply_concentrated_forces(double current_time) {
    for (auto& n : nodes) {
      if (!n.has_cf) continue;
      const double scalar = n.cfn >= 0 ? functions[n.cfn].value(current_time) : 1.0;
      for (int k = 0; k < dim; ++k) n.fext[k] = n.fext[k] * 1. + scalar * n.cforce[k];
    }
  }

  // Mesh::update_nodal_acceleration_constraints (mesh.tcc:1414-1428) at
  // current_time = step * dt (mpm_scheme.tcc:219): each constraint's value is
  // acceleration * f(t) with f its math function, or * 1.0 without one
  // (acceleration_constraint.h:37-43)
  void update_acceleration_constraints(double current_time) {
    std::vector<double> fv(functions.size());
    for (size_t f = 0; f < functions.size(); ++f) fv[f] = functions[f].value(current_time);
    for (auto& n : nodes)
      for (const auto& c : n.acc_c) {
        const auto it = n.acc_fn.find(c.first);
        const double scalar = (it != n.acc_fn.end() && it->second >= 0) ? fv[it->second] : 1.0;
        n.acc_now[c.first % dim] = c.second * scalar;
      }
  }

  // node.tcc:309-346
  bool compute_acceleration_velocity_cundall(Node& n, double dt, double damping) {
    const double tolerance = 1.0E-15;
    if (!(n.mass > tolerance)) return false;
    double f[3] = {0, 0, 0};
    for (int k = 0; k < dim; ++k) f[k] = n.fext[k] + n.fint[k];
    double nrm2 = f[0] * f[0];
    for (int k = 1; k < dim; ++k) nrm2 = nrm2 + f[k] * f[k];
    const double dn = damping * std::sqrt(nrm2);
    for (int k = 0; k < dim; ++k) {
      const double sg = (n.velocity[k] > 0.) ? 1. : ((n.velocity[k] < 0.) ? -1. : 0.);
      n.acceleration[k] = (f[k] - dn * sg) / n.mass;
    }
    apply_friction(n, dt);
    apply_acceleration_constraints(n);
    for (int k = 0; k < dim; ++k) n.velocity[k] += n.acceleration[k] * dt;
    apply_velocity_constraints(n);
    for (int k = 0; k < dim; ++k)
      if (std::abs(n.velocity[k]) < tolerance) n.velocity[k] = 0.;
    for (int k = 0; k < dim; ++k)
      if (std::abs(n.acceleration[k]) < tolerance) n.acceleration[k] = 0.;
    return true;
  }

  // particle.tcc:994-1047 (FLIP), :1050-1070 (PIC), :1073-1112 (ASFLIP),
  // :1115-1126 (APIC: mapping matrix, then PIC), :1129-1140 (TPIC: velocity
  // gradient, then PIC)
  void compute_updated_position(Particle& p, double dt) {
    const Cell& c = cells[p.cell];
    if (vu == VU_APIC || vu == VU_ASFLIP) mapping_matrix(p, false);
    if (vu == VU_TPIC) mapping_matrix(p, true);
    double nv[3] = {0, 0, 0}, na[3] = {0, 0, 0};
    for (int i = 0; i < nn; ++i) {
      const Node& n = nodes[c.nodes[i]];
      for (int k = 0; k < dim; ++k) {
        nv[k] += p.N[i] * n.velocity[k];
        na[k] += p.N[i] * n.acceleration[k];
      }
    }
    if (vu == VU_ASFLIP) {
      const double beta = asflip_beta(p, dt);
      double flip[3], sep[3];
      for (int k = 0; k < dim; ++k) flip[k] = p.v[k] + na[k] * dt;
      for (int k = 0; k < dim; ++k) p.v[k] = blending * flip[k] + (1.0 - blending) * nv[k];
      for (int k = 0; k < dim; ++k)
        sep[k] = beta * blending * flip[k] + (1.0 - beta * blending) * nv[k];
      for (int k = 0; k < dim; ++k) p.x[k] += sep[k] * dt;
      for (int k = 0; k < dim; ++k) p.u[k] += sep[k] * dt;
      return;
    }
    if (vu == VU_FLIP) {
      for (int k = 0; k < dim; ++k) p.v[k] += na[k] * dt;
      for (int k = 0; k < dim; ++k) p.v[k] = blending * p.v[k] + (1.0 - blending) * nv[k];
    } else {
      for (int k = 0; k < dim; ++k) p.v[k] = nv[k];
    }
    for (int k = 0; k < dim; ++k) p.x[k] += nv[k] * dt;
    for (int k = 0; k < dim; ++k) p.u[k] += nv[k] * dt;
  }

  // particle velocity constraints (particle or -1 for all, dir, value), in
  // the order given: mesh.tcc:1454-1468, particle.tcc:1190-1195
  struct PVel {
    Index particle;
    int dir;
    double value;
  };
  std::vector<PVel> pvel;
  void apply_particle_velocity_constraints() {
    for (const auto& c : pvel) {
      if (c.particle < 0) {
        for (auto& p : particles) p.v[c.dir] = c.value;
      } else if (c.particle < (Index)particles.size()) {
        particles[c.particle].v[c.dir] = c.value;
      }
    }
  }

  void node_reset(Node& n) {  // node.tcc:26-43
    n.mass = 0;
    for (int k = 0; k < 3; ++k) {
      n.momentum[k] = n.velocity[k] = n.acceleration[k] = n.fext[k] = n.fint[k] = 0;
    }
    n.status = false;
  }

  // mpm_explicit.tcc:112-190 for one step
  void one_step(double dt, double damping) {
    const Index np = (Index)particles.size();
    const Index nnd = (Index)nodes.size();
    const Index ncl = (Index)cells.size();
    // MPMScheme::initialise (mpm_scheme.tcc:19-40)
    ORC_PRAGMA(omp parallel for schedule(static))
    for (Index i = 0; i < nnd; ++i) node_reset(nodes[i]);
    for (Index c = 0; c < ncl; ++c)
      if (cells[c].nparticles > 0)
        for (int i = 0; i < nn; ++i) nodes[cells[c].nodes[i]].status = true;
    ORC_PRAGMA(omp parallel for schedule(static))
    for (Index i = 0; i < np; ++i) compute_shapefn(particles[i]);
    // compute_nodal_kinematics (mpm_scheme.tcc:44-71)
    ORC_PRAGMA(omp parallel for schedule(static))
    for (Index i = 0; i < np; ++i) map_mass_momentum(particles[i]);
    ORC_PRAGMA(omp parallel for schedule(static))
    for (Index i = 0; i < nnd; ++i)
      if (nodes[i].status) compute_velocity(nodes[i]);
    // precompute_stress_strain (USF)
    if (scheme == SCHEME_USF) compute_stress_strain(dt);
    // gravity ramp (mpm_explicit.tcc:143-153)
    double gm = 1.0;
    if (ramp > 0.0) gm = std::min(1.0, static_cast<double>(step) * dt / ramp);
    double g[3];
    for (int k = 0; k < dim; ++k) g[k] = gm * gravity[k];
    // compute_forces (mpm_scheme.tcc:131-190)
    ORC_PRAGMA(omp parallel for schedule(static))
    for (Index i = 0; i < np; ++i) map_body_force(particles[i], g);
    apply_tractions(static_cast<double>(step) * dt);
    apply_concentrated_forces(static_cast<double>(step) * dt);
    ORC_PRAGMA(omp parallel for schedule(static))
    for (Index i = 0; i < np; ++i) map_internal_force(particles[i]);
    // compute_particle_kinematics (mpm_scheme.tcc:214-241), always Cundall
    update_acceleration_constraints(static_cast<double>(step) * dt);
    ORC_PRAGMA(omp parallel for schedule(static))
    for (Index i = 0; i < nnd; ++i)
      if (nodes[i].status) compute_acceleration_velocity_cundall(nodes[i], dt, damping);
    ORC_PRAGMA(omp parallel for schedule(static))
    for (Index i = 0; i < np; ++i) compute_updated_position(particles[i], dt);
    apply_particle_velocity_constraints();  // mpm_scheme.tcc:239-240
    // postcompute_nodal_kinematics (MUSL, mpm_scheme_musl.tcc:21-35)
    if (scheme == SCHEME_MUSL) {
      for (Index i = 0; i < nnd; ++i)
        if (nodes[i].status) {
          nodes[i].mass = 0.;
          for (int k = 0; k < 3; ++k) nodes[i].momentum[k] = 0.;
        }
      ORC_PRAGMA(omp parallel for schedule(static))
      for (Index i = 0; i < np; ++i) map_mass_momentum(particles[i]);
      ORC_PRAGMA(omp parallel for schedule(static))
      for (Index i = 0; i < nnd; ++i)
        if (nodes[i].status) compute_velocity(nodes[i]);
    }
    // postcompute_stress_strain (USL / MUSL)
    if (scheme != SCHEME_USF) compute_stress_strain(dt);
    // locate_particles
    locate_particles();
    ++step;
  }
};

}  // namespace orc

// ============================ C interface ===================================
using orc::Index;
static bool is_point_in_cell_ok(orc::Solver& s, Index c, const double* x, double* xi) {
  return s.is_point_in_cell(s.cells[c], x, xi);
}
struct orc_ctx {
  orc::Solver s;
};

extern "C" {

const char* orc_last_error(orc_ctx* c) { return c->s.error.c_str(); }

orc_ctx* orc_create(int dim, int scheme, int velocity_update, double blending_ratio) {
  if (dim != 2 && dim != 3) return nullptr;
  orc_ctx* c = new orc_ctx();
  c->s.dim = dim;
  c->s.nn = dim == 2 ? 4 : 8;
  c->s.scheme = scheme;
  c->s.vu = velocity_update;
  c->s.blending = blending_ratio;
  return c;
}

void orc_destroy(orc_ctx* c) { delete c; }

int orc_threads() {
#ifdef ORC_OMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

#define ORC_TRY(...)                  \
  try {                               \
    __VA_ARGS__;                      \
    return 0;                         \
  } catch (std::exception & e) {      \
    c->s.error = e.what();            \
    return -1;                        \
  }

int orc_set_mesh(orc_ctx* c, Index nnodes, const double* coords, Index ncells, const Index* cn) {
  ORC_TRY(c->s.set_mesh(nnodes, coords, ncells, cn))
}

// Mesh::compute_nodal_rotation_matrices (mesh.tcc:1320-1340): node[i] gets the
// rotation of Euler angles angles[i * dim ...] and its constraints act on a
// general boundary.
int orc_set_nodal_euler_angles(orc_ctx* c, Index n, const Index* node, const double* angles) {
  ORC_TRY({
    auto& s = c->s;
    for (Index i = 0; i < n; ++i) {
      if (node[i] < 0 || node[i] >= (Index)s.nodes.size())
        throw std::runtime_error("node id out of range");
      orc::Node& nd = s.nodes[node[i]];
      orc::rotation_matrix(s.dim, angles + i * s.dim, nd.R);
      nd.generic = true;
    }
  })
}

// The mesh's "isoparametric" key (mpm_base.tcc:768-783).
int orc_set_isoparametric(orc_ctx* c, int iso) {
  c->s.iso = iso != 0;
  return 0;
}

int orc_set_materials(orc_ctx* c, int n, const double* params) {
  ORC_TRY({
    c->s.materials.assign(n, orc::Material());
    for (int i = 0; i < n; ++i) orc::material_init(c->s.materials[i], params + i * orc::kMatParams);
  })
}

// Velocity constraints: first assignment per (node, dir) wins (std::map::insert)
int orc_set_velocity_constraints(orc_ctx* c, Index n, const Index* node, const int* dir,
                                 const double* val) {
  ORC_TRY({
    for (Index i = 0; i < n; ++i) {
      if ((unsigned)dir[i] >= (unsigned)c->s.dim) throw std::runtime_error("bad dir");
      c->s.nodes.at(node[i]).vel_c.insert(std::make_pair((unsigned)dir[i], val[i]));
    }
  })
}

// fn: per constraint the index of its math function, -1 or a null array: none
int orc_set_acceleration_constraints_fn(orc_ctx* c, Index n, const Index* node, const int* dir,
                                        const double* val, const int* fn) {
  ORC_TRY({
    for (Index i = 0; i < n; ++i) {
      if ((unsigned)dir[i] >= (unsigned)c->s.dim) throw std::runtime_error("bad dir");
      if (fn && (fn[i] < -1 || fn[i] >= (int)c->s.functions.size()))
        throw std::runtime_error("bad math function index");
      auto& nd = c->s.nodes.at(node[i]);
      if (nd.vel_c.count((unsigned)dir[i]))
        throw std::runtime_error("Acceleration and velocity constraints set in the same direction");
      if (nd.acc_c.insert(std::make_pair((unsigned)dir[i], val[i])).second)
        nd.acc_fn[(unsigned)dir[i]] = fn ? fn[i] : -1;
    }
  })
}

// Concentrated nodal forces (mesh.tcc:1471-1497): node -1 = every node; a
// later assignment replaces the force in its direction and the node's
// function (node.tcc:82-101). fn null: none.
int orc_set_nodal_forces(orc_ctx* c, Index n, const Index* node, const int* dir,
                         const double* force, const int* fn) {
  ORC_TRY({
    for (Index i = 0; i < n; ++i) {
      if ((unsigned)dir[i] >= (unsigned)c->s.dim) throw std::runtime_error("bad dir");
      if (fn && (fn[i] < -1 || fn[i] >= (int)c->s.functions.size()))
        throw std::runtime_error("bad math function index");
      const Index b = node[i] < 0 ? 0 : node[i];
      const Index e = node[i] < 0 ? (Index)c->s.nodes.size() : node[i] + 1;
      for (Index k = b; k < e; ++k) {
        auto& nd = c->s.nodes.at(k);
        nd.has_cf = true;
        nd.cforce[dir[i]] = force[i];
        nd.cfn = fn ? fn[i] : -1;
      }
    }
  })
}

// Particle tractions in order (replacing any earlier list): particle -1 =
// every particle; fn: math function index, -1 or a null array: none.
int orc_set_particle_tractions(orc_ctx* c, Index n, const Index* particle, const int* dir,
                               const double* value, const int* fn) {
  ORC_TRY({
    std::vector<orc::TractionLoad> t;
    for (Index i = 0; i < n; ++i) {
      if ((unsigned)dir[i] >= (unsigned)c->s.dim) throw std::runtime_error("bad dir");
      if (fn && (fn[i] < -1 || fn[i] >= (int)c->s.functions.size()))
        throw std::runtime_error("bad math function index");
      t.push_back({particle[i], dir[i], value[i], fn ? fn[i] : -1});
    }
    c->s.tractions = t;
  })
}

int orc_set_acceleration_constraints(orc_ctx* c, Index n, const Index* node, const int* dir,
                                     const double* val) {
  return orc_set_acceleration_constraints_fn(c, n, node, dir, val, nullptr);
}

// Math functions, in index order: type[f] FN_LINEAR with npts[f] points
// (params: x_0..x_{n-1}, f_0..f_{n-1}) or FN_SIN (params: x0, a, xmin, xmax).
int orc_set_math_functions(orc_ctx* c, int nf, const int* type, const int* npts,
                           const double* params) {
  ORC_TRY({
    std::vector<orc::MathFn> fs;
    size_t off = 0;
    for (int f = 0; f < nf; ++f) {
      orc::MathFn m;
      m.type = type[f];
      if (m.type == orc::FN_LINEAR) {
        if (npts[f] < 1) throw std::runtime_error("linear function needs points");
        m.x.assign(params + off, params + off + npts[f]);
        m.fx.assign(params + off + npts[f], params + off + 2 * npts[f]);
        off += 2 * (size_t)npts[f];
      } else if (m.type == orc::FN_SIN) {
        m.x0 = params[off];
        m.a = params[off + 1];
        m.xmin = params[off + 2];
        m.xmax = params[off + 3];
        off += 4;
      } else {
        throw std::runtime_error("unknown math function type");
      }
      fs.push_back(m);
    }
    c->s.functions = fs;
  })
}

// Friction: last assignment per node wins (std::tuple assignment)
int orc_set_friction_constraints(orc_ctx* c, Index n, const Index* node, const int* dir,
                                 const int* sign_n, const double* mu) {
  ORC_TRY({
    for (Index i = 0; i < n; ++i) {
      if ((unsigned)dir[i] >= (unsigned)c->s.dim) throw std::runtime_error("bad dir");
      auto& nd = c->s.nodes.at(node[i]);
      nd.friction = true;
      nd.f_dir = (unsigned)dir[i];
      nd.f_sign = sign_n[i];
      nd.f_mu = mu[i];
    }
  })
}

int orc_set_particle_velocity_constraints(orc_ctx* c, Index n, const Index* particle,
                                          const int* dir, const double* value) {
  ORC_TRY({
    c->s.pvel.clear();
    for (Index i = 0; i < n; ++i) {
      if (dir[i] < 0 || dir[i] >= c->s.dim)
        throw std::runtime_error("Invalid direction of velocity constraint");
      c->s.pvel.push_back({particle[i], dir[i], value[i]});
    }
  })
}

int orc_compute_volume_mass(orc_ctx* c) { ORC_TRY(c->s.compute_volume_mass()) }

// Particle::assign_material (particle.tcc:499-517) for the listed particles
// (material_sets, mpm_base.tcc:449-468).
int orc_assign_materials(orc_ctx* c, Index n, const Index* particle, int mat) {
  ORC_TRY({
    auto& s = c->s;
    if (mat < 0 || mat >= (int)s.materials.size()) throw std::runtime_error("bad material id");
    for (Index i = 0; i < n; ++i) {
      if (particle[i] < 0 || particle[i] >= (Index)s.particles.size())
        throw std::runtime_error("particle index out of range");
      orc::Particle& p = s.particles[particle[i]];
      p.mat = mat;
      for (int k = 0; k < orc::kNsv; ++k) p.sv[k] = 0.;
      if (s.materials[mat].type == orc::kMC) orc::mc_initialise_state(s.materials[mat], p.sv);
    }
  })
}

int orc_set_locate_particles(orc_ctx* c, int locate_particles) {
  c->s.locate_all = locate_particles != 0;
  return 0;
}

int orc_set_velocity_update(orc_ctx* c, int velocity_update, double blending_ratio) {
  c->s.vu = velocity_update;
  c->s.blending = blending_ratio;
  return 0;
}

int orc_set_gravity(orc_ctx* c, const double* g, double ramp) {
  for (int k = 0; k < c->s.dim; ++k) c->s.gravity[k] = g[k];
  c->s.ramp = ramp;
  return 0;
}

int orc_generate_gauss(orc_ctx* c, int nppd, Index nset, const Index* cset, int mat) {
  ORC_TRY(c->s.generate_gauss(nppd, nset, cset, mat))
}

// Particles from full state (HDF5-resume style, particle.tcc:35-177).
// Arrays are in particle order; null pointers leave the field at its default.
// cell < 0 means "locate by search". volume/mass/density must be given.
int orc_set_particles(orc_ctx* c, Index n, const double* x, const double* u, const double* v,
                      const double* stress, const double* strain, const double* F,
                      const double* mass, const double* volume, const double* density,
                      const int* mat, const Index* cell, const double* sv) {
  ORC_TRY({
    auto& s = c->s;
    const int d = s.dim;
    const Index base = (Index)s.particles.size();
    for (Index i = 0; i < n; ++i) {
      orc::Particle p;
      p.id = base + i;
      for (int k = 0; k < d; ++k) {
        p.x[k] = x[i * d + k];
        p.u[k] = u ? u[i * d + k] : 0.;
        p.v[k] = v ? v[i * d + k] : 0.;
      }
      for (int k = 0; k < 6; ++k) {
        p.stress[k] = stress ? stress[i * 6 + k] : 0.;
        p.strain[k] = strain ? strain[i * 6 + k] : 0.;
      }
      if (F)
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) p.F[a][b] = F[i * 9 + a * 3 + b];
      p.mass = mass[i];
      p.volume = volume[i];
      p.size = std::pow(p.volume, static_cast<double>(1. / c->s.dim));
      p.density = density[i];
      p.mat = mat[i];
      if (p.mat < 0 || p.mat >= (int)s.materials.size()) throw std::runtime_error("bad material id");
      if (s.materials[p.mat].type == orc::kMC) {
        if (sv)
          for (int k = 0; k < orc::kNsv; ++k) p.sv[k] = sv[i * orc::kNsv + k];
        else
          orc::mc_initialise_state(s.materials[p.mat], p.sv);
      }
      p.cell = orc::kNoCell;
      s.particles.push_back(p);
      if (cell && cell[i] >= 0) {
        double xi[3];
        orc::Particle& q = s.particles.back();
        if (cell[i] >= (Index)s.cells.size()) throw std::runtime_error("bad cell id");
        // The cell id is held without a cell (Particle::assign_cell_id,
        // particle.tcc:470-485); locate_particle_cells first tries to assign
        // that cell (assign_cell, particle.tcc:386-416, succeeds only if the
        // point is in it), else searches that cell's neighbours, then every
        // cell (mesh.tcc:1118-1151).
        if (s.is_point_in_cell(s.cells[cell[i]], q.x, xi))
          s.assign_cell_xi(q, cell[i], xi);
        else
          q.hint = cell[i];
      }
    }
    s.locate_particles();
  })
}

int orc_assign_velocity(orc_ctx* c, Index n, const double* v) {
  ORC_TRY({
    auto& s = c->s;
    if (n != (Index)s.particles.size()) throw std::runtime_error("one velocity per particle");
    for (Index i = 0; i < n; ++i)
      for (int k = 0; k < s.dim; ++k) s.particles[i].v[k] = v[i * s.dim + k];
  })
}

int orc_assign_stresses(orc_ctx* c, Index n, const double* st) {
  ORC_TRY({
    auto& s = c->s;
    if (n != (Index)s.particles.size()) throw std::runtime_error("one stress per particle");
    for (Index i = 0; i < n; ++i)
      for (int k = 0; k < 6; ++k) s.particles[i].stress[k] = st[i * 6 + k];
  })
}

int orc_nparticles(orc_ctx* c, Index* n) {
  *n = (Index)c->s.particles.size();
  return 0;
}

int orc_get_particles(orc_ctx* c, double* x, double* u, double* v, double* stress, double* strain,
                      double* F, double* mass, double* volume, double* density, int* mat,
                      Index* cell, double* xi, double* sv) {
  auto& s = c->s;
  const int d = s.dim;
  for (size_t i = 0; i < s.particles.size(); ++i) {
    const auto& p = s.particles[i];
    for (int k = 0; k < d; ++k) {
      if (x) x[i * d + k] = p.x[k];
      if (u) u[i * d + k] = p.u[k];
      if (v) v[i * d + k] = p.v[k];
      if (xi) xi[i * d + k] = p.xi[k];
    }
    for (int k = 0; k < 6; ++k) {
      if (stress) stress[i * 6 + k] = p.stress[k];
      if (strain) strain[i * 6 + k] = p.strain[k];
    }
    if (F)
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) F[i * 9 + a * 3 + b] = p.F[a][b];
    if (mass) mass[i] = p.mass;
    if (volume) volume[i] = p.volume;
    if (density) density[i] = p.density;
    if (mat) mat[i] = p.mat;
    if (cell) cell[i] = p.cell;
    if (sv)
      for (int k = 0; k < orc::kNsv; ++k) sv[i * orc::kNsv + k] = p.sv[k];
  }
  return 0;
}

int orc_get_nodes(orc_ctx* c, double* mass, double* momentum, double* velocity,
                  double* acceleration, double* fext, double* fint, int* status) {
  auto& s = c->s;
  const int d = s.dim;
  for (size_t i = 0; i < s.nodes.size(); ++i) {
    const auto& n = s.nodes[i];
    if (mass) mass[i] = n.mass;
    for (int k = 0; k < d; ++k) {
      if (momentum) momentum[i * d + k] = n.momentum[k];
      if (velocity) velocity[i * d + k] = n.velocity[k];
      if (acceleration) acceleration[i * d + k] = n.acceleration[k];
      if (fext) fext[i * d + k] = n.fext[k];
      if (fint) fint[i * d + k] = n.fint[k];
    }
    if (status) status[i] = n.status ? 1 : 0;
  }
  return 0;
}

int orc_get_cells(orc_ctx* c, double* volume, double* mean_length, double* centroid) {
  auto& s = c->s;
  for (size_t i = 0; i < s.cells.size(); ++i) {
    if (volume) volume[i] = s.cells[i].volume;
    if (mean_length) mean_length[i] = s.cells[i].mean_length;
    if (centroid)
      for (int k = 0; k < s.dim; ++k) centroid[i * s.dim + k] = s.cells[i].centroid[k];
  }
  return 0;
}

// Runs n steps. Returns 0, or -1 with the error message set (e.g. a particle
// left the mesh, mpm_scheme.tcc:250-256).
int orc_step(orc_ctx* c, int n, double dt, double damping) {
  ORC_TRY({
    for (int i = 0; i < n; ++i) c->s.one_step(dt, damping);
  })
}

// ------------- phase-level entry points (replay of reference unit tests) -----
// Each phase runs one reference sweep over all particles or nodes, as named.
enum {
  PH_RESET_NODES = 0,     // node.tcc:26-43 on all nodes + cell.tcc:171-178
  PH_SHAPEFN = 1,         // particle.tcc:537-554
  PH_MAP_MASS_MOM = 2,    // particle.tcc:630-658
  PH_COMPUTE_VEL = 3,     // node.tcc:236-252 (active nodes)
  PH_DF = 4,              // particle.tcc:1643-1651
  PH_STRAIN = 5,          // particle.tcc:861-877
  PH_VOLUME = 6,          // particle.tcc:598-606
  PH_STRESS = 7,          // particle.tcc:880-890
  PH_UPDATE_F = 8,        // particle.tcc:1654-1659
  PH_BODY_FORCE = 9,      // particle.tcc:893-898 with g = arg
  PH_INTERNAL_FORCE = 10, // particle.tcc:914-950
  PH_ACC_VEL = 11,        // node.tcc:309-346 (active nodes), damping = arg[0]
  PH_POSITION = 12,       // particle.tcc:994-1070
  PH_LOCATE = 13,         // mesh.tcc:1097-1153
  PH_ACC_VEL_PLAIN = 14   // node.tcc:272-305 (no damping term)
};

int orc_phase(orc_ctx* c, int phase, double dt, const double* arg) {
  ORC_TRY({
    auto& s = c->s;
    switch (phase) {
      case PH_RESET_NODES:
        for (auto& n : s.nodes) s.node_reset(n);
        for (auto& cl : s.cells)
          if (cl.nparticles > 0)
            for (int i = 0; i < s.nn; ++i) s.nodes[cl.nodes[i]].status = true;
        break;
      case PH_SHAPEFN: for (auto& p : s.particles) s.compute_shapefn(p); break;
      case PH_MAP_MASS_MOM: for (auto& p : s.particles) s.map_mass_momentum(p); break;
      case PH_COMPUTE_VEL:
        for (auto& n : s.nodes)
          if (n.status) s.compute_velocity(n);
        break;
      case PH_DF: for (auto& p : s.particles) s.update_dF(p, dt); break;
      case PH_STRAIN: for (auto& p : s.particles) s.compute_strain(p, dt); break;
      case PH_VOLUME: for (auto& p : s.particles) s.update_volume(p); break;
      case PH_STRESS: for (auto& p : s.particles) s.compute_stress(p); break;
      case PH_UPDATE_F: for (auto& p : s.particles) s.update_F(p); break;
      case PH_BODY_FORCE:
        for (auto& p : s.particles) s.map_body_force(p, arg);
        s.apply_tractions(static_cast<double>(s.step) * dt);
        s.apply_concentrated_forces(static_cast<double>(s.step) * dt);
        break;
      case PH_INTERNAL_FORCE: for (auto& p : s.particles) s.map_internal_force(p); break;
      case PH_ACC_VEL:
        s.update_acceleration_constraints(static_cast<double>(s.step) * dt);
        for (auto& n : s.nodes)
          if (n.status) s.compute_acceleration_velocity_cundall(n, dt, arg ? arg[0] : 0.);
        break;
      case PH_ACC_VEL_PLAIN:
        s.update_acceleration_constraints(static_cast<double>(s.step) * dt);
        for (auto& n : s.nodes)
          if (n.status) s.compute_acceleration_velocity_cundall(n, dt, 0.);
        break;
      case PH_POSITION: for (auto& p : s.particles) s.compute_updated_position(p, dt); break;
      case PH_LOCATE: s.locate_particles(); break;
      default: throw std::runtime_error("unknown phase");
    }
  })
}

// Overwrite one node field: 0 mass, 1 momentum, 2 velocity, 3 acceleration,
// 4 fext, 5 fint, 6 status.
int orc_set_node_field(orc_ctx* c, Index node, int field, const double* val) {
  ORC_TRY({
    auto& n = c->s.nodes.at(node);
    const int d = c->s.dim;
    double* dst = nullptr;
    switch (field) {
      case 0: n.mass = val[0]; return 0;
      case 1: dst = n.momentum; break;
      case 2: dst = n.velocity; break;
      case 3: dst = n.acceleration; break;
      case 4: dst = n.fext; break;
      case 5: dst = n.fint; break;
      case 6: n.status = val[0] != 0.; return 0;
      default: throw std::runtime_error("bad field");
    }
    for (int k = 0; k < d; ++k) dst[k] = val[k];
  })
}

// Overwrite one particle field: 0 x, 1 v, 2 stress(6), 3 volume, 4 mass,
// 5 density, 6 u, 7 strain(6), 8 sv(9).
int orc_set_particle_field(orc_ctx* c, Index pid, int field, const double* val) {
  ORC_TRY({
    auto& p = c->s.particles.at(pid);
    const int d = c->s.dim;
    switch (field) {
      case 0: for (int k = 0; k < d; ++k) p.x[k] = val[k]; break;
      case 1: for (int k = 0; k < d; ++k) p.v[k] = val[k]; break;
      case 2: for (int k = 0; k < 6; ++k) p.stress[k] = val[k]; break;
      case 3: p.volume = val[0]; break;
      case 4: p.mass = val[0]; break;
      case 5: p.density = val[0]; break;
      case 6: for (int k = 0; k < d; ++k) p.u[k] = val[k]; break;
      case 7: for (int k = 0; k < 6; ++k) p.strain[k] = val[k]; break;
      case 8: for (int k = 0; k < orc::kNsv; ++k) p.sv[k] = val[k]; break;
      default: throw std::runtime_error("bad field");
    }
  })
}

// Per-particle scratch fields: N (nn), dndx (nn*d), dstrain (6), dF (9), dvol.
// The particles' mapping matrices of the affine velocity updates
// (mapping_matrix_), [np][dim][dim] row-major, in particle order.
int orc_get_mapping_matrix(orc_ctx* c, double* B) {
  ORC_TRY({
    const int d = c->s.dim;
    Index q = 0;
    for (const auto& p : c->s.particles) {
      for (int a = 0; a < d; ++a)
        for (int b = 0; b < d; ++b) B[(q * d + a) * d + b] = p.B[a][b];
      ++q;
    }
  })
}

int orc_set_mapping_matrix(orc_ctx* c, Index n, const double* B) {
  ORC_TRY({
    const int d = c->s.dim;
    if (n != (Index)c->s.particles.size()) throw std::runtime_error("particle count differs");
    for (Index q = 0; q < n; ++q)
      for (int a = 0; a < d; ++a)
        for (int b = 0; b < d; ++b) c->s.particles[q].B[a][b] = B[(q * d + a) * d + b];
  })
}

int orc_get_particle_scratch(orc_ctx* c, Index pid, double* N, double* dndx, double* dstrain,
                             double* dF, double* dvol) {
  ORC_TRY({
    const auto& p = c->s.particles.at(pid);
    const int d = c->s.dim, nn = c->s.nn;
    for (int i = 0; i < nn; ++i) {
      if (N) N[i] = p.N[i];
      for (int k = 0; k < d; ++k)
        if (dndx) dndx[i * d + k] = p.dndx[i][k];
    }
    for (int k = 0; k < 6; ++k)
      if (dstrain) dstrain[k] = p.dstrain[k];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b)
        if (dF) dF[a * 3 + b] = p.dF[a][b];
    if (dvol) *dvol = p.dvol;
  })
}

void orc_kat_jacobian(int dim, const double* xi, const double* X, double* J) {
  const int nn = dim == 2 ? 4 : 8;
  double g[24];
  orc::grad_shapefn(dim, xi, g);
  for (int a = 0; a < dim; ++a)
    for (int b = 0; b < dim; ++b) {
      double s = g[0 * 3 + a] * X[0 * dim + b];
      for (int i = 1; i < nn; ++i) s = s + g[i * 3 + a] * X[i * dim + b];
      J[a * dim + b] = s;
    }
}

// dF/dsigma and dP/dsigma for a given yield type (mohr_coulomb.tcc:195-344).
void orc_kat_mc_df_dp(const double* params, int yield_type, const double* sv, const double* stress,
                      double* df, double* dp, double* dp_dq, double* softening) {
  orc::Material m;
  orc::material_init(m, params);
  orc::mc_df_dp(m, yield_type, sv, stress, df, dp, dp_dq, softening);
}

// Stress invariants into a state vector (mohr_coulomb.tcc:124-134).
void orc_kat_mc_invariants(const double* stress, double* sv) { orc::mc_invariants(stress, sv); }

void orc_kat_mc_init_state(const double* params, double* sv) {
  orc::Material m;
  orc::material_init(m, params);
  orc::mc_initialise_state(m, sv);
}

// ------------------- per-function entry points for the KATs ------------------
// Cell of one quad4 / hex8 with vertices X[nn * dim]: transform_real_to_unit_cell
// (isoparametric) of point p into xi; returns 0, or -1 when it throws.
int orc_kat_transform_real_to_unit_cell(int dim, const double* Xf, const double* p, double* xi) {
  orc::Solver s;
  s.dim = dim;
  s.nn = dim == 2 ? 4 : 8;
  orc::Cell c;
  for (int i = 0; i < s.nn; ++i)
    for (int k = 0; k < 3; ++k) c.X[i][k] = k < dim ? Xf[i * dim + k] : 0.;
  s.init_cell(c);
  try {
    orc::transform_real_to_unit_cell(dim, c.X, c.mean_length, p, xi);
  } catch (std::exception&) {
    return -1;
  }
  return 0;
}

// is_point_in_cell (cell.tcc:335-380) of the same one-cell setup; iso 0 uses
// the Cartesian local coordinates. Returns 1 inside, 0 outside.
int orc_kat_point_in_cell(int dim, int iso, const double* Xf, const double* p, double* xi) {
  orc::Solver s;
  s.dim = dim;
  s.nn = dim == 2 ? 4 : 8;
  s.iso = iso != 0;
  orc::Cell c;
  for (int i = 0; i < s.nn; ++i)
    for (int k = 0; k < 3; ++k) c.X[i][k] = k < dim ? Xf[i * dim + k] : 0.;
  s.init_cell(c);
  return s.is_point_in_cell(c, p, xi) ? 1 : 0;
}

void orc_kat_shapefn(int dim, const double* xi, double* N, double* grad) {
  orc::shapefn(dim, xi, N);
  if (grad) {
    double g[24];
    orc::grad_shapefn(dim, xi, g);
    const int nn = dim == 2 ? 4 : 8;
    for (int i = 0; i < nn; ++i)
      for (int a = 0; a < dim; ++a) grad[i * dim + a] = g[i * 3 + a];
  }
}

void orc_kat_dn_dx(int dim, const double* xi, const double* X, double* out) {
  const int nn = dim == 2 ? 4 : 8;
  double Xl[8][3] = {{0}};
  for (int i = 0; i < nn; ++i)
    for (int k = 0; k < dim; ++k) Xl[i][k] = X[i * dim + k];
  double o[8][3];
  orc::dn_dx(dim, xi, Xl, o);
  for (int i = 0; i < nn; ++i)
    for (int k = 0; k < dim; ++k) out[i * dim + k] = o[i][k];
}

double orc_kat_cell_volume(int dim, const double* X) {
  orc::Solver s;
  s.dim = dim;
  s.nn = dim == 2 ? 4 : 8;
  double Xl[8][3] = {{0}};
  for (int i = 0; i < s.nn; ++i)
    for (int k = 0; k < dim; ++k) Xl[i][k] = X[i * dim + k];
  return s.element_volume(Xl);
}

// One constitutive update. params: material parameter vector; sv (9) is read
// and updated for MohrCoulomb (initialised from the material when init != 0).
int orc_kat_stress(const double* params, const double* stress, const double* dstrain,
                   double density, double* sv, int init_sv, double* out) {
  orc::Material m;
  orc::material_init(m, params);
  if (m.type == orc::kLE) {
    double ds[6];
    orc::de_mul(m.de, dstrain, ds);
    for (int k = 0; k < 6; ++k) out[k] = stress[k] + ds[k];
  } else {
    if (init_sv) orc::mc_initialise_state(m, sv);
    orc::mc_compute_stress(m, stress, dstrain, density, sv, out);
  }
  return 0;
}

// One LinearElastic update with its objective stress rate, given the
// deformation gradient increment dF (3x3 row-major).
// One LinearElastic update with the material's stress rate for the
// increment dF (3x3) at the step-start deformation gradient F (3x3, null:
// identity; the Green-Naghdi rate needs it).
int orc_kat_le_rate(const double* params, const double* stress, const double* dstrain,
                    const double* dF, const double* Fin, double* out) {
  orc::Material m;
  orc::material_init(m, params);
  if (m.type != orc::kLE) return -1;
  double G[3][3], F[3][3];
  for (int i = 0; i < 9; ++i) {
    G[i / 3][i % 3] = dF[i];
    F[i / 3][i % 3] = Fin ? Fin[i] : (i % 4 == 0 ? 1. : 0.);
  }
  if (m.rate != 0) {
    double L[3][3], LL[3][3];
    orc::rate_rotations(m.rate, G, F, L, LL);
    orc::rate_stress(m.de, stress, dstrain, L, LL, out);
  } else {
    double ds[6];
    orc::de_mul(m.de, dstrain, ds);
    for (int k = 0; k < 6; ++k) out[k] = stress[k] + ds[k];
  }
  return 0;
}

// Invariants of a Voigt stress: p, q, j2, j3, lode angle; dtheta/dsigma (6).
void orc_kat_invariants(const double* s, double* out, double* dtheta) {
  out[0] = orc::mu_p(s);
  out[1] = orc::mu_q(s);
  out[2] = orc::mu_j2(s);
  out[3] = orc::mu_j3(s);
  out[4] = orc::mu_lode_angle(s);
  if (dtheta) orc::mu_dtheta_dsigma(s, orc::kEps, dtheta);
}

// Yield functions and state from MC state variables (epsilon, rho, theta,
// phi, cohesion, tension_cutoff are read from sv).
int orc_kat_mc_yield(const double* sv, double* yf) { return orc::mc_yield_state(yf, sv); }

// Node acceleration/velocity update for one node (node.tcc:309-346).
// vel: in/out, acc: out. Constraint lists are dir-indexed; NaN = none.
int orc_kat_node_update(int dim, double mass, const double* fext, const double* fint,
                        double* vel, double* acc, double dt, double damping,
                        const double* vel_c, const double* acc_c, int f_dir, int f_sign,
                        double f_mu) {
  orc::Solver s;
  s.dim = dim;
  orc::Node n;
  n.mass = mass;
  for (int k = 0; k < dim; ++k) {
    n.fext[k] = fext[k];
    n.fint[k] = fint[k];
    n.velocity[k] = vel[k];
    n.acceleration[k] = acc[k];
    if (vel_c && !std::isnan(vel_c[k])) n.vel_c.insert(std::make_pair((unsigned)k, vel_c[k]));
    if (acc_c && !std::isnan(acc_c[k])) {
      n.acc_c.insert(std::make_pair((unsigned)k, acc_c[k]));
      n.acc_now[k] = acc_c[k];
    }
  }
  if (f_dir >= 0) {
    n.friction = true;
    n.f_dir = f_dir;
    n.f_sign = f_sign;
    n.f_mu = f_mu;
  }
  bool st = s.compute_acceleration_velocity_cundall(n, dt, damping);
  for (int k = 0; k < dim; ++k) {
    vel[k] = n.velocity[k];
    acc[k] = n.acceleration[k];
  }
  return st ? 1 : 0;
}

// Value of one math function (type, npts, params as orc_set_math_functions)
// at x; NaN where the Linear table has no segment for x.
double orc_kat_math_function(int type, int npts, const double* params, double x) {
  orc::MathFn m;
  m.type = type;
  if (type == orc::FN_LINEAR) {
    m.x.assign(params, params + npts);
    m.fx.assign(params + npts, params + 2 * npts);
  } else {
    m.x0 = params[0];
    m.a = params[1];
    m.xmin = params[2];
    m.xmax = params[3];
  }
  return m.value(x);
}

// Friction alone on given acceleration/velocity (node.tcc:510-657).
// Elementary functions of this build: which 0 sin, 1 cos, 2 tan, 3 acos,
// 4 x^1.5, 5 x^-1.5, 6 x^-2.5.
double orc_kat_math(int which, double x) {
  switch (which) {
    case 0: return orc::mth::sin(x);
    case 1: return orc::mth::cos(x);
    case 2: return orc::mth::tan(x);
    case 3: return orc::mth::acos(x);
    case 4: return orc::mth::pow(x, 1.5);
    case 5: return orc::mth::pow(x, -1.5);
    default: return orc::mth::pow(x, -2.5);
  }
}

// One node's constraints on a general boundary (angles: Euler angles, null:
// Cartesian): which 0 apply_velocity_constraints (nv constraints vdir/vval),
// 1 apply_acceleration_constraints (na, adir/aval), 2
// apply_friction_constraints(dt) with (f_dir, f_sign, f_mu). vel and acc in
// and out.
void orc_kat_node_constraints(int dim, const double* angles, double* vel, double* acc, int nv,
                              const int* vdir, const double* vval, int na, const int* adir,
                              const double* aval, int f_dir, int f_sign, double f_mu, double dt,
                              int which) {
  orc::Solver s;
  s.dim = dim;
  orc::Node n;
  for (int k = 0; k < dim; ++k) {
    n.velocity[k] = vel[k];
    n.acceleration[k] = acc[k];
  }
  if (angles) {
    orc::rotation_matrix(dim, angles, n.R);
    n.generic = true;
  }
  for (int i = 0; i < nv; ++i) n.vel_c.insert(std::make_pair((unsigned)vdir[i], vval[i]));
  for (int i = 0; i < na; ++i) {
    n.acc_c.insert(std::make_pair((unsigned)adir[i], aval[i]));
    n.acc_now[adir[i] % dim] = aval[i];
  }
  if (which == 0) s.apply_velocity_constraints(n);
  if (which == 1) s.apply_acceleration_constraints(n);
  if (which == 2) {
    n.friction = true;
    n.f_dir = f_dir;
    n.f_sign = f_sign;
    n.f_mu = f_mu;
    s.apply_friction(n, dt);
  }
  for (int k = 0; k < dim; ++k) {
    vel[k] = n.velocity[k];
    acc[k] = n.acceleration[k];
  }
}

// The rotation of the polar decomposition of F (3x3 row-major).
void orc_kat_polar_rotation(const double* F, double* R) {
  double A[3][3], B[3][3];
  for (int i = 0; i < 9; ++i) A[i / 3][i % 3] = F[i];
  orc::polar_rotation(A, B);
  for (int i = 0; i < 9; ++i) R[i] = B[i / 3][i % 3];
}

// One MohrCoulomb update with the material's stress rate (increment dF at the
// step-start F): new stress and state variables (sv in/out; init: start from
// the material's initial state).
int orc_kat_mc_rate(const double* params, const double* stress, const double* dstrain,
                    const double* dF, const double* Fin, double density, double* sv, int init,
                    double* out) {
  orc::Material m;
  orc::material_init(m, params);
  if (m.type == orc::kLE) return -1;
  if (init) orc::mc_initialise_state(m, sv);
  double G[3][3], F[3][3];
  for (int i = 0; i < 9; ++i) {
    G[i / 3][i % 3] = dF[i];
    F[i / 3][i % 3] = Fin ? Fin[i] : (i % 4 == 0 ? 1. : 0.);
  }
  double trial[6];
  const double* tp = nullptr;
  if (m.rate != 0) {
    double L[3][3], LL[3][3];
    orc::rate_rotations(m.rate, G, F, L, LL);
    orc::rate_stress(m.de, stress, dstrain, L, LL, trial);
    tp = trial;
  }
  orc::mc_compute_stress(m, stress, dstrain, density, sv, out, tp);
  return 0;
}

// The rotation matrix of Euler angles (row-major dim x dim).
void orc_kat_rotation_matrix(int dim, const double* angles, double* R) {
  double M[3][3];
  orc::rotation_matrix(dim, angles, M);
  for (int i = 0; i < dim; ++i)
    for (int j = 0; j < dim; ++j) R[i * dim + j] = M[i][j];
}

void orc_kat_friction(int dim, double* acc, const double* vel, double dt, int f_dir, int f_sign,
                      double f_mu) {
  orc::Solver s;
  s.dim = dim;
  orc::Node n;
  for (int k = 0; k < dim; ++k) {
    n.acceleration[k] = acc[k];
    n.velocity[k] = vel[k];
  }
  n.friction = true;
  n.f_dir = f_dir;
  n.f_sign = f_sign;
  n.f_mu = f_mu;
  s.apply_friction(n, dt);
  for (int k = 0; k < dim; ++k) acc[k] = n.acceleration[k];
}

}  // extern "C"
