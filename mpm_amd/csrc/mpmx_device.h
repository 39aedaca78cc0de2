// Device-side math of the explicit MPM step: shape functions, Jacobian,
// constitutive updates, node update, point location. Each function evaluates
// the reference's expression in the reference's order (cited per function) so
// that, with -ffp-contract=off, the kernels reproduce the single-threaded
// reference arithmetic operation for operation. Where the reference's order
// lives in Eigen products, the sequential k-order of a non-vectorised product
// is used throughout.
#pragma once

#include "mpmx_internal.h"
#include "mpmx_math.h"

namespace mpmx {

#define MPMX_DEV __device__ __forceinline__

static constexpr double kDblEps = 2.220446049250313e-16;  // DBL_EPSILON

// Particle state is streamed: a kernel reads and writes each particle's
// position, displacement, velocity, stress, strain, F, volume, density,
// state variables and material once, so these accesses carry the
// nontemporal hint and leave the L2 to the node and cell data that
// neighbouring particles gather again (collapse step -2 %). Local
// coordinates, cell ids and masses keep the default caching, which
// measured faster for the scatter kernels that read them through the
// cell-sorted permutation; so do the ordered-mode gathers, which read each
// particle once per node of its cell.
template <class T>
MPMX_DEV T stream_ld(const T* a) {
  return __builtin_nontemporal_load(a);
}
template <class T>
MPMX_DEV void stream_st(T* a, T v) {
  __builtin_nontemporal_store(v, a);
}

template <int D> struct Elem;
template <> struct Elem<2> { static constexpr int NN = 4; };
template <> struct Elem<3> { static constexpr int NN = 8; };

// quadrilateral_element.tcc:11-43, hexahedron_element.tcc:19-73
template <int D> MPMX_DEV void shapefn(const double* xi, double* N);
template <> MPMX_DEV void shapefn<2>(const double* xi, double* N) {
  N[0] = 0.25 * (1 - xi[0]) * (1 - xi[1]);
  N[1] = 0.25 * (1 + xi[0]) * (1 - xi[1]);
  N[2] = 0.25 * (1 + xi[0]) * (1 + xi[1]);
  N[3] = 0.25 * (1 - xi[0]) * (1 + xi[1]);
}
template <> MPMX_DEV void shapefn<3>(const double* xi, double* N) {
  N[0] = 0.125 * (1 - xi[0]) * (1 - xi[1]) * (1 - xi[2]);
  N[1] = 0.125 * (1 + xi[0]) * (1 - xi[1]) * (1 - xi[2]);
  N[2] = 0.125 * (1 + xi[0]) * (1 + xi[1]) * (1 - xi[2]);
  N[3] = 0.125 * (1 - xi[0]) * (1 + xi[1]) * (1 - xi[2]);
  N[4] = 0.125 * (1 - xi[0]) * (1 - xi[1]) * (1 + xi[2]);
  N[5] = 0.125 * (1 + xi[0]) * (1 - xi[1]) * (1 + xi[2]);
  N[6] = 0.125 * (1 + xi[0]) * (1 + xi[1]) * (1 + xi[2]);
  N[7] = 0.125 * (1 - xi[0]) * (1 + xi[1]) * (1 + xi[2]);
}

// One shape function N_i(xi), same expression as shapefn<D>.
template <int D> MPMX_DEV double shapefn_i(const double* xi, int i);
template <> MPMX_DEV double shapefn_i<2>(const double* xi, int i) {
  const double a = (i == 0 || i == 3) ? (1 - xi[0]) : (1 + xi[0]);
  const double b = (i < 2) ? (1 - xi[1]) : (1 + xi[1]);
  return 0.25 * a * b;
}
template <> MPMX_DEV double shapefn_i<3>(const double* xi, int i) {
  const int l = i & 3;
  const double a = (l == 0 || l == 3) ? (1 - xi[0]) : (1 + xi[0]);
  const double b = (l < 2) ? (1 - xi[1]) : (1 + xi[1]);
  const double c = (i < 4) ? (1 - xi[2]) : (1 + xi[2]);
  return 0.125 * a * b * c;
}

// Gradient dN_i/dxi (same expressions as the element files).
template <int D> MPMX_DEV void grad_i(const double* xi, int i, double* g);
template <> MPMX_DEV void grad_i<2>(const double* xi, int i, double* g) {
  switch (i) {
    case 0: g[0] = -0.25 * (1 - xi[1]); g[1] = -0.25 * (1 - xi[0]); break;
    case 1: g[0] = 0.25 * (1 - xi[1]);  g[1] = -0.25 * (1 + xi[0]); break;
    case 2: g[0] = 0.25 * (1 + xi[1]);  g[1] = 0.25 * (1 + xi[0]);  break;
    default: g[0] = -0.25 * (1 + xi[1]); g[1] = 0.25 * (1 - xi[0]); break;
  }
}
template <> MPMX_DEV void grad_i<3>(const double* xi, int i, double* g) {
  const double m0 = 1 - xi[0], p0 = 1 + xi[0];
  const double m1 = 1 - xi[1], p1 = 1 + xi[1];
  const double m2 = 1 - xi[2], p2 = 1 + xi[2];
  switch (i) {
    case 0: g[0] = -0.125 * m1 * m2; g[1] = -0.125 * m0 * m2; g[2] = -0.125 * m0 * m1; break;
    case 1: g[0] = 0.125 * m1 * m2;  g[1] = -0.125 * p0 * m2; g[2] = -0.125 * p0 * m1; break;
    case 2: g[0] = 0.125 * p1 * m2;  g[1] = 0.125 * p0 * m2;  g[2] = -0.125 * p0 * p1; break;
    case 3: g[0] = -0.125 * p1 * m2; g[1] = 0.125 * m0 * m2;  g[2] = -0.125 * m0 * p1; break;
    case 4: g[0] = -0.125 * m1 * p2; g[1] = -0.125 * m0 * p2; g[2] = 0.125 * m0 * m1;  break;
    case 5: g[0] = 0.125 * m1 * p2;  g[1] = -0.125 * p0 * p2; g[2] = 0.125 * p0 * m1;  break;
    case 6: g[0] = 0.125 * p1 * p2;  g[1] = 0.125 * p0 * p2;  g[2] = 0.125 * p0 * p1;  break;
    default: g[0] = -0.125 * p1 * p2; g[1] = 0.125 * m0 * p2; g[2] = 0.125 * m0 * p1;  break;
  }
}

template <int D>
MPMX_DEV void load_cell_X(const Mesh& M, int c, double (*X)[D]) {
  constexpr int NN = Elem<D>::NN;
#pragma unroll
  for (int i = 0; i < NN; ++i) {
    const int n = M.cell_nodes[(int64_t)c * NN + i];
#pragma unroll
    for (int k = 0; k < D; ++k) X[i][k] = M.node_x[k * M.nnodes + n];
  }
}

template <int D>
MPMX_DEV void mat_inverse(const double (*J)[D], double (*Ji)[D]);

// J = grad^T X (sequential over nodes), J^-1 by adjugate / determinant.
// quadrilateral_element.tcc:295-311, hexahedron_element.tcc:307-321
template <int D>
MPMX_DEV void jacobian_inverse(const double* xi, const double (*X)[D], double (*Ji)[D]) {
  constexpr int NN = Elem<D>::NN;
  double J[D][D];
  {
    double g[D];
    grad_i<D>(xi, 0, g);
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
      for (int b = 0; b < D; ++b) J[a][b] = g[a] * X[0][b];
#pragma unroll
    for (int i = 1; i < NN; ++i) {
      grad_i<D>(xi, i, g);
#pragma unroll
      for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = 0; b < D; ++b) J[a][b] = J[a][b] + g[a] * X[i][b];
    }
  }
  mat_inverse<D>(J, Ji);
}

// Inverse of a DxD matrix by the adjugate over the determinant (cofactor
// expansion along the first column).
template <int D>
MPMX_DEV void mat_inverse(const double (*J)[D], double (*Ji)[D]) {
  if constexpr (D == 2) {
    const double det = J[0][0] * J[1][1] - J[1][0] * J[0][1];
    const double invdet = 1.0 / det;
    Ji[0][0] = J[1][1] * invdet;
    Ji[1][0] = -J[1][0] * invdet;
    Ji[0][1] = -J[0][1] * invdet;
    Ji[1][1] = J[0][0] * invdet;
  } else {
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
    const double det = (J[0][0] * c[0][0] + J[1][0] * c[0][1]) + J[2][0] * c[0][2];
    const double invdet = 1.0 / det;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) Ji[i][j] = c[i][j] * invdet;
  }
}

// dN_i/dx = grad_i * (J^-1)^T (sequential over b)
template <int D>
MPMX_DEV void dndx_i(const double* xi, int i, const double (*Ji)[D], double* out) {
  double g[D];
  grad_i<D>(xi, i, g);
#pragma unroll
  for (int a = 0; a < D; ++a) {
    double s = g[0] * Ji[a][0];
#pragma unroll
    for (int b = 1; b < D; ++b) s = s + g[b] * Ji[a][b];
    out[a] = s;
  }
}

// ------------------ isoparametric local coordinates -------------------------
// Cell::transform_real_to_unit_cell (cell.tcc:617-828) for quad4 / hex8, in
// the same operation order as the oracle's restatement.

// Vertex i of the unit cell (quad counter-clockwise from (-1,-1); hex the
// same square at zeta = -1, then +1), coordinate a.
MPMX_DEV double unit_vertex(int i, int a) {
  const int l = i & 3;
  if (a == 0) return (l == 0 || l == 3) ? -1. : 1.;
  if (a == 1) return l < 2 ? -1. : 1.;
  return i < 4 ? -1. : 1.;
}

// r = X^T N(xi) - p (sequential over nodes)
template <int D>
MPMX_DEV void iso_residual(const double (*X)[D], const double* xi, const double* p, double* r) {
  constexpr int NN = Elem<D>::NN;
  double N[NN];
  shapefn<D>(xi, N);
#pragma unroll
  for (int k = 0; k < D; ++k) {
    double s = X[0][k] * N[0];
#pragma unroll
    for (int i = 1; i < NN; ++i) s = s + X[i][k] * N[i];
    r[k] = s - p[k];
  }
}

template <int D>
MPMX_DEV double sqnorm(const double* v) {
  double s = v[0] * v[0];
#pragma unroll
  for (int k = 1; k < D; ++k) s = s + v[k] * v[k];
  return s;
}

// Closed-form inverse of the quad4 bilinear map (after Zhao et al. 1999,
// QuadrilateralElement<2,4>::natural_coordinates_analytical,
// quadrilateral_element.tcc:523-754): with 4x = a1 + a2 xi + a3 eta +
// a4 xi eta (b for y) the point solves a2 xi + a3 eta + a4 xi eta = c1,
// b2 xi + b3 eta + b4 xi eta = c2. A parallelogram (a4 = b4 = 0) is a linear
// system; otherwise eliminating the product gives p xi + q eta = k and a
// quadratic in xi (q != 0), or xi itself (q = 0). Returns the root with both
// coordinates in [-1, 1] (smaller xi first), else DBL_MAX.
MPMX_DEV void q4_natural_coordinates(const double (*X)[2], const double* pt, double* xi) {
  const double big = 1.7976931348623157e308;
  xi[0] = xi[1] = big;
  const double x1 = X[0][0], y1 = X[0][1], x2 = X[1][0], y2 = X[1][1];
  const double x3 = X[2][0], y3 = X[2][1], x4 = X[3][0], y4 = X[3][1];
  const double a1 = x1 + x2 + x3 + x4, a2 = -x1 + x2 + x3 - x4;
  const double a3 = -x1 - x2 + x3 + x4, a4 = x1 - x2 + x3 - x4;
  const double b1 = y1 + y2 + y3 + y4, b2 = -y1 + y2 + y3 - y4;
  const double b3 = -y1 - y2 + y3 + y4, b4 = y1 - y2 + y3 - y4;
  const double c1 = 4. * pt[0] - a1, c2 = 4. * pt[1] - b1;
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
    if (u >= -1. && u <= 1. && v >= -1. && v <= 1.) {
      xi[0] = u;
      xi[1] = v;
    }
    return;
  }
  const double A = -a4 * pl, B = a2 * ql + a4 * kl - a3 * pl, C = a3 * kl - c1 * ql;
  double r[2];
  int nr = 0;
  if (A == 0.) {
    r[nr++] = -C / B;
  } else {
    const double disc = B * B - 4. * A * C;
    if (!(disc >= 0.)) return;
    const double sd = sqrt(disc);
    const double qq = -0.5 * (B + (B >= 0. ? sd : -sd));
    const double ra = qq / A;
    const double rb = qq != 0. ? C / qq : ra;
    r[nr++] = ra < rb ? ra : rb;
    r[nr++] = ra < rb ? rb : ra;
  }
  for (int j = 0; j < nr; ++j) {
    const double u = r[j];
    const double v = (kl - pl * u) / ql;
    if (u >= -1. && u <= 1. && v >= -1. && v <= 1.) {
      xi[0] = u;
      xi[1] = v;
      return;
    }
  }
}

// The isoparametric branch of transform_real_to_unit_cell: the closed form
// (2D) with its early exit, the affine guess A^-1 (p - b) of the
// least-squares affine map (A = X^T KA, b = X^T Kb with KA = unit vertex /
// nn, Kb = 1 / nn) with its early exit, then Newton-Raphson from the better
// of the two with a halving line search (at most 10000 iterations; the
// iteration matrix is jacobian_local on the unit cell's own vertices,
// cell.tcc:772). Returns false where the reference throws (NaN).
template <int D>
MPMX_DEV bool transform_real_to_unit_cell(const double (*X)[D], double mean_length,
                                          const double* p, double* xi) {
  constexpr int NN = Elem<D>::NN;
  const double big = 1.7976931348623157e308;
  double an_xi[D], an_res[D];
#pragma unroll
  for (int k = 0; k < D; ++k) an_xi[k] = an_res[k] = big;
  const double tol2 = 1.0E-16 * mean_length * mean_length;
  if constexpr (D == 2) {
    q4_natural_coordinates(X, p, an_xi);
    bool in_cell = true;
#pragma unroll
    for (int k = 0; k < D; ++k)
      if (an_xi[k] < -1. || an_xi[k] > 1. || isnan(an_xi[k])) in_cell = false;
    iso_residual<D>(X, an_xi, p, an_res);
    if (sqnorm<D>(an_res) < tol2 && in_cell) {
#pragma unroll
      for (int k = 0; k < D; ++k) xi[k] = an_xi[k];
      return true;
    }
  }
  const double inv_nn = 1.0 / NN;
  double A[D][D], Ai[D][D], b[D];
#pragma unroll
  for (int k = 0; k < D; ++k) {
#pragma unroll
    for (int a = 0; a < D; ++a) {
      double s = X[0][k] * (unit_vertex(0, a) * inv_nn);
#pragma unroll
      for (int i = 1; i < NN; ++i) s = s + X[i][k] * (unit_vertex(i, a) * inv_nn);
      A[k][a] = s;
    }
    double s = X[0][k] * inv_nn;
#pragma unroll
    for (int i = 1; i < NN; ++i) s = s + X[i][k] * inv_nn;
    b[k] = p[k] - s;
  }
  mat_inverse<D>(A, Ai);
  double af_xi[D], af_res[D];
  bool af_nan = false;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    double s = Ai[k][0] * b[0];
#pragma unroll
    for (int a = 1; a < D; ++a) s = s + Ai[k][a] * b[a];
    af_xi[k] = s;
    af_res[k] = big;
    if (isnan(s)) af_nan = true;
  }
  if (!af_nan) {
    iso_residual<D>(X, af_xi, p, af_res);
    if (sqnorm<D>(af_res) < tol2) {
#pragma unroll
      for (int k = 0; k < D; ++k) xi[k] = af_xi[k];
      return true;
    }
  }
  double g_xi[D], g_res[D];
#pragma unroll
  for (int k = 0; k < D; ++k) {
    g_xi[k] = 0.;
    g_res[k] = big;
  }
  if (sqrt(sqnorm<D>(an_res)) < sqrt(sqnorm<D>(af_res))) {
#pragma unroll
    for (int k = 0; k < D; ++k) {
      g_xi[k] = an_xi[k];
      g_res[k] = an_res[k];
    }
  } else if (!af_nan) {
#pragma unroll
    for (int k = 0; k < D; ++k) {
      g_xi[k] = af_xi[k];
      g_res[k] = af_res[k];
    }
  }
  double nr[D], nr_res[D];
#pragma unroll
  for (int k = 0; k < D; ++k) {
    nr[k] = g_xi[k];
    if (nr[k] < -1. && nr[k] > -1.001)
      nr[k] = -1. + 1.E-12;
    else if (nr[k] > 1. && nr[k] < 1.001)
      nr[k] = 1. - 1.E-12;
    nr_res[k] = big;
  }
  for (int iter = 0; iter < 10000; ++iter) {
    double J[D][D], Ji[D][D];
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
      for (int c = 0; c < D; ++c) J[a][c] = 0.;
    {
      double g[D];
      grad_i<D>(nr, 0, g);
#pragma unroll
      for (int a = 0; a < D; ++a)
#pragma unroll
        for (int c = 0; c < D; ++c) J[a][c] = g[a] * unit_vertex(0, c);
#pragma unroll
      for (int i = 1; i < NN; ++i) {
        grad_i<D>(nr, i, g);
#pragma unroll
        for (int a = 0; a < D; ++a)
#pragma unroll
          for (int c = 0; c < D; ++c) J[a][c] = J[a][c] + g[a] * unit_vertex(i, c);
      }
    }
    double det;
    if constexpr (D == 2)
      det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    else
      det = J[0][0] * (J[1][1] * J[2][2] - J[2][1] * J[1][2]) -
            J[1][0] * (J[0][1] * J[2][2] - J[2][1] * J[0][2]) +
            J[2][0] * (J[0][1] * J[1][2] - J[1][1] * J[0][2]);
    if (fabs(det) < 1.0E-10) {
#pragma unroll
      for (int k = 0; k < D; ++k) nr[k] = 0.;
    }
    iso_residual<D>(X, nr, p, nr_res);
    mat_inverse<D>(J, Ji);
    double delta[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
      double s = Ji[k][0] * nr_res[0];
#pragma unroll
      for (int a = 1; a < D; ++a) s = s + Ji[k][a] * nr_res[a];
      delta[k] = s;
    }
    double step = 1.;
    const double rn = sqrt(sqnorm<D>(nr_res));
    for (int t = 0; t < 10; ++t) {
      double trial[D], tres[D];
#pragma unroll
      for (int k = 0; k < D; ++k) trial[k] = nr[k] - step * delta[k];
      iso_residual<D>(X, trial, p, tres);
      if (sqrt(sqnorm<D>(tres)) < rn) {
#pragma unroll
        for (int k = 0; k < D; ++k) {
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
    double sdel[D];
#pragma unroll
    for (int k = 0; k < D; ++k) sdel[k] = step * delta[k];
    if (sqrt(sqnorm<D>(sdel)) < 1.0E-10) break;
    if (isnan(nr[0]) || isnan(nr[1])) {
#pragma unroll
      for (int k = 0; k < D; ++k) nr[k] = 0.;
    }
  }
  const bool geo = sqrt(sqnorm<D>(g_res)) < sqrt(sqnorm<D>(nr_res));
#pragma unroll
  for (int k = 0; k < D; ++k) xi[k] = geo ? g_xi[k] : nr[k];
  return !(isnan(xi[0]) || isnan(xi[1]));
}

// --------------------------- locate (cell.tcc:324-380) ----------------------
// ISO: the mesh is isoparametric (local coordinates by
// transform_real_to_unit_cell); else Cartesian (local_coordinates_point).
template <int D, bool ISO = false>
MPMX_DEV bool point_in_cell(const Mesh& M, int c, const double* p, double* xi) {
  {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) {
      const double d = p[k] - M.centroid[k * M.ncells + c];
      s = (k == 0) ? d * d : s + d * d;
    }
    if (!(sqrt(s) < M.two_mean[c])) return false;
  }
  bool status = true;
  if constexpr (ISO) {
    constexpr int NN = Elem<D>::NN;
    double X[NN][D];
    load_cell_X<D>(M, c, X);
    if (!transform_real_to_unit_cell<D>(X, M.two_mean[c] * 0.5, p, xi)) return false;
  }
#pragma unroll
  for (int k = 0; k < D; ++k) {
    double x;
    if constexpr (ISO)
      x = xi[k];
    else
      x = 2. * (p[k] - M.centroid[k * M.ncells + c]) / M.len[k * M.ncells + c];
    if (x < -1. || x > 1. || isnan(x)) {
      status = false;
    } else {
      if (x < -1. + kDblEps) x = -1. + kDblEps;
      if (x > 1. - kDblEps) x = 1. - kDblEps;
    }
    xi[k] = x;
  }
  return status;
}

// ----------------------- material_utility.tcc ------------------------------
MPMX_DEV double sq(double x) { return x * x; }
MPMX_DEV double mu_p(const double* s) { return (1. / 3. * (s[0] + s[1] + s[2])); }
MPMX_DEV double mu_j2(const double* s) {
  return (sq(s[0] - s[1]) + sq(s[1] - s[2]) + sq(s[0] - s[2])) / 6.0 + sq(s[3]) + sq(s[4]) +
         sq(s[5]);
}
MPMX_DEV void mu_dev(const double* s, double* d) {
  const double p = mu_p(s);
#pragma unroll
  for (int i = 0; i < 6; ++i) d[i] = s[i];
#pragma unroll
  for (int i = 0; i < 3; ++i) d[i] -= p;
}
MPMX_DEV double mu_j3(const double* s) {
  double d[6];
  mu_dev(s, d);
  return (d[0] * d[1] * d[2]) - (d[2] * sq(d[3])) +
         ((2 * d[3] * d[4] * d[5]) - (d[0] * sq(d[4])) - (d[1] * sq(d[5])));
}
MPMX_DEV double mu_q(const double* s) { return sqrt(3 * mu_j2(s)); }
MPMX_DEV double mu_lode(const double* s) {
  const double j2 = mu_j2(s);
  const double j3 = mu_j3(s);
  double v = 0.0;
  if (fabs(j2) > kDblEps) v = (3. * sqrt(3.) / 2.) * (j3 / dmath::pow_1p5(j2));
  if (v > 1.0) v = 1.0;
  if (v < -1.0) v = -1.0;
  return (1. / 3.) * dmath::acos(v);
}
MPMX_DEV void mu_dq_dsigma(const double* s, double* out) {
  const double q = mu_q(s);
  double d[6];
  mu_dev(s, d);
#pragma unroll
  for (int i = 0; i < 6; ++i) out[i] = 0.;
  if (fabs(q) > kDblEps) {
    out[0] = 3. / (2. * q) * d[0];
    out[1] = 3. / (2. * q) * d[1];
    out[2] = 3. / (2. * q) * d[2];
    out[3] = 3. / q * d[3];
    out[4] = 3. / q * d[4];
    out[5] = 3. / q * d[5];
  }
}
MPMX_DEV double dot3(const double* a, const double* b) {
  return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}
MPMX_DEV void mu_dtheta_dsigma(const double* s, double* out) {
  const double tol = kDblEps;
  const double j2 = mu_j2(s);
  const double j3 = mu_j3(s);
  double dj2[6];
  mu_dev(s, dj2);
  dj2[3] *= 2.0;
  dj2[4] *= 2.0;
  dj2[5] *= 2.0;
  double dv[6];
  mu_dev(s, dv);
  const double d1[3] = {dv[0], dv[3], dv[5]};
  const double d2[3] = {dv[3], dv[1], dv[4]};
  const double d3[3] = {dv[5], dv[4], dv[2]};
  double dj3[6];
  dj3[0] = dot3(d1, d1) - (2. / 3.) * j2;
  dj3[1] = dot3(d2, d2) - (2. / 3.) * j2;
  dj3[2] = dot3(d3, d3) - (2. / 3.) * j2;
  dj3[3] = 2.0 * dot3(d1, d2);
  dj3[4] = 2.0 * dot3(d2, d3);
  dj3[5] = 2.0 * dot3(d1, d3);
  double dr_dj2 = -9.0 / 4.0 * sqrt(3.0) * j3;
  double dr_dj3 = 3.0 / 2.0 * sqrt(3.0);
  double dtheta_dr = -1.0 / 3.0;
  if (fabs(j2) > tol) {
    double r = j3 / 2.0 * dmath::pow_m1p5(j2 / 3.0);
    dr_dj2 *= dmath::pow_m2p5(j2);
    dr_dj3 *= dmath::pow_m1p5(j2);
    const double factor = (fabs(1 - r * r) < tol) ? tol : fabs(1 - r * r);
    dtheta_dr = -1.0 / (3.0 * sqrt(factor));
  }
#pragma unroll
  for (int i = 0; i < 6; ++i) out[i] = dtheta_dr * ((dr_dj2 * dj2[i]) + (dr_dj3 * dj3[i]));
}

// --------------------------- elastic tensor products ------------------------
// D x with D = [[a1,a2,a2],[a2,a1,a2],[a2,a2,a1]] (+) G I3; the zero entries of
// the reference's 6x6 product only add exact zeros and are skipped.
MPMX_DEV void de_mul(const Material& m, const double* x, double* y) {
  y[0] = (m.a1 * x[0] + m.a2 * x[1]) + m.a2 * x[2];
  y[1] = (m.a2 * x[0] + m.a1 * x[1]) + m.a2 * x[2];
  y[2] = (m.a2 * x[0] + m.a2 * x[1]) + m.a1 * x[2];
  y[3] = m.G * x[3];
  y[4] = m.G * x[4];
  y[5] = m.G * x[5];
}
// exp(A) of the skew matrix A with A01 = a[0], A02 = a[1], A12 = a[2], by
// Rodrigues' formula I + sin(t)/t A + (1 - cos t)/t^2 A^2 (t = |a|); below
// t = 1e-2 the two coefficients by their Taylor series to t^6 (truncation
// below 1e-21). Stands for Eigen's MatrixExponential in
// linear_elastic.tcc:106-107; both are exact to rounding for the rotations
// of a time step.
MPMX_DEV void skew_exp(const double* a, double (*Q)[3]) {
  const double t2 = (a[0] * a[0] + a[1] * a[1]) + a[2] * a[2];
  double c1, c2;
  if (t2 < 1e-4) {
    c1 = 1. - t2 * (1. / 6. - t2 * (1. / 120. - t2 * (1. / 5040.)));
    c2 = 0.5 - t2 * (1. / 24. - t2 * (1. / 720. - t2 * (1. / 40320.)));
  } else {
    const double t = sqrt(t2);
    double sn, cs;
    dmath::sincos(t, &sn, &cs);
    c1 = sn / t;
    c2 = (1. - cs) / t2;
  }
  const double A[3][3] = {{0., a[0], a[1]}, {-a[0], 0., a[2]}, {-a[1], -a[2], 0.}};
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double a2 = A[i][0] * A[0][j];
#pragma unroll
      for (int k = 1; k < 3; ++k) a2 = a2 + A[i][k] * A[k][j];
      Q[i][j] = ((i == j ? 1. : 0.) + c1 * A[i][j]) + c2 * a2;
    }
}

// Q X Q^T of a symmetric X, in Voigt order (00, 11, 22, 01, 12, 02);
// sequential k-order in both products.
MPMX_DEV void rotate_sym(const double (*Q)[3], const double (*X)[3], double* y) {
  double T[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double t = Q[i][0] * X[0][j];
#pragma unroll
      for (int k = 1; k < 3; ++k) t = t + Q[i][k] * X[k][j];
      T[i][j] = t;
    }
  constexpr int I[6] = {0, 1, 2, 0, 1, 0}, J[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
  for (int v = 0; v < 6; ++v) {
    double t = T[I[v]][0] * Q[J[v]][0];
#pragma unroll
    for (int k = 1; k < 3; ++k) t = t + T[I[v]][k] * Q[J[v]][k];
    y[v] = t;
  }
}

// C = A B and C = A B^T of 3x3 matrices (sequential k)
MPMX_DEV void mat3_mul(const double (*A)[3], const double (*B)[3], double (*C)[3]) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double t = A[i][0] * B[0][j];
#pragma unroll
      for (int k = 1; k < 3; ++k) t = t + A[i][k] * B[k][j];
      C[i][j] = t;
    }
}
MPMX_DEV void mat3_mul_t(const double (*A)[3], const double (*B)[3], double (*C)[3]) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double t = A[i][0] * B[j][0];
#pragma unroll
      for (int k = 1; k < 3; ++k) t = t + A[i][k] * B[j][k];
      C[i][j] = t;
    }
}

// The rotation R of the polar decomposition F = R U (det F > 0), which
// mpm::materials::compute_rotation_tensor (material_utility.tcc) takes from
// the eigen-decomposition of F^T F: by Newton's iteration
// X <- (X + X^-T) / 2 from X = F (Higham 1986) until an iteration changes no
// entry by more than 1e-15 of the largest (at most 100).
MPMX_DEV void polar_rotation(const double (*F)[3], double (*R)[3]) {
  double X[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) X[i][j] = F[i][j];
  for (int it = 0; it < 100; ++it) {
    double Xi[3][3];
    mat_inverse<3>(X, Xi);
    double diff = 0., big = 0.;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const double n = 0.5 * (X[i][j] + Xi[j][i]);
        const double d = fabs(n - X[i][j]), a = fabs(n);
        if (d > diff) diff = d;
        if (a > big) big = a;
        X[i][j] = n;
      }
    if (diff <= 1.0E-15 * big) break;
  }
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) R[i][j] = X[i][j];
}

// The two rotations of the Green-Naghdi rate (linear_elastic.tcc:134-150):
// Lambda_delta = R_new R_half^T (strain) and Lambda_Delta = R_new R_current^T
// (stress), R the rotations of F, (F + F_new) / 2 and F_new = dF F.
static __device__ __attribute__((noinline)) void green_naghdi_rotations(
    const double (*F)[3], const double (*dF)[3], double (*L)[3], double (*LL)[3]) {
  double Fn[3][3], Fh[3][3], Rc[3][3], Rh[3][3], Rn[3][3];
  mat3_mul(dF, F, Fn);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) Fh[i][j] = 0.5 * F[i][j] + 0.5 * Fn[i][j];
  polar_rotation(F, Rc);
  polar_rotation(Fh, Rh);
  polar_rotation(Fn, Rn);
  mat3_mul_t(Rn, Rh, L);
  mat3_mul_t(Rn, Rc, LL);
}

// The trial stress of an objective rate (linear_elastic.tcc:109-124,
// 152-164; infinitesimal_elasto_plastic.tcc:1-25): the stress s rotated by
// LL, the strain increment ds (tensor shears) by L, plus the elastic
// increment of the rotated strain. s: in/out.
static __device__ __attribute__((noinline)) void rate_trial(const Material& m,
                                                           const double (*L)[3],
                                                           const double (*LL)[3],
                                                           const double* ds, double* s) {
  const double S[3][3] = {{s[0], s[3], s[5]}, {s[3], s[1], s[4]}, {s[5], s[4], s[2]}};
  const double E[3][3] = {{ds[0], 0.5 * ds[3], 0.5 * ds[5]},
                          {0.5 * ds[3], ds[1], 0.5 * ds[4]},
                          {0.5 * ds[5], 0.5 * ds[4], ds[2]}};
  double rs[6], re[6], d[6];
  rotate_sym(LL, S, rs);
  rotate_sym(L, E, re);
#pragma unroll
  for (int v = 3; v < 6; ++v) re[v] = 2.0 * re[v];
  de_mul(m, re, d);
#pragma unroll
  for (int v = 0; v < 6; ++v) s[v] = rs[v] + d[v];
}

// The Jaumann rotations (linear_elastic.tcc:102-107): L = exp(W/2) from the
// entries a = 01, 02, 12 of W/2 (W the skew part of dF - I) and LL = L L.
MPMX_DEV void jaumann_rotations(const double* a, double (*L)[3], double (*LL)[3]) {
  skew_exp(a, L);
  mat3_mul(L, L, LL);
}

// x^T D (D symmetric in structure; column sums in k order)
MPMX_DEV void de_tmul(const Material& m, const double* x, double* y) {
  y[0] = (x[0] * m.a1 + x[1] * m.a2) + x[2] * m.a2;
  y[1] = (x[0] * m.a2 + x[1] * m.a1) + x[2] * m.a2;
  y[2] = (x[0] * m.a2 + x[1] * m.a2) + x[2] * m.a1;
  y[3] = x[3] * m.G;
  y[4] = x[4] * m.G;
  y[5] = x[5] * m.G;
}
MPMX_DEV double dot6(const double* a, const double* b) {
  double s = a[0] * b[0];
#pragma unroll
  for (int k = 1; k < 6; ++k) s = s + a[k] * b[k];
  return s;
}

// ------------------------------ Mohr-Coulomb --------------------------------
// mohr_coulomb.tcc:85-533. Every sin/cos/tan/acos/pow of the reference is
// evaluated with the deterministic functions of mpmx_math.h; each distinct
// argument is evaluated once and reused (the functions are pure, so the bits
// equal those of the reference's repeated calls).
enum { FS_ELASTIC = 0, FS_SHEAR = 1, FS_TENSILE = 2 };

// trigonometry of the friction and dilation angles
struct McAngles {
  double sphi, cphi, tphi, tpsi;
};
MPMX_DEV McAngles mc_angles(double phi, double psi) {
  McAngles a;
  dmath::sincos(phi, &a.sphi, &a.cphi);
  a.tphi = a.sphi / a.cphi;  // = dmath::tan(phi)
  a.tpsi = dmath::tan(psi);
  return a;
}

// The angles' trigonometry for state-variable angles phi, psi: the
// material's precomputed values when they are its peak angles (always
// without softening), the same bits as mc_angles(phi, psi).
MPMX_DEV McAngles mc_angles_of(const Material& m, double phi, double psi) {
  if (phi == m.phi_peak && psi == m.psi_peak) {
    McAngles a;
    a.sphi = m.sphi;
    a.cphi = m.cphi;
    a.tphi = m.tphi;
    a.tpsi = m.tpsi;
    return a;
  }
  return mc_angles(phi, psi);
}

// Haigh-Westergaard invariants (the state variables epsilon, rho, theta)
// plus the trigonometry of the Lode angle.
struct McInv {
  double eps, rho, theta;
  double st, ct;    // sin, cos of theta
  double s3, c3;    // sin, cos of theta + pi/3
};

// compute_stress_invariants (mohr_coulomb.tcc:124-134), without the trigonometry
MPMX_DEV void mc_invariants_only(const double* s, McInv& I) {
  I.eps = mu_p(s) * sqrt(3.);
  I.theta = mu_lode(s);
  I.rho = sqrt(2. * mu_j2(s));
}

MPMX_DEV void mc_invariants(const double* s, McInv& I) {
  mc_invariants_only(s, I);
  dmath::sincos(I.theta, &I.st, &I.ct);
  dmath::sincos(I.theta + M_PI / 3., &I.s3, &I.c3);
}

// compute_yield_state (mohr_coulomb.tcc:139-191)
MPMX_DEV int mc_yield(double* yf, const McInv& I, const McAngles& A, double cohesion,
                      double tension_cutoff) {
  const double Tolerance = -1E-1;
  const double epsilon = I.eps, rho = I.rho, theta = I.theta;
  yf[0] = sqrt(2. / 3.) * I.ct * rho + epsilon / sqrt(3.) - tension_cutoff;
  yf[1] = sqrt(1.5) * rho * ((I.s3 / (sqrt(3.) * A.cphi)) + (I.c3 * A.tphi / 3.)) +
          (epsilon / sqrt(3.)) * A.tphi - cohesion;
  int t = FS_ELASTIC;
  if (yf[0] > Tolerance && yf[1] > Tolerance) {
    const double n_phi = (1. + A.sphi) / (1. - A.sphi);
    const double sigma_p = tension_cutoff * n_phi - 2. * cohesion * sqrt(n_phi);
    const double alpha_p = sqrt(1. + n_phi * n_phi) + n_phi;
    const double h =
        yf[0] + alpha_p * (sqrt(2. / 3.) * dmath::cos(theta - 4. * M_PI / 3.) * rho +
                           epsilon / sqrt(3.) - sigma_p);
    t = (h > kDblEps) ? FS_TENSILE : FS_SHEAR;
  }
  if (yf[0] < Tolerance && yf[1] > Tolerance) t = FS_SHEAR;
  if (yf[0] > Tolerance && yf[1] < Tolerance) t = FS_TENSILE;
  return t;
}

// compute_df_dp (mohr_coulomb.tcc:195-344); I holds the invariants of `stress`
MPMX_DEV void mc_df_dp(const Material& m, int yt, double tension_cutoff, double pdstrain,
                       const double* stress, const McInv& I, const McAngles& A, double* df,
                       double* dp, double* dp_dq, double* softening) {
  const double rho = I.rho;
  const double ct = I.ct, st = I.st;
  double df_depsilon, df_drho, df_dtheta;
  if (yt == FS_TENSILE) {
    df_depsilon = 1. / sqrt(3.);
    df_drho = sqrt(2. / 3.) * ct;
    df_dtheta = -sqrt(2. / 3.) * rho * st;
  } else {
    df_depsilon = A.tphi / sqrt(3.);
    df_drho = sqrt(1.5) * ((I.s3 / (sqrt(3.) * A.cphi)) + (I.c3 * A.tphi / 3.));
    df_dtheta = sqrt(1.5) * rho * ((I.c3 / (sqrt(3.) * A.cphi)) - (I.s3 * A.tphi / 3.));
  }
  const double depsilon_dsigma = 1. / 3. * sqrt(3.);  // entries 0..2 (3..5 are 0)
  double drho_dsigma[6];
  mu_dq_dsigma(stress, drho_dsigma);
  for (int i = 0; i < 6; ++i) drho_dsigma[i] = drho_dsigma[i] * sqrt(2. / 3.);
  double dtheta_dsigma[6];
  mu_dtheta_dsigma(stress, dtheta_dsigma);
  for (int i = 0; i < 6; ++i)
    df[i] = ((df_depsilon * (i < 3 ? depsilon_dsigma : 0.0)) + (df_drho * drho_dsigma[i])) +
            (df_dtheta * dtheta_dsigma[i]);
  if (yt == FS_TENSILE) {
    const double et_value = 0.6;
    const double xit = 0.1;
    double sqpart = 4. * (1 - et_value * et_value) * ct * ct + 5. * et_value * et_value -
                    4. * et_value;
    if (sqpart < kDblEps) sqpart = 1.E-5;
    double rt_den = 2. * (1 - et_value * et_value) * ct + (2. * et_value - 1) * sqrt(sqpart);
    const double rt_num = 4. * (1 - et_value * et_value) * ct * ct +
                          (2. * et_value - 1) * (2. * et_value - 1);
    if (fabs(rt_den) < kDblEps) rt_den = 1.E-5;
    const double rt = rt_num / (3. * rt_den);
    const double den = sqrt(xit * xit * tension_cutoff * tension_cutoff + 1.5 * rt * rt * rho * rho);
    const double cl = den > 1.0e-15 ? den : 1.0e-15;
    const double dp_drt = 1.5 * rho * rho * rt / cl;
    const double dp_drho = 1.5 * rho * rt * rt / cl;
    const double dp_depsilon = 1. / sqrt(3.);
    const double drtden_dtheta =
        -2. * (1 - et_value * et_value) * st -
        (2. * et_value - 1) * 4. * (1 - et_value * et_value) * ct * st /
            sqrt(4. * (1 - et_value * et_value) * ct * ct + 5. * et_value * et_value -
                 4. * et_value);
    const double drtnum_dtheta = -8. * (1 - et_value * et_value) * ct * st;
    const double drt_dtheta =
        (drtnum_dtheta * rt_den - drtden_dtheta * rt_num) / (3. * rt_den * rt_den);
    for (int i = 0; i < 6; ++i)
      dp[i] = ((dp_depsilon * (i < 3 ? depsilon_dsigma : 0.0)) + (dp_drho * drho_dsigma[i])) +
              (dp_drt * drt_dtheta * dtheta_dsigma[i]);
    *dp_dq = dp_drho * sqrt(2. / 3.);
  } else {
    const double r_mc = (3. - A.sphi) / (6 * A.cphi);
    double e_val = (3. - A.sphi) / (3. + A.sphi);
    if (e_val <= 0.5) e_val = 0.5 + 1.E-10;
    if (e_val > 1.) e_val = 1.;
    double sqpart = (4. * (1 - e_val * e_val) * sq(ct)) + (5 * e_val * e_val) - (4. * e_val);
    if (sqpart < kDblEps) sqpart = 1.E-5;
    double mm = (2. * (1 - e_val * e_val) * ct) + ((2. * e_val - 1) * sqrt(sqpart));
    if (fabs(mm) < kDblEps) mm = 1.E-5;
    const double l = (4. * (1. - e_val * e_val) * sq(ct)) + sq(2. * e_val - 1.);
    const double r_mw = (l / mm) * r_mc;
    const double xi = 0.1;
    double omega = sq(xi * m.c_peak * A.tpsi) + sq(r_mw * sqrt(1.5) * rho);
    if (omega < kDblEps) omega = 1.E-5;
    const double dl_dtheta = -8. * (1. - e_val * e_val) * ct * st;
    const double dm_dtheta = (-2. * (1. - e_val * e_val) * st) +
                             (0.5 * (2. * e_val - 1.) * dl_dtheta) / sqrt(sqpart);
    const double drmw_dtheta = ((mm * dl_dtheta) - (l * dm_dtheta)) / (mm * mm);
    const double dp_depsilon = A.tpsi / sqrt(3.);
    const double dp_drho = 3. * rho * r_mw * r_mw / (2. * sqrt(omega));
    const double dp_dtheta = (3. * rho * rho * r_mw * r_mc * drmw_dtheta) / (2. * sqrt(omega));
    for (int i = 0; i < 6; ++i)
      dp[i] = ((dp_depsilon * (i < 3 ? depsilon_dsigma : 0.0)) + (dp_drho * drho_dsigma[i])) +
              (dp_dtheta * dtheta_dsigma[i]);
    *dp_dq = dp_drho * sqrt(2. / 3.);
  }
  *softening = 0.;
  if (m.softening && pdstrain > m.pd_peak && pdstrain < m.pd_res) {
    const double dphi_dpstrain = (m.phi_res - m.phi_peak) / (m.pd_res - m.pd_peak);
    const double dc_dpstrain = (m.c_res - m.c_peak) / (m.pd_res - m.pd_peak);
    const double df_dphi = sqrt(1.5) * rho *
                               ((A.sphi * I.s3 / (sqrt(3.) * A.cphi * A.cphi)) +
                                (I.c3 / (3. * A.cphi * A.cphi))) +
                           (mu_p(stress) / (A.cphi * A.cphi));
    const double df_dc = -1.;
    *softening = (-1.) * ((df_dphi * dphi_dpstrain) + (df_dc * dc_dpstrain)) * (*dp_dq);
  }
}

// compute_stress (mohr_coulomb.tcc:348-513) in three phases, so that a
// caller can run them in separate kernels over compacted particle lists:
//   mc_begin   : softening, elastic predictor, separated state, invariants and
//                yield state of the trial stress (complete if elastic);
//   mc_first   : the first plastic correction;
//   mc_iterate : the correction loop, resumable.
// mc_stress composes them. sv holds the 9 state variables; density is the
// particle's mass density.

// What the correction needs from the trial state.
struct McTrial {
  McInv I;             // invariants of the trial stress
  double yield_trial;  // its yield function value (tensile or shear)
  int yt;              // its yield type (FS_TENSILE or FS_SHEAR)
};

// Loop state of the plastic correction.
struct McIter {
  double upd[6];
  double dpdstrain;
  double yield_trial;
  int itr;  // corrections done by the reference's loop (0..100)
};

// Phase 1 (mohr_coulomb.tcc:356-411). Returns true when the update is
// complete (elastic or separated; stress and sv final). Otherwise `trial` and
// T describe the plastic trial state, and sv holds the updated strength
// parameters and yield_state.
// trial_in: the trial stress of an objective rate
// (InfinitesimalElastoPlastic::compute_trial_stress); null: stress + D dstrain.
MPMX_DEV bool mc_begin(const Material& m, double* stress, const double* dstrain, double density,
                       double* sv, double* trial, McTrial& T,
                       const double* trial_in = nullptr) {
  const double pdstrain = sv[SV_PDS];
  if (m.softening && pdstrain > m.pd_peak) {
    if (pdstrain < m.pd_res) {
      sv[SV_PHI] = m.phi_res + ((m.phi_peak - m.phi_res) * (pdstrain - m.pd_res) / (m.pd_peak - m.pd_res));
      sv[SV_PSI] = m.psi_res + ((m.psi_peak - m.psi_res) * (pdstrain - m.pd_res) / (m.pd_peak - m.pd_res));
      sv[SV_COH] = m.c_res + ((m.c_peak - m.c_res) * (pdstrain - m.pd_res) / (m.pd_peak - m.pd_res));
    } else {
      sv[SV_PHI] = m.phi_res;
      sv[SV_PSI] = m.psi_res;
      sv[SV_COH] = m.c_res;
    }
    const double apex = sv[SV_COH] / dmath::tan(sv[SV_PHI]);
    if (sv[SV_TCUT] > apex) sv[SV_TCUT] = (apex > 1.0e-15 ? apex : 1.0e-15);
  }
  sv[SV_YIELD] = 0;
  if (trial_in) {
    for (int i = 0; i < 6; ++i) trial[i] = trial_in[i];
  } else {
    double ds[6];
    de_mul(m, dstrain, ds);
    for (int i = 0; i < 6; ++i) trial[i] = stress[i] + ds[i];
  }
  if (density <= m.critical_density) {
    sv[SV_PDS] += mu_q(trial) / 3.0 / m.shear;
    sv[SV_YIELD] = 3;
    for (int i = 0; i < 6; ++i) stress[i] = 0.;
    return true;
  }
  const McAngles A = mc_angles_of(m, sv[SV_PHI], sv[SV_PSI]);
  mc_invariants(trial, T.I);
  double yf[2];
  const int yt = mc_yield(yf, T.I, A, sv[SV_COH], sv[SV_TCUT]);
  if (yt == FS_ELASTIC) {
    sv[SV_EPS] = T.I.eps;
    sv[SV_THETA] = T.I.theta;
    sv[SV_RHO] = T.I.rho;
    for (int i = 0; i < 6; ++i) stress[i] = trial[i];
    return true;
  }
  sv[SV_YIELD] = (yt == FS_TENSILE) ? 2 : 1;
  T.yt = yt;
  T.yield_trial = (yt == FS_TENSILE) ? yf[0] : yf[1];
  return false;
}

// upd += -(lam * D * dp): each entry sums (lam*D_ik)*dp_k over k in order
MPMX_DEV void mc_correct(const Material& m, double lam, const double* dp, double* upd) {
  const double la1 = lam * m.a1, la2 = lam * m.a2, lg = lam * m.G;
  upd[0] += -(((la1 * dp[0] + la2 * dp[1]) + la2 * dp[2]));
  upd[1] += -(((la2 * dp[0] + la1 * dp[1]) + la2 * dp[2]));
  upd[2] += -(((la2 * dp[0] + la2 * dp[1]) + la1 * dp[2]));
  upd[3] += -(lg * dp[3]);
  upd[4] += -(lg * dp[4]);
  upd[5] += -(lg * dp[5]);
}

// Phase 2 (mohr_coulomb.tcc:412-470). The reference evaluates df/dsigma and
// dp/dsigma at both the trial and the current stress and corrects with one
// pair: the current one if |yield(current)| < 0.1, else the trial one. Only
// the used pair is evaluated (compute_df_dp is pure, so the bits are the
// reference's). `stress` is the start-of-step stress.
MPMX_DEV void mc_first(const Material& m, const double* stress, const double* trial,
                       const double* dstrain, const double* sv, const McTrial& T, McIter& R) {
  const double Tolerance = 1E-1;
  const McAngles A = mc_angles_of(m, sv[SV_PHI], sv[SV_PSI]);
  McInv Ic;
  mc_invariants(stress, Ic);
  double yc[2];
  const int ytc = mc_yield(yc, Ic, A, sv[SV_COH], sv[SV_TCUT]);
  double yield = 1.7976931348623157e308;
  if (ytc == FS_TENSILE) yield = yc[0];
  if (ytc == FS_SHEAR) yield = yc[1];
  const bool use_current = fabs(yield) < Tolerance;
  double x[6];
  for (int i = 0; i < 6; ++i) x[i] = use_current ? stress[i] : trial[i];
  double df[6], dp[6], tmp[6], dp_dq = 0., softening = 0.;
  mc_df_dp(m, use_current ? ytc : T.yt, sv[SV_TCUT], sv[SV_PDS], x, use_current ? Ic : T.I, A,
           df, dp, &dp_dq, &softening);
  de_tmul(m, df, tmp);
  // current: lambda = (df^T D).dstrain / ((df^T D).dp + softening)
  // trial:   lambda_trial = yield_trial / ((df^T D).dp + softening)
  const double lam =
      (use_current ? dot6(tmp, dstrain) : T.yield_trial) / (dot6(tmp, dp) + softening);
  for (int i = 0; i < 6; ++i) R.upd[i] = trial[i];
  mc_correct(m, lam, dp, R.upd);
  R.dpdstrain = lam * dp_dq;
  R.yield_trial = T.yield_trial;
  R.itr = 0;
}

// Phase 3 (mohr_coulomb.tcc:472-512): the correction loop from R.itr, at most
// max_corr corrections in this call. Returns true when the update is complete
// (converged, or 100 corrections; stress and sv final), false with R advanced
// to a yield check that a later call repeats (the check is pure).
MPMX_DEV bool mc_iterate(const Material& m, double* stress, double* sv, McIter& R, int max_corr) {
  const double Tolerance = 1E-1;
  const McAngles A = mc_angles_of(m, sv[SV_PHI], sv[SV_PSI]);
  const double cohesion = sv[SV_COH], tcut = sv[SV_TCUT], pdstrain = sv[SV_PDS];
  McInv I;
  bool converged = false;
  int n = 0;
#pragma nounroll
  while (R.itr < 100) {
    mc_invariants(R.upd, I);
    double yf[2];
    const int yt = mc_yield(yf, I, A, cohesion, tcut);
    if (yf[0] < Tolerance && yf[1] < Tolerance) {
      converged = true;
      break;
    }
    if (n == max_corr) return false;
    if (yt == FS_TENSILE) R.yield_trial = yf[0];
    if (yt == FS_SHEAR) R.yield_trial = yf[1];
    double df[6], dp[6], tmp[6], dp_dq = 0., softening = 0.;
    mc_df_dp(m, yt, tcut, pdstrain, R.upd, I, A, df, dp, &dp_dq, &softening);
    de_tmul(m, df, tmp);
    const double lambda_trial = R.yield_trial / (dot6(tmp, dp) + softening);
    mc_correct(m, lambda_trial, dp, R.upd);
    R.dpdstrain += lambda_trial * dp_dq;
    ++R.itr;
    ++n;
  }
  // final invariants (mohr_coulomb.tcc:507): after a converged loop they are
  // those just computed from the unchanged `upd`
  if (!converged) mc_invariants_only(R.upd, I);
  sv[SV_EPS] = I.eps;
  sv[SV_THETA] = I.theta;
  sv[SV_RHO] = I.rho;
  sv[SV_PDS] += R.dpdstrain;
  for (int i = 0; i < 6; ++i) stress[i] = R.upd[i];
  return true;
}

// The whole update in one call. Updates `stress` in place.
MPMX_DEV void mc_stress(const Material& m, double* stress, const double* dstrain, double density,
                        double* sv, const double* trial_in = nullptr) {
  double trial[6];
  McTrial T;
  if (mc_begin(m, stress, dstrain, density, sv, trial, T, trial_in)) return;
  McIter R;
  mc_first(m, stress, trial, dstrain, sv, T, R);
  mc_iterate(m, stress, sv, R, 100);
}

// Value of a math function at x: LinearFunction::value
// (src/functions/linear_function.cc:19-41), SinFunction::value
// (src/functions/sin_function.cc:17-20).
MPMX_DEV double math_fn_value(const MathFn& f, const double* data, double x) {
  const double* d = data + f.off;
  if (f.type == FN_SIN) {
    if ((x < d[2]) || (x > d[3])) return 0.0;
    return dmath::sin(d[1] * (x - d[0]));
  }
  const int n = f.n;
  const double* fx = d + n;
  if (x < d[0]) return fx[0];
  if (x > d[n - 1]) return fx[n - 1];
  for (int i = 0; i + 1 < n; ++i) {
    const double t = (x - d[i]) / (d[i + 1] - d[i]);
    if (t > -kDblEps && t < (1.0 + kDblEps)) return (fx[i] + t * (fx[i + 1] - fx[i]));
  }
  return __builtin_nan("");
}

// ------------------------------ node update ---------------------------------
// node.tcc:370-404 (Cartesian branch)
// y = A x for the leading DxD block of a 3x3 matrix (sequential over k)
template <int D>
MPMX_DEV void rot_apply(const double (*A)[3], const double* x, double* y) {
#pragma unroll
  for (int i = 0; i < D; ++i) {
    double s = A[i][0] * x[0];
#pragma unroll
    for (int k = 1; k < D; ++k) s = s + A[i][k] * x[k];
    y[i] = s;
  }
}

template <int D>
MPMX_DEV void apply_velocity_constraints(const NodeCons& c, double* v, double* a) {
  if (c.rot) {
    // general boundary: per constraint to the local frame and back
    // (node.tcc:385-401)
#pragma unroll
    for (int k = 0; k < D; ++k)
      if (c.vmask & (1 << k)) {
        double lv[D], la[D];
        rot_apply<D>(c.Ri, v, lv);
        rot_apply<D>(c.Ri, a, la);
        lv[k] = c.vval[k];
        la[k] = 0.;
        rot_apply<D>(c.R, lv, v);
        rot_apply<D>(c.R, la, a);
      }
    return;
  }
#pragma unroll
  for (int k = 0; k < D; ++k)
    if (c.vmask & (1 << k)) {
      v[k] = c.vval[k];
      a[k] = 0.;
    }
}

// node.tcc:510-657 (Cartesian branches)
template <int D>
MPMX_DEV void apply_friction(const NodeCons& c, double* acc, const double* vel, double dt) {
  auto sign = [](double value) { return (value > 0.) ? 1. : -1.; };
  const int dir_n = c.fdir;
  const double sign_n = c.fsign;
  const double mu = c.fmu;
  if constexpr (D == 2) {
    const int dir_t = (D - 1) - dir_n;
    double acc_n, acc_t, vel_t;
    if (!c.rot) {
      acc_n = acc[dir_n];
      acc_t = acc[dir_t];
      vel_t = vel[dir_t];
    } else {  // general boundary: in the local frame (node.tcc:540-556)
      double la[D], lv[D];
      rot_apply<D>(c.Ri, acc, la);
      rot_apply<D>(c.Ri, vel, lv);
      acc_n = la[dir_n];
      acc_t = la[dir_t];
      vel_t = lv[dir_t];
    }
    if ((acc_n * sign_n) > 0.0) {
      if (vel_t != 0.0) {
        const double vel_net = dt * acc_t + vel_t;
        const double vel_frictional = dt * mu * fabs(acc_n);
        if (fabs(vel_net) <= vel_frictional)
          acc_t = -vel_t / dt;
        else
          acc_t -= sign(vel_net) * mu * fabs(acc_n);
      } else {
        if (fabs(acc_t) <= mu * fabs(acc_n))
          acc_t = 0.0;
        else
          acc_t -= sign(acc_t) * mu * fabs(acc_n);
      }
      if (!c.rot) {
        acc[dir_t] = acc_t;
      } else {
        double la[D];
        la[dir_t] = acc_t;
        la[dir_n] = acc_n;
        rot_apply<D>(c.R, la, acc);
      }
    }
  } else {
    const int t0 = dir_n == 0 ? 1 : 0;
    const int t1 = dir_n == 2 ? 1 : 2;
    double gacc[3], gvel[3];  // the global values on a general boundary
    if (c.rot) {
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        gacc[k] = acc[k];
        gvel[k] = vel[k];
      }
      double lacc[3], lvel[3];
      rot_apply<3>(c.Ri, gacc, lacc);
      rot_apply<3>(c.Ri, gvel, lvel);
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        acc[k] = lacc[k];
        gvel[k] = lvel[k];
      }
      vel = gvel;
    }
    const double acc_n = acc[dir_n];
    double acc_t = sqrt(acc[t0] * acc[t0] + acc[t1] * acc[t1]);
    const double vel_t = sqrt(vel[t0] * vel[t0] + vel[t1] * vel[t1]);
    if (acc_n * sign_n > 0.0) {
      if (vel_t != 0.0) {
        const double vn0 = vel[t0] + acc[t0] * dt;
        const double vn1 = vel[t1] + acc[t1] * dt;
        const double vel_net_t = sqrt(vn0 * vn0 + vn1 * vn1);
        const double vel_fricion = mu * fabs(acc_n) * dt;
        if (vel_net_t <= vel_fricion) {
          acc[t0] = -vel[t0] / dt;
          acc[t1] = -vel[t1] / dt;
        } else {
          acc[t0] -= mu * fabs(acc_n) * (vn0 / vel_net_t);
          acc[t1] -= mu * fabs(acc_n) * (vn1 / vel_net_t);
        }
      } else {
        if (acc_t <= mu * fabs(acc_n)) {
          acc[t0] = 0;
          acc[t1] = 0;
        } else {
          acc_t -= mu * fabs(acc_n);
          acc[t0] -= mu * fabs(acc_n) * (acc[t0] / acc_t);
          acc[t1] -= mu * fabs(acc_n) * (acc[t1] / acc_t);
        }
      }
      if (c.rot) {
        double la[3] = {acc[0], acc[1], acc[2]};
        rot_apply<3>(c.R, la, acc);
      }
    } else if (c.rot) {  // no friction force: the global acceleration stands
#pragma unroll
      for (int k = 0; k < 3; ++k) acc[k] = gacc[k];
    }
  }
}

// node.tcc:236-252 compute_velocity (vel in/out, acc in/out for constraints)
template <int D>
MPMX_DEV void node_compute_velocity(double mass, const double* mom, double* v, double* a,
                                    const NodeCons* c) {
  if (mass > 1.E-16) {
#pragma unroll
    for (int k = 0; k < D; ++k) v[k] = mom[k] / mass;
#pragma unroll
    for (int k = 0; k < D; ++k)
      if (fabs(v[k]) < 1.E-15) v[k] = 0.;
  }
  if (c) apply_velocity_constraints<D>(*c, v, a);
}

// Traction vector of particle p at the step's time (Traction::traction,
// traction.h; Particle::assign_traction, particle.tcc:962-979):
// ((value * f(t)) * volume) / size per direction, 0 where none is set.
template <int D>
MPMX_DEV void particle_traction(const Particles& P, const double* fnval, int64_t p, double* t) {
  const int64_t np = P.np;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    const int fn = P.tr_fn[k * np + p];
    t[k] = 0.;
    if (fn >= -1) {
      const double s = fn >= 0 ? fnval[fn] : 1.0;
      t[k] = ((P.tr_val[k * np + p] * s) * P.vol[p]) / P.psize[p];
    }
  }
}

// Node::apply_concentrated_force (node.tcc:103-111) at the step's time:
// external force += f(t) * concentrated force, every direction.
template <int D>
MPMX_DEV void add_concentrated_force(const NodeCons* c, const double* fnval, double* fe) {
  if (!c || !c->has_cf) return;
  const double s = c->cfn >= 0 ? fnval[c->cfn] : 1.0;
#pragma unroll
  for (int k = 0; k < D; ++k) fe[k] = fe[k] + s * c->cf[k];
}

// node.tcc:309-346 compute_acceleration_velocity_cundall
template <int D>
MPMX_DEV bool node_acc_vel(double mass, const double* fext, const double* fint, double* v,
                           double* a, const NodeCons* c, double dt, double damping,
                           const double* fnval) {
  const double tolerance = 1.0E-15;
  if (!(mass > tolerance)) return false;
  double f[D];
#pragma unroll
  for (int k = 0; k < D; ++k) f[k] = fext[k] + fint[k];
  double nrm2 = f[0] * f[0];
#pragma unroll
  for (int k = 1; k < D; ++k) nrm2 = nrm2 + f[k] * f[k];
  const double dn = damping * sqrt(nrm2);
#pragma unroll
  for (int k = 0; k < D; ++k) {
    const double sg = (v[k] > 0.) ? 1. : ((v[k] < 0.) ? -1. : 0.);
    a[k] = (f[k] - dn * sg) / mass;
  }
  if (c) {
    if (c->fdir >= 0) apply_friction<D>(*c, a, v, dt);
#pragma unroll
    for (int k = 0; k < D; ++k)
      if (c->amask & (1 << k)) {  // acceleration_constraint.h:37-43
        const double val = c->afn[k] >= 0 ? c->aval[k] * fnval[c->afn[k]] : c->aval[k];
        if (!c->rot) {
          a[k] = val;
        } else {  // general boundary (node.tcc:469-482)
          double la[D];
          rot_apply<D>(c->Ri, a, la);
          la[k] = val;
          rot_apply<D>(c->R, la, a);
        }
      }
  }
#pragma unroll
  for (int k = 0; k < D; ++k) v[k] += a[k] * dt;
  if (c) apply_velocity_constraints<D>(*c, v, a);
#pragma unroll
  for (int k = 0; k < D; ++k)
    if (fabs(v[k]) < tolerance) v[k] = 0.;
#pragma unroll
  for (int k = 0; k < D; ++k)
    if (fabs(a[k]) < tolerance) a[k] = 0.;
  return true;
}

}  // namespace mpmx
