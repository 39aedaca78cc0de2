// Reverse-mode adjoint of the explicit MPM step.
//
// The forward step (mpm_explicit.tcc:112-190) maps the particle state
// z_t = (x, v, stress, volume) to z_{t+1}; the masses, material parameters
// and, without softening, the Mohr-Coulomb strength state are constant. Its
// other outputs (u, strain, F, density, yield state, invariants) do not feed
// back. The adjoint of one step takes lambda_{t+1} = dL/dz_{t+1} to
// lambda_t and accumulates dL/dE, dL/dphi and dL/dm per particle.
//
// Inputs of a backward step: the recorded particle state of step t (and of
// t+1 for USF and MUSL), and the nodal quantities of step t that the replayed
// forward step recorded (mass, momentum, force, acceleration, updated
// velocity; MUSL: second momentum and its velocity; StepArgs::ad_nodes).
// The kernels then transpose the step phase by phase, per scheme:
//
//   USL : b1 [stress (v'), G2P] -> node [update, velocity] -> b3 [forces(t), P2G]
//   USF : b1 [G2P] -> node [update] -> b3 [forces(t+1), stress (v)]
//         -> node [velocity] -> b5 [P2G]
//   MUSL: b1 [stress (v'')] -> node [velocity of P2G-2] -> b3 [P2G-2, G2P]
//         -> node [update, velocity] -> b5 [forces(t), P2G]
//
// Transposed gathers (G2P, strain rate) scatter node adjoints with FP64
// atomics; transposed scatters (P2G, forces) gather them per particle.
// Shape functions depend on the position through the local coordinate
// xi = 2 (x - centroid) / len (cell.tcc:423-606, Cartesian cells); their
// adjoints are folded into lxq and added to dL/dx at the end of the step.
// Cell membership is piecewise constant and has no derivative.
//
// Non-smooth points follow the forward branch taken (yield type, iteration
// count, constraints: a constrained component passes no derivative), with one
// exception: the reference's snapping of tiny values to zero (|v|, |a|,
// strain rate < 1e-15, node.tcc:245-248, 340-343, particle.tcc:829-857) only
// removes round-off, and its derivative is taken as that of the unsnapped
// value. Otherwise a body at rest in discrete equilibrium (nodal
// acceleration at round-off level) would have no derivative with respect to
// its mass or loads, though any finite perturbation leaves the snapped
// branch.
#include <hip/hip_runtime.h>

#include "mpmx_ad.h"
#include "mpmx_adjoint.h"
#include "mpmx_device.h"

namespace mpmx {

namespace {

constexpr int kAdBlock = 256;
#ifndef MPMX_AD_TANGENTS
#define MPMX_AD_TANGENTS 3
#endif
#ifndef MPMX_AD_ATTR
#define MPMX_AD_ATTR
#endif

MPMX_DEV void ad_atomic(double* p, double v) {
  if (v != 0.) unsafeAtomicAdd(p, v);
}

// local-coordinate signs of node i: N_i = prod_k (1 + s_k xi_k) / 2^D
template <int D>
MPMX_DEV void node_signs(int i, double* s) {
  const int l = i & 3;
  s[0] = (l == 0 || l == 3) ? -1. : 1.;
  s[1] = (l < 2) ? -1. : 1.;
  if constexpr (D == 3) s[2] = (i < 4) ? -1. : 1.;
}

// Shape data of one particle at its recorded position.
template <int D>
struct Shape {
  static constexpr int NN = Elem<D>::NN;
  int c;
  int node[NN];
  double xi[D];
  double Ji[D][D];
  double dxi_dx[D];  // 2 / len, 0 where xi was clamped
};

template <int D>
MPMX_DEV bool load_shape(const Mesh& M, const Slot& s, int64_t np, int64_t p, Shape<D>& S) {
  constexpr int NN = Elem<D>::NN;
  const int c = stream_ld(&s.cell[p]);
  if (c < 0 || (int64_t)c >= M.ncells) return false;
  S.c = c;
#pragma unroll
  for (int i = 0; i < NN; ++i) S.node[i] = M.cell_nodes[(int64_t)c * NN + i];
#pragma unroll
  for (int k = 0; k < D; ++k) {
    S.xi[k] = stream_ld(&s.xi[k * np + p]);
    const bool clamped = S.xi[k] <= -1. + kDblEps || S.xi[k] >= 1. - kDblEps;
    S.dxi_dx[k] = clamped ? 0. : 2. / M.len[k * M.ncells + c];
  }
  double X[NN][D];
  load_cell_X<D>(M, c, X);
  jacobian_inverse<D>(S.xi, X, S.Ji);
  return true;
}

// lxi += lN dN_i/dxi + lG . d(dN_i/dx)/dxi (J constant on Cartesian cells)
template <int D>
MPMX_DEV void fold_xi(const Shape<D>& S, int i, double lN, const double* lG, double* lxi) {
  double s[D], g[D];
  node_signs<D>(i, s);
  grad_i<D>(S.xi, i, g);
#pragma unroll
  for (int a = 0; a < D; ++a) lxi[a] += lN * g[a];
  if (!lG) return;
  // dN_i/dx_a = sum_b g_b Ji[a][b]  ->  lg_b = sum_a lG_a Ji[a][b]
  double lg[D];
#pragma unroll
  for (int b = 0; b < D; ++b) {
    double t = 0.;
#pragma unroll
    for (int a = 0; a < D; ++a) t += lG[a] * S.Ji[a][b];
    lg[b] = t;
  }
  constexpr double inv = D == 2 ? 0.25 : 0.125;
#pragma unroll
  for (int cc = 0; cc < D; ++cc)
#pragma unroll
    for (int a = 0; a < D; ++a) {
      if (a == cc) continue;
      double h = inv * s[cc] * s[a];
      if constexpr (D == 3) {
        const int k = 3 - a - cc;
        h *= 1. + s[k] * S.xi[k];
      }
      lxi[a] += lg[cc] * h;
    }
}

template <int D>
MPMX_DEV void add_lxq(const AdjParticles& L, int64_t np, int64_t p, const Shape<D>& S,
                      const double* lxi) {
#pragma unroll
  for (int k = 0; k < D; ++k) L.lxq[k * np + p] += lxi[k] * S.dxi_dx[k];
}

// ------------------------- constitutive adjoint ------------------------------
// Tangent seeds: stress components, strain-increment components, E, phi.
template <int D> struct Seeds;
template <> struct Seeds<2> {
  static constexpr int NS = 4, NE = 3;
  static MPMX_DEV int s(int q) { return q; }
  static MPMX_DEV int e(int q) { return q == 2 ? 3 : q; }
};
template <> struct Seeds<3> {
  static constexpr int NS = 6, NE = 6;
  static MPMX_DEV int s(int q) { return q; }
  static MPMX_DEV int e(int q) { return q; }
};

// Mohr-Coulomb: ls_in = J_s^T ls_out, lde = J_de^T ls_out, gE += dsigma/dE .
// ls_out, gphi += dsigma/dphi . ls_out, by forward-mode passes of KC tangents.
template <int D, int KC>
MPMX_DEV void mc_vjp(const Material& m, double phi, double psi, double coh, double tcut,
                     const double* s_in, const double* de, const double* ls_out, double* ls_in,
                     double* lde, double& gE, double& gphi, bool frozen_flow) {
  using DT = ad::Dual<KC>;
  using Sd = Seeds<D>;
  constexpr int NQ = Sd::NS + Sd::NE + 2;
#pragma unroll 1
  for (int base = 0; base < NQ; base += KC) {
    DT S[6], E6[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      S[k] = ad::cst<DT>(s_in[k]);
      E6[k] = ad::cst<DT>(de[k]);
    }
    ad::McParams<DT> P{ad::cst<DT>(m.a1), ad::cst<DT>(m.a2), ad::cst<DT>(m.G), ad::cst<DT>(phi),
                       psi, coh, ad::cst<DT>(tcut), m.c_peak};
#pragma unroll
    for (int j = 0; j < KC; ++j) {
      const int q = base + j;
      if (q >= NQ) break;
      if (q < Sd::NS) {
#pragma unroll
        for (int k = 0; k < 6; ++k)
          if (k == Sd::s(q)) S[k].d[j] = 1.;
      } else if (q < Sd::NS + Sd::NE) {
#pragma unroll
        for (int k = 0; k < 6; ++k)
          if (k == Sd::e(q - Sd::NS)) E6[k].d[j] = 1.;
      } else if (q == Sd::NS + Sd::NE) {
        P.a1.d[j] = m.a1 / m.E;
        P.a2.d[j] = m.a2 / m.E;
        P.G.d[j] = m.G / m.E;
      } else {
        P.phi.d[j] = 1.;
        P.tcut.d[j] = m.dtcut_dphi;
      }
    }
    DT O[6];
    ad::mc_update(P, S, E6, O, frozen_flow);
#pragma unroll
    for (int j = 0; j < KC; ++j) {
      const int q = base + j;
      if (q >= NQ) break;
      double acc = 0.;
#pragma unroll
      for (int i = 0; i < 6; ++i) acc += ls_out[i] * O[i].d[j];
      if (q < Sd::NS) {
#pragma unroll
        for (int k = 0; k < 6; ++k)
          if (k == Sd::s(q)) ls_in[k] = acc;
      } else if (q < Sd::NS + Sd::NE) {
#pragma unroll
        for (int k = 0; k < 6; ++k)
          if (k == Sd::e(q - Sd::NS)) lde[k] = acc;
      } else if (q == Sd::NS + Sd::NE) {
        gE += acc;
      } else {
        gphi += acc;
      }
    }
  }
}

// Near the apex of the Mohr-Coulomb surface (cohesionless material at low
// stress, e.g. at a free surface) the directions of the yield and plastic
// potential gradients are set by the Lode angle of a stress whose deviatoric
// part is small against the step's plastic correction; their derivatives
// scale with 1 / rho and the step's Jacobian gains up to 1e4, which the
// backward pass multiplies over hundreds of steps. For returns whose plastic
// correction exceeds apex_ratio times the returned deviatoric norm
// sqrt(2 J2), the adjoint takes the gradient directions as constants (the
// derivatives through the yield-function values, the multipliers and the
// elastic predictor are kept); apex_ratio <= 0 differentiates every return
// exactly.

// One constitutive update's adjoint (particle.tcc:880-890 -> material).
// yield_new and s_out are the update's result as the recorded run computed
// it (Mohr-Coulomb yield state and stress after the update). A plastic
// Mohr-Coulomb update is not differentiated here: it is appended to Q
// (strain increment, apex flag) for k_ad_mc, and the function returns true
// with ls_in and lde untouched.
template <int D>
MPMX_DEV bool material_vjp(const Material& m, const double* s_in, const double* de,
                           double yield_new, const double* s_out, const double* ls_out,
                           double* ls_in, double* lde, double& gE, double apex_ratio,
                           int64_t p, const AdjList& Q) {
#pragma unroll
  for (int k = 0; k < 6; ++k) ls_in[k] = lde[k] = 0.;
  double dde[6];
  de_mul(m, de, dde);
  if (m.type == MC) {
    if (yield_new == 3.) return false;  // separated: stress = 0
    if (yield_new != 0.) {
      double corr = 0.;
#pragma unroll
      for (int k = 0; k < 6; ++k) corr = fmax(corr, fabs((s_in[k] + dde[k]) - s_out[k]));
      const bool apex = apex_ratio > 0. && corr > apex_ratio * sqrt(2. * ad::mu_j2(s_out));
      // tensile returns from the back of the list, shear ones from the
      // front, so that a wave of k_ad_mc mostly runs one branch; one atomic
      // per wave and kind on the counts
      const bool tensile = yield_new == FS_TENSILE;
      const unsigned long long tmask = __ballot(tensile);
      const unsigned long long mine = tensile ? tmask : (__ballot(1) & ~tmask);
      const int lane = threadIdx.x & 63;
      const int leader = __ffsll((long long)mine) - 1;
      int base = 0;
      if (lane == leader) base = atomicAdd(Q.count + (tensile ? 1 : 0), __popcll(mine));
      base = __shfl(base, leader, 64);
      const int k = base + __popcll(mine & ((1ull << lane) - 1ull));
      const int64_t slot = tensile ? Q.cap - 1 - k : k;
      Q.ids[slot] = (int)p;
      Q.apex[slot] = apex ? 1 : 0;
#pragma unroll
      for (int k = 0; k < 6; ++k) Q.buf[k * Q.cap + slot] = de[k];
      return true;
    }
  }
  // elastic: s' = s + D de: ls_in = ls_out, lde = D ls_out, dE: ls_out . (D de) / E
  de_mul(m, ls_out, lde);
  double t = 0.;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    ls_in[k] = ls_out[k];
    t += ls_out[k] * dde[k];
  }
  gE += t / m.E;
  return false;
}

// Velocity-gradient adjoint of one particle's strain increment: adds
// dL/d(nodal velocity) = sum_b lL[a][b] dN_i/dx_b (+ the centroid term of the
// volume change, lsc) to acc and folds dL/d(dN_i/dx) into lxi.
template <int D>
MPMX_DEV void strain_rate_vjp(const Mesh& M, const Shape<D>& S, const double (*v)[D],
                              const double (*G)[D], const double* lde, const double* lsc,
                              double dt, double (*acc)[D], double* lxi) {
  constexpr int NN = Elem<D>::NN;
  const double* dc = M.dndx_c + (int64_t)S.c * NN * D;
  double lsr[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) lsr[k] = lde[k] * dt;
  double lL[D][D];
  if constexpr (D == 2) {
    lL[0][0] = lsr[0];
    lL[1][1] = lsr[1];
    lL[0][1] = lL[1][0] = lsr[3];
  } else {
    lL[0][0] = lsr[0];
    lL[1][1] = lsr[1];
    lL[2][2] = lsr[2];
    lL[0][1] = lL[1][0] = lsr[3];
    lL[1][2] = lL[2][1] = lsr[4];
    lL[0][2] = lL[2][0] = lsr[5];
  }
#pragma unroll
  for (int i = 0; i < NN; ++i) {
    const double* gc = dc + i * D;
    double lG[D];
#pragma unroll
    for (int a = 0; a < D; ++a) {
      double t = lsc ? lsc[a] * gc[a] : 0.;
#pragma unroll
      for (int b = 0; b < D; ++b) t += lL[a][b] * G[i][b];
      acc[i][a] += t;
    }
#pragma unroll
    for (int b = 0; b < D; ++b) {
      double t = 0.;
#pragma unroll
      for (int a = 0; a < D; ++a) t += lL[a][b] * v[i][a];
      lG[b] = t;
    }
    fold_xi<D>(S, i, 0., lG, lxi);
  }
}

template <int D>
MPMX_DEV void load_kinematics(const Mesh& M, const Shape<D>& S, const double* nv,
                              double (*v)[D], double (*G)[D]) {
  constexpr int NN = Elem<D>::NN;
  const int64_t nn = M.nnodes;
#pragma unroll
  for (int i = 0; i < NN; ++i) {
#pragma unroll
    for (int k = 0; k < D; ++k) v[i][k] = nv[k * nn + S.node[i]];
    dndx_i<D>(S.xi, i, S.Ji, G[i]);
  }
}

// Adjoint of compute_stress_strain for one particle (mpm_scheme.tcc:75-103:
// strain rate from the nodal velocity nv, strain increment, volume and
// density update, constitutive update). Returns the adjoints of the input
// stress and volume, adds dL/dE, adds dL/d(nodal velocity) to acc and
// folds the shape-gradient adjoint into lxi. Returns true when the
// constitutive adjoint was deferred to k_ad_mc (ls_in not set; the volume
// part is done).
template <int D>
MPMX_DEV bool stress_vjp(const Mesh& M, const Material& mat, int64_t p, const Shape<D>& S,
                         const double* nv, double dt, double apex_ratio, const double* s_in,
                         double V_in, double yield_new, const double* s_out,
                         const double* ls_out, double lV_out, double* ls_in, double& lV_in,
                         double& gE, double (*acc)[D], double* lxi, const AdjList& Q) {
  constexpr int NN = Elem<D>::NN;
  double v[NN][D], G[NN][D];
  load_kinematics<D>(M, S, nv, v, G);
  const double* dc = M.dndx_c + (int64_t)S.c * NN * D;
  // strain rate at the particle and at the centroid (particle.tcc:819-858)
  double sr[6], sc[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) sr[k] = sc[k] = 0.;
#pragma unroll
  for (int i = 0; i < NN; ++i) {
    const double* gc = dc + i * D;
    if constexpr (D == 2) {
      sr[0] += G[i][0] * v[i][0];
      sr[1] += G[i][1] * v[i][1];
      sr[3] += G[i][1] * v[i][0] + G[i][0] * v[i][1];
      sc[0] += gc[0] * v[i][0];
      sc[1] += gc[1] * v[i][1];
    } else {
      sr[0] += G[i][0] * v[i][0];
      sr[1] += G[i][1] * v[i][1];
      sr[2] += G[i][2] * v[i][2];
      sr[3] += G[i][1] * v[i][0] + G[i][0] * v[i][1];
      sr[4] += G[i][2] * v[i][1] + G[i][1] * v[i][2];
      sr[5] += G[i][2] * v[i][0] + G[i][0] * v[i][2];
      sc[0] += gc[0] * v[i][0];
      sc[1] += gc[1] * v[i][1];
      sc[2] += gc[2] * v[i][2];
    }
  }
#pragma unroll
  for (int k = 0; k < 6; ++k)
    if (fabs(sr[k]) < 1.E-15) sr[k] = 0.;
#pragma unroll
  for (int k = 0; k < D; ++k)
    if (fabs(sc[k]) < 1.E-15) sc[k] = 0.;
  double de[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) de[k] = sr[k] * dt;
  double ssum = sc[0];
#pragma unroll
  for (int k = 1; k < D; ++k) ssum = ssum + sc[k];
  const double dvol = dt * ssum;
  // constitutive update
  double lde[6];
  const bool deferred = material_vjp<D>(mat, s_in, de, yield_new, s_out, ls_out, ls_in, lde, gE,
                                        apex_ratio, p, Q);
  // volume: V' = V (1 + dvol)
  lV_in = lV_out * (1. + dvol);
  double lsc[3];
#pragma unroll
  for (int k = 0; k < D; ++k) lsc[k] = lV_out * V_in * dt;
  strain_rate_vjp<D>(M, S, v, G, lde, lsc, dt, acc, lxi);
  return deferred;
}

// Adjoint of compute_updated_position (particle.tcc:994-1070) and of the
// particle velocity constraints (pvc: constrained components, mesh.tcc:
// 1454-1468): adds dL/dv' and dL/da of the cell's nodes to accvp and acca,
// returns dL/dv_p (in lv), folds lxi.
template <int D>
MPMX_DEV void g2p_vjp(const Mesh& M, const Shape<D>& S, const AdjNodes& A, const AdjArgs& a,
                      int pvc, const double* lx, double* lv, double (*accvp)[D],
                      double (*acca)[D], double* lxi) {
  constexpr int NN = Elem<D>::NN;
  const int64_t nn = M.nnodes;
  const bool flip = a.vu == FLIP;
  const double beta = a.blending;
  double lnv[D], lna[D];
#pragma unroll
  for (int k = 0; k < D; ++k) {
    if (pvc & (1 << k)) lv[k] = 0.;  // velocity constraint: v'_k is fixed
    lnv[k] = a.dt * lx[k] + (flip ? (1.0 - beta) * lv[k] : lv[k]);
    lna[k] = flip ? beta * a.dt * lv[k] : 0.;
    lv[k] = flip ? beta * lv[k] : 0.;
  }
#pragma unroll
  for (int i = 0; i < NN; ++i) {
    const int n = S.node[i];
    const double N = shapefn_i<D>(S.xi, i);
    double lN = 0.;
#pragma unroll
    for (int k = 0; k < D; ++k) {
      accvp[i][k] += N * lnv[k];
      acca[i][k] += N * lna[k];
      lN += A.vp[k * nn + n] * lnv[k] + (flip ? A.a[k * nn + n] * lna[k] : 0.);
    }
    fold_xi<D>(S, i, lN, nullptr, lxi);
  }
}

// Sum per-node contributions over the lanes of a wave that hold particles of
// the same cell (consecutive lanes with equal key: the cells' particles are
// mostly consecutive in particle order) and add each sum to target with one
// FP64 atomic per node and component. Every lane of the wave must call it;
// lanes without a particle pass key < 0 and zero contributions.
template <int D>
MPMX_DEV void wave_flush(int key, const int* node, double (*acc)[D], double* target,
                         int64_t nn) {
  constexpr int NN = Elem<D>::NN;
  const int lane = __lane_id();
  const int prev = __shfl_up(key, 1);
  const bool head = lane == 0 || prev != key;
  const unsigned long long heads = __ballot(head);
  const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
  const int start = 63 - __clzll(heads & upto);
  const bool tail = lane == 63 || ((heads >> (lane + 1)) & 1ull);
  for (int off = 1; off < 64; off <<= 1) {
    const bool take = lane - start >= off;
    if (!__ballot(take)) break;
#pragma unroll
    for (int i = 0; i < NN; ++i)
#pragma unroll
      for (int k = 0; k < D; ++k) {
        const double y = __shfl_up(acc[i][k], off);
        if (take) acc[i][k] += y;
      }
  }
  if (tail && key >= 0) {
#pragma unroll
    for (int i = 0; i < NN; ++i)
#pragma unroll
      for (int k = 0; k < D; ++k) ad_atomic(&target[k * nn + node[i]], acc[i][k]);
  }
}

template <int D>
MPMX_DEV void zero_acc(double (*acc)[D]) {
#pragma unroll
  for (int i = 0; i < Elem<D>::NN; ++i)
#pragma unroll
    for (int k = 0; k < D; ++k) acc[i][k] = 0.;
}

// Adjoint of the nodal sums a particle contributes to (particle.tcc:630-658,
// 893-950): mass m N_i, momentum m N_i v, external force g m N_i, internal
// force -V sigma.dN_i/dx. lmass/lmom/lf are the node adjoints of those sums
// (null: not part of this pass). Adds to lm_p, lv_p, ls, lV, lxi.
template <int D>
MPMX_DEV void gather_vjp(const Mesh& M, const Shape<D>& S, const double* lmass,
                         const double* lmom, const double* lf, double mp, const double* vp,
                         const double* g, const double* s, double V, double& lm_p, double* lv_p,
                         double* ls, double& lV, double* lxi) {
  constexpr int NN = Elem<D>::NN;
  const int64_t nn = M.nnodes;
#pragma unroll
  for (int i = 0; i < NN; ++i) {
    const int n = S.node[i];
    const double N = shapefn_i<D>(S.xi, i);
    double lN = 0.;
    if (lmass || lmom) {
      double t = lmass ? lmass[n] : 0.;
      if (lmom) {
#pragma unroll
        for (int k = 0; k < D; ++k) {
          const double l = lmom[k * nn + n];
          t += vp[k] * l;
          lv_p[k] += mp * N * l;
        }
      }
      lm_p += N * t;
      lN += mp * t;
    }
    if (lf) {
      double l[D], gl = 0.;
#pragma unroll
      for (int k = 0; k < D; ++k) {
        l[k] = lf[k * nn + n];
        gl += g[k] * l[k];
      }
      lm_p += N * gl;
      lN += mp * gl;
      double Gi[D];
      dndx_i<D>(S.xi, i, S.Ji, Gi);
      double f[D], lG[D];
      if constexpr (D == 2) {
        f[0] = Gi[0] * s[0] + Gi[1] * s[3];
        f[1] = Gi[1] * s[1] + Gi[0] * s[3];
        ls[0] += -V * Gi[0] * l[0];
        ls[1] += -V * Gi[1] * l[1];
        ls[3] += -V * (Gi[1] * l[0] + Gi[0] * l[1]);
        lG[0] = -V * (s[0] * l[0] + s[3] * l[1]);
        lG[1] = -V * (s[3] * l[0] + s[1] * l[1]);
      } else {
        f[0] = Gi[0] * s[0] + Gi[1] * s[3] + Gi[2] * s[5];
        f[1] = Gi[1] * s[1] + Gi[0] * s[3] + Gi[2] * s[4];
        f[2] = Gi[2] * s[2] + Gi[1] * s[4] + Gi[0] * s[5];
        ls[0] += -V * Gi[0] * l[0];
        ls[1] += -V * Gi[1] * l[1];
        ls[2] += -V * Gi[2] * l[2];
        ls[3] += -V * (Gi[1] * l[0] + Gi[0] * l[1]);
        ls[4] += -V * (Gi[2] * l[1] + Gi[1] * l[2]);
        ls[5] += -V * (Gi[2] * l[0] + Gi[0] * l[2]);
        lG[0] = -V * (s[0] * l[0] + s[3] * l[1] + s[5] * l[2]);
        lG[1] = -V * (s[3] * l[0] + s[1] * l[1] + s[4] * l[2]);
        lG[2] = -V * (s[5] * l[0] + s[4] * l[1] + s[2] * l[2]);
      }
      double fl = 0.;
#pragma unroll
      for (int k = 0; k < D; ++k) fl += f[k] * l[k];
      lV += -fl;
      fold_xi<D>(S, i, lN, lG, lxi);
    } else {
      fold_xi<D>(S, i, lN, nullptr, lxi);
    }
  }
}

template <int D>
MPMX_DEV void load_vec(const double* a, int64_t np, int64_t p, int n, double* out) {
  for (int k = 0; k < n; ++k) out[k] = stream_ld(&a[k * np + p]);
}
template <int D>
MPMX_DEV void store_vec(double* a, int64_t np, int64_t p, int n, const double* in) {
  for (int k = 0; k < n; ++k) stream_st(&a[k * np + p], in[k]);
}

// ------------------------------- kernels -------------------------------------
// Record the state a backward step needs.
template <int D>
__global__ __launch_bounds__(kAdBlock) void k_ad_save(Particles P, SlotW s) {
  const int64_t p = blockIdx.x * (int64_t)kAdBlock + threadIdx.x;
  const int64_t np = P.np;
  if (p >= np) return;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    stream_st(&s.v[k * np + p], stream_ld(&P.v[k * np + p]));
    stream_st(&s.xi[k * np + p], P.xi[k * np + p]);
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) stream_st(&s.stress[k * np + p], stream_ld(&P.stress[k * np + p]));
  stream_st(&s.vol[p], stream_ld(&P.vol[p]));
  stream_st(&s.rho[p], stream_ld(&P.rho[p]));
  if (s.yield) stream_st(&s.yield[p], stream_ld(&P.sv[SV_YIELD * np + p]));
  stream_st(&s.cell[p], P.cell[p]);
}

__global__ void k_ad_fill(double* a, int64_t n, double v) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// Pre-update nodal velocity of step t from its recorded mass and momentum
// (node.tcc:236-252).
template <int D>
// Also clears this step's nodal adjoints and the plastic list length.
__global__ __launch_bounds__(kAdBlock) void k_ad_node_fwd(Mesh M, Nodes Nd, AdjNodes A,
                                                          int* qcount) {
  const int64_t n = blockIdx.x * (int64_t)kAdBlock + threadIdx.x;
  const int64_t nn = M.nnodes;
  if (n == 0) qcount[0] = qcount[1] = 0;
  if (n >= nn) return;
  A.lm[n] = 0.;
  A.lm2[n] = 0.;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    const int64_t i = k * nn + n;
    A.v[i] = A.lvp[i] = A.la[i] = A.lv[i] = A.lvpp[i] = A.lf[i] = A.lP[i] = A.lP2[i] = 0.;
  }
  const double mass = A.m[n];
  if (!(mass > 0.)) return;
  const int ci = Nd.cons[n];
  const NodeCons* c = ci >= 0 ? &Nd.table[ci] : nullptr;
  double v[D], acc[D], mom[D];
#pragma unroll
  for (int k = 0; k < D; ++k) {
    v[k] = acc[k] = 0.;
    mom[k] = A.P[k * nn + n];
  }
  node_compute_velocity<D>(mass, mom, v, acc, c);
#pragma unroll
  for (int k = 0; k < D; ++k) A.v[k * nn + n] = v[k];
}

// Node-side adjoints.
//   ND_AD_UPD: of compute_acceleration_velocity_cundall (node.tcc:309-346)
//              from lvp (updated velocity) and la (acceleration): lf, lm,
//              and lv (pre-update velocity) accumulated;
//   ND_AD_VEL: of compute_velocity (node.tcc:236-252) from lv: lP, lm;
//   ND_AD_VEL2: the same for MUSL's second momentum: lvpp -> lP2, lm2.
enum { ND_AD_UPD = 1, ND_AD_VEL = 2, ND_AD_VEL2 = 4 };

template <int D>
MPMX_DEV void velocity_vjp(double mass, const double* mom, const NodeCons* c, const double* lv,
                           double* lP, double& lm) {
#pragma unroll
  for (int k = 0; k < D; ++k) lP[k] = 0.;
  if (!(mass > 1.E-16)) return;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    if (c && (c->vmask & (1 << k))) continue;
    const double vk = mom[k] / mass;
    lP[k] = lv[k] / mass;
    lm += -vk * lv[k] / mass;
  }
}

template <int D, int MODE>
__global__ __launch_bounds__(kAdBlock) void k_ad_node(Mesh M, Nodes Nd, AdjNodes A, AdjArgs a) {
  const int64_t n = blockIdx.x * (int64_t)kAdBlock + threadIdx.x;
  const int64_t nn = M.nnodes;
  if (n >= nn) return;
  const double mass = A.m[n];
  if (!(mass > 0.)) return;
  const int ci = Nd.cons[n];
  const NodeCons* c = ci >= 0 ? &Nd.table[ci] : nullptr;
  if constexpr ((MODE & ND_AD_VEL2) != 0) {
    double mom[D], lv[D], lP[D];
    double lm = 0.;
#pragma unroll
    for (int k = 0; k < D; ++k) {
      mom[k] = A.P2[k * nn + n];
      lv[k] = A.lvpp[k * nn + n];
    }
    velocity_vjp<D>(mass, mom, c, lv, lP, lm);
#pragma unroll
    for (int k = 0; k < D; ++k) A.lP2[k * nn + n] = lP[k];
    A.lm2[n] = lm;
  }
  if constexpr ((MODE & ND_AD_UPD) != 0) {
    const double tol = 1.0E-15;
    double lv[D];
#pragma unroll
    for (int k = 0; k < D; ++k) lv[k] = A.lv[k * nn + n];
    double lm = 0.;
    if (!(mass > tol)) {
      // no update: v' = v, a = 0
#pragma unroll
      for (int k = 0; k < D; ++k) lv[k] += A.lvp[k * nn + n];
    } else {
      double v[D], f[D], ap[D];
#pragma unroll
      for (int k = 0; k < D; ++k) {
        v[k] = A.v[k * nn + n];
        f[k] = A.f[k * nn + n];
      }
      double nrm2 = f[0] * f[0];
#pragma unroll
      for (int k = 1; k < D; ++k) nrm2 = nrm2 + f[k] * f[k];
      const double nrm = sqrt(nrm2);
      const double dn = a.damping * nrm;
      double sg[D];
#pragma unroll
      for (int k = 0; k < D; ++k) {
        sg[k] = (v[k] > 0.) ? 1. : ((v[k] < 0.) ? -1. : 0.);
        ap[k] = (f[k] - dn * sg[k]) / mass;
      }
      double ldn = 0., lf[D];
#pragma unroll
      for (int k = 0; k < D; ++k) lf[k] = 0.;
#pragma unroll
      for (int k = 0; k < D; ++k) {
        if (c && (c->vmask & (1 << k))) continue;  // v'_k and a_k fixed
        const bool acon = c && (c->amask & (1 << k));
        const double lvp = A.lvp[k * nn + n];
        const double la = A.la[k * nn + n];
        lv[k] += lvp;
        if (acon) continue;
        const double lap = la + a.dt * lvp;
        lf[k] += lap / mass;
        ldn += -sg[k] * lap / mass;
        lm += -ap[k] * lap / mass;
      }
      if (a.damping != 0. && nrm > 0.) {
#pragma unroll
        for (int k = 0; k < D; ++k) lf[k] += a.damping * ldn * f[k] / nrm;
      }
#pragma unroll
      for (int k = 0; k < D; ++k) A.lf[k * nn + n] = lf[k];
    }
#pragma unroll
    for (int k = 0; k < D; ++k) A.lv[k * nn + n] = lv[k];
    A.lm[n] += lm;
  }
  if constexpr ((MODE & ND_AD_VEL) != 0) {
    double mom[D], lv[D], lP[D];
    double lm = 0.;
#pragma unroll
    for (int k = 0; k < D; ++k) {
      mom[k] = A.P[k * nn + n];
      lv[k] = A.lv[k * nn + n];
    }
    velocity_vjp<D>(mass, mom, c, lv, lP, lm);
#pragma unroll
    for (int k = 0; k < D; ++k) A.lP[k * nn + n] = lP[k];
    A.lm[n] += lm;
  }
}

// First particle pass of a backward step.
//   USL : stress update with v' (-> lvp), then G2P (-> lvp, la)
//   USF : G2P (-> lvp, la)
//   MUSL: stress update with v'' (-> lvpp)
template <int D, int SCH>
__global__ __launch_bounds__(kAdBlock) MPMX_AD_ATTR void k_ad_b1(Mesh M, Particles P,
                                                                 const Material* mats, Slot s0,
                                                                 Slot s1, AdjNodes A,
                                                                 AdjParticles L, AdjList Q,
                                                                 AdjArgs a) {
  constexpr int NN = Elem<D>::NN;
  const int64_t p = blockIdx.x * (int64_t)kAdBlock + threadIdx.x;
  const int64_t np = P.np;
  Shape<D> S;
  const bool valid = p < np && load_shape<D>(M, s0, np, p, S);
  double acc0[NN][D], acc1[NN][D];
  zero_acc<D>(acc0);
  zero_acc<D>(acc1);
  if (valid) {
    double lxi[D];
#pragma unroll
    for (int k = 0; k < D; ++k) lxi[k] = 0.;
    if (SCH == USL || SCH == MUSL) {
      const Material mat = mats[stream_ld(&P.mat[p])];
      double s_in[6], s_out[6], ls_out[6], ls_in[6];
      load_vec<D>(s0.stress, np, p, 6, s_in);
      load_vec<D>(s1.stress, np, p, 6, s_out);
      load_vec<D>(L.ls, np, p, 6, ls_out);
      double lV_in = 0., gE = stream_ld(&L.gE[p]);
      const double ynew = s1.yield ? stream_ld(&s1.yield[p]) : 0.;
      const bool deferred = stress_vjp<D>(M, mat, p, S, SCH == USL ? A.vp : A.vpp, a.dt,
                                          a.apex_ratio, s_in, s0.vol[p], ynew, s_out, ls_out,
                                          stream_ld(&L.lvol[p]), ls_in, lV_in, gE, acc0, lxi, Q);
      if (!deferred) store_vec<D>(L.ls, np, p, 6, ls_in);
      stream_st(&L.lvol[p], lV_in);
      stream_st(&L.gE[p], gE);
    }
    if (SCH == USL || SCH == USF) {
      double lx[D], lv[D];
      load_vec<D>(L.lx, np, p, D, lx);
      load_vec<D>(L.lv, np, p, D, lv);
      g2p_vjp<D>(M, S, A, a, P.pvc_mask ? P.pvc_mask[p] : 0, lx, lv, acc0, acc1, lxi);
      store_vec<D>(L.lv, np, p, D, lv);
    }
    add_lxq<D>(L, np, p, S, lxi);
  }
  const int key = valid ? S.c : -1;
  wave_flush<D>(key, S.node, acc0, SCH == MUSL ? A.lvpp : A.lvp, M.nnodes);
  if (SCH != MUSL && a.vu == FLIP) wave_flush<D>(key, S.node, acc1, A.la, M.nnodes);
}

// Second particle pass.
//   USL : forces of step t and P2G (gather lf, lm, lP); end of step
//   USF : forces of step t+1 (gather lf), then the stress update with v
//         (-> lv)
//   MUSL: second P2G (gather lm2, lP2), then G2P (-> lvp, la)
template <int D, int SCH>
__global__ __launch_bounds__(kAdBlock) MPMX_AD_ATTR void k_ad_b3(Mesh M, Particles P,
                                                                 const Material* mats, Slot s0,
                                                                 Slot s1, AdjNodes A,
                                                                 AdjParticles L, AdjList Q,
                                                                 AdjArgs a) {
  constexpr int NN = Elem<D>::NN;
  const int64_t p = blockIdx.x * (int64_t)kAdBlock + threadIdx.x;
  const int64_t np = P.np;
  Shape<D> S;
  const bool valid = p < np && load_shape<D>(M, s0, np, p, S);
  double acc0[NN][D], acc1[NN][D];
  zero_acc<D>(acc0);
  zero_acc<D>(acc1);
  if (valid) {
    double lxi[D];
#pragma unroll
    for (int k = 0; k < D; ++k) lxi[k] = 0.;
    const double mp = P.mass[p];
    double lm_p = 0.;
    double lv[D];
    load_vec<D>(L.lv, np, p, D, lv);
    if (SCH == USL) {
      double v0[D], st[6], ls[6];
      load_vec<D>(s0.v, np, p, D, v0);
      load_vec<D>(s0.stress, np, p, 6, st);
      load_vec<D>(L.ls, np, p, 6, ls);
      double lV = stream_ld(&L.lvol[p]);
      gather_vjp<D>(M, S, A.lm, A.lP, A.lf, mp, v0, a.g, st, s0.vol[p], lm_p, lv, ls, lV, lxi);
      store_vec<D>(L.ls, np, p, 6, ls);
      stream_st(&L.lvol[p], lV);
    } else if (SCH == USF) {
      // forces of the updated stress and volume
      double st1[6], ls1[6];
      load_vec<D>(s1.stress, np, p, 6, st1);
      load_vec<D>(L.ls, np, p, 6, ls1);
      double lV1 = stream_ld(&L.lvol[p]);
      gather_vjp<D>(M, S, nullptr, nullptr, A.lf, mp, nullptr, a.g, st1, s1.vol[p], lm_p, lv,
                    ls1, lV1, lxi);
      // stress update from the pre-update nodal velocity
      const Material mat = mats[stream_ld(&P.mat[p])];
      double s_in[6], ls_in[6];
      load_vec<D>(s0.stress, np, p, 6, s_in);
      double lV_in = 0., gE = stream_ld(&L.gE[p]);
      const double ynew = s1.yield ? stream_ld(&s1.yield[p]) : 0.;
      const bool deferred = stress_vjp<D>(M, mat, p, S, A.v, a.dt, a.apex_ratio, s_in,
                                          s0.vol[p], ynew, st1, ls1, lV1, ls_in, lV_in, gE, acc0,
                                          lxi, Q);
      store_vec<D>(L.ls, np, p, 6, deferred ? ls1 : ls_in);
      stream_st(&L.lvol[p], lV_in);
      stream_st(&L.gE[p], gE);
    } else {
      // second P2G with the updated particle velocity, then G2P
      double v1[D], dummy_s[6] = {0, 0, 0, 0, 0, 0}, dummy_ls[6] = {0, 0, 0, 0, 0, 0};
      double dummy_lV = 0.;
      load_vec<D>(s1.v, np, p, D, v1);
      gather_vjp<D>(M, S, A.lm2, A.lP2, nullptr, mp, v1, a.g, dummy_s, 0., lm_p, lv, dummy_ls,
                    dummy_lV, lxi);
      double lx[D];
      load_vec<D>(L.lx, np, p, D, lx);
      g2p_vjp<D>(M, S, A, a, P.pvc_mask ? P.pvc_mask[p] : 0, lx, lv, acc0, acc1, lxi);
    }
    store_vec<D>(L.lv, np, p, D, lv);
    L.gm[p] += lm_p;
    add_lxq<D>(L, np, p, S, lxi);
    if (SCH == USL) {
#pragma unroll
      for (int k = 0; k < D; ++k) {
        L.lx[k * np + p] += stream_ld(&L.lxq[k * np + p]);
        stream_st(&L.lxq[k * np + p], 0.);
      }
    }
  }
  if (SCH == USF || SCH == MUSL) {
    const int key = valid ? S.c : -1;
    wave_flush<D>(key, S.node, acc0, SCH == USF ? A.lv : A.lvp, M.nnodes);
    if (SCH == MUSL && a.vu == FLIP) wave_flush<D>(key, S.node, acc1, A.la, M.nnodes);
  }
}

// Third particle pass (USF, MUSL): P2G of step t (and MUSL: forces of step
// t); end of step.
template <int D, int SCH>
__global__ __launch_bounds__(kAdBlock) void k_ad_b5(Mesh M, Particles P, Slot s0, AdjNodes A,
                                                    AdjParticles L, AdjArgs a) {
  const int64_t p = blockIdx.x * (int64_t)kAdBlock + threadIdx.x;
  const int64_t np = P.np;
  if (p >= np) return;
  Shape<D> S;
  if (!load_shape<D>(M, s0, np, p, S)) return;
  double lxi[D];
#pragma unroll
  for (int k = 0; k < D; ++k) lxi[k] = 0.;
  const double mp = P.mass[p];
  double lm_p = 0.;
  double lv[D], v0[D], st[6], ls[6];
  load_vec<D>(L.lv, np, p, D, lv);
  load_vec<D>(s0.v, np, p, D, v0);
  load_vec<D>(s0.stress, np, p, 6, st);
  load_vec<D>(L.ls, np, p, 6, ls);
  double lV = stream_ld(&L.lvol[p]);
  gather_vjp<D>(M, S, A.lm, A.lP, SCH == MUSL ? A.lf : nullptr, mp, v0, a.g, st, s0.vol[p],
                lm_p, lv, ls, lV, lxi);
  store_vec<D>(L.ls, np, p, 6, ls);
  stream_st(&L.lvol[p], lV);
  store_vec<D>(L.lv, np, p, D, lv);
  L.gm[p] += lm_p;
  add_lxq<D>(L, np, p, S, lxi);
#pragma unroll
  for (int k = 0; k < D; ++k) {
    L.lx[k * np + p] += stream_ld(&L.lxq[k * np + p]);
    stream_st(&L.lxq[k * np + p], 0.);
  }
}

// The deferred plastic Mohr-Coulomb updates: constitutive adjoint by
// forward-mode passes (stress adjoint and parameter gradients written; the
// strain-increment adjoint left in Q.buf).
// One wave per workgroup: the waves' run times vary with their particles'
// iteration counts, and a finished wave frees its slot at once.
constexpr int kAdMcBlock = 64;
template <int D>
__global__ __launch_bounds__(kAdMcBlock) MPMX_AD_ATTR void k_ad_mc(Particles P,
                                                                   const Material* mats, Slot s0,
                                                                   AdjParticles L, AdjList Q) {
  const int64_t e = blockIdx.x * (int64_t)kAdMcBlock + threadIdx.x;
  const int ns = Q.count[0];
  if (e >= ns + Q.count[1]) return;
  const int64_t slot = e < ns ? e : Q.cap - 1 - (e - ns);
  const int64_t np = P.np;
  const int p = Q.ids[slot];
  const Material m = mats[stream_ld(&P.mat[p])];
  double s_in[6], de[6], ls_out[6], ls_in[6], lde[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    s_in[k] = stream_ld(&s0.stress[k * np + p]);
    de[k] = Q.buf[k * Q.cap + slot];
    ls_out[k] = stream_ld(&L.ls[k * np + p]);
    ls_in[k] = lde[k] = 0.;
  }
  const double phi = stream_ld(&P.sv[SV_PHI * np + p]), psi = stream_ld(&P.sv[SV_PSI * np + p]);
  const double coh = stream_ld(&P.sv[SV_COH * np + p]), tcut = stream_ld(&P.sv[SV_TCUT * np + p]);
  double gE = 0., gphi = 0.;
  mc_vjp<D, MPMX_AD_TANGENTS>(m, phi, psi, coh, tcut, s_in, de, ls_out, ls_in, lde, gE, gphi,
                              Q.apex[slot] != 0);
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    stream_st(&L.ls[k * np + p], ls_in[k]);
    Q.buf[k * Q.cap + slot] = lde[k];
  }
  L.gE[p] += gE;
  L.gphi[p] += gphi;
}

// Their strain-increment adjoints through the velocity gradient: scatter
// into lnv, shape adjoint into lxq.
template <int D>
__global__ __launch_bounds__(kAdBlock) void k_ad_mc_tail(Mesh M, Particles P, Slot s0,
                                                         const double* nv, double* lnv,
                                                         AdjParticles L, AdjList Q,
                                                         AdjArgs a) {
  constexpr int NN = Elem<D>::NN;
  const int64_t e = blockIdx.x * (int64_t)kAdBlock + threadIdx.x;
  const int ns = Q.count[0];
  if (e >= ns + Q.count[1]) return;
  const int64_t slot = e < ns ? e : Q.cap - 1 - (e - ns);
  const int64_t np = P.np;
  const int p = Q.ids[slot];
  Shape<D> S;
  if (!load_shape<D>(M, s0, np, p, S)) return;
  double v[NN][D], G[NN][D], lde[6], lxi[D];
  load_kinematics<D>(M, S, nv, v, G);
#pragma unroll
  for (int k = 0; k < 6; ++k) lde[k] = Q.buf[k * Q.cap + slot];
#pragma unroll
  for (int k = 0; k < D; ++k) lxi[k] = 0.;
  double acc[NN][D];
  zero_acc<D>(acc);
  strain_rate_vjp<D>(M, S, v, G, lde, nullptr, a.dt, acc, lxi);
  add_lxq<D>(L, np, p, S, lxi);
#pragma unroll
  for (int i = 0; i < NN; ++i)
#pragma unroll
    for (int k = 0; k < D; ++k) ad_atomic(&lnv[k * M.nnodes + S.node[i]], acc[i][k]);
}


inline dim3 ad_grid(int64_t n) { return dim3((unsigned)((n + kAdBlock - 1) / kAdBlock)); }

}  // namespace

template <int D>
void launch_ad_save(const Particles& P, const SlotW& s, hipStream_t st) {
  if (P.np > 0) hipLaunchKernelGGL(k_ad_save<D>, ad_grid(P.np), dim3(kAdBlock), 0, st, P, s);
}

void launch_ad_fill(double* a, int64_t n, double v, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_ad_fill, ad_grid(n), dim3(kAdBlock), 0, st, a, n, v);
}

namespace {
__global__ void k_rows_to_soa(const double* rows, double* soa, int64_t n, int comps) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int k = 0; k < comps; ++k) soa[k * n + i] = rows[i * comps + k];
}
__global__ void k_soa_to_rows(const double* soa, double* rows, int64_t n, int comps,
                              double scale, const double* mass, const int* mat,
                              const Material* mats) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double per = mass ? mass[i] / mats[mat[i]].density : 1.0;
  for (int k = 0; k < comps; ++k) rows[i * comps + k] = soa[k * n + i] * scale * per;
}
}  // namespace

void launch_rows_to_soa(const double* rows, double* soa, int64_t n, int comps, hipStream_t st) {
  if (n > 0)
    hipLaunchKernelGGL(k_rows_to_soa, ad_grid(n), dim3(kAdBlock), 0, st, rows, soa, n, comps);
}
void launch_soa_to_rows(const double* soa, double* rows, int64_t n, int comps, double scale,
                        const double* mass, const int* mat, const Material* mats,
                        hipStream_t st) {
  if (n > 0)
    hipLaunchKernelGGL(k_soa_to_rows, ad_grid(n), dim3(kAdBlock), 0, st, soa, rows, n, comps,
                       scale, mass, mat, mats);
}

template <int D, int SCH>
static void backward_scheme(const Mesh& M, const Particles& P, const Nodes& Nd,
                            const Material* mats, const Slot& s0, const Slot& s1,
                            const AdjNodes& A, const AdjParticles& L, const AdjList& Q,
                            const AdjArgs& a, hipStream_t st) {
  const dim3 gp = ad_grid(P.np), gn = ad_grid(M.nnodes), b(kAdBlock);
  // the plastic Mohr-Coulomb updates queued by the stress pass
  auto plastic = [&](const double* nv, double* lnv) {
    hipLaunchKernelGGL((k_ad_mc<D>), dim3((unsigned)((P.np + kAdMcBlock - 1) / kAdMcBlock)),
                       dim3(kAdMcBlock), 0, st, P, mats, s0, L, Q);
    hipLaunchKernelGGL((k_ad_mc_tail<D>), gp, b, 0, st, M, P, s0, nv, lnv, L, Q, a);
  };
  hipLaunchKernelGGL((k_ad_node_fwd<D>), gn, b, 0, st, M, Nd, A, Q.count);
  hipLaunchKernelGGL((k_ad_b1<D, SCH>), gp, b, 0, st, M, P, mats, s0, s1, A, L, Q, a);
  if (SCH == USL) {
    plastic(A.vp, A.lvp);
    hipLaunchKernelGGL((k_ad_node<D, ND_AD_UPD | ND_AD_VEL>), gn, b, 0, st, M, Nd, A, a);
    hipLaunchKernelGGL((k_ad_b3<D, SCH>), gp, b, 0, st, M, P, mats, s0, s1, A, L, Q, a);
  } else if (SCH == USF) {
    hipLaunchKernelGGL((k_ad_node<D, ND_AD_UPD>), gn, b, 0, st, M, Nd, A, a);
    hipLaunchKernelGGL((k_ad_b3<D, SCH>), gp, b, 0, st, M, P, mats, s0, s1, A, L, Q, a);
    plastic(A.v, A.lv);
    hipLaunchKernelGGL((k_ad_node<D, ND_AD_VEL>), gn, b, 0, st, M, Nd, A, a);
    hipLaunchKernelGGL((k_ad_b5<D, SCH>), gp, b, 0, st, M, P, s0, A, L, a);
  } else {
    plastic(A.vpp, A.lvpp);
    hipLaunchKernelGGL((k_ad_node<D, ND_AD_VEL2>), gn, b, 0, st, M, Nd, A, a);
    hipLaunchKernelGGL((k_ad_b3<D, SCH>), gp, b, 0, st, M, P, mats, s0, s1, A, L, Q, a);
    hipLaunchKernelGGL((k_ad_node<D, ND_AD_UPD | ND_AD_VEL>), gn, b, 0, st, M, Nd, A, a);
    hipLaunchKernelGGL((k_ad_b5<D, SCH>), gp, b, 0, st, M, P, s0, A, L, a);
  }
}

template <int D>
void launch_ad_backward_step(int scheme, const Mesh& M, const Particles& P, const Nodes& Nd,
                             const Material* mats, const Slot& s0, const Slot& s1,
                             const AdjNodes& A, const AdjParticles& L, const AdjList& Q,
                             const AdjArgs& a, hipStream_t st) {
  if (P.np == 0) return;
  if (scheme == USL)
    backward_scheme<D, USL>(M, P, Nd, mats, s0, s1, A, L, Q, a, st);
  else if (scheme == USF)
    backward_scheme<D, USF>(M, P, Nd, mats, s0, s1, A, L, Q, a, st);
  else
    backward_scheme<D, MUSL>(M, P, Nd, mats, s0, s1, A, L, Q, a, st);
}

template void launch_ad_save<2>(const Particles&, const SlotW&, hipStream_t);
template void launch_ad_save<3>(const Particles&, const SlotW&, hipStream_t);
template void launch_ad_backward_step<2>(int, const Mesh&, const Particles&, const Nodes&,
                                         const Material*, const Slot&, const Slot&,
                                         const AdjNodes&, const AdjParticles&, const AdjList&,
                                         const AdjArgs&, hipStream_t);
template void launch_ad_backward_step<3>(int, const Mesh&, const Particles&, const Nodes&,
                                         const Material*, const Slot&, const Slot&,
                                         const AdjNodes&, const AdjParticles&, const AdjList&,
                                         const AdjArgs&, hipStream_t);

}  // namespace mpmx
