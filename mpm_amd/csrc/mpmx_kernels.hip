// HIP kernels of the explicit MPM step for gfx950 (MI355X).
//
// Particle state is SoA in particle (reference) order. Every step the
// particles are radix-sorted by cell (stable, so each cell keeps ascending
// particle ids); scatter kernels walk that order in 256-particle chunks, stage
// per-particle data in LDS, reduce each cell segment's contributions per
// local node in particle order, and add the segment sums to the nodes with
// FP64 global atomics (one atomic per cell segment, node and component
// instead of one per particle). In ordered mode the nodes instead gather:
// each node merges its cells' particle lists by particle id and sums in that
// order, which is the single-threaded reference's order.
#include <hip/hip_runtime.h>

#include "mpmx_device.h"
#include "mpmx_kernels.h"

namespace mpmx {

static constexpr int kBlock = 256;
// Workgroup size of the stress kernel k_stress: two waves (the 2M-particle
// collapse's stress kernel runs 3.5 % faster than with four; the 2D G2P with
// the stress update, k_g2p, is 15 % slower with two and keeps kBlock).
static constexpr int kStressBlock = 128;

// Occupancy floor (waves per SIMD) for the kernels that run the constitutive
// update; 0 leaves the register budget to the compiler. Left alone, the
// compiler schedules the Mohr-Coulomb return mapping for 1 wave per SIMD
// (>256 registers); 2 fits in 256 registers without spills and halves the
// kernel time on the 2M-particle collapse, 3 and 4 spill to scratch and lose.
#ifndef MPMX_MC_WAVES
#define MPMX_MC_WAVES 2
#endif
#if MPMX_MC_WAVES > 0
#define MPMX_MC_ATTR __attribute__((amdgpu_waves_per_eu(MPMX_MC_WAVES)))
#else
#define MPMX_MC_ATTR
#endif

// Occupancy floor of the kernels that run the kinematics and the elastic
// predictor and defer plastic return mappings (k_stress, k_g2p).
#ifndef MPMX_PRED_WAVES
#define MPMX_PRED_WAVES 2
#endif
#define MPMX_PRED_ATTR __attribute__((amdgpu_waves_per_eu(MPMX_PRED_WAVES)))
// ... and of the deferred return mappings (k_plastic)
#ifndef MPMX_PLASTIC_WAVES
#define MPMX_PLASTIC_WAVES 2
#endif
#define MPMX_PLASTIC_ATTR __attribute__((amdgpu_waves_per_eu(MPMX_PLASTIC_WAVES)))

MPMX_DEV void atomic_add(double* p, double v) { unsafeAtomicAdd(p, v); }

template <int D>
MPMX_DEV bool valid_cell(const Mesh& M, int c) {
  return c >= 0 && (int64_t)c < M.ncells;
}

// A particle could not be located in an earlier step of the queued run: the
// reference throws at the end of that step (mpm_scheme.tcc:250-256), so the
// kernels of every later step return at once and leave the state as that
// step ended it. The same value for every workgroup of a launch.
MPMX_DEV bool run_stopped(const Status* S) { return S->first_bad_step < S->step; }

// ------------------------- per-particle building blocks ----------------------
// Nodal velocity (and acceleration) of the particle's cell nodes.
template <int D>
MPMX_DEV void gather_nodes(const Mesh& M, const Nodes& Nd, int c, double (*nv)[D],
                           double (*na)[D]) {
  constexpr int NN = Elem<D>::NN;
#pragma unroll
  for (int i = 0; i < NN; ++i) {
    const int n = M.cell_nodes[(int64_t)c * NN + i];
#pragma unroll
    for (int k = 0; k < D; ++k) {
      nv[i][k] = Nd.vel[k * M.nnodes + n];
      if (na) na[i][k] = Nd.acc[k * M.nnodes + n];
    }
  }
}

// The kinematic half of compute_stress_strain for one particle
// (mpm_scheme.tcc:75-103: ΔF, strain rate, strain, volume, density, F).
// dndx: the particle's dN/dx at its xi. Writes the particle's strain, volume,
// density and F (and with ROT P.spin: the rotation of a Jaumann-rate
// update, linear_elastic.tcc:102-106); returns the strain increment ds and
// the new density.
template <int D, bool ROT = false>
MPMX_DEV void kinematics(const Mesh& M, const Particles& P, int p, int c, const double (*dndx)[D],
                         const double (*nv)[D], double dt, double* ds, double* rho_out,
                         double* vol_out) {
  constexpr int NN = Elem<D>::NN;
  const int64_t np = P.np;
  // The particle fields updated here are all loaded before the first store:
  // the compiler must assume that a store may alias any array, so a load
  // placed after a store waits for the round trip of the loads before it.
  double eps[6];
#pragma unroll
  for (int k = 0; k < 6; ++k)
    eps[k] = (D == 2 && (k == 2 || k > 3)) ? 0. : stream_ld(&P.strain[k * np + p]);
  double vol = stream_ld(&P.vol[p]);
  const double rho0 = stream_ld(&P.rho[p]);
  double Fr[D][3];
#pragma unroll
  for (int a = 0; a < D; ++a)
#pragma unroll
    for (int j = 0; j < 3; ++j) Fr[a][j] = stream_ld(&P.F[(a * 3 + j) * np + p]);
  // update_deformation_gradient_increment (particle.tcc:1578-1651)
  double R[D][D];
#pragma unroll
  for (int a = 0; a < D; ++a)
#pragma unroll
    for (int b = 0; b < D; ++b) R[a][b] = (a == b) ? 1.0 : 0.0;
#pragma unroll
  for (int i = 0; i < NN; ++i)
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
      for (int b = 0; b < D; ++b) R[a][b] = R[a][b] + (nv[i][a] * dndx[i][b]) * dt;
#pragma unroll
  for (int a = 0; a < D; ++a)
#pragma unroll
    for (int b = 0; b < D; ++b) {
      if (a != b && fabs(R[a][b]) < 1.E-15) R[a][b] = 0.;
      if (a == b && fabs(R[a][b] - 1.) < 1.E-15) R[a][b] = 1.;
    }
  // compute_strain_rate (particle.tcc:819-858) at the particle and centroid
  double sr[6] = {0, 0, 0, 0, 0, 0}, sc[6] = {0, 0, 0, 0, 0, 0};
  const double* dc = M.dndx_c + (int64_t)c * NN * D;
#pragma unroll
  for (int i = 0; i < NN; ++i) {
    const double* v = nv[i];
    const double* d = dndx[i];
    const double* e = dc + i * D;
    if constexpr (D == 2) {
      sr[0] += d[0] * v[0];
      sr[1] += d[1] * v[1];
      sr[3] += d[1] * v[0] + d[0] * v[1];
      sc[0] += e[0] * v[0];
      sc[1] += e[1] * v[1];
    } else {
      sr[0] += d[0] * v[0];
      sr[1] += d[1] * v[1];
      sr[2] += d[2] * v[2];
      sr[3] += d[1] * v[0] + d[0] * v[1];
      sr[4] += d[2] * v[1] + d[1] * v[2];
      sr[5] += d[2] * v[0] + d[0] * v[2];
      sc[0] += e[0] * v[0];
      sc[1] += e[1] * v[1];
      sc[2] += e[2] * v[2];
    }
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    if (fabs(sr[k]) < 1.E-15) sr[k] = 0.;
    if (fabs(sc[k]) < 1.E-15) sc[k] = 0.;
  }
  // compute_strain (particle.tcc:861-877)
#pragma unroll
  for (int k = 0; k < 6; ++k) ds[k] = sr[k] * dt;
  if constexpr (D == 2) {
    stream_st(&P.strain[0 * np + p], eps[0] + ds[0]);
    stream_st(&P.strain[1 * np + p], eps[1] + ds[1]);
    stream_st(&P.strain[3 * np + p], eps[3] + ds[3]);
  } else {
#pragma unroll
    for (int k = 0; k < 6; ++k) stream_st(&P.strain[k * np + p], eps[k] + ds[k]);
  }
  double svol = sc[0];
#pragma unroll
  for (int k = 1; k < D; ++k) svol = svol + sc[k];
  const double dvol = dt * svol;
  // update_volume (particle.tcc:598-606)
  vol *= (1. + dvol);
  const double rho = rho0 / (1. + dvol);
  stream_st(&P.vol[p], vol);
  stream_st(&P.rho[p], rho);
  *vol_out = vol;
  *rho_out = rho;
  // update_deformation_gradient (particle.tcc:1654-1659): rows 0..D-1 of
  // F <- ΔF F; row 2 of a 2D ΔF is (0,0,1) and leaves F's row 2 unchanged.
  {
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        double t = R[a][0] * Fr[0][j];
#pragma unroll
        for (int b = 1; b < D; ++b) t = t + R[a][b] * Fr[b][j];
        stream_st(&P.F[(a * 3 + j) * np + p], t);
      }
  }
  if constexpr (ROT) {  // stored rather than kept live across the stress update
    if (P.rotF) {  // Green-Naghdi: the step-start F rows and the dF block
#pragma unroll
      for (int a = 0; a < D; ++a)
#pragma unroll
        for (int j = 0; j < 3; ++j) P.rotF[(a * 3 + j) * np + p] = Fr[a][j];
#pragma unroll
      for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = 0; b < D; ++b) P.rotF[(D * 3 + a * D + b) * np + p] = R[a][b];
    }
    P.spin[p] = 0.5 * (0.5 * (R[0][1] - R[1][0]));
    if constexpr (D == 3) {
      P.spin[np + p] = 0.5 * (0.5 * (R[0][2] - R[2][0]));
      P.spin[2 * np + p] = 0.5 * (0.5 * (R[1][2] - R[2][1]));
    }
  }
}

// Mohr-Coulomb state variables written back after an update
// (mohr_coulomb.tcc:348-513): the yield state always; epsilon, rho and theta
// unless the particle separated (the update returns before the invariants);
// pdstrain when it changed (separated or plastic); friction, dilation,
// cohesion and tension cutoff only with softening (:360-384).
MPMX_DEV void store_sv(const Particles& P, const Material& m, int64_t p, const double* sv) {
  const int64_t np = P.np;
  const double y = sv[SV_YIELD];
  stream_st(&P.sv[SV_YIELD * np + p], y);
  if (y != 3.) {
    stream_st(&P.sv[SV_EPS * np + p], sv[SV_EPS]);
    stream_st(&P.sv[SV_RHO * np + p], sv[SV_RHO]);
    stream_st(&P.sv[SV_THETA * np + p], sv[SV_THETA]);
  }
  if (y != 0.) stream_st(&P.sv[SV_PDS * np + p], sv[SV_PDS]);
  if (m.softening) {
    stream_st(&P.sv[SV_PHI * np + p], sv[SV_PHI]);
    stream_st(&P.sv[SV_PSI * np + p], sv[SV_PSI]);
    stream_st(&P.sv[SV_COH * np + p], sv[SV_COH]);
    stream_st(&P.sv[SV_TCUT * np + p], sv[SV_TCUT]);
  }
}

// The state variables an update reads: friction, dilation, cohesion, tension
// cutoff (from the material when uniform) and pdstrain; the yield state,
// epsilon, rho and theta are outputs only.
MPMX_DEV void load_sv(const Particles& P, const Material& m, int64_t p, double* sv) {
  const int64_t np = P.np;
  if (m.uniform) {
    sv[SV_PHI] = m.sv0[SV_PHI];
    sv[SV_PSI] = m.sv0[SV_PSI];
    sv[SV_COH] = m.sv0[SV_COH];
    sv[SV_TCUT] = m.sv0[SV_TCUT];
  } else {
    sv[SV_PHI] = stream_ld(&P.sv[SV_PHI * np + p]);
    sv[SV_PSI] = stream_ld(&P.sv[SV_PSI * np + p]);
    sv[SV_COH] = stream_ld(&P.sv[SV_COH * np + p]);
    sv[SV_TCUT] = stream_ld(&P.sv[SV_TCUT * np + p]);
  }
  sv[SV_PDS] = stream_ld(&P.sv[SV_PDS * np + p]);
  sv[SV_YIELD] = sv[SV_EPS] = sv[SV_RHO] = sv[SV_THETA] = 0.;
}

// Appends particle p to the deferred plastic list with its record: strain
// increment and trial state. Shear returns fill the list from the front,
// tensile ones from the back, so that k_plastic's waves mostly run one
// branch of the flow rule (one atomic per wave and kind on the counts).
MPMX_DEV void defer_plastic(const Particles& P, Status* S, int p, const double* ds,
                            const McTrial& T, const double* trial = nullptr) {
  const bool tensile = T.yt == FS_TENSILE;
  const unsigned long long mask = __ballot(tensile);
  const unsigned long long all = __ballot(1);
  const unsigned long long mine = tensile ? mask : (all & ~mask);
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((long long)mine) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(tensile ? &S->nplastic_t : &S->nplastic, __popcll(mine));
  base = __shfl(base, leader, 64);
  const int k = base + __popcll(mine & ((1ull << lane) - 1ull));
  const int64_t np = P.np;
  const int slot = tensile ? (int)(np - 1 - k) : k;
  P.plastic_list[slot] = p;
  double* r = P.plastic_rec + slot;
#pragma unroll
  for (int c = 0; c < 6; ++c) r[(PR_DS + c) * np] = ds[c];
  r[PR_EPS * np] = T.I.eps;
  r[PR_RHO * np] = T.I.rho;
  r[PR_THETA * np] = T.I.theta;
  r[PR_ST * np] = T.I.st;
  r[PR_CT * np] = T.I.ct;
  r[PR_S3 * np] = T.I.s3;
  r[PR_C3 * np] = T.I.c3;
  r[PR_YTRIAL * np] = T.yield_trial;
  if (P.plastic_trial) {  // a stress rate: the (rotated) trial stress itself
#pragma unroll
    for (int c = 0; c < 6; ++c) P.plastic_trial[c * np + slot] = trial[c];
  }
}

// compute_stress_strain for one particle: kinematics, then the constitutive
// update (particle.tcc:880-890 -> material). Returns the new stress in `s`.
// DEFER: a Mohr-Coulomb particle whose trial stress yields is appended to
// the plastic list instead (stress and state variables left as they were,
// `s` not the new stress); k_plastic completes it with the same inputs.
// ROT: kernels built for a Jaumann-rate material (P.spin set); without it
// the rate is not looked at.
template <int D, bool DEFER = false, bool ROT = false>
MPMX_DEV void stress_strain(const Mesh& M, const Particles& P, const Material* mats, int p, int c,
                            const double (*dndx)[D], const double (*nv)[D], double dt,
                            double* s, double* vol_out, Status* S = nullptr) {
  const int64_t np = P.np;
  double ds[6], rho;
  kinematics<D, ROT>(M, P, p, c, dndx, nv, dt, ds, &rho, vol_out);
  // one grouped round trip after the kinematic stores
#pragma unroll
  for (int k = 0; k < 6; ++k) s[k] = stream_ld(&P.stress[k * np + p]);
  const Material& m = mats[stream_ld(&P.mat[p])];
  double sv[NSV];
  if (m.type != LE) load_sv(P, m, p, sv);
  // an objective stress rate: the rotated trial stress (ROT builds only)
  int rate = RATE_NONE;
  if constexpr (ROT) rate = m.rate;
  double rtrial[6];
  if (rate != RATE_NONE) {
    double L[3][3], LL[3][3];
    if (rate == RATE_JAUMANN) {
      double a[3] = {P.spin[p], 0., 0.};
      if constexpr (D == 3) {
        a[1] = P.spin[np + p];
        a[2] = P.spin[2 * np + p];
      }
      jaumann_rotations(a, L, LL);
    } else {  // Green-Naghdi: F (step start) and dF as 3x3
      double F0[3][3], G[3][3];
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          F0[a][j] = a < D ? P.rotF[(a * 3 + j) * np + p] : stream_ld(&P.F[(a * 3 + j) * np + p]);
          G[a][j] = (a < D && j < D) ? P.rotF[(D * 3 + a * D + j) * np + p] : (a == j ? 1. : 0.);
        }
      green_naghdi_rotations(F0, G, L, LL);
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) rtrial[k] = s[k];
    rate_trial(m, L, LL, ds, rtrial);
  }
  if (m.type == LE && rate != RATE_NONE) {
#pragma unroll
    for (int k = 0; k < 6; ++k) s[k] = rtrial[k];
  } else if (m.type == LE) {
    double t[6];
    de_mul(m, ds, t);
#pragma unroll
    for (int k = 0; k < 6; ++k) s[k] = s[k] + t[k];
  } else if constexpr (DEFER) {
    double trial[6];
    McTrial T;
    if (!mc_begin(m, s, ds, rho, sv, trial, T, rate != RATE_NONE ? rtrial : nullptr)) {
      // the yield state (and softened strength parameters) now, the rest of
      // the state variables and the stress in k_plastic
      stream_st(&P.sv[SV_YIELD * np + p], sv[SV_YIELD]);
      if (m.softening) {
        stream_st(&P.sv[SV_PHI * np + p], sv[SV_PHI]);
        stream_st(&P.sv[SV_PSI * np + p], sv[SV_PSI]);
        stream_st(&P.sv[SV_COH * np + p], sv[SV_COH]);
        stream_st(&P.sv[SV_TCUT * np + p], sv[SV_TCUT]);
      }
      defer_plastic(P, S, p, ds, T, trial);
      return;
    }
    store_sv(P, m, p, sv);
  } else {
    mc_stress(m, s, ds, rho, sv, rate != RATE_NONE ? rtrial : nullptr);
    store_sv(P, m, p, sv);
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) stream_st(&P.stress[k * np + p], s[k]);
}

// Particle-level dN/dx at its xi.
template <int D>
MPMX_DEV void particle_dndx(const Mesh& M, int c, const double* xi, double (*dndx)[D],
                            double (*Ji)[D]) {
  constexpr int NN = Elem<D>::NN;
  double X[NN][D];
  load_cell_X<D>(M, c, X);
  jacobian_inverse<D>(xi, X, Ji);
#pragma unroll
  for (int i = 0; i < NN; ++i) dndx_i<D>(xi, i, Ji, dndx[i]);
}

// locate_particle_cells (mesh.tcc:1116-1153): current cell, then its
// neighbours in ascending id; otherwise queue the particle for the full scan.
template <int D, bool ISO = false>
MPMX_DEV void locate_one(const Mesh& M, const Particles& P, Status* S, int* scan_list, int p,
                         const double* x, int cur, int hint) {
  const int64_t np = P.np;
  double xi[D];
  if (cur >= 0 && point_in_cell<D, ISO>(M, cur, x, xi)) {
#pragma unroll
    for (int k = 0; k < D; ++k) P.xi[k * np + p] = xi[k];
    return;
  }
  const int base = cur >= 0 ? cur : hint;
  if (base >= 0) {
    const int e = M.nb_ptr[base + 1];
    for (int j = M.nb_ptr[base]; j < e; ++j) {
      const int nb = M.nb_idx[j];
      if (point_in_cell<D, ISO>(M, nb, x, xi)) {
        P.cell[p] = nb;
#pragma unroll
        for (int k = 0; k < D; ++k) P.xi[k * np + p] = xi[k];
        return;
      }
    }
  }
  const int slot = atomicAdd(&S->nscan, 1);
  scan_list[slot] = p;
}

// ------------------------------- kernels ------------------------------------

// Locate every particle (initial placement). hint[p] >= 0 is a cell to try
// first (kept only if the point lies in it), -1: none.
template <int D>
__global__ __launch_bounds__(kBlock) void k_locate_init(Mesh M, Particles P, Status* S,
                                                        int* scan_list, const int* hint) {
  const int64_t p = blockIdx.x * (int64_t)kBlock + threadIdx.x;
  if (p >= P.np) return;
  double x[D];
#pragma unroll
  for (int k = 0; k < D; ++k) x[k] = stream_ld(&P.x[k * P.np + p]);
  const int h = hint ? hint[p] : -1;
  double xi[D];
  const bool in = h >= 0 && (M.iso ? point_in_cell<D, true>(M, h, x, xi)
                                   : point_in_cell<D, false>(M, h, x, xi));
  if (in) {
    P.cell[p] = h;
#pragma unroll
    for (int k = 0; k < D; ++k) P.xi[k * P.np + p] = xi[k];
    return;
  }
  P.cell[p] = -1;
  if (M.iso)
    locate_one<D, true>(M, P, S, scan_list, (int)p, x, -1, h);
  else
    locate_one<D, false>(M, P, S, scan_list, (int)p, x, -1, h);
}

// Per-step locate of an isoparametric mesh (the step kernels locate Cartesian
// meshes themselves): current cell, then neighbours, else the scan queue.
template <int D>
__global__ __launch_bounds__(kBlock) void k_locate_step(Mesh M, Particles P, Status* S,
                                                        int* scan_list) {
  if (run_stopped(S)) return;
  const int64_t p = blockIdx.x * (int64_t)kBlock + threadIdx.x;
  if (p >= P.np) return;
  const int c = P.cell[p];
  if (!valid_cell<D>(M, c)) return;
  double x[D];
#pragma unroll
  for (int k = 0; k < D; ++k) x[k] = stream_ld(&P.x[k * P.np + p]);
  locate_one<D, true>(M, P, S, scan_list, (int)p, x, c, -1);
}

// Full-mesh scan for queued particles: first cell in ascending id order that
// contains the point (the single-threaded reading of mesh.tcc:1142-1151).
// With instances, the scan covers the cells of the particle's instance (that
// of its previous cell, or of its hint at initial placement). Persistent
// grid; every block exits after the queue is drained. A particle found in no
// cell gets cell -1 (every step kernel skips it from then on) and either
// stops the run (the reference throws, mpm_scheme.tcc:250-256) or, with
// remove_lost, is counted for removal (mpm_scheme.tcc:258-260).
template <int D>
__global__ __launch_bounds__(kBlock) void k_locate_scan(Mesh M, Particles P, Status* S,
                                                        const int* scan_list, const int* hint,
                                                        int remove_lost) {
  if (run_stopped(S)) return;
  __shared__ int best;
  const int n = S->nscan;
  for (int q = blockIdx.x; q < n; q += gridDim.x) {
    const int p = scan_list[q];
    double x[D];
#pragma unroll
    for (int k = 0; k < D; ++k) x[k] = stream_ld(&P.x[k * P.np + p]);
    int64_t lo = 0, hi = M.ncells;
    if (M.cell_inst) {
      int prev = P.cell[p];
      if (prev < 0 && hint) prev = hint[p];
      if (prev >= 0 && prev < M.ncells) {
        const int inst = M.cell_inst[prev];
        lo = M.inst_begin[inst];
        hi = M.inst_begin[inst + 1];
      }
    }
    if (threadIdx.x == 0) best = 0x7fffffff;
    __syncthreads();
    for (int64_t c0 = lo; c0 < hi; c0 += kBlock) {
      const int64_t c = c0 + threadIdx.x;
      double xi[D];
      const bool in = c < hi && (M.iso ? point_in_cell<D, true>(M, (int)c, x, xi)
                                       : point_in_cell<D, false>(M, (int)c, x, xi));
      if (in) atomicMin(&best, (int)c);
      __syncthreads();
      if (best != 0x7fffffff) break;
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      if (best != 0x7fffffff) {
        double xi[D];
        if (M.iso)
          point_in_cell<D, true>(M, best, x, xi);
        else
          point_in_cell<D, false>(M, best, x, xi);
        P.cell[p] = best;
#pragma unroll
        for (int k = 0; k < D; ++k) P.xi[k * P.np + p] = xi[k];
      } else {
        P.cell[p] = -1;
        if (remove_lost) {
          atomicAdd(&S->nlost, 1);
        } else {
          const long long step = (long long)S->step;
          atomicMin((unsigned long long*)&S->first_bad_step, (unsigned long long)step);
          atomicExch(&S->bad_particle, p);
        }
      }
    }
    __syncthreads();
  }
}

// End of step: advance the counter, clear the scan queue.
// Mesh::update_nodal_acceleration_constraints (mesh.tcc:1414-1428): the math
// functions' values at current_time = step * dt (mpm_scheme.tcc:219).
__global__ void k_fn_eval(const MathFn* fns, const double* data, int nf, const Status* S,
                          double dt, double* out) {
  const int f = threadIdx.x;
  if (f < nf) out[f] = math_fn_value(fns[f], data, static_cast<double>(S->step) * dt);
}

__global__ void k_step_end(Status* S) {
  if (run_stopped(S)) return;
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    S->step += 1;
    S->nscan = 0;
    S->nplastic = 0;
    S->nplastic_t = 0;
    S->ad_slot += 1;
  }
}

// ---------------- fast scatter: LDS-staged segmented reduction --------------
// What the staging carries per particle. SC_AFF: the momentum is mapped with
// the affine velocity v + C (x_i - x_p) (APIC, ASFLIP, TPIC).
enum { SC_MV = 1, SC_F = 2, SC_MOM = 4, SC_AFF = 8 };

// CH: particles per workgroup (one per thread), a multiple of 64.
template <int D, int WHAT, int CH>
struct Stage {
  static constexpr int NN = Elem<D>::NN;
  int key[CH];
  double xi[D][CH];
  double m[CH];
  double v[D][CH];
  double B[(WHAT & SC_F) ? D * D : 1][CH];  // force matrix, see force_matrix()
  double C[(WHAT & SC_AFF) ? D * D : 1][CH];  // affine velocity matrix
  double xp[(WHAT & SC_AFF) ? D : 1][CH];     // particle position
  int seg_start[CH + 1];
  int nseg;
  int wave_heads[CH / 64];
};

// The affine velocity of a particle at node position xn: v + C (xn - xp)
// (map_mass_momentum_to_nodes_affine / _taylor, particle.tcc:661-724), with
// C = B D^-1 (APIC, ASFLIP) or B (TPIC) precomputed per transfer.
template <int D>
MPMX_DEV void affine_velocity(const double* v, const double* C, const double* xp,
                              const double* xn, double* mv) {
  double r[D];
#pragma unroll
  for (int k = 0; k < D; ++k) r[k] = xn[k] - xp[k];
#pragma unroll
  for (int a = 0; a < D; ++a) {
    double t = C[a * D] * r[0];
#pragma unroll
    for (int k = 1; k < D; ++k) t = t + C[a * D + k] * r[k];
    mv[a] = v[a] + t;
  }
}

// Internal force of a particle on node i (particle.tcc:914-950) is
// f_i = -V (sigma . dN_i/dx) with dN_i/dx = J^-1 grad_i(xi), i.e.
// f_i = B grad_i(xi) with B = -V sigma J^-1 (a DxD matrix per particle).
// Staging B instead of sigma, V and J^-1 (9 doubles instead of 16) lets three
// scatter workgroups share a CU's LDS; the reassociation changes only the
// rounding of each contribution (fast mode; ordered mode keeps the
// reference's order).
template <int D>
MPMX_DEV void force_matrix(const double* s, double vol, const double (*Ji)[D], double* B) {
  double sig[D][D];
  if constexpr (D == 2) {
    sig[0][0] = s[0]; sig[0][1] = s[3];
    sig[1][0] = s[3]; sig[1][1] = s[1];
  } else {
    sig[0][0] = s[0]; sig[0][1] = s[3]; sig[0][2] = s[5];
    sig[1][0] = s[3]; sig[1][1] = s[1]; sig[1][2] = s[4];
    sig[2][0] = s[5]; sig[2][1] = s[4]; sig[2][2] = s[2];
  }
  const double mv = -1. * vol;
#pragma unroll
  for (int a = 0; a < D; ++a)
#pragma unroll
    for (int c = 0; c < D; ++c) {
      double t = sig[a][0] * Ji[0][c];
#pragma unroll
      for (int b = 1; b < D; ++b) t = t + sig[a][b] * Ji[b][c];
      B[a * D + c] = t * mv;
    }
}

// Nodal components a scatter adds: mass and momentum (SC_MV), body and
// internal force (SC_F), momentum only (SC_MOM).
template <int D, int WHAT>
struct Comp {
  static constexpr int MOM = (WHAT & SC_MV) ? 1 : 0;
  static constexpr int FE = (WHAT & SC_MV) ? 1 + D : 0;
  static constexpr int FI = FE + D;
  static constexpr int NC = ((WHAT & SC_MV) ? 1 + D : 0) + ((WHAT & SC_F) ? 2 * D : 0) +
                            ((WHAT & SC_MOM) ? D : 0);
};

// Tasks of one round: lane l of a wave holds (cell segment s, local node i)
// with l = 2^D s + i. Cells adjacent along x in the sorted order share a
// face: node 1 (5) of a cell is node 0 (4) of the next, node 2 (6) is node 3
// (7) of the next (quad: 1 -> 0, 2 -> 3), i.e. the lanes 2^D - 1 and 2^D + 1
// further on. Where the node ids agree, the lower lane adds its neighbour's
// sums and the neighbour skips its atomics: about 1.8x fewer FP64 atomics
// for cell-sorted particles, exact for any mesh (only equal ids merge).
template <int D, int WHAT>
MPMX_DEV void merge_face_neighbours(int i, int& n, double* vals) {
  constexpr int NN = Elem<D>::NN;
  constexpr int NC = Comp<D, WHAT>::NC;
  const int lane = threadIdx.x & 63;
  bool absorbed = false;
#pragma unroll
  for (int side = 0; side < 2; ++side) {
    const int off = side == 0 ? NN - 1 : NN + 1;
    const int pn = __shfl_down(n, off, 64);
    const bool give = D == 3 ? (side == 0 ? (i == 1 || i == 5) : (i == 2 || i == 6))
                             : (side == 0 ? i == 1 : i == 2);
    const bool take = n >= 0 && give && lane + off < 64 && pn == n;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const double pv = __shfl_down(vals[c], off, 64);
      if (take) vals[c] = vals[c] + pv;
    }
    const int qn = __shfl_up(n, off, 64);
    const bool recv = D == 3 ? (side == 0 ? (i == 0 || i == 4) : (i == 3 || i == 7))
                             : (side == 0 ? i == 0 : i == 3);
    if (n >= 0 && recv && lane >= off && qn == n) absorbed = true;
  }
  if (absorbed) n = -1;
}

// Phase 2 of every scatter kernel: given staged particles [0, cnt) of this
// chunk, sorted by cell, add each cell segment's per-node sums to the nodes.
template <int D, int WHAT, int CH>
MPMX_DEV void scatter_segments(Stage<D, WHAT, CH>& st, int cnt, const Mesh& M, const Nodes& Nd,
                               const double* g) {
  constexpr int NN = Elem<D>::NN;
  const int tid = threadIdx.x;
  // segment heads (valid particles only; invalid ones carry key -1 at the end)
  const bool valid = tid < cnt && st.key[tid] >= 0;
  const bool head = valid && (tid == 0 || st.key[tid] != st.key[tid - 1]);
  const unsigned long long mask = __ballot(head);
  const int lane = tid & 63, wave = tid >> 6;
  if (lane == 0) st.wave_heads[wave] = __popcll(mask);
  __syncthreads();
  int before = 0;
#pragma unroll
  for (int w = 0; w < CH / 64; ++w) before += (w < wave) ? st.wave_heads[w] : 0;
  const int rank = before + __popcll(mask & ((1ull << lane) - 1ull));
  if (head) st.seg_start[rank] = tid;
  if (tid == 0) {
    int tot = 0;
#pragma unroll
    for (int w = 0; w < CH / 64; ++w) tot += st.wave_heads[w];
    st.nseg = tot;
  }
  __syncthreads();
  const int nseg = st.nseg;
  // A segment runs from its head to the next head or to the first invalid
  // slot before it, whichever comes first. Between full sorts the keys are
  // an earlier permutation's with current cells, so a particle that left the
  // mesh (key -1, nothing staged) can sit inside a run of valid keys; the
  // valid slot after it is a head (its key differs from -1), so invalid
  // slots only ever trail a segment.
  using CP = Comp<D, WHAT>;
  const int ntask = nseg * NN;
  for (int t0 = 0; t0 < ntask; t0 += CH) {  // same trip count for every lane
    const int task = t0 + tid;
    const int sidx = task / NN;
    const int i = task - sidx * NN;
    int n = -1;
    double vals[CP::NC];
#pragma unroll
    for (int k = 0; k < CP::NC; ++k) vals[k] = 0.;
    if (task < ntask) {
    const int b = st.seg_start[sidx];
    const int e = (sidx + 1 < nseg) ? st.seg_start[sidx + 1] : cnt;
    const int c = st.key[b];
    double am = 0., amom[D], afe[D], afi[D];
#pragma unroll
    for (int k = 0; k < D; ++k) amom[k] = afe[k] = afi[k] = 0.;
    double xn[D];  // the node's position (affine transfers)
    if constexpr ((WHAT & SC_AFF) != 0) {
      const int nid = M.cell_nodes[(int64_t)st.key[b] * NN + i];
#pragma unroll
      for (int k = 0; k < D; ++k) xn[k] = M.node_x[k * M.nnodes + nid];
    }
    for (int t = b; t < e && st.key[t] >= 0; ++t) {
      double xi[D];
#pragma unroll
      for (int k = 0; k < D; ++k) xi[k] = st.xi[k][t];
      const double N = shapefn_i<D>(xi, i);
      const double m = st.m[t];
      if constexpr ((WHAT & (SC_MV | SC_MOM)) != 0) {
        const double mN = m * N;
        if constexpr ((WHAT & SC_MV) != 0) am = am + mN;
        double vel[D];
#pragma unroll
        for (int k = 0; k < D; ++k) vel[k] = st.v[k][t];
        if constexpr ((WHAT & SC_AFF) != 0) {
          double Cp[D * D], xp[D];
#pragma unroll
          for (int k = 0; k < D * D; ++k) Cp[k] = st.C[k][t];
#pragma unroll
          for (int k = 0; k < D; ++k) xp[k] = st.xp[k][t];
          affine_velocity<D>(vel, Cp, xp, xn, vel);
        }
#pragma unroll
        for (int k = 0; k < D; ++k) amom[k] = amom[k] + mN * vel[k];
      }
      if constexpr ((WHAT & SC_F) != 0) {
#pragma unroll
        for (int k = 0; k < D; ++k) afe[k] = afe[k] + g[k] * m * N;
        double gr[D];
        grad_i<D>(xi, i, gr);
#pragma unroll
        for (int a = 0; a < D; ++a) {
          double f = st.B[a * D][t] * gr[0];
#pragma unroll
          for (int c = 1; c < D; ++c) f = f + st.B[a * D + c][t] * gr[c];
          afi[a] = afi[a] + f;
        }
      }
    }
    n = M.cell_nodes[(int64_t)c * NN + i];
    if constexpr ((WHAT & SC_MV) != 0) vals[0] = am;
    if constexpr ((WHAT & (SC_MV | SC_MOM)) != 0) {
#pragma unroll
      for (int k = 0; k < D; ++k) vals[CP::MOM + k] = amom[k];
    }
    if constexpr ((WHAT & SC_F) != 0) {
#pragma unroll
      for (int k = 0; k < D; ++k) {
        vals[CP::FE + k] = afe[k];
        vals[CP::FI + k] = afi[k];
      }
    }
    }  // task < ntask
    merge_face_neighbours<D, WHAT>(i, n, vals);
    if (n < 0) continue;
    if constexpr ((WHAT & SC_MV) != 0) {
      atomic_add(&Nd.mass[n], vals[0]);
      Nd.active[n] = 1;
    }
    if constexpr ((WHAT & (SC_MV | SC_MOM)) != 0) {
#pragma unroll
      for (int k = 0; k < D; ++k) atomic_add(&Nd.mom[k * M.nnodes + n], vals[CP::MOM + k]);
    }
    if constexpr ((WHAT & SC_F) != 0) {
#pragma unroll
      for (int k = 0; k < D; ++k) {
        atomic_add(&Nd.fext[k * M.nnodes + n], vals[CP::FE + k]);
        atomic_add(&Nd.fint[k * M.nnodes + n], vals[CP::FI + k]);
      }
    }
  }
}

MPMX_DEV void gravity_now(const StepArgs& A, const Status* S, int dim, double* g) {
  double gm = 1.0;
  if (A.ramp > 0.0) {
    const double r = static_cast<double>(S->step) * A.dt / A.ramp;
    gm = r < 1.0 ? r : 1.0;  // std::min(1.0, r)
  }
  for (int k = 0; k < dim; ++k) g[k] = gm * A.g[k];
}

// Mass/momentum (and, for USL/MUSL, body and internal forces from the
// stress of the previous step) scattered from sorted particles.
// mpm_scheme.tcc:44-71 and 131-190; particle.tcc:630-658, 893-950.
template <int D, int WHAT, int CH>
__global__ __launch_bounds__(CH) void k_p2g(Mesh M, Particles P, Nodes Nd, StepArgs A,
                                                const Status* S, const int* skey,
                                                const int* perm) {
  if (run_stopped(S)) return;
  __shared__ Stage<D, WHAT, CH> st;
  const int64_t base = blockIdx.x * (int64_t)CH;
  const int cnt = (int)((P.np - base) < CH ? (P.np - base) : CH);
  const int tid = threadIdx.x;
  const int64_t np = P.np;
  if (tid < cnt) {
    const int key = skey[base + tid];
    const int p = perm[base + tid];
    const bool ok = valid_cell<D>(M, key);
    st.key[tid] = ok ? key : -1;
    if (ok) {
      double xi[D];
#pragma unroll
      for (int k = 0; k < D; ++k) {
        xi[k] = P.xi[k * np + p];
        st.xi[k][tid] = xi[k];
        st.v[k][tid] = stream_ld(&P.v[k * np + p]);
      }
      st.m[tid] = P.mass[p];
      if constexpr ((WHAT & SC_AFF) != 0) {
#pragma unroll
        for (int k = 0; k < D * D; ++k) st.C[k][tid] = P.Cmat[k * np + p];
#pragma unroll
        for (int k = 0; k < D; ++k) st.xp[k][tid] = stream_ld(&P.x[k * np + p]);
      }
      if constexpr ((WHAT & SC_F) != 0) {
        double sg[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) sg[k] = stream_ld(&P.stress[k * np + p]);
        const double vol = stream_ld(&P.vol[p]);
        constexpr int NN = Elem<D>::NN;
        double X[NN][D], Ji[D][D], Bm[D * D];
        load_cell_X<D>(M, key, X);
        jacobian_inverse<D>(xi, X, Ji);
        force_matrix<D>(sg, vol, Ji, Bm);
#pragma unroll
        for (int k = 0; k < D * D; ++k) st.B[k][tid] = Bm[k];
      }
    }
  }
  double g[3] = {0, 0, 0};
  if constexpr ((WHAT & SC_F) != 0) gravity_now(A, S, D, g);
  __syncthreads();
  scatter_segments<D, WHAT, CH>(st, cnt, M, Nd, g);
}

// USF stress update in sorted order, then the force scatter of the new stress.
template <int D, int CH, bool ROT>
__global__ __launch_bounds__(CH) MPMX_MC_ATTR void k_stress_scatter(Mesh M, Particles P, Nodes Nd,
                                                           const Material* mats, StepArgs A,
                                                           const Status* S, const int* skey,
                                                           const int* perm) {
  if (run_stopped(S)) return;
  constexpr int NN = Elem<D>::NN;
  __shared__ Stage<D, SC_F, CH> st;
  const int64_t base = blockIdx.x * (int64_t)CH;
  const int cnt = (int)((P.np - base) < CH ? (P.np - base) : CH);
  const int tid = threadIdx.x;
  const int64_t np = P.np;
  if (tid < cnt) {
    const int key = skey[base + tid];
    const int p = perm[base + tid];
    const bool ok = valid_cell<D>(M, key);
    st.key[tid] = ok ? key : -1;
    if (ok) {
      double xi[D];
#pragma unroll
      for (int k = 0; k < D; ++k) {
        xi[k] = P.xi[k * np + p];
        st.xi[k][tid] = xi[k];
      }
      double dndx[NN][D], Ji[D][D], nv[NN][D];
      particle_dndx<D>(M, key, xi, dndx, Ji);
      gather_nodes<D>(M, Nd, key, nv, nullptr);
      double s[6], vol;
      stress_strain<D, false, ROT>(M, P, mats, p, key, dndx, nv, A.dt, s, &vol);
      st.m[tid] = P.mass[p];
      double Bm[D * D];
      force_matrix<D>(s, vol, Ji, Bm);
#pragma unroll
      for (int k = 0; k < D * D; ++k) st.B[k][tid] = Bm[k];
    }
  }
  double g[3] = {0, 0, 0};
  gravity_now(A, S, D, g);
  __syncthreads();
  scatter_segments<D, SC_F, CH>(st, cnt, M, Nd, g);
}

// Stress update in particle order (USL/MUSL postcompute, and USF in ordered
// mode); optionally locate afterwards.
template <int D, bool LOCATE, bool ROT>
__global__ __launch_bounds__(kStressBlock) MPMX_PRED_ATTR void k_stress(Mesh M, Particles P,
                                                                      Nodes Nd,
                                                                      const Material* mats,
                                                                      StepArgs A, Status* S,
                                                                      int* scan_list) {
  if (run_stopped(S)) return;
  constexpr int NN = Elem<D>::NN;
  const int64_t p = blockIdx.x * (int64_t)kStressBlock + threadIdx.x;
  if (p >= P.np) return;
  const int64_t np = P.np;
  const int c = P.cell[p];
  if (!valid_cell<D>(M, c)) return;
  double xi[D];
#pragma unroll
  for (int k = 0; k < D; ++k) xi[k] = P.xi[k * np + p];
  double dndx[NN][D], Ji[D][D], nv[NN][D];
  particle_dndx<D>(M, c, xi, dndx, Ji);
  gather_nodes<D>(M, Nd, c, nv, nullptr);
  double s[6], vol;
  stress_strain<D, true, ROT>(M, P, mats, (int)p, c, dndx, nv, A.dt, s, &vol, S);
  if constexpr (LOCATE) {
    double x[D];
#pragma unroll
    for (int k = 0; k < D; ++k) x[k] = stream_ld(&P.x[k * np + p]);
    locate_one<D>(M, P, S, scan_list, (int)p, x, c, -1);
  }
}

// G2P: compute_updated_position (particle.tcc:994-1070) in particle order,
// optionally followed by the USL stress update and by locate.
template <int D, bool STRESS, bool LOCATE, bool ROT = false>
__global__ __launch_bounds__(kBlock) MPMX_PRED_ATTR void k_g2p(Mesh M, Particles P, Nodes Nd,
                                                const Material* mats, StepArgs A, int vu,
                                                Status* S, int* scan_list) {
  if (run_stopped(S)) return;
  constexpr int NN = Elem<D>::NN;
  const int64_t p = blockIdx.x * (int64_t)kBlock + threadIdx.x;
  if (p >= P.np) return;
  const int64_t np = P.np;
  const int c = P.cell[p];
  if (!valid_cell<D>(M, c)) return;
  double xi[D];
#pragma unroll
  for (int k = 0; k < D; ++k) xi[k] = P.xi[k * np + p];
  double N[NN];
  shapefn<D>(xi, N);
  double nv[NN][D], na[NN][D];
  gather_nodes<D>(M, Nd, c, nv, na);
  double sv[D], sa[D];
#pragma unroll
  for (int k = 0; k < D; ++k) sv[k] = sa[k] = 0.;
#pragma unroll
  for (int i = 0; i < NN; ++i)
#pragma unroll
    for (int k = 0; k < D; ++k) {
      sv[k] += N[i] * nv[i][k];
      sa[k] += N[i] * na[i][k];
    }
  double x[D], v0[D], u0[D];
#pragma unroll
  for (int k = 0; k < D; ++k) {  // loads before stores (see kinematics)
    v0[k] = stream_ld(&P.v[k * np + p]);
    x[k] = stream_ld(&P.x[k * np + p]);
    u0[k] = stream_ld(&P.u[k * np + p]);
  }
  const int pvc = P.pvc_mask ? P.pvc_mask[p] : 0;
  const double beta = vu == ASFLIP ? P.beta[p] : 0.;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    double v = v0[k];
    double dx = sv[k];  // position increment velocity
    if (vu == FLIP) {
      v += sa[k] * A.dt;
      v = A.blending * v + (1.0 - A.blending) * sv[k];
    } else if (vu == ASFLIP) {  // particle.tcc:1097-1111
      const double flip = v + sa[k] * A.dt;
      v = A.blending * flip + (1.0 - A.blending) * sv[k];
      dx = beta * A.blending * flip + (1.0 - beta * A.blending) * sv[k];
    } else {
      v = sv[k];
    }
    if (pvc & (1 << k)) v = P.pvc_val[k * np + p];  // particle.tcc:1190-1195
    stream_st(&P.v[k * np + p], v);
    x[k] = x[k] + dx * A.dt;
    stream_st(&P.x[k * np + p], x[k]);
    stream_st(&P.u[k * np + p], u0[k] + dx * A.dt);
  }
  if constexpr (STRESS) {
    double dndx[NN][D], Ji[D][D];
    particle_dndx<D>(M, c, xi, dndx, Ji);
    double s[6], vol;
    stress_strain<D, true, ROT>(M, P, mats, (int)p, c, dndx, nv, A.dt, s, &vol, S);
  }
  if constexpr (LOCATE) locate_one<D>(M, P, S, scan_list, (int)p, x, c, -1);
}

// The mapping matrices of the affine velocity updates at the start of G2P
// (before the positions move), |B_ij| < 1e-15 -> 0:
// compute_affine_mapping_matrix (particle.tcc:1687-1708),
// B = sum_i N_i v_i (x_i - x_p)^T for APIC and ASFLIP, and
// compute_velocity_gradient (:1662-1684), B = sum_i v_i dN_i/dx^T, for TPIC.
// ASFLIP also gets its beta (compute_asflip_beta, :1711-1730): 0 when a node
// of the cell has a velocity constraint or the deformation gradient
// increment of the current nodal velocities compresses (det < 1), else 1.
template <int D>
__global__ __launch_bounds__(kBlock) void k_mapping_matrix(Mesh M, Particles P, Nodes Nd,
                                                           StepArgs A, int vu,
                                                           const Status* S) {
  if (run_stopped(S)) return;
  constexpr int NN = Elem<D>::NN;
  const int64_t p = blockIdx.x * (int64_t)kBlock + threadIdx.x;
  if (p >= P.np) return;
  const int64_t np = P.np;
  const int c = P.cell[p];
  if (!valid_cell<D>(M, c)) return;
  double xi[D], xp[D];
#pragma unroll
  for (int k = 0; k < D; ++k) {
    xi[k] = P.xi[k * np + p];
    xp[k] = stream_ld(&P.x[k * np + p]);
  }
  double nv[NN][D];
  gather_nodes<D>(M, Nd, c, nv, nullptr);
  double B[D][D];
#pragma unroll
  for (int a = 0; a < D; ++a)
#pragma unroll
    for (int b = 0; b < D; ++b) B[a][b] = 0.;
  double dndx[NN][D];
  if (vu == TPIC || vu == ASFLIP) {
    double Ji[D][D];
    particle_dndx<D>(M, c, xi, dndx, Ji);
  }
  if (vu == TPIC) {
#pragma unroll
    for (int i = 0; i < NN; ++i)
#pragma unroll
      for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = 0; b < D; ++b) B[a][b] = B[a][b] + nv[i][a] * dndx[i][b];
  } else {
    double N[NN], X[NN][D];
    shapefn<D>(xi, N);
    load_cell_X<D>(M, c, X);
#pragma unroll
    for (int i = 0; i < NN; ++i)
#pragma unroll
      for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = 0; b < D; ++b) B[a][b] = B[a][b] + (N[i] * nv[i][a]) * (X[i][b] - xp[b]);
  }
#pragma unroll
  for (int a = 0; a < D; ++a)
#pragma unroll
    for (int b = 0; b < D; ++b) P.Bmat[(a * D + b) * np + p] = fabs(B[a][b]) < 1.E-15 ? 0. : B[a][b];
  if (vu != ASFLIP) return;
  double beta = 1.0;
#pragma unroll
  for (int i = 0; i < NN; ++i) {
    const int n = M.cell_nodes[(int64_t)c * NN + i];
    const int ci = Nd.cons ? Nd.cons[n] : -1;
    if (ci >= 0 && Nd.table[ci].vmask != 0) beta = 0.0;
  }
  if (beta != 0.0) {
    double R[3][3] = {{1., 0., 0.}, {0., 1., 0.}, {0., 0., 1.}};
#pragma unroll
    for (int i = 0; i < NN; ++i)
#pragma unroll
      for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = 0; b < D; ++b) R[a][b] = R[a][b] + (nv[i][a] * dndx[i][b]) * A.dt;
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
      for (int b = 0; b < D; ++b) {
        if (a != b && fabs(R[a][b]) < 1.E-15) R[a][b] = 0.;
        if (a == b && fabs(R[a][b] - 1.) < 1.E-15) R[a][b] = 1.;
      }
    const double det = R[0][0] * (R[1][1] * R[2][2] - R[2][1] * R[1][2]) -
                       R[1][0] * (R[0][1] * R[2][2] - R[2][1] * R[0][2]) +
                       R[2][0] * (R[0][1] * R[1][2] - R[1][1] * R[0][2]);
    if (det < 1.0) beta = 0.0;
  }
  P.beta[p] = beta;
}

// The matrix C of the next affine particle-to-grid transfer, momentum
// m N_i (v + C (x_i - x_p)): C = B D^-1 with the shape tensor
// D = sum_i N_i (x_i - x_p)(x_i - x_p)^T for APIC and ASFLIP
// (map_mass_momentum_to_nodes_affine, particle.tcc:661-696), C = B for TPIC
// (_taylor, :699-724). N at the particle's local coordinates, x_p its
// current position.
template <int D>
__global__ __launch_bounds__(kBlock) void k_affine_matrix(Mesh M, Particles P, int vu,
                                                          const Status* S) {
  if (run_stopped(S)) return;
  constexpr int NN = Elem<D>::NN;
  const int64_t p = blockIdx.x * (int64_t)kBlock + threadIdx.x;
  if (p >= P.np) return;
  const int64_t np = P.np;
  const int c = P.cell[p];
  if (!valid_cell<D>(M, c)) return;
  double B[D][D];
#pragma unroll
  for (int a = 0; a < D; ++a)
#pragma unroll
    for (int b = 0; b < D; ++b) B[a][b] = P.Bmat[(a * D + b) * np + p];
  if (vu == TPIC) {
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
      for (int b = 0; b < D; ++b) P.Cmat[(a * D + b) * np + p] = B[a][b];
    return;
  }
  double xi[D], xp[D];
#pragma unroll
  for (int k = 0; k < D; ++k) {
    xi[k] = P.xi[k * np + p];
    xp[k] = stream_ld(&P.x[k * np + p]);
  }
  double N[NN], X[NN][D];
  shapefn<D>(xi, N);
  load_cell_X<D>(M, c, X);
  double Dm[D][D], Di[D][D];
#pragma unroll
  for (int a = 0; a < D; ++a)
#pragma unroll
    for (int b = 0; b < D; ++b) Dm[a][b] = 0.;
#pragma unroll
  for (int i = 0; i < NN; ++i) {
    double r[D];
#pragma unroll
    for (int k = 0; k < D; ++k) r[k] = X[i][k] - xp[k];
#pragma unroll
    for (int a = 0; a < D; ++a)
#pragma unroll
      for (int b = 0; b < D; ++b) Dm[a][b] = Dm[a][b] + (N[i] * r[a]) * r[b];
  }
  mat_inverse<D>(Dm, Di);
#pragma unroll
  for (int a = 0; a < D; ++a)
#pragma unroll
    for (int b = 0; b < D; ++b) {
      double t = B[a][0] * Di[0][b];
#pragma unroll
      for (int k = 1; k < D; ++k) t = t + B[a][k] * Di[k][b];
      P.Cmat[(a * D + b) * np + p] = t;
    }
}

// The deferred Mohr-Coulomb return mappings of this step: the rest of
// compute_stress (mohr_coulomb.tcc:412-512) after the elastic predictor of
// stress_strain, i.e. mc_first and mc_iterate on the same start-of-step
// stress, state variables and strain increment, with the trial state from
// the list record. Grid-stride loop over the shear entries (front of the
// list), then the tensile ones (back).
constexpr int kPlasticBlock = 64;  // one wave per workgroup (run times vary by wave)
__global__ __launch_bounds__(kPlasticBlock) MPMX_PLASTIC_ATTR void k_plastic(Particles P,
                                                                             const Material* mats,
                                                                             const Status* S) {
  if (run_stopped(S)) return;
  const int ns = S->nplastic, nt = S->nplastic_t;
  const int64_t np = P.np;
  for (int c = blockIdx.x * kPlasticBlock + threadIdx.x; c < ns + nt;
       c += gridDim.x * kPlasticBlock) {
    const int slot = c < ns ? c : (int)(np - 1 - (c - ns));
    const int p = P.plastic_list[slot];
    const double* r = P.plastic_rec + slot;
    double ds[6], s[6], sv[NSV];
#pragma unroll
    for (int k = 0; k < 6; ++k) ds[k] = r[(PR_DS + k) * np];
    McTrial T;
    T.I.eps = r[PR_EPS * np];
    T.I.rho = r[PR_RHO * np];
    T.I.theta = r[PR_THETA * np];
    T.I.st = r[PR_ST * np];
    T.I.ct = r[PR_CT * np];
    T.I.s3 = r[PR_S3 * np];
    T.I.c3 = r[PR_C3 * np];
    T.yield_trial = r[PR_YTRIAL * np];
    T.yt = c < ns ? FS_SHEAR : FS_TENSILE;
#pragma unroll
    for (int k = 0; k < 6; ++k) s[k] = stream_ld(&P.stress[k * np + p]);
    const Material& m = mats[stream_ld(&P.mat[p])];
    load_sv(P, m, p, sv);
    double trial[6];
    if (P.plastic_trial) {  // stress rates: the trial stress the predictor used
#pragma unroll
      for (int k = 0; k < 6; ++k) trial[k] = P.plastic_trial[k * np + slot];
    } else {
      double t[6];
      de_mul(m, ds, t);
#pragma unroll
      for (int k = 0; k < 6; ++k) trial[k] = s[k] + t[k];
    }
    McIter R;
    mc_first(m, s, trial, ds, sv, T, R);
    mc_iterate(m, s, sv, R, 100);
    stream_st(&P.sv[SV_EPS * np + p], sv[SV_EPS]);
    stream_st(&P.sv[SV_THETA * np + p], sv[SV_THETA]);
    stream_st(&P.sv[SV_RHO * np + p], sv[SV_RHO]);
    stream_st(&P.sv[SV_PDS * np + p], sv[SV_PDS]);
#pragma unroll
    for (int k = 0; k < 6; ++k) stream_st(&P.stress[k * np + p], s[k]);
  }
}

// MUSL G2P in sorted order with the momentum of the updated particle
// velocities scattered back (mpm_scheme_musl.tcc:21-35; mass is unchanged
// because the shape functions are not recomputed).
template <int D, int CH>
__global__ __launch_bounds__(CH) void k_g2p_scatter(Mesh M, Particles P, Nodes Nd, StepArgs A,
                                                        int vu, const Status* S, const int* skey,
                                                        const int* perm) {
  if (run_stopped(S)) return;
  constexpr int NN = Elem<D>::NN;
  __shared__ Stage<D, SC_MOM, CH> st;
  const int64_t base = blockIdx.x * (int64_t)CH;
  const int cnt = (int)((P.np - base) < CH ? (P.np - base) : CH);
  const int tid = threadIdx.x;
  const int64_t np = P.np;
  if (tid < cnt) {
    const int key = skey[base + tid];
    const int p = perm[base + tid];
    const bool ok = valid_cell<D>(M, key);
    st.key[tid] = ok ? key : -1;
    if (ok) {
      double xi[D];
#pragma unroll
      for (int k = 0; k < D; ++k) xi[k] = P.xi[k * np + p];
      double N[NN];
      shapefn<D>(xi, N);
      double nv[NN][D], na[NN][D];
      gather_nodes<D>(M, Nd, key, nv, na);
      double sv[D], sa[D];
#pragma unroll
      for (int k = 0; k < D; ++k) sv[k] = sa[k] = 0.;
#pragma unroll
      for (int i = 0; i < NN; ++i)
#pragma unroll
        for (int k = 0; k < D; ++k) {
          sv[k] += N[i] * nv[i][k];
          sa[k] += N[i] * na[i][k];
        }
      double v0[D], x0[D], u0[D];
#pragma unroll
      for (int k = 0; k < D; ++k) {  // loads before stores (see kinematics)
        v0[k] = stream_ld(&P.v[k * np + p]);
        x0[k] = stream_ld(&P.x[k * np + p]);
        u0[k] = stream_ld(&P.u[k * np + p]);
      }
      st.m[tid] = P.mass[p];
      const int pvc = P.pvc_mask ? P.pvc_mask[p] : 0;
#pragma unroll
      for (int k = 0; k < D; ++k) {
        double v = v0[k];
        if (vu == FLIP) {
          v += sa[k] * A.dt;
          v = A.blending * v + (1.0 - A.blending) * sv[k];
        } else {
          v = sv[k];
        }
        if (pvc & (1 << k)) v = P.pvc_val[k * np + p];  // particle.tcc:1190-1195
        stream_st(&P.v[k * np + p], v);
        stream_st(&P.x[k * np + p], x0[k] + sv[k] * A.dt);
        stream_st(&P.u[k * np + p], u0[k] + sv[k] * A.dt);
        st.v[k][tid] = v;
        st.xi[k][tid] = xi[k];
      }
    }
  }
  __syncthreads();
  scatter_segments<D, SC_MOM, CH>(st, cnt, M, Nd, nullptr);
}

// ------------------------------- node kernels -------------------------------
// Particle tractions mapped to the nodes' external force
// (Particle::map_traction_force, particle.tcc:983-990), one thread per
// particle with a traction, FP64 atomics (fast mode).
template <int D>
__global__ __launch_bounds__(kBlock) void k_traction_scatter(Mesh M, Particles P, Nodes Nd,
                                                             const Status* S) {
  if (run_stopped(S)) return;
  constexpr int NN = Elem<D>::NN;
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= P.ntr) return;
  const int64_t p = P.tr_list[t];
  const int64_t np = P.np;
  const int c = P.cell[p];
  if (!valid_cell<D>(M, c)) return;
  double xi[D], N[NN], tr[D];
#pragma unroll
  for (int k = 0; k < D; ++k) xi[k] = P.xi[k * np + p];
  shapefn<D>(xi, N);
  particle_traction<D>(P, Nd.fnval, p, tr);
#pragma unroll
  for (int i = 0; i < NN; ++i) {
    const int n = M.cell_nodes[(int64_t)c * NN + i];
#pragma unroll
    for (int k = 0; k < D; ++k) atomic_add(&Nd.fext[k * M.nnodes + n], N[i] * tr[k]);
  }
}

// Where this step's nodal record goes (StepArgs::ad_nodes), or null.
MPMX_DEV double* ad_record_base(const StepArgs& A) {
  if (!A.ad_nodes) return nullptr;
  const int slot = *A.ad_slot;
  if (slot < 0 || slot >= A.ad_nslots) return nullptr;
  return A.ad_nodes + (int64_t)slot * A.ad_stride;
}

// Adjoint recording (StepArgs::ad_nodes): the nodal quantities of this step.
// FIRST: after the update (mass, momentum, external + internal force,
// updated velocity, acceleration); SECOND: MUSL's second momentum and its
// velocity.

template <int D>
MPMX_DEV void ad_record_first(const StepArgs& A, int64_t n, int64_t nn, double mass,
                              const double* mom, const double* fe, const double* fi,
                              const double* v, const double* a) {
  double* b = ad_record_base(A);
  if (!b) return;
  b[n] = mass;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    b[(1 + 0 * D + k) * nn + n] = mom[k];
    b[(1 + 1 * D + k) * nn + n] = fe[k] + fi[k];
    b[(1 + 2 * D + k) * nn + n] = v[k];
    b[(1 + 3 * D + k) * nn + n] = a[k];
  }
}
template <int D>
MPMX_DEV void ad_record_second(const StepArgs& A, int64_t n, int64_t nn, const double* mom,
                               const double* v) {
  double* b = ad_record_base(A);
  if (!b) return;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    b[(1 + 4 * D + k) * nn + n] = mom[k];
    b[(1 + 5 * D + k) * nn + n] = v[k];
  }
}

// Flags for what a node kernel does.
// ND_INIT: start from v = a = 0 (Node::initialise, node.tcc:26-43) instead of
// the stored values.
// (flag values are declared in mpmx_kernels.h)

// compute_velocity (ND_VEL) and/or compute_acceleration_velocity_cundall
// (ND_ACC) for active nodes; then clears accumulators for the next pass.
template <int D, int FL>
__global__ __launch_bounds__(kBlock) void k_node(Mesh M, Nodes Nd, StepArgs A, const Status* S) {
  if (run_stopped(S)) return;
  const int64_t n = blockIdx.x * (int64_t)kBlock + threadIdx.x;
  if (n >= M.nnodes) return;
  const int64_t nn = M.nnodes;
  if (!Nd.active[n]) {
    if ((FL & ND_ACC) != 0) {
      double* b = ad_record_base(A);
      if (b) b[n] = 0.;
    }
    return;
  }
  const int ci = Nd.cons[n];
  const NodeCons* c = ci >= 0 ? &Nd.table[ci] : nullptr;
  const double mass = Nd.mass[n];
  double v[D], a[D];
#pragma unroll
  for (int k = 0; k < D; ++k) {
    v[k] = ((FL & ND_INIT) != 0) ? 0. : Nd.vel[k * nn + n];
    a[k] = ((FL & ND_INIT) != 0) ? 0. : Nd.acc[k * nn + n];
  }
  if constexpr ((FL & ND_VEL) != 0) {
    double mom[D];
#pragma unroll
    for (int k = 0; k < D; ++k) mom[k] = Nd.mom[k * nn + n];
    node_compute_velocity<D>(mass, mom, v, a, c);
  }
  if constexpr ((FL & ND_ACC) != 0) {
    double fe[D], fi[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
      fe[k] = Nd.fext[k * nn + n];
      fi[k] = Nd.fint[k * nn + n];
    }
    add_concentrated_force<D>(c, Nd.fnval, fe);
    node_acc_vel<D>(mass, fe, fi, v, a, c, A.dt, A.damping, Nd.fnval);
    if (A.ad_nodes) {
      double mom[D];
#pragma unroll
      for (int k = 0; k < D; ++k) mom[k] = Nd.mom[k * nn + n];
      ad_record_first<D>(A, n, nn, mass, mom, fe, fi, v, a);
    }
  }
  if constexpr ((FL & (ND_VEL | ND_ACC | ND_INIT)) == ND_VEL) {
    if (A.ad_nodes) {
      double mom[D];
#pragma unroll
      for (int k = 0; k < D; ++k) mom[k] = Nd.mom[k * nn + n];
      ad_record_second<D>(A, n, nn, mom, v);
    }
  }
#pragma unroll
  for (int k = 0; k < D; ++k) {
    Nd.vel[k * nn + n] = v[k];
    Nd.acc[k * nn + n] = a[k];
  }
  if constexpr ((FL & (ND_ZERO_MOM | ND_ZERO_ALL)) != 0) {
#pragma unroll
    for (int k = 0; k < D; ++k) Nd.mom[k * nn + n] = 0.;
  }
  if constexpr ((FL & ND_ZERO_ALL) != 0) {
    Nd.mass[n] = 0.;
#pragma unroll
    for (int k = 0; k < D; ++k) {
      Nd.fext[k * nn + n] = 0.;
      Nd.fint[k * nn + n] = 0.;
    }
    Nd.active[n] = 0;
  }
}

// ---------------------- ordered mode: node-centric gather --------------------
// cell_start/cell_end: range of each cell in the sorted order (empty: equal).
template <int D>
__global__ __launch_bounds__(kBlock) void k_cell_ranges(int64_t np, int64_t ncells,
                                                        const int* skey, int* cstart,
                                                        int* cend) {
  const int64_t k = blockIdx.x * (int64_t)kBlock + threadIdx.x;
  if (k >= np) return;
  const int c = skey[k];
  if (c < 0 || c >= ncells) return;
  if (k == 0 || skey[k - 1] != c) cstart[c] = (int)k;
  if (k == np - 1 || skey[k + 1] != c) cend[c] = (int)k + 1;
}

__global__ void k_fill_int(int* a, int64_t n, int v) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// Per node: the particles of its cells summed in particle-id order (the
// single-threaded reference order). WHAT: SC_MV (mass+momentum), SC_F
// (body+internal force), SC_MOM (momentum only, mass kept). Then the node
// update flags FL as in k_node.
//
// One wave per node. Each lane takes slots of the node's concatenated cell
// lists (each sorted by particle id, the lists disjoint), ranks its
// particles by id among all of the node's particles, computes their
// contributions and stores them at their rank in LDS; then one lane per
// sum folds it in rank order, so every sum adds the same terms in the same
// order as one thread walking the particles by id. A node with more than
// kGatherMax particles is summed by one lane merging the cell lists.
constexpr int kGatherMax = 128;

// Slots of a node's sums: mass, momentum, external force, internal force.
template <int D>
struct GatherSlots {
  static constexpr int M = 0, MOM = 1, FE = 1 + D, FI = 1 + 2 * D, N = 1 + 3 * D;
  template <int WHAT>
  static constexpr MPMX_DEV bool used(int v) {
    return v == M ? (WHAT & SC_MV) != 0
                  : (v < FE ? (WHAT & (SC_MV | SC_MOM)) != 0 : (WHAT & SC_F) != 0);
  }
  // the first slot WHAT uses (the used slots are consecutive) and their count
  template <int WHAT>
  static constexpr int first() {
    return (WHAT & SC_MV) != 0 ? M : ((WHAT & SC_MOM) != 0 ? MOM : FE);
  }
  template <int WHAT>
  static constexpr int count() {
    return ((WHAT & SC_MV) != 0 ? 1 + D : ((WHAT & SC_MOM) != 0 ? D : 0)) +
           ((WHAT & SC_F) != 0 ? 2 * D : 0);
  }
};

// Before a force gather: each particle's shape function and internal-force
// term for every node of its cell (Particles::gterm), so that the node
// gathers read them instead of recomputing the cell's Jacobian per node.
// Each term is read once, by the gather of its node: streamed like the
// particle state.
template <int D>
__global__ __launch_bounds__(kBlock) void k_gather_prep(Mesh M, Particles P, const Status* S) {
  if (run_stopped(S)) return;
  constexpr int NN = Elem<D>::NN;
  const int64_t p = blockIdx.x * (int64_t)kBlock + threadIdx.x;
  if (p >= P.np) return;
  const int64_t np = P.np;
  const int c = P.cell[p];
  if (!valid_cell<D>(M, c)) return;
  double xi[D], X[NN][D], Ji[D][D], s[6];
#pragma unroll
  for (int k = 0; k < D; ++k) xi[k] = P.xi[k * np + p];
#pragma unroll
  for (int k = 0; k < 6; ++k) s[k] = stream_ld(&P.stress[k * np + p]);
  const double mv = -1. * stream_ld(&P.vol[p]);
  load_cell_X<D>(M, c, X);
  jacobian_inverse<D>(xi, X, Ji);
#pragma unroll
  for (int i = 0; i < NN; ++i) {
    double d[D], f[D];
    dndx_i<D>(xi, i, Ji, d);
    if constexpr (D == 2) {
      f[0] = d[0] * s[0] + d[1] * s[3];
      f[1] = d[1] * s[1] + d[0] * s[3];
    } else {
      f[0] = d[0] * s[0] + d[1] * s[3] + d[2] * s[5];
      f[1] = d[1] * s[1] + d[0] * s[3] + d[2] * s[4];
      f[2] = d[2] * s[2] + d[1] * s[4] + d[0] * s[5];
    }
    double* t = P.gterm + (int64_t)i * (1 + D) * np + p;
    stream_st(&t[0], shapefn_i<D>(xi, i));
#pragma unroll
    for (int k = 0; k < D; ++k) stream_st(&t[(1 + k) * np], f[k] * mv);
  }
}

// One particle's terms of node n's sums (n is local node i of the
// particle's cell), each computed as the sum it enters takes it: m N,
// m N v_k, g_k m N, and -(dN/dx . sigma)_k V. N and the force terms come
// from Particles::gterm in the force gathers and in the MUSL momentum
// gather (which follows the first gather of the same step at unchanged
// local coordinates).
template <int D, int WHAT>
MPMX_DEV void gather_terms(const Mesh& M, const Particles& P, int64_t n, int p, int i,
                           const double* g, double* out) {
  using GS = GatherSlots<D>;
  const int64_t np = P.np;
  double N;
  const double* t = P.gterm + (int64_t)i * (1 + D) * np + p;
  if constexpr ((WHAT & (SC_F | SC_MOM)) != 0) {
    N = stream_ld(&t[0]);
  } else {
    double xi[D];
#pragma unroll
    for (int k = 0; k < D; ++k) xi[k] = P.xi[k * np + p];
    N = shapefn_i<D>(xi, i);
  }
  const double m = P.mass[p];
  if constexpr ((WHAT & (SC_MV | SC_MOM)) != 0) {
    const double mN = m * N;
    out[GS::M] = mN;
    double vel[D];
#pragma unroll
    for (int k = 0; k < D; ++k) vel[k] = P.v[k * np + p];
    if (P.Cmat) {  // affine velocity updates
      double Cp[D * D], xp[D], xn[D];
#pragma unroll
      for (int k = 0; k < D * D; ++k) Cp[k] = P.Cmat[k * np + p];
#pragma unroll
      for (int k = 0; k < D; ++k) {
        xp[k] = P.x[k * np + p];
        xn[k] = M.node_x[k * M.nnodes + n];
      }
      affine_velocity<D>(vel, Cp, xp, xn, vel);
    }
#pragma unroll
    for (int k = 0; k < D; ++k) out[GS::MOM + k] = mN * vel[k];
  }
  if constexpr ((WHAT & SC_F) != 0) {
#pragma unroll
    for (int k = 0; k < D; ++k) {
      out[GS::FE + k] = g[k] * m * N;
      out[GS::FI + k] = stream_ld(&t[(1 + k) * np]);
    }
  }
}

template <int D, int WHAT, int FL>
__global__ __launch_bounds__(64) void k_gather(Mesh M, Particles P, Nodes Nd, StepArgs A,
                                               const Status* S, const int* perm,
                                               const int* cstart, const int* cend) {
  if (run_stopped(S)) return;
  using GS = GatherSlots<D>;
  constexpr int NN = Elem<D>::NN;
  constexpr int NV = GS::N;
  constexpr int V0 = GS::template first<WHAT>(), NU = GS::template count<WHAT>();
  __shared__ double terms[NU][kGatherMax];  // the used slots V0 .. V0 + NU - 1
  const int lane = threadIdx.x;
  const int64_t n = blockIdx.x;
  if (n >= M.nnodes) return;
  const int64_t nnodes = M.nnodes, np = P.np;
  // lane j < NN: cell j of the node, its range in the sorted order and the
  // node's local index in it
  int cj = -1, hj = 0, ej = 0, lj = 0;
  if (lane < NN) {
    cj = M.node_cells[n * NN + lane];
    if (cj >= 0) {
      hj = cstart[cj];
      ej = cend[cj];
#pragma unroll
      for (int i = 0; i < NN; ++i)
        if (M.cell_nodes[(int64_t)cj * NN + i] == (int)n) lj = i;
    }
  }
  int cells[NN], loc[NN], h[NN], e[NN], off[NN];
  int T = 0;
#pragma unroll
  for (int j = 0; j < NN; ++j) {
    cells[j] = __shfl(cj, j, 64);
    loc[j] = __shfl(lj, j, 64);
    h[j] = __shfl(hj, j, 64);
    e[j] = __shfl(ej, j, 64);
    if (e[j] < h[j]) e[j] = h[j];
    off[j] = T;
    T += e[j] - h[j];
  }
  const bool any = T > 0;
  if ((WHAT & SC_MV) != 0 && lane == 0) Nd.active[n] = any ? 1 : 0;
  if (!any) {
    if ((FL & ND_ACC) != 0 && lane == 0) {
      double* b = ad_record_base(A);
      if (b) b[n] = 0.;
    }
    return;
  }
  double g[3] = {0, 0, 0};
  if constexpr ((WHAT & SC_F) != 0) gravity_now(A, S, D, g);
  double sums[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) sums[v] = 0.;
  if (T <= kGatherMax) {
    // this lane's slots lane and lane + 64: particle and local node
    int ps[2], is[2] = {0, 0};
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int sl = lane + 64 * q;
      ps[q] = 0x7fffffff;
      if (sl < T) {  // the cell holding sl (selects, no indexed register arrays)
        int hs = h[0], os = off[0], ls = loc[0];
#pragma unroll
        for (int jj = 1; jj < NN; ++jj)
          if (sl >= off[jj]) {
            hs = h[jj];
            os = off[jj];
            ls = loc[jj];
          }
        ps[q] = perm[hs + (sl - os)];
        is[q] = ls;
      }
    }
    // rank among the node's particles: each cell list holds ascending ids
    // (stable sort of id-ordered input) and the lists are disjoint, so a
    // particle's rank is the sum over the lists of how many of their ids are
    // below its own (its own list: its position there), each count by a
    // binary search in LDS with the list's (wave-uniform) power-of-two steps
    __shared__ int ids[kGatherMax];
#pragma unroll
    for (int q = 0; q < 2; ++q)
      if (lane + 64 * q < T) ids[lane + 64 * q] = ps[q];
    __syncthreads();
    const int nq = T > 64 ? 2 : 1;
    int rk[2] = {0, 0};
#pragma unroll
    for (int j = 0; j < NN; ++j) {
      const int len = (j + 1 < NN ? off[j + 1] : T) - off[j];
      if (len == 0) continue;
      const int top = 1 << (31 - __builtin_clz(len));  // largest power of two <= len
#pragma unroll 1
      for (int q = 0; q < nq; ++q) {
        const int a = q ? ps[1] : ps[0];
        int below = 0;
        for (int step = top; step > 0; step >>= 1)
          if (below + step <= len && ids[off[j] + below + step - 1] < a) below += step;
        if (q) rk[1] += below;
        else rk[0] += below;
      }
    }
#pragma unroll 1
    for (int q = 0; q < 2; ++q) {  // one slot at a time (register pressure)
      if (lane + 64 * q < T) {
        const int pq = q ? ps[1] : ps[0], iq = q ? is[1] : is[0];
        const int rq = q ? rk[1] : rk[0];
        double out[NV];
        gather_terms<D, WHAT>(M, P, n, pq, iq, g, out);
#pragma unroll
        for (int v = 0; v < NU; ++v) terms[v][rq] = out[V0 + v];
      }
    }
    __syncthreads();
    double acc = 0.;
    if (lane < NU) {
#pragma unroll 8
      for (int r = 0; r < T; ++r) acc = acc + terms[lane][r];
    }
#pragma unroll
    for (int v = 0; v < NU; ++v) sums[V0 + v] = __shfl(acc, v, 64);
  } else if (lane == 0) {
    // too many particles for the LDS slots: merge the cell lists by id
    int hh[NN];
#pragma unroll
    for (int j = 0; j < NN; ++j) hh[j] = h[j];
    while (true) {
      int bj = -1, bp = 0x7fffffff, bl = 0;
#pragma unroll
      for (int j = 0; j < NN; ++j)
        if (hh[j] < e[j]) {
          const int p = perm[hh[j]];
          if (p < bp) {
            bp = p;
            bj = j;
            bl = loc[j];
          }
        }
      if (bj < 0) break;
#pragma unroll
      for (int j = 0; j < NN; ++j)
        if (j == bj) hh[j] += 1;
      double out[NV];
      gather_terms<D, WHAT>(M, P, n, bp, bl, g, out);
#pragma unroll
      for (int v = 0; v < NV; ++v)
        if (GS::template used<WHAT>(v)) sums[v] = sums[v] + out[v];
    }
  }
  if (lane != 0) return;
  double am = sums[GS::M], amom[D], afe[D], afi[D];
#pragma unroll
  for (int k = 0; k < D; ++k) {
    amom[k] = sums[GS::MOM + k];
    afe[k] = sums[GS::FE + k];
    afi[k] = sums[GS::FI + k];
  }
  int h0[NN];
#pragma unroll
  for (int j = 0; j < NN; ++j) h0[j] = h[j];
  // particle tractions after every body force, in particle order
  // (Mesh::apply_traction_on_particles, mesh.tcc:1371-1386)
  if constexpr ((WHAT & SC_F) != 0) {
    if (P.ntr > 0) {
#pragma unroll
      for (int j = 0; j < NN; ++j) h[j] = h0[j];
      while (true) {
        int bj = -1, bp = 0x7fffffff, bl = 0;
#pragma unroll
        for (int j = 0; j < NN; ++j)
          if (h[j] < e[j]) {
            const int p = perm[h[j]];
            if (p < bp) {
              bp = p;
              bj = j;
              bl = loc[j];
            }
          }
        if (bj < 0) break;
#pragma unroll
        for (int j = 0; j < NN; ++j)
          if (j == bj) h[j] += 1;
        const int p = bp, i = bl;
        bool has = false;
#pragma unroll
        for (int k = 0; k < D; ++k) has = has || P.tr_fn[k * np + p] >= -1;
        if (!has) continue;
        double xi[D], t[D];
#pragma unroll
        for (int k = 0; k < D; ++k) xi[k] = P.xi[k * np + p];
        const double N = shapefn_i<D>(xi, i);
        particle_traction<D>(P, Nd.fnval, p, t);
#pragma unroll
        for (int k = 0; k < D; ++k) afe[k] = afe[k] + N * t[k];
      }
    }
  }
  if constexpr ((WHAT & SC_MV) != 0) Nd.mass[n] = am;
  if constexpr ((WHAT & (SC_MV | SC_MOM)) != 0) {
#pragma unroll
    for (int k = 0; k < D; ++k) Nd.mom[k * nnodes + n] = amom[k];
  }
  if constexpr ((WHAT & SC_F) != 0) {
#pragma unroll
    for (int k = 0; k < D; ++k) {
      Nd.fext[k * nnodes + n] = afe[k];
      Nd.fint[k * nnodes + n] = afi[k];
    }
  }
  // node update
  const int ci = Nd.cons[n];
  const NodeCons* cc = ci >= 0 ? &Nd.table[ci] : nullptr;
  const double mass = Nd.mass[n];
  double v[D], a[D];
#pragma unroll
  for (int k = 0; k < D; ++k) {
    v[k] = ((FL & ND_INIT) != 0) ? 0. : Nd.vel[k * nnodes + n];
    a[k] = ((FL & ND_INIT) != 0) ? 0. : Nd.acc[k * nnodes + n];
  }
  if constexpr ((FL & ND_VEL) != 0) {
    double mom[D];
#pragma unroll
    for (int k = 0; k < D; ++k) mom[k] = Nd.mom[k * nnodes + n];
    node_compute_velocity<D>(mass, mom, v, a, cc);
  }
  if constexpr ((FL & ND_ACC) != 0) {
    double fe[D], fi[D];
#pragma unroll
    for (int k = 0; k < D; ++k) {
      fe[k] = Nd.fext[k * nnodes + n];
      fi[k] = Nd.fint[k * nnodes + n];
    }
    add_concentrated_force<D>(cc, Nd.fnval, fe);
    node_acc_vel<D>(mass, fe, fi, v, a, cc, A.dt, A.damping, Nd.fnval);
    if (A.ad_nodes) {
      double mom[D];
#pragma unroll
      for (int k = 0; k < D; ++k) mom[k] = Nd.mom[k * nnodes + n];
      ad_record_first<D>(A, n, nnodes, mass, mom, fe, fi, v, a);
    }
  }
  if constexpr ((FL & (ND_VEL | ND_ACC | ND_INIT)) == ND_VEL) {
    if (A.ad_nodes) {
      double mom[D];
#pragma unroll
      for (int k = 0; k < D; ++k) mom[k] = Nd.mom[k * nnodes + n];
      ad_record_second<D>(A, n, nnodes, mom, v);
    }
  }
#pragma unroll
  for (int k = 0; k < D; ++k) {
    Nd.vel[k * nnodes + n] = v[k];
    Nd.acc[k * nnodes + n] = a[k];
  }
}

// Row compaction of a component-major array: out[k * nnew + i] =
// in[k * nold + idx[i]] for every component k (removal of particles).
template <class T>
__global__ __launch_bounds__(kBlock) void k_take_rows(const T* in, T* out, const int* idx,
                                                      int64_t nnew, int64_t nold, int comps) {
  const int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x;
  if (i >= nnew) return;
  const int64_t j = idx[i];
  for (int k = 0; k < comps; ++k) out[k * nnew + i] = in[k * nold + j];
}

// ------------------------------- launchers ----------------------------------
static inline dim3 grid_for(int64_t n) { return dim3((unsigned)((n + kBlock - 1) / kBlock)); }

// Scatter kernels: one workgroup of CH threads per chunk of CH particles.
#define MPMX_LAUNCH_CH(CH, kernel, n, stream, ...)                                    \
  do {                                                                                \
    if ((n) > 0)                                                                      \
      hipLaunchKernelGGL(kernel, dim3((unsigned)(((n) + (CH)-1) / (CH))), dim3(CH), 0, \
                         stream, __VA_ARGS__);                                        \
  } while (0)
// One 64-lane workgroup per item (the ordered-mode node gathers).
#define MPMX_LAUNCH_WAVES(kernel, n, stream, ...)                                  \
  do {                                                                             \
    if ((n) > 0)                                                                   \
      hipLaunchKernelGGL(kernel, dim3((unsigned)(n)), dim3(64), 0, stream, __VA_ARGS__); \
  } while (0)
#define MPMX_CHUNK_DISPATCH(chunk, LAUNCH) \
  do {                                     \
    if ((chunk) == 128) {                  \
      LAUNCH(128);                         \
    } else if ((chunk) == 512) {           \
      LAUNCH(512);                         \
    } else {                               \
      LAUNCH(256);                         \
    }                                      \
  } while (0)

#define MPMX_LAUNCH(kernel, n, stream, ...)                                       \
  do {                                                                            \
    if ((n) > 0) hipLaunchKernelGGL(kernel, grid_for(n), dim3(kBlock), 0, stream, \
                                    __VA_ARGS__);                                 \
  } while (0)

template <int D>
void launch_locate_init(const Mesh& M, const Particles& P, Status* S, int* scan_list,
                        const int* hint, hipStream_t st) {
  MPMX_LAUNCH((k_locate_init<D>), P.np, st, M, P, S, scan_list, hint);
}

template <int D>
void launch_mapping_matrix(const Mesh& M, const Particles& P, const Nodes& Nd,
                           const StepArgs& A, int vu, const Status* S, hipStream_t st) {
  MPMX_LAUNCH((k_mapping_matrix<D>), P.np, st, M, P, Nd, A, vu, S);
}

template <int D>
void launch_affine_matrix(const Mesh& M, const Particles& P, int vu, const Status* S,
                          hipStream_t st) {
  MPMX_LAUNCH((k_affine_matrix<D>), P.np, st, M, P, vu, S);
}

template <int D>
void launch_locate_step(const Mesh& M, const Particles& P, Status* S, int* scan_list,
                        hipStream_t st) {
  MPMX_LAUNCH((k_locate_step<D>), P.np, st, M, P, S, scan_list);
}

template <int D>
void launch_locate_scan(const Mesh& M, const Particles& P, Status* S, const int* scan_list,
                        const int* hint, int remove_lost, hipStream_t st) {
  hipLaunchKernelGGL((k_locate_scan<D>), dim3(256), dim3(kBlock), 0, st, M, P, S, scan_list,
                     hint, remove_lost);
}

void launch_fn_eval(const MathFn* fns, const double* data, int nf, const Status* S, double dt,
                    double* out, hipStream_t st) {
  if (nf <= 0) return;
  hipLaunchKernelGGL(k_fn_eval, dim3(1), dim3(nf), 0, st, fns, data, nf, S, dt, out);
}

void launch_step_end(Status* S, hipStream_t st) {
  hipLaunchKernelGGL(k_step_end, dim3(1), dim3(64), 0, st, S);
}

template <int D>
void launch_p2g(int what, int chunk, const Mesh& M, const Particles& P, const Nodes& Nd,
                const StepArgs& A, const Status* S, const int* skey, const int* perm,
                hipStream_t st) {
  if (P.Cmat) {  // affine velocity updates: 128-particle chunks
    if (what == SC_MV)
      MPMX_LAUNCH_CH(128, (k_p2g<D, SC_MV | SC_AFF, 128>), P.np, st, M, P, Nd, A, S, skey, perm);
    else if (what == SC_MOM)
      MPMX_LAUNCH_CH(128, (k_p2g<D, SC_MOM | SC_AFF, 128>), P.np, st, M, P, Nd, A, S, skey, perm);
    else
      MPMX_LAUNCH_CH(128, (k_p2g<D, SC_MV | SC_F | SC_AFF, 128>), P.np, st, M, P, Nd, A, S, skey,
                     perm);
    return;
  }
  if (what == SC_MV) {
#define L_(CH) MPMX_LAUNCH_CH(CH, (k_p2g<D, SC_MV, CH>), P.np, st, M, P, Nd, A, S, skey, perm)
    MPMX_CHUNK_DISPATCH(chunk, L_);
#undef L_
  } else {
#define L_(CH) \
  MPMX_LAUNCH_CH(CH, (k_p2g<D, SC_MV | SC_F, CH>), P.np, st, M, P, Nd, A, S, skey, perm)
    MPMX_CHUNK_DISPATCH(chunk, L_);
#undef L_
  }
}

template <int D>
void launch_stress_scatter(int chunk, const Mesh& M, const Particles& P, const Nodes& Nd,
                           const Material* mats, const StepArgs& A, const Status* S,
                           const int* skey, const int* perm, hipStream_t st) {
#define L_(CH)                                                                                 \
  if (P.spin)                                                                                  \
    MPMX_LAUNCH_CH(CH, (k_stress_scatter<D, CH, true>), P.np, st, M, P, Nd, mats, A, S, skey, \
                   perm);                                                                      \
  else                                                                                         \
    MPMX_LAUNCH_CH(CH, (k_stress_scatter<D, CH, false>), P.np, st, M, P, Nd, mats, A, S, skey, \
                   perm)
  MPMX_CHUNK_DISPATCH(chunk, L_);
#undef L_
}

template <int D>
void launch_stress(bool locate, const Mesh& M, const Particles& P, const Nodes& Nd,
                   const Material* mats, const StepArgs& A, Status* S, int* scan_list,
                   hipStream_t st) {
  // the Jaumann-rate build only when such a material exists (P.spin set)
  if (locate && P.spin)
    MPMX_LAUNCH_CH(kStressBlock, (k_stress<D, true, true>), P.np, st, M, P, Nd, mats, A, S,
                   scan_list);
  else if (locate)
    MPMX_LAUNCH_CH(kStressBlock, (k_stress<D, true, false>), P.np, st, M, P, Nd, mats, A, S,
                   scan_list);
  else if (P.spin)
    MPMX_LAUNCH_CH(kStressBlock, (k_stress<D, false, true>), P.np, st, M, P, Nd, mats, A, S,
                   scan_list);
  else
    MPMX_LAUNCH_CH(kStressBlock, (k_stress<D, false, false>), P.np, st, M, P, Nd, mats, A, S,
                   scan_list);
}

template <int D>
void launch_traction_scatter(const Mesh& M, const Particles& P, const Nodes& Nd, const Status* S,
                             hipStream_t st) {
  MPMX_LAUNCH((k_traction_scatter<D>), P.ntr, st, M, P, Nd, S);
}

void launch_plastic(const Particles& P, const Material* mats, const Status* S, hipStream_t st) {
  // enough workgroups to fill the device twice over; each loops over the list
  const int64_t blocks = std::min<int64_t>((P.np + kPlasticBlock - 1) / kPlasticBlock, 8192);
  if (blocks < 1) return;
  hipLaunchKernelGGL(k_plastic, dim3((unsigned)blocks), dim3(kPlasticBlock), 0, st, P, mats, S);
}

template <int D>
void launch_g2p(bool stress, bool locate, const Mesh& M, const Particles& P, const Nodes& Nd,
                const Material* mats, const StepArgs& A, int vu, Status* S, int* scan_list,
                hipStream_t st) {
  if (stress && locate && P.spin)
    MPMX_LAUNCH((k_g2p<D, true, true, true>), P.np, st, M, P, Nd, mats, A, vu, S, scan_list);
  else if (stress && locate)
    MPMX_LAUNCH((k_g2p<D, true, true>), P.np, st, M, P, Nd, mats, A, vu, S, scan_list);
  else if (stress && P.spin)  // isoparametric meshes locate in a kernel of their own
    MPMX_LAUNCH((k_g2p<D, true, false, true>), P.np, st, M, P, Nd, mats, A, vu, S, scan_list);
  else if (stress)
    MPMX_LAUNCH((k_g2p<D, true, false>), P.np, st, M, P, Nd, mats, A, vu, S, scan_list);
  else if (locate)
    MPMX_LAUNCH((k_g2p<D, false, true>), P.np, st, M, P, Nd, mats, A, vu, S, scan_list);
  else
    MPMX_LAUNCH((k_g2p<D, false, false>), P.np, st, M, P, Nd, mats, A, vu, S, scan_list);
}

template <int D>
void launch_g2p_scatter(int chunk, const Mesh& M, const Particles& P, const Nodes& Nd,
                        const StepArgs& A, int vu, const Status* S, const int* skey,
                        const int* perm, hipStream_t st) {
#define L_(CH) MPMX_LAUNCH_CH(CH, (k_g2p_scatter<D, CH>), P.np, st, M, P, Nd, A, vu, S, skey, perm)
  MPMX_CHUNK_DISPATCH(chunk, L_);
#undef L_
}

template <int D>
void launch_node(int fl, const Mesh& M, const Nodes& Nd, const StepArgs& A, const Status* S,
                 hipStream_t st) {
  switch (fl) {
#define MPMX_NODE_CASE(F) \
  case (F): MPMX_LAUNCH((k_node<D, (F)>), M.nnodes, st, M, Nd, A, S); break;
    MPMX_NODE_CASE(ND_INIT | ND_VEL)
    MPMX_NODE_CASE(ND_ACC | ND_ZERO_ALL)
    MPMX_NODE_CASE(ND_INIT | ND_VEL | ND_ACC | ND_ZERO_ALL)
    MPMX_NODE_CASE(ND_INIT | ND_VEL | ND_ACC | ND_ZERO_MOM)
    MPMX_NODE_CASE(ND_VEL | ND_ZERO_ALL)
#undef MPMX_NODE_CASE
    default: break;
  }
}

template <int D>
void launch_cell_ranges(int64_t np, int64_t ncells, const int* skey, int* cstart, int* cend,
                        hipStream_t st) {
  if (ncells > 0) {
    hipLaunchKernelGGL(k_fill_int, grid_for(ncells), dim3(kBlock), 0, st, cstart, ncells, 0);
    hipLaunchKernelGGL(k_fill_int, grid_for(ncells), dim3(kBlock), 0, st, cend, ncells, 0);
  }
  MPMX_LAUNCH((k_cell_ranges<D>), np, st, np, ncells, skey, cstart, cend);
}

template <int D>
void launch_gather(int which, const Mesh& M, const Particles& P, const Nodes& Nd,
                   const StepArgs& A, const Status* S, const int* perm, const int* cstart,
                   const int* cend, hipStream_t st) {
  switch (which) {
    case GK_MV_VEL:  // USF first pass
      MPMX_LAUNCH_WAVES((k_gather<D, SC_MV, ND_INIT | ND_VEL>), M.nnodes, st, M, P, Nd, A, S,
                     perm, cstart, cend);
      break;
    case GK_F_ACC:  // USF force pass
      MPMX_LAUNCH((k_gather_prep<D>), P.np, st, M, P, S);
      MPMX_LAUNCH_WAVES((k_gather<D, SC_F, ND_ACC>), M.nnodes, st, M, P, Nd, A, S, perm, cstart,
                     cend);
      break;
    case GK_ALL_VEL_ACC:  // USL / MUSL first pass
      MPMX_LAUNCH((k_gather_prep<D>), P.np, st, M, P, S);
      MPMX_LAUNCH_WAVES((k_gather<D, SC_MV | SC_F, ND_INIT | ND_VEL | ND_ACC>), M.nnodes, st, M,
                     P, Nd, A, S, perm, cstart, cend);
      break;
    case GK_MOM_VEL:  // MUSL second momentum pass
      MPMX_LAUNCH_WAVES((k_gather<D, SC_MOM, ND_VEL>), M.nnodes, st, M, P, Nd, A, S, perm, cstart,
                     cend);
      break;
  }
}

template <class T>
void take_rows(const T* in, T* out, const int* idx, int64_t nnew, int64_t nold, int comps,
               hipStream_t st) {
  MPMX_LAUNCH((k_take_rows<T>), nnew, st, in, out, idx, nnew, nold, comps);
}
template void take_rows<double>(const double*, double*, const int*, int64_t, int64_t, int,
                                hipStream_t);
template void take_rows<int>(const int*, int*, const int*, int64_t, int64_t, int, hipStream_t);

#define MPMX_INSTANTIATE(D)                                                                   \
  template void launch_locate_init<D>(const Mesh&, const Particles&, Status*, int*,          \
                                      const int*, hipStream_t);                              \
  template void launch_locate_scan<D>(const Mesh&, const Particles&, Status*, const int*,    \
                                      const int*, int, hipStream_t);                         \
  template void launch_locate_step<D>(const Mesh&, const Particles&, Status*, int*,          \
                                      hipStream_t);                                          \
  template void launch_mapping_matrix<D>(const Mesh&, const Particles&, const Nodes&,        \
                                         const StepArgs&, int, const Status*, hipStream_t);  \
  template void launch_affine_matrix<D>(const Mesh&, const Particles&, int, const Status*,   \
                                        hipStream_t);                                        \
  template void launch_p2g<D>(int, int, const Mesh&, const Particles&, const Nodes&,         \
                              const StepArgs&, const Status*, const int*, const int*,        \
                              hipStream_t);                                                  \
  template void launch_stress_scatter<D>(int, const Mesh&, const Particles&, const Nodes&,   \
                                         const Material*, const StepArgs&, const Status*,    \
                                         const int*, const int*, hipStream_t);               \
  template void launch_stress<D>(bool, const Mesh&, const Particles&, const Nodes&,           \
                                 const Material*, const StepArgs&, Status*, int*,            \
                                 hipStream_t);                                               \
  template void launch_g2p<D>(bool, bool, const Mesh&, const Particles&, const Nodes&,       \
                              const Material*, const StepArgs&, int, Status*, int*,          \
                              hipStream_t);                                                  \
  template void launch_g2p_scatter<D>(int, const Mesh&, const Particles&, const Nodes&,      \
                                      const StepArgs&, int, const Status*, const int*,       \
                                      const int*, hipStream_t);                              \
  template void launch_node<D>(int, const Mesh&, const Nodes&, const StepArgs&,             \
                               const Status*, hipStream_t);                                  \
  template void launch_cell_ranges<D>(int64_t, int64_t, const int*, int*, int*, hipStream_t);\
  template void launch_gather<D>(int, const Mesh&, const Particles&, const Nodes&,           \
                                 const StepArgs&, const Status*, const int*, const int*,     \
                                 const int*, hipStream_t);                                   \
  template void launch_traction_scatter<D>(const Mesh&, const Particles&, const Nodes&,      \
                                           const Status*, hipStream_t);

MPMX_INSTANTIATE(2)
MPMX_INSTANTIATE(3)

}  // namespace mpmx
