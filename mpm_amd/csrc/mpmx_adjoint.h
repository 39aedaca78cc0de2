// Device-side data of the adjoint (reverse-mode) step and its launchers
// (mpmx_adjoint.hip). Layouts are SoA with the particle or node count as
// stride, like the forward state.
#pragma once

#include <hip/hip_runtime.h>

#include "mpmx_internal.h"

namespace mpmx {

// The recorded state of one step that the backward pass reads.
struct Slot {
  const double* v;       // [dim][np]
  const double* stress;  // [6][np]
  const double* vol;     // [np]
  const double* rho;     // [np]
  const double* xi;      // [dim][np]
  const double* yield;   // [np] Mohr-Coulomb yield state (null: no Mohr-Coulomb)
  const int* cell;       // [np]
};
struct SlotW {
  double* v;
  double* stress;
  double* vol;
  double* rho;
  double* xi;
  double* yield;
  int* cell;
};

// Node quantities of the step being transposed (recorded by the replayed
// forward step; v computed from mass and momentum) and their adjoints.
struct AdjNodes {
  double* m;    // mass [nn]
  double* P;    // momentum [dim][nn]
  double* f;    // external + internal force
  double* v;    // velocity from the momentum (before the update)
  double* a;    // acceleration
  double* vp;   // updated velocity
  double* P2;   // MUSL: momentum of the updated particle velocities
  double* vpp;  // MUSL: velocity from P2
  double* lvp;  // adjoints of the above
  double* la;
  double* lv;
  double* lvpp;
  double* lf;
  double* lm;
  double* lP;
  double* lm2;
  double* lP2;
};

// Per-particle adjoints of the state (x, v, stress, volume), the shape
// function adjoint of the current step (lxq), and the accumulated parameter
// gradients (Young's modulus, friction angle in rad, mass).
struct AdjParticles {
  double* lx;    // [dim][np]
  double* lv;    // [dim][np]
  double* ls;    // [6][np]
  double* lvol;  // [np]
  double* lxq;   // [dim][np]
  double* gE;
  double* gphi;
  double* gm;
};

// Plastic Mohr-Coulomb updates of a backward step, whose constitutive
// adjoint runs in a separate pass over this compacted list: shear returns
// from the front (count[0] entries), tensile ones from the back (count[1]
// entries, slot cap - 1 down); per entry the particle id, the apex flag and
// (SoA, stride cap) the strain increment, replaced by its adjoint.
struct AdjList {
  int* count;
  int* ids;
  int* apex;
  double* buf;  // [6][cap]
  int64_t cap;
};

struct AdjArgs {
  double dt, damping, blending;
  double g[3];        // gravity of this step (ramp applied)
  double apex_ratio;  // Mohr-Coulomb returns near the apex: see mpmx_adjoint_set_apex_ratio
  int vu;
};

template <int D>
void launch_ad_save(const Particles& P, const SlotW& s, hipStream_t st);
void launch_ad_fill(double* a, int64_t n, double v, hipStream_t st);
// Row-major n x comps <-> component-major [comps][n] copies on the device;
// to rows, each value times `scale` and, with per_mass, times the
// particle's mass over its material's density (the density gradient
// through m = V rho).
void launch_rows_to_soa(const double* rows, double* soa, int64_t n, int comps, hipStream_t st);
void launch_soa_to_rows(const double* soa, double* rows, int64_t n, int comps, double scale,
                        const double* mass, const int* mat, const Material* mats,
                        hipStream_t st);
// One backward step: transpose step t from the recorded particle states s0
// (step t) and s1 (step t+1) and the step's recorded nodal quantities.
template <int D>
void launch_ad_backward_step(int scheme, const Mesh& M, const Particles& P, const Nodes& Nd,
                             const Material* mats, const Slot& s0, const Slot& s1,
                             const AdjNodes& A, const AdjParticles& L, const AdjList& Q,
                             const AdjArgs& a, hipStream_t st);

}  // namespace mpmx
