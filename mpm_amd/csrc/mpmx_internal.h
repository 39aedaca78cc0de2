// Internal device-side data layout of the mpmx library (shared by the host
// driver and the HIP kernels). Everything is structure-of-arrays in HBM:
// component k of a per-particle vector lives at base + k * np, so a wave of
// 64 consecutive particles reads 512 contiguous bytes per component.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mpmx {

enum { LE = 0, MC = 1 };
enum { SV_YIELD = 0, SV_PHI, SV_PSI, SV_COH, SV_TCUT, SV_EPS, SV_RHO, SV_THETA, SV_PDS, NSV };
enum { USF = 0, USL = 1, MUSL = 2 };
// velocity updates, the order of mpm::VelocityUpdate (data_types.h:50)
enum { FLIP = 0, PIC = 1, ASFLIP = 2, TPIC = 3, APIC = 4 };

// Material constants, derived once on the host with the reference's
// expressions (linear_elastic.tcc:20-69, mohr_coulomb.tcc:9-77, 517-533).
struct Material {
  int type;
  int softening;
  int rate;  // objective stress rate: RATE_NONE, RATE_JAUMANN or RATE_GREEN_NAGHDI
  double density;
  double a1, a2, G;  // elastic tensor entries: D = [a1 a2 a2; a2 a1 a2; a2 a2 a1] (+) G*I3
  double shear;      // MohrCoulomb shear_modulus_ (pdstrain of separated state)
  double phi_peak, psi_peak, c_peak, phi_res, psi_res, c_res, pd_peak, pd_res;
  double critical_density;  // minimum_packing_fraction * grain_density
  double sv0[NSV];          // initial state variables
  double E;                 // Young's modulus (the elastic constants are proportional to it)
  double dtcut_dphi;        // d(initial tension cutoff)/d(friction angle): -c/sin^2(phi)
                            // when the cutoff is the apex c/tan(phi), else 0
  // sin, cos, tan of the peak friction angle and tan of the peak dilation
  // angle, by the same functions as the per-particle evaluation
  double sphi, cphi, tphi, tpsi;
  // MohrCoulomb without softening whose particles all hold the initial
  // friction, dilation, cohesion and tension cutoff (state variables 1-4,
  // which nothing changes then): the kernels take them from sv0
  int uniform;
};

enum { RATE_NONE = 0, RATE_JAUMANN = 1, RATE_GREEN_NAGHDI = 2 };

// Nodal constraint record, referenced by Nodes::cons[node] (-1: none).
struct NodeCons {
  int vmask;     // bit k: velocity constraint in direction k
  int amask;     // bit k: acceleration constraint in direction k
  int fdir;      // friction normal direction, -1: none
  double fsign;  // sign(sign_n): +1 if sign_n > 0 else -1 (node.tcc:514)
  double fmu;
  double vval[3], aval[3];
  int afn[3];    // math function scaling aval[k] (index into Nodes::fnval), -1: none
  // concentrated nodal force (Node::concentrated_force_, node.tcc:82-111):
  // added to the external force as cfs * cf, cfs = f(t) of math function
  // cfn (-1: 1.0); has_cf = 0: none
  int has_cf;
  int cfn;
  double cf[3];
  // general (inclined) boundary (Node::assign_rotation_matrix,
  // node.h:281-285): the constraints act in the frame of R (row-major) and
  // Ri is its inverse; rot = 0: Cartesian
  int rot;
  double R[3][3];
  double Ri[3][3];
};

// A math function of time-dependent boundary conditions: Linear (a table of
// n points (x, f(x)) at data[off, off + 2n): x then f) or Sin (data[off..off+3]
// = x0, a, xmin, xmax).
enum { FN_LINEAR = 0, FN_SIN = 1 };
struct MathFn {
  int type;
  int n;
  int off;
};

struct Mesh {
  int64_t nnodes, ncells;
  const int* cell_nodes;     // [ncells * NN]
  const double* node_x;      // SoA [dim][nnodes]
  const double* centroid;    // SoA [dim][ncells]
  const double* len;         // SoA [dim][ncells] side lengths used by the local coordinates
  const double* two_mean;    // [ncells] mean_length * 2.
  const double* dndx_c;      // [ncells][NN][dim] dN/dx at the cell centroid
  const int* nb_ptr;         // neighbour CSR (ascending ids)
  const int* nb_idx;
  const int* node_cells;     // [nnodes][NN] adjacent cells (-1 pad), ascending
  // Independent instances sharing one context (batched runs): cells
  // [inst_begin[i], inst_begin[i+1]) form instance i, cell_inst[c] its index.
  // Null for a single mesh.
  const int* cell_inst;
  const int64_t* inst_begin;
  // isoparametric mesh (the reference's "isoparametric" mesh key): local
  // coordinates by the inverse of the element map, else the Cartesian formula
  int iso;
};

struct Particles {
  int64_t np;
  double* x;  // [dim][np]
  double* u;
  double* v;
  double* stress;  // [6][np]
  double* strain;  // [6][np]
  double* F;       // [9][np] row-major
  double* mass;
  double* vol;
  double* rho;
  double* xi;      // [dim][np]
  double* sv;      // [NSV][np] or null when no MohrCoulomb material
  int* mat;
  int* cell;       // current cell (-1: none)
  // Particle velocity constraints (mesh.tcc:1454-1468): bit k of pvc_mask[p]
  // sets v_k = pvc_val[k][p] after the velocity update. Null when none.
  const unsigned char* pvc_mask;
  const double* pvc_val;
  // Deferred Mohr-Coulomb return mapping: the particles whose trial stress
  // yields this step (plastic_list[0, Status::nplastic) in shear,
  // plastic_list[np - Status::nplastic_t, np) in tension) and their records
  // (PR_* below), by list slot. Null when no MohrCoulomb material.
  // [3][np] entries 01, 02, 12 of W/2, W the skew part of the step's
  // ΔF - I, from the kinematics to a Jaumann-rate stress update of the same
  // thread (per-step scratch; 2D uses entry 01 only). Null without a
  // Jaumann-rate material.
  double* spin;
  // Green-Naghdi rate: [D*3 + D*D][np] the rows 0..D-1 of the step-start F
  // and the block of dF, stored by the kinematics for the constitutive
  // update of the same thread (per-step scratch; null without such a
  // material). Mohr-Coulomb with a stress rate: [6][np] the rotated trial
  // stress of each deferred return mapping, by list slot (null without).
  double* rotF;
  double* plastic_trial;
  // Ordered mode: each particle's shape function and internal-force term
  // for each node of its cell, [NN][1 + dim][np] (N_i, then
  // -(dN_i/dx . sigma)_k V), written before a force gather (null in fast
  // mode).
  double* gterm;
  int* plastic_list;
  double* plastic_rec;  // [PR_N][np] by list slot
  // Particle tractions (Particle::assign_traction / map_traction_force,
  // particle.tcc:962-990): the ntr particles tr_list[] carry a traction;
  // per direction k its value tr_val[k][p] times math function tr_fn[k][p]
  // (-1: 1.0; -2: no traction in that direction, 0), times the volume,
  // over the particle size psize[p]. Null / 0 when none.
  const int* tr_list;
  int ntr;
  const double* tr_val;
  const int* tr_fn;
  const double* psize;
  // Affine velocity updates (APIC, ASFLIP, TPIC): the mapping matrix B
  // (mapping_matrix_, [D*D][np], row-major per particle) kept between steps;
  // the matrix C of the next particle-to-grid transfer ([D*D][np], momentum
  // m N_i (v + C (x_i - x_p))); ASFLIP's beta of the step ([np]). Null for
  // FLIP and PIC.
  double* Bmat;
  double* Cmat;
  double* beta;
};

// Record of a deferred return mapping: the strain increment and what the
// correction needs from the trial state (its invariants with the Lode-angle
// trigonometry and its yield function value).
enum { PR_DS = 0, PR_EPS = 6, PR_RHO, PR_THETA, PR_ST, PR_CT, PR_S3, PR_C3, PR_YTRIAL, PR_N };

struct Nodes {
  double* mass;   // [nnodes]
  double* mom;    // [dim][nnodes]
  double* vel;
  double* acc;
  double* fext;
  double* fint;
  unsigned char* active;
  const int* cons;  // [nnodes] index into constraint table or -1
  const NodeCons* table;
  const double* fnval;  // the math functions' values at the current step's time
};

// Per-step scalars
struct StepArgs {
  double dt;
  double damping;
  double blending;
  double g[3];
  double ramp;  // gravity ramping time (<= 0: none)
  // Adjoint recording: when set, the node update kernels also store the
  // step's nodal quantities at ad_nodes + (*ad_slot) * ad_stride (layout:
  // mass [nn], then momentum, force, updated velocity, acceleration, MUSL
  // second momentum and its velocity, [dim][nn] each); *ad_slot is the
  // device-side step index within the recorded segment (Status::ad_slot).
  // Null in ordinary steps.
  double* ad_nodes;
  const int* ad_slot;
  int64_t ad_stride;
  int ad_nslots;  // slots at ad_nodes; an index outside [0, ad_nslots) records nothing
};

// Device-side status written by kernels and read after a run of steps.
struct Status {
  int64_t step;           // step counter (drives the gravity ramp)
  int64_t first_bad_step;  // first step with an unlocatable particle (INT64_MAX: none)
  int bad_particle;        // one such particle (pid)
  int nscan;               // particles needing the full-mesh scan this step
  int nplastic;            // shear entries of the deferred plastic list (its front)
  int nplastic_t;          // tensile entries (its back)
  int ad_slot;             // step index within an adjoint segment (recording)
  int nlost;               // particles removed because they left the mesh (remove mode)
};

}  // namespace mpmx
