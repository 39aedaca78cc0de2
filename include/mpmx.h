/*
 * mpmx — C-ABI of the MI355X explicit MPM step.
 *
 * The reference (geomechanics/mpm) has no FFI: its explicit step is reached
 * through factory keys and virtual C++ interfaces. This header is the
 * drop-in boundary that replaces, for one solver instance, the phase
 * interface mpm::MPMScheme<Tdim> (include/solvers/mpm_scheme/mpm_scheme.h:18-150)
 * as driven by mpm::MPMExplicit<Tdim>::solve()'s time loop
 * (include/solvers/mpm_explicit.tcc:112-190), together with the state that
 * loop advances (particles, nodes, cells of mpm::Mesh<Tdim>).
 *
 * Conventions (mirroring the reference):
 *  - One opaque context per solver instance; calls on one context are not
 *    thread-safe (the reference's caller is single-threaded).
 *  - Every call returns MPMX_OK (0) or a negative error class; the message is
 *    available from mpmx_last_error(). No C++ exception crosses the ABI.
 *  - Host arrays are caller-owned and only read/written during the call.
 *    Device buffers are library-owned. Each context owns one HIP stream.
 *  - Particle arrays are in the reference's particle order (creation order,
 *    mpm::Mesh::particles_ vector, include/mesh/mesh.tcc:1156-1161); the
 *    library keeps that order at its boundary whatever order it uses inside.
 *  - Floating point is IEEE FP64 throughout.
 */
#ifndef MPMX_H_
#define MPMX_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mpmx_ctx mpmx_ctx;

/* status codes */
#define MPMX_OK 0
#define MPMX_E_ARG (-1)         /* invalid argument or size */
#define MPMX_E_STATE (-2)       /* call out of order (e.g. step before mesh) */
#define MPMX_E_HIP (-3)         /* HIP runtime error */
#define MPMX_E_UNLOCATABLE (-4) /* particle left the mesh (mpm_scheme.tcc:250-256) */
#define MPMX_E_UNSUPPORTED (-5) /* option not available in this build */

/* stress update schemes (mpm_explicit.tcc:8-13) */
#define MPMX_USF 0
#define MPMX_USL 1
#define MPMX_MUSL 2

/* velocity update (mpm::VelocityUpdate, data_types.h:50; particle.tcc:994-1140;
 * the affine ones map momentum with map_mass_momentum_to_nodes_affine/_taylor,
 * particle.tcc:661-724) */
#define MPMX_FLIP 0
#define MPMX_PIC 1
#define MPMX_ASFLIP 2
#define MPMX_TPIC 3
#define MPMX_APIC 4

/* creation flags */
#define MPMX_ORDERED 0x1u /* nodal sums in particle order: bitwise equal to a
                             single-threaded reference run (slower) */
#define MPMX_NO_GRAPH 0x2u /* launch kernels directly instead of replaying a
                              captured hipGraph per step */

/* Number of doubles in one material parameter vector (layout documented in
 * mpm_amd/materials.py; replaces Factory<Material<Tdim>, unsigned, const Json&>
 * construction from the JSON "materials" array, src/material.cc:31-53). */
#define MPMX_MATERIAL_PARAMS 17
/* MohrCoulomb state variables per particle (mohr_coulomb.tcc:85-111) */
#define MPMX_STATE_VARS 9

/* Create a solver for dim 2 (quad4, "ED2Q4") or 3 (hex8, "ED3H8").
 * Replaces Factory<mpm::MPM>::create("MPMExplicit2D"/"MPMExplicit3D")
 * (src/mpm.cc:26-32) plus the scheme/velocity-update choice read by
 * MPMBase's constructor (mpm_base.tcc:57-100). */
int mpmx_create(mpmx_ctx** ctx, int dim, int scheme, int velocity_update,
                double blending_ratio, unsigned flags);
int mpmx_destroy(mpmx_ctx* ctx);
const char* mpmx_last_error(const mpmx_ctx* ctx);

/* Mesh: node coordinates [nnodes*dim] and cell connectivity [ncells*2^dim]
 * in the element's local node order. Replaces Mesh::create_nodes /
 * create_cells / find_cell_neighbours (mesh.tcc:17-49, 211-257, 287-307).
 * isoparametric = 0: local coordinates by the Cartesian formula
 * (Cell::local_coordinates_point, cell.tcc:423-606); 1: by the inverse of the
 * element map (Cell::transform_real_to_unit_cell, cell.tcc:617-828), the
 * reference's default when the mesh's "isoparametric" key is absent
 * (mpm_base.tcc:768-783). */
int mpmx_set_mesh(mpmx_ctx* ctx, int64_t nnodes, const double* coords,
                  int64_t ncells, const int64_t* cell_nodes, int isoparametric);

/* Batched runs: the mesh given to mpmx_set_mesh is the union of n independent
 * instances; instance i owns cells [cell_begin[i], cell_begin[i+1])
 * (cell_begin[0] = 0, cell_begin[n] = ncells) and no node is shared between
 * instances (checked). One mpmx_step then advances every instance, each as
 * the reference's MPMExplicit<Tdim>::solve would advance it alone
 * (mpm_explicit.tcc:112-190): nodal sums never mix instances, and the
 * full-scan fallback of locate (mesh.tcc:1142-1151) searches only the
 * particle's own instance (that of its previous cell, or of its cell hint in
 * mpmx_set_particles). Instances share dt, gravity and scheme; materials,
 * constraints and particles are per instance through the usual calls with
 * union-mesh ids. No counterpart in the reference, which runs one mesh per
 * analysis. Call after mpmx_set_mesh and before adding particles. */
int mpmx_set_instances(mpmx_ctx* ctx, int64_t n, const int64_t* cell_begin);

/* Materials (params: n * MPMX_MATERIAL_PARAMS). Material id = index.
 * For LinearElastic, params[4] is the objective stress rate
 * (linear_elastic.tcc:8-17, 96-124): MPMX_RATE_NONE or MPMX_RATE_JAUMANN;
 * MPMX_RATE_GREEN_NAGHDI returns MPMX_E_UNSUPPORTED. */
#define MPMX_RATE_NONE 0
#define MPMX_RATE_JAUMANN 1
#define MPMX_RATE_GREEN_NAGHDI 2
int mpmx_set_materials(mpmx_ctx* ctx, int n, const double* params);

/* Nodal constraints, Cartesian form. Velocity/acceleration: per (node, dir)
 * the first value given wins, as std::map::insert does in
 * Node::assign_velocity_constraint (node.tcc:348-366, 407-437).
 * Friction: one (dir, sign_n, mu) per node, last one wins
 * (Node::assign_friction_constraint, node.tcc:488-507). */
int mpmx_set_velocity_constraints(mpmx_ctx* ctx, int64_t n, const int64_t* node,
                                  const int* dir, const double* value);
int mpmx_set_acceleration_constraints(mpmx_ctx* ctx, int64_t n, const int64_t* node,
                                      const int* dir, const double* value);
int mpmx_set_friction_constraints(mpmx_ctx* ctx, int64_t n, const int64_t* node,
                                  const int* dir, const int* sign_n, const double* mu);

/* General (inclined) boundaries: the velocity, acceleration and friction
 * constraints of node[i] act in the frame rotated by the Euler angles
 * angles[i*dim ...] (radians; alpha, beta in 2D, alpha, beta, gamma in 3D),
 * directions given in that frame. Replaces the "nodal_euler_angles" file
 * (MPMBase::node_euler_angles, mpm_base.tcc:1041-1063) and
 * Mesh::compute_nodal_rotation_matrices (mesh.tcc:1320-1340) with
 * mpm::geometry::rotation_matrix (src/geometry.cc:6-56); applied by
 * node.tcc:385-401, 469-482, 540-583, 602-655. */
int mpmx_set_nodal_euler_angles(mpmx_ctx* ctx, int64_t n, const int64_t* node,
                                const double* angles);

/* Math functions of time-dependent boundary conditions, indexed 0..nf-1
 * (the reference's "math_functions", MPMBase::initialise_math_functions
 * mpm_base.tcc:958-1010, functions stored by id). type[f] MPMX_FN_LINEAR:
 * npts[f] points, params x_0..x_{n-1} then f(x_0)..f(x_{n-1})
 * (LinearFunction, src/functions/linear_function.cc:3-41: f of the first
 * point before the table, of the last after it, linear in between);
 * MPMX_FN_SIN: params x0, a, xmin, xmax (SinFunction,
 * src/functions/sin_function.cc:3-20: sin(a (x - x0)) on [xmin, xmax], 0
 * elsewhere). params holds the functions one after another. Replaces any
 * earlier set (at most 256 functions). */
#define MPMX_FN_LINEAR 0
#define MPMX_FN_SIN 1
int mpmx_set_math_functions(mpmx_ctx* ctx, int nf, const int* type, const int* npts,
                            const double* params);

/* Acceleration constraints with a math function each (fn[i]: index into the
 * math functions, -1 none): in the step at time t = step * dt the
 * constraint's acceleration is value[i] * f(t) (AccelerationConstraint::
 * acceleration, acceleration_constraint.h:37-43, updated by
 * Mesh::update_nodal_acceleration_constraints, mesh.tcc:1414-1428, at
 * mpm_scheme.tcc:219). mpmx_set_acceleration_constraints is this call with
 * no functions. */
int mpmx_set_acceleration_constraints_fn(mpmx_ctx* ctx, int64_t n, const int64_t* node,
                                         const int* dir, const double* value, const int* fn);

/* Concentrated nodal forces (the reference's concentrated_nodal_forces,
 * MPMBase::initialise_loads mpm_base.tcc:913-953, Mesh::
 * assign_nodal_concentrated_forces mesh.tcc:1471-1497, 2155-2181): force[i]
 * in direction dir[i] of node node[i] (-1: every node), scaled at time
 * t = step * dt by math function fn[i] (-1 or a null array: none). As in
 * Node::assign_concentrated_force (node.tcc:82-101) a later assignment
 * replaces the node's force in that direction and the node's function for
 * all directions. Added to the nodes' external force after the particles'
 * body forces (Node::apply_concentrated_force, node.tcc:103-111;
 * mpm_scheme.tcc:156-160). */
int mpmx_set_nodal_forces(mpmx_ctx* ctx, int64_t n, const int64_t* node, const int* dir,
                          const double* force, const int* fn);

/* Particle surface tractions (the reference's particle_surface_traction,
 * MPMBase::initialise_loads mpm_base.tcc:882-911, Mesh::
 * create_particles_tractions mesh.tcc:1349-1367): load i sets, at time
 * t = step * dt, traction value[i] * f(t) (math function fn[i]; -1 or a null
 * array: none) in direction dir[i] of particle particle[i] (-1: every
 * particle), the loads applied in order so a later one wins per particle
 * and direction (Mesh::apply_traction_on_particles, mesh.tcc:1371-1386).
 * The particle's traction vector is that value times its volume over its
 * size pow(V0, 1/dim), V0 its volume when it was created
 * (Particle::assign_traction, particle.tcc:962-979; assign_volume
 * :556-568); it is mapped to the nodes' external force as N_i * traction
 * after the body forces (Particle::map_traction_force, particle.tcc:
 * 983-990). Replaces any earlier list; n = 0 removes them. */
int mpmx_set_particle_tractions(mpmx_ctx* ctx, int64_t n, const int64_t* particle,
                                const int* dir, const double* value, const int* fn);

/* Particle velocity constraints: after the velocity update of every step,
 * velocity component dir[i] of particle particle[i] (-1: every particle)
 * is set to value[i], constraints applied in the order given (the reference's
 * particles_velocity_constraints with particle sets expanded to ids,
 * MPMBase::particle_velocity_constraints mpm_base.tcc:1549-1585,
 * Mesh::apply_particle_velocity_constraints mesh.tcc:1454-1468,
 * Particle::apply_particle_velocity_constraints particle.tcc:1190-1195).
 * Replaces any earlier set; n = 0 removes them. Particle ids in reference
 * order; particles added later are unconstrained unless the set uses -1. */
int mpmx_set_particle_velocity_constraints(mpmx_ctx* ctx, int64_t n, const int64_t* particle,
                                           const int* dir, const double* value);

/* Gravity and its ramp (MPMBase::initialise_loads, mpm_base.tcc:788-830;
 * multiplier min(1, step*dt/T), mpm_explicit.tcc:143-153). */
int mpmx_set_gravity(mpmx_ctx* ctx, const double* g, double ramping_time);

/* Particles at the Gauss points of cells ("gauss" generator,
 * Mesh::generate_material_points + Cell::generate_points, mesh.tcc:508-576,
 * cell.tcc:91-143), then located, volume = cell volume / particles in the
 * cell and mass = volume * density (particle.tcc:589-626).
 * nset < 0 means all cells. nparticles_per_dir: 1 to 4, the reference's
 * quadratures QQUAD1-4 / QHEX1-4 in their point order
 * (quadrilateral_quadrature.tcc, hexahedron_quadrature.tcc). */
int mpmx_generate_gauss(mpmx_ctx* ctx, int nparticles_per_dir, int64_t nset,
                        const int64_t* cell_set, int material_id);

/* Append particles from full state, as the HDF5 resume path does
 * (Particle::initialise_particle(POD), particle.tcc:35-177). Arrays of n
 * rows; x,u,v: dim per row; stress, strain: 6 (Voigt xx,yy,zz,xy,yz,xz);
 * F: 9 (row-major); statevars: MPMX_STATE_VARS (MohrCoulomb only).
 * u, v, stress, strain, F, cell, statevars may be NULL (zero / identity /
 * locate / material initial state). cell[i] >= 0 is a hint that is kept when
 * the point lies in that cell (locate_particle_cells, mesh.tcc:1116-1153). */
int mpmx_set_particles(mpmx_ctx* ctx, int64_t n, const double* x, const double* u,
                       const double* v, const double* stress, const double* strain,
                       const double* F, const double* mass, const double* volume,
                       const double* density, const int* material_id,
                       const int64_t* cell, const double* statevars);

/* Every particle's volume = volume of its cell / number of particles in the
 * cell, then mass = volume * density of its material (the passes
 * MPMBase::initialise_particles runs over all particles after generation and
 * locate, mpm_base.tcc:413-415 with Particle::compute_volume,
 * particle.tcc:589-595, and MPMExplicit::solve's compute_mass,
 * mpm_explicit.tcc:93-95 with particle.tcc:618-626). */
int mpmx_compute_volume_mass(mpmx_ctx* ctx);

/* Assign every particle's velocity (n must equal the particle count; rows of
 * dim, reference particle order). Replaces Particle::assign_velocity
 * (particle.tcc:771-777) as applied to all particles, e.g. for initial
 * velocities that the reference receives through its HDF5 resume path. */
int mpmx_assign_velocity(mpmx_ctx* ctx, int64_t n, const double* v);

/* Material of the listed particles (positions in the particle order) becomes
 * material_id, their state variables that material's initial ones
 * (Particle::assign_material, particle.tcc:499-517), as the deck's
 * "material_sets" does for a particle set (mpm_base.tcc:449-468). Call before
 * mpmx_compute_volume_mass, which takes the density of each particle's
 * material. */
int mpmx_assign_materials(mpmx_ctx* ctx, int64_t n, const int64_t* particle, int material_id);

/* Assign every particle's stress (n rows of 6 Voigt components, reference
 * particle order). Replaces Mesh::assign_particles_stresses
 * (mesh.tcc:1502-1523), the "particles_stresses" input (mpm_base.tcc:1595-1632). */
int mpmx_assign_stresses(mpmx_ctx* ctx, int64_t n, const double* stress);

int mpmx_nparticles(const mpmx_ctx* ctx, int64_t* n);

/* Copy particle state back in reference particle order; NULL skips a field.
 * xi: natural coordinates in the current cell (dim per row). */
int mpmx_get_particles(mpmx_ctx* ctx, double* x, double* u, double* v, double* stress,
                       double* strain, double* F, double* mass, double* volume,
                       double* density, int* material_id, int64_t* cell, double* xi,
                       double* statevars);

/* Nodal state after the last step (for checks): mass, velocity [n*dim],
 * acceleration [n*dim], active flag. NULL skips a field. */
int mpmx_get_nodes(mpmx_ctx* ctx, double* mass, double* velocity, double* acceleration,
                   int* active);

/* Mapping matrices of the affine velocity updates (Particle::mapping_matrix_,
 * particle.h; the PODParticle mapping_matrix_ij fields, pod_particle.h):
 * [n][dim][dim] row-major in particle order; zero at creation, set by every
 * step with MPMX_APIC, MPMX_ASFLIP or MPMX_TPIC. Set on resume. */
int mpmx_get_mapping_matrix(mpmx_ctx* ctx, double* B);
int mpmx_set_mapping_matrix(mpmx_ctx* ctx, int64_t n, const double* B);

/* Advance nsteps explicit steps (one iteration of the loop in
 * mpm_explicit.tcc:112-190 each: MPMScheme::initialise,
 * compute_nodal_kinematics, pre/postcompute_stress_strain, compute_forces,
 * compute_particle_kinematics with Cundall damping, postcompute_nodal_kinematics,
 * locate_particles). Returns MPMX_E_UNLOCATABLE if a particle left the mesh;
 * mpmx_last_error() then names it and the step. */
int mpmx_step(mpmx_ctx* ctx, int nsteps, double dt, double damping_factor);

/* mpmx_step without any device allocation. No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same steps and results as mpmx_step; the kernels launch directly (no
 * graph is instantiated) and take their scratch (sort keys and permutation,
 * radix-sort temporaries, locate queue, deferred Mohr-Coulomb list and
 * records, Jaumann spin, math-function values, ordered-mode cell ranges
 * and gather terms) from one 64 MiB file-scope device array per GPU, held
 * until the process ends. A context whose step needs more (144 bytes per
 * particle with a Mohr-Coulomb material, 32 without, plus 24 with the
 * Jaumann rate, plus 256 in 3D and 96 in 2D in ordered mode: about 460k
 * Mohr-Coulomb particles in fast mode) gets MPMX_E_ARG; particle removal
 * (mpmx_set_locate_particles(ctx, 0)) gets MPMX_E_UNSUPPORTED. Every call on
 * one GPU shares that array and the first step of each call sorts fully:
 * calls that overlap on one GPU (contexts used from several threads) race. */
int mpmx_step_noalloc(mpmx_ctx* ctx, int nsteps, double dt, double damping_factor);

/* Fast mode: a full radix sort of the particles by cell every `steps` steps
 * (default 8; 1 = every step). In between, the last sort's order is reused
 * with the current cells; particles that changed cell add scatter atomics
 * but not error. Ordered mode sorts every step regardless. No counterpart in
 * the reference, whose scatter works in particle order (mpm_scheme.tcc:44-71). */
int mpmx_set_sort_interval(mpmx_ctx* ctx, int steps);

/* locate_particles (the analysis key read at mpm_base.tcc:53-54, default
 * true): with 1, a particle that no cell contains after a step stops the run
 * at that step with MPMX_E_UNLOCATABLE (MPMScheme::locate_particles throws,
 * mpm_scheme.tcc:250-256); the state returned is the state at the end of
 * that step, later queued steps do nothing. With 0 such particles are
 * removed (mpm_scheme.tcc:258-260 -> Mesh::remove_particle, mesh.tcc:658-665):
 * they take no part in later steps and are dropped from the particle arrays,
 * the others keeping their order; particle-indexed loads follow their
 * particles. */
int mpmx_set_locate_particles(mpmx_ctx* ctx, int locate_particles);

/* The reference's particle ids (ParticleBase::id()) in particle order: a
 * particle created by mpmx_generate_gauss / mpmx_set_particles gets the
 * number of particles before it (pid = particles_.size(), mesh.tcc:531-536);
 * mpmx_set_particle_ids replaces them (n = particle count), e.g. with the ids
 * of a resumed checkpoint (particle.tcc:35-40). */
int mpmx_get_particle_ids(mpmx_ctx* ctx, int64_t* ids);
int mpmx_set_particle_ids(mpmx_ctx* ctx, int64_t n, const int64_t* ids);

/* Fast mode: particles per workgroup of the LDS-staged scatter kernels (the
 * chunk of cell-sorted particles whose per-node sums are formed in LDS
 * before the FP64 atomics): 128, 256 (default) or 512. No counterpart in the
 * reference. */
int mpmx_set_chunk(mpmx_ctx* ctx, int particles_per_workgroup);

/* Fast mode: time trial_steps steps (after one warm-up step) for every pair
 * of scatter chunk {128, 256, 512} and full-sort interval {4, 8, 16} from the
 * current state, keep the fastest pair and restore the state (particles and
 * step counter are as before the call). Writes the chosen pair. Ordered mode
 * has nothing to tune and returns the current settings. No counterpart in
 * the reference. */
int mpmx_autotune(mpmx_ctx* ctx, double dt, double damping_factor, int trial_steps,
                  int* chunk, int* sort_every);

/* Step counter (drives the gravity ramp); set on resume (mpm_base.tcc:509-551). */
int mpmx_set_step(mpmx_ctx* ctx, int64_t step);
int mpmx_get_step(const mpmx_ctx* ctx, int64_t* step);

/* Block until all work queued on the context's stream has finished. */
int mpmx_synchronize(mpmx_ctx* ctx);

/* Per-kernel timing with HIP events recorded on the context's stream
 * (no counterpart in the reference, which only logs the solver duration,
 * mpm_explicit.tcc:191-196). While enabled, mpmx_step launches kernels
 * directly with an event pair around each launch. */
int mpmx_profile(mpmx_ctx* ctx, int enable);
/* Fills up to cap entries: names (NUL-terminated, <= 47 chars each, 48-byte
 * stride), total milliseconds and launch counts since profiling was enabled.
 * *n receives the number of kernels. */
int mpmx_kernel_times(mpmx_ctx* ctx, int cap, char* names, double* total_ms,
                      int64_t* launches, int* n);

/* ------------------------------ adjoint -------------------------------------
 * Reverse-mode derivatives of a run of explicit steps, for material-parameter
 * inversion. No counterpart in the reference, which has no differentiation.
 *
 * mpmx_adjoint_forward advances nsteps steps exactly as mpmx_step does and
 * records a checkpoint of the full particle state every checkpoint_every
 * steps (device memory: about (nsteps / checkpoint_every + 1) * np * 40
 * doubles). mpmx_adjoint_backward then takes the gradient of a loss L with
 * respect to the final positions and velocities (dL_dx, dL_dv: rows of dim,
 * NULL = zero) back through the recorded steps, recomputing each segment
 * from its checkpoint (memory: checkpoint_every + 1 step records of about
 * (2 dim + 9) * np values), and returns per particle:
 *   dL_dE   : Young's modulus of the particle's constitutive update;
 *   dL_drho : material density, through the mass m = V rho (m/rho fixed);
 *   dL_dphi : Mohr-Coulomb friction angle in degrees (0 for LinearElastic),
 *             including the tension cutoff where it is the apex c / tan(phi);
 *   dL_dx0, dL_dv0 : initial positions and velocities.
 * Any output may be NULL. The context is left at the state after the
 * recorded run. Derivatives follow the branches the run took (yield states,
 * return-mapping iterations, cell membership, values snapped to zero).
 * Supported: USF, USL, MUSL; FLIP, PIC; Cundall damping; velocity and
 * acceleration constraints; gravity ramp; LinearElastic, MohrCoulomb without
 * softening; axis-aligned rectangular/box cells. Friction constraints and
 * softening return MPMX_E_UNSUPPORTED. */
int mpmx_adjoint_forward(mpmx_ctx* ctx, int nsteps, double dt, double damping_factor,
                         int checkpoint_every);
int mpmx_adjoint_backward(mpmx_ctx* ctx, const double* dL_dx, const double* dL_dv,
                          double* dL_dE, double* dL_drho, double* dL_dphi, double* dL_dx0,
                          double* dL_dv0);
/* mpmx_adjoint_backward with every array in the memory of the context's
 * device (e.g. the storage of torch CUDA tensors), same layouts and results:
 * no host staging. Returns when the gradients are written. No counterpart
 * in the reference (which has no derivatives). */
int mpmx_adjoint_backward_device(mpmx_ctx* ctx, const double* dL_dx, const double* dL_dv,
                                 double* dL_dE, double* dL_drho, double* dL_dphi,
                                 double* dL_dx0, double* dL_dv0);
/* Particle positions and velocities (rows of dim, the reference's particle
 * order) copied into arrays in the memory of the context's device; either
 * may be null. The device-side counterpart of mpmx_get_particles' x and v. */
int mpmx_get_particles_device(mpmx_ctx* ctx, double* x, double* v);
/* Mohr-Coulomb returns near the apex of the yield surface (cohesionless
 * material at low stress, e.g. at a free surface): the directions of the
 * yield and plastic-potential gradients there depend on a Lode angle that is
 * ill-conditioned, the exact derivative of one update can exceed 1e3, and
 * over hundreds of steps the backward pass overflows. For returns whose
 * plastic correction exceeds `ratio` times the returned stress's deviatoric
 * norm sqrt(2 J2), the backward pass takes those directions as constants
 * (keeping the derivatives through the yield-function values, the plastic
 * multipliers and the elastic predictor). Default 0.3; ratio <= 0
 * differentiates every return exactly. */
int mpmx_adjoint_set_apex_ratio(mpmx_ctx* ctx, double ratio);
/* Free the adjoint's checkpoints and buffers. */
int mpmx_adjoint_release(mpmx_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif /* MPMX_H_ */
