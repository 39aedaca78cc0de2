// Launchers of the step kernels (defined in mpmx_kernels.hip and mpmx_sort.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "mpmx_internal.h"

namespace mpmx {

// node-kernel flags
enum { ND_VEL = 1, ND_ACC = 2, ND_ZERO_MOM = 4, ND_ZERO_ALL = 8, ND_INIT = 16 };
// scatter contents
enum { P2G_MV = 1, P2G_MV_F = 3, P2G_MOM = 4 };  // P2G_MOM: affine velocity updates only
// ordered-mode gathers
enum { GK_MV_VEL = 0, GK_F_ACC = 1, GK_ALL_VEL_ACC = 2, GK_MOM_VEL = 3 };

template <int D>
void launch_locate_init(const Mesh& M, const Particles& P, Status* S, int* scan_list,
                        const int* hint, hipStream_t st);
// Per-step locate of an isoparametric mesh (current cell, neighbours, else the
// scan queue); the step kernels locate Cartesian meshes themselves.
template <int D>
void launch_locate_step(const Mesh& M, const Particles& P, Status* S, int* scan_list,
                        hipStream_t st);
// Affine velocity updates (APIC, ASFLIP, TPIC): the mapping matrices (and
// ASFLIP's beta) before G2P; the matrix of the next particle-to-grid
// transfer before each P2G.
template <int D>
void launch_mapping_matrix(const Mesh& M, const Particles& P, const Nodes& Nd,
                           const StepArgs& A, int vu, const Status* S, hipStream_t st);
template <int D>
void launch_affine_matrix(const Mesh& M, const Particles& P, int vu, const Status* S,
                          hipStream_t st);
// remove_lost: a particle found in no cell is counted in Status::nlost for
// removal instead of stopping the run (Status::first_bad_step).
template <int D>
void launch_locate_scan(const Mesh& M, const Particles& P, Status* S, const int* scan_list,
                        const int* hint, int remove_lost, hipStream_t st);
void launch_step_end(Status* S, hipStream_t st);
// values of the nf (<= kMaxMathFns) math functions at the current step's time
constexpr int kMaxMathFns = 256;
void launch_fn_eval(const MathFn* fns, const double* data, int nf, const Status* S, double dt,
                    double* out, hipStream_t st);
// chunk: particles per workgroup of the LDS-staged scatters (128, 256, 512)
template <int D>
void launch_p2g(int what, int chunk, const Mesh& M, const Particles& P, const Nodes& Nd,
                const StepArgs& A, const Status* S, const int* skey, const int* perm,
                hipStream_t st);
template <int D>
void launch_stress_scatter(int chunk, const Mesh& M, const Particles& P, const Nodes& Nd,
                           const Material* mats, const StepArgs& A, const Status* S,
                           const int* skey, const int* perm, hipStream_t st);
template <int D>
void launch_stress(bool locate, const Mesh& M, const Particles& P, const Nodes& Nd,
                   const Material* mats, const StepArgs& A, Status* S, int* scan_list,
                   hipStream_t st);
// Particle tractions onto the nodes' external force (fast mode; P.ntr > 0).
template <int D>
void launch_traction_scatter(const Mesh& M, const Particles& P, const Nodes& Nd, const Status* S,
                             hipStream_t st);
// The deferred Mohr-Coulomb return mappings queued by launch_stress and
// launch_g2p (stress) this step; runs after them, before anything reads the
// stress or state variables.
void launch_plastic(const Particles& P, const Material* mats, const Status* S, hipStream_t st);
template <int D>
void launch_g2p(bool stress, bool locate, const Mesh& M, const Particles& P, const Nodes& Nd,
                const Material* mats, const StepArgs& A, int vu, Status* S, int* scan_list,
                hipStream_t st);
template <int D>
void launch_g2p_scatter(int chunk, const Mesh& M, const Particles& P, const Nodes& Nd,
                        const StepArgs& A, int vu, const Status* S, const int* skey,
                        const int* perm, hipStream_t st);
template <int D>
void launch_node(int fl, const Mesh& M, const Nodes& Nd, const StepArgs& A, const Status* S,
                 hipStream_t st);
template <int D>
void launch_cell_ranges(int64_t np, int64_t ncells, const int* skey, int* cstart, int* cend,
                        hipStream_t st);
template <int D>
void launch_gather(int which, const Mesh& M, const Particles& P, const Nodes& Nd,
                   const StepArgs& A, const Status* S, const int* perm, const int* cstart,
                   const int* cend, hipStream_t st);

// out[k * nnew + i] = in[k * nold + idx[i]] for the comps components of a
// component-major array (T = double or int).
template <class T>
void take_rows(const T* in, T* out, const int* idx, int64_t nnew, int64_t nold, int comps,
               hipStream_t st);

// Stable LSD radix sort of (cell key, particle id) pairs. keys_in: cell of
// each particle in particle order (values are the particle ids 0..n-1).
// Sorted keys/values end in keys_out/vals_out. Temporary buffers are sized by
// radix_sort_temp_ints(n).
int64_t radix_sort_temp_ints(int64_t n);
void radix_sort_by_cell(const int* cell, int64_t n, int max_key, int* keys_out, int* vals_out,
                        int* tmp, hipStream_t st);
// keys[i] = cell[perm[i]] for an earlier sort's permutation (cells changed
// since then leave the keys only nearly sorted).
void refresh_sorted_keys(const int* cell, const int* perm, int64_t n, int max_key, int* keys,
                         hipStream_t st);

}  // namespace mpmx
