/*
 * mpmx_pod — the reference's particle checkpoint file: an HDF5 table named
 * "table" of one record per particle with the 74 fields of mpm::PODParticle
 * (include/particles/pod_particles/pod_particle.h:12-62, field names and
 * HDF5 types in src/pod_particle.cc), written by Mesh::write_particles_hdf5
 * (include/mesh/mesh.tcc:1573-1611) at the analysis' output_steps and read
 * back by Mesh::read_particles_hdf5 (mesh.tcc:1659-1721) on a checkpoint
 * resume (mpm_base.tcc:509-551). Files written here are read by the
 * reference and the other way round.
 *
 * Host-side (no GPU): the columns are host arrays, usually filled from
 * mpmx_get_particles / mpmx_get_particle_ids. Vector fields are rows of 3
 * (x, y, z; a 2D particle writes 0 in z), stress and strain rows of 6
 * (xx, yy, zz, xy, yz, xz), the deformation gradient and the mapping matrix
 * rows of 9 (row-major), the state variables rows of 20 (the material's
 * state_variables() order; MohrCoulomb: yield_state, phi, psi, cohesion,
 * tension_cutoff, epsilon, rho, theta, pdstrain). Booleans are one byte.
 * Every call returns 0 or -1 with the message in mpmx_pod_last_error().
 */
#ifndef MPMX_POD_H_
#define MPMX_POD_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPMX_POD_NFIELDS 74
#define MPMX_POD_NSVARS 20

typedef struct mpmx_pod_columns {
  int64_t n;                  /* records */
  int64_t* id;                /* Particle::id() */
  double* mass;
  double* volume;
  double* pressure;           /* the "pressure" state variable, 0 without one */
  double* coord;              /* n x 3 */
  double* displacement;       /* n x 3 */
  double* nsize;              /* n x 3: natural particle size */
  double* velocity;           /* n x 3 */
  double* acceleration;       /* n x 3 */
  double* stress;             /* n x 6 */
  double* strain;             /* n x 6 */
  double* defgrad;            /* n x 9 */
  unsigned char* initialise_mapping_matrix;
  double* mapping_matrix;     /* n x 9 */
  unsigned char* status;
  int64_t* cell_id;           /* the particle's cell (max uint64 as -1: none) */
  uint32_t* material_id;
  uint32_t* nstate_vars;
  double* svars;              /* n x 20 */
} mpmx_pod_columns;

/* Write the table (replaces the file), as mesh.tcc:1573-1611: H5TBmake_table
 * "Table Title" / "table", chunk 10000, no compression. */
int mpmx_pod_write(const char* filename, const mpmx_pod_columns* cols);

/* Number of records of the table, after checking that it has the 74 fields
 * of PODParticle (mesh.tcc:1673-1677). */
int mpmx_pod_count(const char* filename, int64_t* n);

/* Read the whole table into columns sized by cols->n (mesh.tcc:1679-1683). */
int mpmx_pod_read(const char* filename, mpmx_pod_columns* cols);

/* The table's field names, NUL-separated in field order, into buf (cap
 * bytes); *nfields receives their number. */
int mpmx_pod_field_names(const char* filename, char* buf, int64_t cap, int64_t* nfields);

const char* mpmx_pod_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* MPMX_POD_H_ */
