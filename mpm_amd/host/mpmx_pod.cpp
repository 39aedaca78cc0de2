// The reference's particle checkpoint table (PODParticle records in an HDF5
// table, include/particles/pod_particles/pod_particle.h:12-62,
// src/pod_particle.cc) over the image's HDF5 high-level library.
//
// The record below keeps the reference's 74 fields in their order with the
// same names and HDF5 types, so a table written here reads back in the
// reference (mesh.tcc:1659-1721, which reads by field position) and the
// other way round; its in-memory layout is this file's own (the offsets are
// passed to H5TB explicitly).
#include <cstdio>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "hdf5.h"
#include "hdf5_hl.h"

#include "../../include/mpmx_pod.h"

namespace {

struct Record {
  long long id;
  double mass, volume, pressure;
  double coord[3], displacement[3], nsize[3], velocity[3], acceleration[3];
  double stress[6], strain[6], defgrad[9];
  hbool_t initialise_mapping_matrix;
  double mapping_matrix[9];
  hbool_t status;
  long long cell_id;
  unsigned material_id;
  unsigned nstate_vars;
  double svars[MPMX_POD_NSVARS];
};

struct Fields {
  std::vector<std::string> names;
  std::vector<size_t> offsets, sizes;
  std::vector<hid_t> types;

  void add(const std::string& name, size_t off, size_t size, hid_t type) {
    names.push_back(name);
    offsets.push_back(off);
    sizes.push_back(size);
    types.push_back(type);
  }
  void vec(const std::string& base, const char* const* suffix, int n, size_t off) {
    for (int k = 0; k < n; ++k)
      add(base + suffix[k], off + k * sizeof(double), sizeof(double), H5T_NATIVE_DOUBLE);
  }
};

const Fields& fields() {
  static Fields f = [] {
    Fields t;
    static const char* xyz[] = {"_x", "_y", "_z"};
    static const char* mat[] = {"_00", "_01", "_02", "_10", "_11", "_12", "_20", "_21", "_22"};
    const size_t d = sizeof(double);
    t.add("id", offsetof(Record, id), sizeof(long long), H5T_NATIVE_LLONG);
    t.add("mass", offsetof(Record, mass), d, H5T_NATIVE_DOUBLE);
    t.add("volume", offsetof(Record, volume), d, H5T_NATIVE_DOUBLE);
    t.add("pressure", offsetof(Record, pressure), d, H5T_NATIVE_DOUBLE);
    t.vec("coord", xyz, 3, offsetof(Record, coord));
    t.vec("displacement", xyz, 3, offsetof(Record, displacement));
    t.vec("nsize", xyz, 3, offsetof(Record, nsize));
    t.vec("velocity", xyz, 3, offsetof(Record, velocity));
    t.vec("acceleration", xyz, 3, offsetof(Record, acceleration));
    static const char* stress[] = {"stress_xx", "stress_yy", "stress_zz", "tau_xy", "tau_yz", "tau_xz"};
    static const char* strain[] = {"strain_xx", "strain_yy", "strain_zz",
                                   "gamma_xy",  "gamma_yz",  "gamma_xz"};
    for (int k = 0; k < 6; ++k)
      t.add(stress[k], offsetof(Record, stress) + k * d, d, H5T_NATIVE_DOUBLE);
    for (int k = 0; k < 6; ++k)
      t.add(strain[k], offsetof(Record, strain) + k * d, d, H5T_NATIVE_DOUBLE);
    t.vec("defgrad", mat, 9, offsetof(Record, defgrad));
    t.add("initialise_mapping_matrix", offsetof(Record, initialise_mapping_matrix),
          sizeof(hbool_t), H5T_NATIVE_HBOOL);
    t.vec("mapping_matrix", mat, 9, offsetof(Record, mapping_matrix));
    t.add("status", offsetof(Record, status), sizeof(hbool_t), H5T_NATIVE_HBOOL);
    t.add("cell_id", offsetof(Record, cell_id), sizeof(long long), H5T_NATIVE_LLONG);
    t.add("material_id", offsetof(Record, material_id), sizeof(unsigned), H5T_NATIVE_UINT);
    t.add("nstate_vars", offsetof(Record, nstate_vars), sizeof(unsigned), H5T_NATIVE_UINT);
    for (int k = 0; k < MPMX_POD_NSVARS; ++k)
      t.add("svars_" + std::to_string(k), offsetof(Record, svars) + k * d, d, H5T_NATIVE_DOUBLE);
    if (t.names.size() != MPMX_POD_NFIELDS) throw std::logic_error("POD field table size");
    return t;
  }();
  return f;
}

thread_local std::string g_error;

template <class T>
const T* need(const T* p, const char* what) {
  if (!p) throw std::runtime_error(std::string("null column: ") + what);
  return p;
}

// Opens the file and checks the table's field count (mesh.tcc:1663-1677).
hid_t open_table(const char* filename, hsize_t* nrecords) {
  if (!filename) throw std::runtime_error("null file name");
  const hid_t file = H5Fopen(filename, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (file < 0) throw std::runtime_error(std::string("HDF5 particle file is not found: ") + filename);
  hsize_t nfields = 0;
  *nrecords = 0;
  if (H5TBget_table_info(file, "table", &nfields, nrecords) < 0) {
    H5Fclose(file);
    throw std::runtime_error("no particle table in " + std::string(filename));
  }
  if (nfields != MPMX_POD_NFIELDS) {
    H5Fclose(file);
    throw std::runtime_error("HDF5 table has incorrect number of fields");
  }
  return file;
}

}  // namespace

extern "C" {

const char* mpmx_pod_last_error(void) { return g_error.c_str(); }

int mpmx_pod_write(const char* filename, const mpmx_pod_columns* c) {
  try {
    if (!filename || !c || c->n < 0) throw std::runtime_error("bad arguments");
    const Fields& f = fields();
    std::vector<Record> rec((size_t)c->n);
    for (int64_t i = 0; i < c->n; ++i) {
      Record& r = rec[i];
      std::memset(&r, 0, sizeof(Record));
      r.id = need(c->id, "id")[i];
      r.mass = need(c->mass, "mass")[i];
      r.volume = need(c->volume, "volume")[i];
      r.pressure = c->pressure ? c->pressure[i] : 0.0;
      for (int k = 0; k < 3; ++k) {
        r.coord[k] = need(c->coord, "coord")[3 * i + k];
        r.displacement[k] = c->displacement ? c->displacement[3 * i + k] : 0.0;
        r.nsize[k] = c->nsize ? c->nsize[3 * i + k] : 0.0;
        r.velocity[k] = c->velocity ? c->velocity[3 * i + k] : 0.0;
        r.acceleration[k] = c->acceleration ? c->acceleration[3 * i + k] : 0.0;
      }
      for (int k = 0; k < 6; ++k) {
        r.stress[k] = c->stress ? c->stress[6 * i + k] : 0.0;
        r.strain[k] = c->strain ? c->strain[6 * i + k] : 0.0;
      }
      for (int k = 0; k < 9; ++k) {
        r.defgrad[k] = c->defgrad ? c->defgrad[9 * i + k] : ((k % 4) == 0 ? 1.0 : 0.0);
        r.mapping_matrix[k] = c->mapping_matrix ? c->mapping_matrix[9 * i + k] : 0.0;
      }
      r.initialise_mapping_matrix = c->initialise_mapping_matrix
                                        ? c->initialise_mapping_matrix[i] != 0 : false;
      r.status = c->status ? c->status[i] != 0 : true;
      r.cell_id = c->cell_id ? c->cell_id[i] : -1;
      r.material_id = need(c->material_id, "material_id")[i];
      r.nstate_vars = c->nstate_vars ? c->nstate_vars[i] : 0;
      if (r.nstate_vars > MPMX_POD_NSVARS)
        throw std::runtime_error("# of state variables cannot be more than 20");
      for (int k = 0; k < MPMX_POD_NSVARS; ++k)
        r.svars[k] = c->svars ? c->svars[MPMX_POD_NSVARS * i + k] : 0.0;
    }
    std::vector<const char*> names;
    for (const auto& s : f.names) names.push_back(s.c_str());
    const hid_t file = H5Fcreate(filename, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (file < 0) throw std::runtime_error(std::string("cannot create ") + filename);
    const herr_t e = H5TBmake_table("Table Title", file, "table", MPMX_POD_NFIELDS, (hsize_t)c->n,
                                    sizeof(Record), names.data(), f.offsets.data(),
                                    f.types.data(), 10000, nullptr, 0,
                                    rec.empty() ? nullptr : rec.data());
    H5Fclose(file);
    if (e < 0) throw std::runtime_error("H5TBmake_table failed");
    return 0;
  } catch (std::exception& ex) {
    g_error = ex.what();
    return -1;
  }
}

int mpmx_pod_count(const char* filename, int64_t* n) {
  try {
    if (!n) throw std::runtime_error("bad arguments");
    hsize_t nrec = 0;
    const hid_t file = open_table(filename, &nrec);
    H5Fclose(file);
    *n = (int64_t)nrec;
    return 0;
  } catch (std::exception& ex) {
    g_error = ex.what();
    return -1;
  }
}

int mpmx_pod_read(const char* filename, mpmx_pod_columns* c) {
  try {
    if (!c) throw std::runtime_error("bad arguments");
    hsize_t nrec = 0;
    const hid_t file = open_table(filename, &nrec);
    if ((int64_t)nrec != c->n) {
      H5Fclose(file);
      throw std::runtime_error("columns sized for " + std::to_string(c->n) + " records, the table has " +
                               std::to_string((long long)nrec));
    }
    const Fields& f = fields();
    std::vector<Record> rec(nrec);
    const herr_t e = H5TBread_table(file, "table", sizeof(Record), f.offsets.data(),
                                    f.sizes.data(), rec.empty() ? nullptr : rec.data());
    H5Fclose(file);
    if (e < 0) throw std::runtime_error("H5TBread_table failed");
    for (int64_t i = 0; i < c->n; ++i) {
      const Record& r = rec[i];
      if (c->id) c->id[i] = r.id;
      if (c->mass) c->mass[i] = r.mass;
      if (c->volume) c->volume[i] = r.volume;
      if (c->pressure) c->pressure[i] = r.pressure;
      for (int k = 0; k < 3; ++k) {
        if (c->coord) c->coord[3 * i + k] = r.coord[k];
        if (c->displacement) c->displacement[3 * i + k] = r.displacement[k];
        if (c->nsize) c->nsize[3 * i + k] = r.nsize[k];
        if (c->velocity) c->velocity[3 * i + k] = r.velocity[k];
        if (c->acceleration) c->acceleration[3 * i + k] = r.acceleration[k];
      }
      for (int k = 0; k < 6; ++k) {
        if (c->stress) c->stress[6 * i + k] = r.stress[k];
        if (c->strain) c->strain[6 * i + k] = r.strain[k];
      }
      for (int k = 0; k < 9; ++k) {
        if (c->defgrad) c->defgrad[9 * i + k] = r.defgrad[k];
        if (c->mapping_matrix) c->mapping_matrix[9 * i + k] = r.mapping_matrix[k];
      }
      if (c->initialise_mapping_matrix) c->initialise_mapping_matrix[i] = r.initialise_mapping_matrix;
      if (c->status) c->status[i] = r.status;
      if (c->cell_id) c->cell_id[i] = r.cell_id;
      if (c->material_id) c->material_id[i] = r.material_id;
      if (c->nstate_vars) c->nstate_vars[i] = r.nstate_vars;
      if (c->svars)
        for (int k = 0; k < MPMX_POD_NSVARS; ++k) c->svars[MPMX_POD_NSVARS * i + k] = r.svars[k];
    }
    return 0;
  } catch (std::exception& ex) {
    g_error = ex.what();
    return -1;
  }
}

int mpmx_pod_field_names(const char* filename, char* buf, int64_t cap, int64_t* nfields) {
  try {
    if (!buf || !nfields || cap <= 0) throw std::runtime_error("bad arguments");
    hsize_t nrec = 0;
    const hid_t file = open_table(filename, &nrec);
    std::vector<std::vector<char>> store(MPMX_POD_NFIELDS, std::vector<char>(256, 0));
    std::vector<char*> names;
    for (auto& s : store) names.push_back(s.data());
    std::vector<size_t> sizes(MPMX_POD_NFIELDS), offsets(MPMX_POD_NFIELDS);
    size_t type_size = 0;
    const herr_t e = H5TBget_field_info(file, "table", names.data(), sizes.data(), offsets.data(),
                                        &type_size);
    H5Fclose(file);
    if (e < 0) throw std::runtime_error("H5TBget_field_info failed");
    int64_t pos = 0;
    for (int k = 0; k < MPMX_POD_NFIELDS; ++k) {
      const size_t len = std::strlen(names[k]);
      if (pos + (int64_t)len + 1 > cap) throw std::runtime_error("name buffer too small");
      std::memcpy(buf + pos, names[k], len + 1);
      pos += (int64_t)len + 1;
    }
    *nfields = MPMX_POD_NFIELDS;
    return 0;
  } catch (std::exception& ex) {
    g_error = ex.what();
    return -1;
  }
}

}  // extern "C"
