// mpmx_run: the explicit MPM analysis driver over the mpmx C ABI.
//
// Reads the reference's input deck -- the JSON file and the ASCII mesh,
// particle, constraint, stress and entity-set files it names (src/io/io.cc,
// include/io/io_mesh_ascii.tcc, include/solvers/mpm_base.tcc) -- and runs the
// analysis types MPMExplicit2D / MPMExplicit3D (mpm_explicit.tcc:24-199) with
// the time step on the GPU through include/mpmx.h. Particle states are
// written at the deck's output_steps as the reference's HDF5 particle table
// (include/mpmx_pod.h), from which an analysis with "resume" restarts.
//
//   mpmx_run -f <working directory> [-i <input file, default mpm.json>]
//            [--ordered] [--dry-run]
//
// --ordered: particle-order nodal sums (bit-identical to a single-threaded CPU
// run of the reference algorithm); --dry-run: read and check the deck, print
// a JSON summary, touch no GPU.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <iostream>
#include <limits>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "../../include/mpmx.h"
#include "../../include/mpmx_pod.h"
#include "json.h"

namespace mpm {

// ----------------------------------------------------------------- IO ------
// The reference's mpm::IO (src/io/io.cc): working directory, input file, JSON.
class IO {
 public:
  IO(int argc, char** argv) {
    for (int a = 1; a < argc; ++a) {
      const std::string s = argv[a];
      auto next = [&]() -> std::string {
        if (a + 1 >= argc) throw std::runtime_error("missing value after " + s);
        return argv[++a];
      };
      if (s == "-f" || s == "--working_dir") working_dir_ = next();
      else if (s == "-i" || s == "--input_file") input_file_ = next();
      else if (s == "--ordered") ordered_ = true;
      else if (s == "--dry-run") dry_run_ = true;
      else if (s == "-p" || s == "--parallel") next();  // thread count: unused on the GPU
      else throw std::runtime_error("unknown argument " + s);
    }
    if (working_dir_.empty()) throw std::runtime_error("working directory (-f) not given");
    if (working_dir_.back() != '/') working_dir_ += '/';
    std::ifstream f(working_dir_ + input_file_);
    if (!f) throw std::runtime_error("cannot open input file " + working_dir_ + input_file_);
    std::stringstream ss;
    ss << f.rdbuf();
    json_ = Json::parse(ss.str());
  }
  // file path inside the working directory; empty if it does not exist
  // (io.cc file_name)
  std::string file_name(const std::string& f) const {
    const std::string p = working_dir_ + f;
    struct stat st;
    return stat(p.c_str(), &st) == 0 ? p : std::string();
  }
  const Json& json() const { return json_; }
  const Json& object(const std::string& k) const { return json_.at(k); }
  std::string analysis_type() const { return json_.at("analysis").at("type").str(); }
  const std::string& working_dir() const { return working_dir_; }
  bool ordered() const { return ordered_; }
  bool dry_run() const { return dry_run_; }

 private:
  std::string working_dir_, input_file_ = "mpm.json";
  bool ordered_ = false, dry_run_ = false;
  Json json_;
};

// ------------------------------------------------------- ASCII readers ------
// Lines with '#' or '!' and blank lines are skipped, as in io_mesh_ascii.tcc.
static std::vector<std::vector<double>> read_rows(const std::string& path) {
  if (path.empty()) throw std::runtime_error("input file not found");
  std::ifstream f(path);
  if (!f) throw std::runtime_error("cannot open " + path);
  std::vector<std::vector<double>> rows;
  std::string line;
  while (std::getline(f, line)) {
    if (line.find('#') != std::string::npos || line.find('!') != std::string::npos) continue;
    std::istringstream is(line);
    std::vector<double> r;
    double v;
    while (is >> v) r.push_back(v);
    if (!r.empty()) rows.push_back(std::move(r));
  }
  return rows;
}

// Mesh file: "nnodes ncells", nnodes coordinate rows, ncells node-id rows
// (io_mesh_ascii.tcc:1-139).
static void read_mesh(const std::string& path, int dim, std::vector<double>& coords,
                      std::vector<int64_t>& cells, int nn) {
  const auto rows = read_rows(path);
  if (rows.empty() || rows[0].size() < 2) throw std::runtime_error("mesh file header missing");
  const int64_t nnodes = (int64_t)rows[0][0], ncells = (int64_t)rows[0][1];
  if ((int64_t)rows.size() < 1 + nnodes + ncells) throw std::runtime_error("mesh file truncated");
  coords.resize(nnodes * dim);
  for (int64_t i = 0; i < nnodes; ++i) {
    if ((int)rows[1 + i].size() < dim) throw std::runtime_error("short node row");
    for (int k = 0; k < dim; ++k) coords[i * dim + k] = rows[1 + i][k];
  }
  cells.resize(ncells * nn);
  for (int64_t c = 0; c < ncells; ++c) {
    const auto& r = rows[1 + nnodes + c];
    if ((int)r.size() != nn) throw std::runtime_error("cell row does not have 2^dim nodes");
    for (int q = 0; q < nn; ++q) cells[c * nn + q] = (int64_t)r[q];
  }
}

// Particles file: "nparticles", then coordinate rows (io_mesh_ascii.tcc:141-197).
static std::vector<double> read_particles(const std::string& path, int dim) {
  const auto rows = read_rows(path);
  if (rows.empty()) throw std::runtime_error("particles file empty");
  const int64_t n = (int64_t)rows[0][0];
  if ((int64_t)rows.size() < 1 + n) throw std::runtime_error("particles file truncated");
  std::vector<double> x(n * dim);
  for (int64_t i = 0; i < n; ++i)
    for (int k = 0; k < dim; ++k) x[i * dim + k] = rows[1 + i].at(k);
  return x;
}

// Entity sets file: {"node_sets": [{"id": i, "set": [...]}, ...], ...}
// (io.cc:159-195).
static std::map<int64_t, std::vector<int64_t>> read_sets(const std::string& path,
                                                         const std::string& type) {
  std::map<int64_t, std::vector<int64_t>> sets;
  if (path.empty()) return sets;
  std::ifstream f(path);
  std::stringstream ss;
  ss << f.rdbuf();
  const Json j = Json::parse(ss.str());
  if (!j.contains(type)) return sets;
  for (const Json& s : j.at(type).elements()) {
    std::vector<int64_t> ids;
    for (const Json& e : s.at("set").elements()) ids.push_back(e.integer());
    sets[s.at("id").integer()] = std::move(ids);
  }
  return sets;
}

// ----------------------------------------------------------- materials ------
// One entry of the "materials" array -> the 16-double mpmx parameter vector
// (LinearElastic: linear_elastic.tcc:3-49; MohrCoulomb: mohr_coulomb.tcc:3-81;
// every property the reference reads with .at() is required).
static std::vector<double> material_params(const Json& m, int dim) {
  const std::string type = m.at("type").str();
  const std::string sfx = std::to_string(dim) + "D";
  std::vector<double> p(MPMX_MATERIAL_PARAMS, 0.);
  // objective stress rate (linear_elastic.tcc:8-17, inherited by MohrCoulomb
  // through InfinitesimalElastoPlastic): "jaumann", "green_naghdi", else none
  std::string rate = "none";
  if (m.contains("stress_rate") && m.at("stress_rate").is_string()) rate = m.at("stress_rate").str();
  const double rate_id = rate == "jaumann" ? 1. : (rate == "green_naghdi" ? 2. : 0.);
  if (type == "LinearElastic" + sfx) {
    p[0] = 0;
    p[1] = m.at("density").number();
    p[2] = m.at("youngs_modulus").number();
    p[3] = m.at("poisson_ratio").number();
    p[4] = rate_id;
    return p;
  }
  if (type == "MohrCoulomb" + sfx) {
    p[0] = 1;
    p[1] = m.at("density").number();
    p[2] = m.at("youngs_modulus").number();
    p[3] = m.at("poisson_ratio").number();
    p[4] = m.at("friction").number();
    p[5] = m.at("dilation").number();
    p[6] = m.at("cohesion").number();
    p[7] = m.at("residual_friction").number();
    p[8] = m.at("residual_dilation").number();
    p[9] = m.at("residual_cohesion").number();
    p[10] = m.at("peak_pdstrain").number();
    p[11] = m.at("residual_pdstrain").number();
    p[12] = m.at("tension_cutoff").number();
    p[13] = m.at("softening").boolean() ? 1. : 0.;
    p[14] = m.get("packing_fraction_minimum", 0.);
    if (m.contains("packing_fraction")) p[15] = m.at("packing_fraction").number();
    else if (m.contains("porosity")) p[15] = 1. - m.at("porosity").number();
    p[16] = rate_id;
    return p;
  }
  throw std::runtime_error("material type " + type + " is not supported in " + sfx);
}

static void check(mpmx_ctx* ctx, int rc, const char* what) {
  if (rc != MPMX_OK)
    throw std::runtime_error(std::string(what) + ": " +
                             (ctx ? mpmx_last_error(ctx) : "mpmx error " + std::to_string(rc)));
}

// -------------------------------------------------------- MPMExplicit -------
template <unsigned Tdim>
class MPMExplicit {
 public:
  explicit MPMExplicit(std::shared_ptr<IO> io) : io_(std::move(io)) {}
  ~MPMExplicit() {
    if (ctx_) mpmx_destroy(ctx_);
  }
  bool solve();

 private:
  static constexpr int NN = Tdim == 2 ? 4 : 8;
  std::shared_ptr<IO> io_;
  mpmx_ctx* ctx_ = nullptr;
  // deck content
  std::string scheme_ = "usf", uuid_ = "mpmx", out_path_ = "results/";
  int vu_ = MPMX_FLIP;
  double blending_ = 1.0, dt_ = 0., damping_ = 0., ramp_ = 0.;
  int64_t nsteps_ = 0, output_steps_ = 1;
  double gravity_[3] = {0, 0, 0};
  std::vector<double> coords_;
  std::vector<int64_t> cells_;
  bool isoparametric_ = true;  // the mesh key "isoparametric"
  std::map<int64_t, std::vector<int64_t>> node_sets_, cell_sets_, particle_sets_;
  std::vector<double> mat_params_;
  std::map<int64_t, int> mat_index_;
  std::vector<int64_t> vnode_, fnode_;
  std::vector<int64_t> pvc_particle_;
  std::vector<int> pvc_dir_;
  std::vector<double> pvc_val_;
  std::vector<int> vdir_, fdir_, fsign_;
  std::vector<double> vval_, fmu_;
  std::vector<int64_t> enode_;  // nodal Euler angles: nodes, dim angles each
  std::vector<double> eang_;
  // acceleration constraints and the math functions they may refer to
  std::vector<int64_t> anode_;
  std::vector<int> adir_, afn_;
  std::vector<double> aval_;
  std::map<int64_t, int> fn_index_;  // math function id -> index
  std::vector<int64_t> cfnode_;      // concentrated nodal forces
  std::vector<int> cfdir_, cffn_;
  std::vector<double> cfval_;
  std::vector<int> fn_type_, fn_npts_;
  std::vector<double> fn_params_;
  int64_t expected_particles_ = 0;
  // checkpoint resume (analysis "resume": {resume, uuid, step}) and
  // locate_particles (mpm_base.tcc:53-54)
  bool resume_ = false, locate_particles_ = true;
  int64_t step_ = 0;
  // natural particle size by particle id (Particle::assign_volume,
  // particle.tcc:556-583), written to the particle table
  std::map<int64_t, double> nsize_;
  // particle_cells (mpm_base.tcc:1495-1520): particle id -> cell id
  std::map<int64_t, int64_t> particle_cells_;

  void read_analysis();
  void initialise_materials();
  void initialise_math_functions();
  void initialise_mesh();
  void initialise_particles();
  void resume_particles();
  void particle_loads();
  void initialise_loads();
  std::string output_file(const std::string& attribute, const std::string& ext,
                          int64_t step) const;
  std::vector<int64_t> ids_to_indices(const std::vector<int64_t>& ids) const;
  void write_outputs(int64_t step);
  void summary() const;
  int material_index(const Json& id) const {
    const Json& v = id.is_array() ? id.at(0) : id;
    auto it = mat_index_.find(v.integer());
    if (it == mat_index_.end()) throw std::runtime_error("material id not defined");
    return it->second;
  }
};

// analysis keys (mpm_base.tcc:40-140)
template <unsigned Tdim>
void MPMExplicit<Tdim>::read_analysis() {
  const Json& a = io_->object("analysis");
  dt_ = a.at("dt").number();
  nsteps_ = a.at("nsteps").integer();
  if (a.contains("mpm_scheme")) scheme_ = a.at("mpm_scheme").str();
  if (a.contains("velocity_update")) {
    const Json& v = a.at("velocity_update");
    const std::string t = v.is_boolean() ? (v.boolean() ? "pic" : "flip") : v.str();
    // VelocityUpdateType (src/mpm.cc:19-24)
    if (t == "pic") vu_ = MPMX_PIC;
    else if (t == "flip") vu_ = MPMX_FLIP;
    else if (t == "asflip") vu_ = MPMX_ASFLIP;
    else if (t == "tpic") vu_ = MPMX_TPIC;
    else if (t == "apic") vu_ = MPMX_APIC;
    else throw std::runtime_error("velocity update " + t + " is not supported");
  }
  if (a.contains("velocity_update_settings") &&
      a.at("velocity_update_settings").contains("blending_ratio")) {
    blending_ = a.at("velocity_update_settings").at("blending_ratio").number();
    if (blending_ < 0. || blending_ > 1.) blending_ = 1.0;
  }
  if (a.contains("damping")) {
    const Json& d = a.at("damping");
    const std::string t = d.at("type").str();
    if (t == "Cundall") damping_ = d.at("damping_factor").number();
    else if (t != "None") throw std::runtime_error("damping type " + t + " is not supported");
  }
  if (a.contains("uuid")) uuid_ = a.at("uuid").str();
  if (a.contains("locate_particles")) locate_particles_ = a.at("locate_particles").boolean();
  // checkpoint resume (mpm_explicit.tcc:43-46, mpm_base.tcc:509-551): the
  // particle table of step "step" of analysis "uuid"
  if (a.contains("resume") && a.at("resume").contains("resume") &&
      a.at("resume").at("resume").boolean()) {
    resume_ = true;
    uuid_ = a.at("resume").at("uuid").str();
    step_ = a.at("resume").at("step").integer();
  }
  if (io_->json().contains("post_processing")) {
    const Json& pp = io_->object("post_processing");
    output_steps_ = pp.contains("output_steps") ? pp.at("output_steps").integer() : nsteps_;
    if (pp.contains("path")) out_path_ = pp.at("path").str();
  }
  if (output_steps_ <= 0) output_steps_ = nsteps_ > 0 ? nsteps_ : 1;
}

template <unsigned Tdim>
void MPMExplicit<Tdim>::initialise_materials() {
  for (const Json& m : io_->object("materials").elements()) {
    const int64_t id = m.at("id").integer();
    if (mat_index_.count(id)) throw std::runtime_error("duplicate material id");
    mat_index_[id] = (int)mat_index_.size();
    const auto p = material_params(m, Tdim);
    mat_params_.insert(mat_params_.end(), p.begin(), p.end());
  }
}

// mesh, entity sets, nodal constraints (mpm_base.tcc:211-330, 1126-1283)
template <unsigned Tdim>
void MPMExplicit<Tdim>::initialise_mesh() {
  const Json& mp = io_->object("mesh");
  // MPMBase::is_isoparametric (mpm_base.tcc:768-783): true unless the key says
  // otherwise
  isoparametric_ = mp.contains("isoparametric") ? mp.at("isoparametric").boolean() : true;
  read_mesh(io_->file_name(mp.at("mesh").str()), Tdim, coords_, cells_, NN);
  if (mp.contains("entity_sets")) {
    const std::string es = io_->file_name(mp.at("entity_sets").str());
    node_sets_ = read_sets(es, "node_sets");
    cell_sets_ = read_sets(es, "cell_sets");
    particle_sets_ = read_sets(es, "particle_sets");
  }
  if (!mp.contains("boundary_conditions") || !mp.at("boundary_conditions").is_object()) return;
  const Json& bc = mp.at("boundary_conditions");
  auto set_of = [&](int64_t id) -> const std::vector<int64_t>& {
    auto it = node_sets_.find(id);
    if (it == node_sets_.end()) throw std::runtime_error("node set " + std::to_string(id) + " not defined");
    return it->second;
  };
  // nodal_euler_angles (MPMBase::node_euler_angles, mpm_base.tcc:1041-1063):
  // a file of "node angle_1 .. angle_dim" rows in radians
  // (IOMeshAscii::read_euler_angles, io_mesh_ascii.tcc:351-391; a node's
  // first row counts)
  if (bc.contains("nodal_euler_angles")) {
    std::map<int64_t, std::vector<double>> ang;
    for (const auto& r : read_rows(io_->file_name(bc.at("nodal_euler_angles").str())))
      for (size_t q = 0; q + Tdim < r.size(); q += 1 + Tdim)
        ang.emplace((int64_t)r[q], std::vector<double>(r.begin() + q + 1, r.begin() + q + 1 + Tdim));
    for (const auto& kv : ang) {
      enode_.push_back(kv.first);
      eang_.insert(eang_.end(), kv.second.begin(), kv.second.end());
    }
  }
  if (bc.contains("velocity_constraints")) {
    for (const Json& c : bc.at("velocity_constraints").elements()) {
      if (c.contains("file")) {
        for (const auto& r : read_rows(io_->file_name(c.at("file").str()))) {
          vnode_.push_back((int64_t)r.at(0));
          vdir_.push_back((int)r.at(1));
          vval_.push_back(r.at(2));
        }
      } else {
        for (int64_t n : set_of(c.at("nset_id").integer())) {
          vnode_.push_back(n);
          vdir_.push_back((int)c.at("dir").integer());
          vval_.push_back(c.at("velocity").number());
        }
      }
    }
  }
  // nodal_acceleration_constraints (mpm_base.tcc:1067-1124): from a file
  // (rows node dir acceleration, io_mesh_ascii.tcc:550-595) or per node set,
  // with an optional math_function_id
  if (bc.contains("acceleration_constraints")) {
    for (const Json& c : bc.at("acceleration_constraints").elements()) {
      if (c.contains("file")) {
        for (const auto& r : read_rows(io_->file_name(c.at("file").str()))) {
          anode_.push_back((int64_t)r.at(0));
          adir_.push_back((int)r.at(1));
          aval_.push_back(r.at(2));
          afn_.push_back(-1);
        }
      } else {
        int fn = -1;
        if (c.contains("math_function_id")) {
          const auto it = fn_index_.find(c.at("math_function_id").integer());
          if (it == fn_index_.end()) throw std::runtime_error("math function not defined");
          fn = it->second;
        }
        for (int64_t n : set_of(c.at("nset_id").integer())) {
          anode_.push_back(n);
          adir_.push_back((int)c.at("dir").integer());
          aval_.push_back(c.at("acceleration").number());
          afn_.push_back(fn);
        }
      }
    }
  }
  if (bc.contains("friction_constraints")) {
    for (const Json& c : bc.at("friction_constraints").elements()) {
      if (c.contains("file")) {
        for (const auto& r : read_rows(io_->file_name(c.at("file").str()))) {
          fnode_.push_back((int64_t)r.at(0));
          fdir_.push_back((int)r.at(1));
          fsign_.push_back((int)r.at(2));
          fmu_.push_back(r.at(3));
        }
      } else {
        for (int64_t n : set_of(c.at("nset_id").integer())) {
          fnode_.push_back(n);
          fdir_.push_back((int)c.at("dir").integer());
          fsign_.push_back((int)c.at("sign_n").integer());
          fmu_.push_back(c.at("friction").number());
        }
      }
    }
  }
}

// math_functions (MPMBase::initialise_math_functions, mpm_base.tcc:958-1010):
// Linear with xvalues/fxvalues or a file of "x, f(x)" rows
// (io_mesh_ascii.tcc:798-816), Sin with x0, a and an optional xrange.
template <unsigned Tdim>
void MPMExplicit<Tdim>::initialise_math_functions() {
  if (!io_->json().contains("math_functions")) return;
  for (const Json& f : io_->json().at("math_functions").elements()) {
    const int64_t id = f.at("id").integer();
    const std::string type = f.at("type").str();
    if (fn_index_.count(id)) throw std::runtime_error("math function id defined twice");
    if (type == "Linear") {
      std::vector<double> x, fx;
      if (f.contains("file")) {
        std::ifstream in(io_->file_name(f.at("file").str()));
        if (!in) throw std::runtime_error("cannot open math function file");
        std::string line;
        while (std::getline(in, line)) {
          for (char& ch : line)
            if (ch == ',') ch = ' ';
          std::istringstream is(line);
          double a, b;
          if (is >> a >> b) {
            x.push_back(a);
            fx.push_back(b);
          }
        }
      } else {
        for (const Json& v : f.at("xvalues").elements()) x.push_back(v.number());
        for (const Json& v : f.at("fxvalues").elements()) fx.push_back(v.number());
      }
      if (x.empty() || x.size() != fx.size())
        throw std::runtime_error("Cannot initialise linear function; x and f(x) are invalid");
      fn_type_.push_back(MPMX_FN_LINEAR);
      fn_npts_.push_back((int)x.size());
      fn_params_.insert(fn_params_.end(), x.begin(), x.end());
      fn_params_.insert(fn_params_.end(), fx.begin(), fx.end());
    } else if (type == "Sin") {
      double lo = std::numeric_limits<double>::min(), hi = std::numeric_limits<double>::max();
      if (f.contains("xrange") && f.at("xrange").is_array() && f.at("xrange").elements().size() == 2) {
        lo = f.at("xrange").at(0).number();
        hi = f.at("xrange").at(1).number();
      }
      fn_type_.push_back(MPMX_FN_SIN);
      fn_npts_.push_back(4);
      fn_params_.insert(fn_params_.end(), {f.at("x0").number(), f.at("a").number(), lo, hi});
    } else {
      throw std::runtime_error("math function type " + type + " is not supported");
    }
    fn_index_[id] = (int)fn_type_.size() - 1;
  }
}

// gravity and its ramp (mpm_base.tcc:788-830)
template <unsigned Tdim>
void MPMExplicit<Tdim>::initialise_loads() {
  if (!io_->json().contains("external_loading_conditions")) return;
  const Json& l = io_->object("external_loading_conditions");
  // rotation_forces (centrifugal and Coriolis body forces, mpm_base.tcc:
  // 845-870, Particle::map_rotation_force) are not part of this build's step
  if (l.contains("rotation_forces"))
    throw std::runtime_error("rotation_forces are not supported by this build");
  // concentrated_nodal_forces (mpm_base.tcc:913-953): from a file (rows node
  // dir force, io_mesh_ascii.tcc:752-795, no function) or per node set (-1:
  // every node) with an optional math_function_id
  if (l.contains("concentrated_nodal_forces")) {
    for (const Json& c : l.at("concentrated_nodal_forces").elements()) {
      if (c.contains("file")) {
        for (const auto& r : read_rows(io_->file_name(c.at("file").str()))) {
          cfnode_.push_back((int64_t)r.at(0));
          cfdir_.push_back((int)r.at(1));
          cfval_.push_back(r.at(2));
          cffn_.push_back(-1);
        }
      } else {
        int fn = -1;
        if (c.contains("math_function_id")) {
          const auto it = fn_index_.find(c.at("math_function_id").integer());
          if (it == fn_index_.end()) throw std::runtime_error("math function not defined");
          fn = it->second;
        }
        const int64_t sid = c.at("nset_id").integer();
        std::vector<int64_t> all{-1};
        const std::vector<int64_t>* nodes = &all;
        if (sid != -1) {
          const auto it = node_sets_.find(sid);
          if (it == node_sets_.end())
            throw std::runtime_error("node set " + std::to_string(sid) + " not defined");
          nodes = &it->second;
        }
        for (int64_t n : *nodes) {
          cfnode_.push_back(n);
          cfdir_.push_back((int)c.at("dir").integer());
          cfval_.push_back(c.at("force").number());
          cffn_.push_back(fn);
        }
      }
    }
  }
  if (!l.contains("gravity")) return;
  const Json& g = l.at("gravity");
  const Json& vec = g.is_array() ? g : g.at("gravity");
  if (vec.size() != Tdim) throw std::runtime_error("gravity must have one value per dimension");
  for (unsigned k = 0; k < Tdim; ++k) gravity_[k] = vec.at(k).number();
  if (g.is_object() && g.contains("ramping_time")) ramp_ = std::max(0., g.at("ramping_time").number());
}

// particle generators, locate, stresses, volume and mass
// (mpm_base.tcc:345-440, mesh.tcc:508-576, 1958-2046; mpm_explicit.tcc:90-96)
template <unsigned Tdim>
void MPMExplicit<Tdim>::initialise_particles() {
  const Json& mp = io_->object("mesh");
  // particle_cells: "pid cid" rows (io_mesh_ascii.tcc:442-480); they give the
  // cell a particle without a cell is located from first (assign_cell_id,
  // mesh.tcc:1528-1548), i.e. particles read from a file; a generated
  // particle already has its cell and the assignment fails
  if (mp.contains("particle_cells")) {
    const std::string f = io_->file_name(mp.at("particle_cells").str());
    if (!f.empty())
      for (const auto& r : read_rows(f)) {
        if (r.size() < 2) throw std::runtime_error("particle_cells: two columns per row");
        particle_cells_[(int64_t)r[0]] = (int64_t)r[1];
      }
  }
  int64_t created = 0;
  for (const Json& pj : io_->object("particles").elements()) {
    const Json& g = pj.at("generator");
    const std::string type = g.at("type").str();
    const int mid = material_index(g.at("material_id"));
    if (type == "gauss") {
      const int nppd = (int)g.at("nparticles_per_dir").integer();
      const int64_t cset = g.contains("cset_id") ? g.at("cset_id").integer() : -1;
      std::vector<int64_t> cells;
      if (cset >= 0) {
        auto it = cell_sets_.find(cset);
        if (it == cell_sets_.end()) throw std::runtime_error("cell set " + std::to_string(cset) + " not defined");
        cells = it->second;
      }
      const int64_t nc = cset >= 0 ? (int64_t)cells.size() : (int64_t)(cells_.size() / NN);
      const int64_t ng = nc * (int64_t)std::pow(nppd, (int)Tdim);
      expected_particles_ += ng;
      if (particle_cells_.lower_bound(created) != particle_cells_.end() &&
          particle_cells_.lower_bound(created)->first < created + ng)
        throw std::runtime_error("Particle cells are not properly assigned to particles");
      created += ng;
      if (ctx_)
        check(ctx_, mpmx_generate_gauss(ctx_, nppd, cset >= 0 ? (int64_t)cells.size() : -1,
                                        cset >= 0 ? cells.data() : nullptr, mid),
              "generate gauss particles");
    } else if (type == "file") {
      const auto x = read_particles(io_->file_name(g.at("location").str()), Tdim);
      const int64_t n = (int64_t)(x.size() / Tdim);
      expected_particles_ += n;
      std::vector<int64_t> hint(n, -1);
      for (int64_t i = 0; i < n; ++i) {
        const auto it = particle_cells_.find(created + i);
        if (it != particle_cells_.end()) hint[i] = it->second;
      }
      created += n;
      if (ctx_) {
        const std::vector<double> one(n, 1.0);
        const std::vector<int> mat(n, mid);
        check(ctx_, mpmx_set_particles(ctx_, n, x.data(), nullptr, nullptr, nullptr, nullptr,
                                       nullptr, one.data(), one.data(), one.data(), mat.data(),
                                       hint.data(), nullptr),
              "particles from file");
      }
    } else {
      throw std::runtime_error("particle generator " + type + " is not supported (gauss, file)");
    }
  }
  // material_sets (mpm_base.tcc:449-468): the particles of a set take the
  // material (and its initial state variables), before the mass is computed
  if (ctx_ && io_->json().contains("material_sets")) {
    for (const Json& ms : io_->object("material_sets").elements()) {
      const int mid = material_index(ms.at("material_id"));
      const int64_t pset = ms.at("pset_id").integer();
      std::vector<int64_t> idx;
      if (pset == -1) {
        int64_t n = 0;
        check(ctx_, mpmx_nparticles(ctx_, &n), "particle count");
        for (int64_t i = 0; i < n; ++i) idx.push_back(i);
      } else {
        auto it = particle_sets_.find(pset);
        if (it == particle_sets_.end()) throw std::runtime_error("material set: particle set not defined");
        idx = ids_to_indices(it->second);
      }
      check(ctx_, mpmx_assign_materials(ctx_, (int64_t)idx.size(), idx.data(), mid),
            "material sets");
    }
  }
  if (ctx_) {
    check(ctx_, mpmx_compute_volume_mass(ctx_), "particle volume and mass");
    // natural size: unit element length 2 over nparticles_in_cell^(1/dim)
    // when the volume was assigned (particle.tcc:570-580)
    int64_t n = 0;
    check(ctx_, mpmx_nparticles(ctx_, &n), "particle count");
    std::vector<int64_t> cell(n), ids(n);
    check(ctx_, mpmx_get_particles(ctx_, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                                   nullptr, nullptr, nullptr, nullptr, cell.data(), nullptr,
                                   nullptr),
          "particle cells");
    check(ctx_, mpmx_get_particle_ids(ctx_, ids.data()), "particle ids");
    std::map<int64_t, int64_t> count;
    for (int64_t c : cell) count[c] += 1;
    for (int64_t i = 0; i < n; ++i)
      nsize_[ids[i]] = 2.0 / std::pow((double)count[cell[i]], 1.0 / Tdim);
  }
  particle_loads();
  if (ctx_ && mp.contains("particles_stresses")) {
    const Json& ps = mp.at("particles_stresses");
    const std::string path = io_->file_name(ps.at("location").str());
    const auto rows = read_rows(path);
    const int64_t n = (int64_t)rows.at(0).at(0);
    std::vector<double> s(n * 6);
    for (int64_t i = 0; i < n; ++i)
      for (int k = 0; k < 6; ++k) s[i * 6 + k] = rows.at(1 + i).at(k);
    check(ctx_, mpmx_assign_stresses(ctx_, n, s.data()), "particles stresses");
  }
}

// Particle sets hold particle ids; the C ABI takes positions in the particle
// order (equal until particles are removed or read back from a checkpoint).
// Ids of no particle are skipped, as Mesh::iterate_over_particle_set does
// (mesh.tcc:1175-1189); -1 (every particle) stays -1.
template <unsigned Tdim>
std::vector<int64_t> MPMExplicit<Tdim>::ids_to_indices(const std::vector<int64_t>& ids) const {
  int64_t n = 0;
  check(ctx_, mpmx_nparticles(ctx_, &n), "particle count");
  std::vector<int64_t> pid(n);
  check(ctx_, mpmx_get_particle_ids(ctx_, pid.data()), "particle ids");
  std::map<int64_t, int64_t> at;
  for (int64_t i = 0; i < n; ++i) at[pid[i]] = i;
  std::vector<int64_t> out;
  for (int64_t id : ids) {
    if (id < 0) {
      out.push_back(-1);
      continue;
    }
    const auto it = at.find(id);
    if (it != at.end()) out.push_back(it->second);
  }
  return out;
}

// The loads and constraints given per particle (particle velocity
// constraints, particle surface tractions), for generated or resumed
// particles (mpm_explicit.tcc:75-91: particle_entity_sets and
// particle_velocity_constraints also run on resume; initialise_loads after).
template <unsigned Tdim>
void MPMExplicit<Tdim>::particle_loads() {
  const Json& mp = io_->object("mesh");
  // particles_velocity_constraints (mpm_base.tcc:1549-1585): a particle set
  // id (-1: all particles), a direction and a velocity; applied in order
  if (mp.contains("boundary_conditions") && mp.at("boundary_conditions").is_object() &&
      mp.at("boundary_conditions").contains("particles_velocity_constraints")) {
    for (const Json& c : mp.at("boundary_conditions").at("particles_velocity_constraints").elements()) {
      const int64_t pset = c.at("pset_id").integer();
      const int dir = (int)c.at("dir").integer();
      const double vel = c.at("velocity").number();
      if (dir < 0 || dir >= (int)Tdim) throw std::runtime_error("Invalid direction of velocity constraint");
      if (pset == -1) {
        pvc_particle_.push_back(-1);
        pvc_dir_.push_back(dir);
        pvc_val_.push_back(vel);
        continue;
      }
      auto it = particle_sets_.find(pset);
      if (it == particle_sets_.end())
        throw std::runtime_error("No particle set found to assign velocity constraint");
      for (int64_t p : it->second) {
        pvc_particle_.push_back(p);
        pvc_dir_.push_back(dir);
        pvc_val_.push_back(vel);
      }
    }
  }
  if (!ctx_) return;  // a dry run counts the constraints only
  // particle_surface_traction (mpm_base.tcc:882-911): a particle set id (-1:
  // all particles), a direction, a traction and an optional math function;
  // applied in order
  if (io_->json().contains("external_loading_conditions") &&
      io_->object("external_loading_conditions").contains("particle_surface_traction")) {
    std::vector<int64_t> tp;
    std::vector<int> td, tf;
    std::vector<double> tv;
    for (const Json& c :
         io_->object("external_loading_conditions").at("particle_surface_traction").elements()) {
      const int64_t pset = c.at("pset_id").integer();
      int fn = -1;
      if (c.contains("math_function_id")) {
        const auto it = fn_index_.find(c.at("math_function_id").integer());
        if (it == fn_index_.end()) throw std::runtime_error("math function not defined");
        fn = it->second;
      }
      std::vector<int64_t> all{-1};
      const std::vector<int64_t>* ids = &all;
      if (pset != -1) {
        auto it = particle_sets_.find(pset);
        if (it == particle_sets_.end())
          throw std::runtime_error("No particle set found to assign traction");
        ids = &it->second;
      }
      for (int64_t p : *ids) {
        tp.push_back(p);
        td.push_back((int)c.at("dir").integer());
        tv.push_back(c.at("traction").number());
        tf.push_back(fn);
      }
    }
    tp = ids_to_indices(tp);
    check(ctx_, mpmx_set_particle_tractions(ctx_, (int64_t)tp.size(), tp.data(), td.data(),
                                            tv.data(), tf.data()),
          "particle surface tractions");
  }
  if (!pvc_particle_.empty()) {
    const std::vector<int64_t> idx = ids_to_indices(pvc_particle_);
    check(ctx_, mpmx_set_particle_velocity_constraints(ctx_, (int64_t)idx.size(), idx.data(),
                                                       pvc_dir_.data(), pvc_val_.data()),
          "particle velocity constraints");
  }
}

// Checkpoint resume (mpm_base.tcc:509-551, mesh.tcc:1659-1721,
// particle.tcc:35-177): the particles of the table written at step step_,
// in its order and with its ids, located from their stored cells; the mass
// density is mass / volume; the step counter continues from step_.
template <unsigned Tdim>
void MPMExplicit<Tdim>::resume_particles() {
  const std::string f = output_file("particles", ".h5", step_);
  int64_t n = 0;
  if (mpmx_pod_count(f.c_str(), &n) != 0) throw std::runtime_error(mpmx_pod_last_error());
  std::vector<int64_t> id(n), cell(n);
  std::vector<double> mass(n), vol(n), pres(n), co(3 * n), di(3 * n), ns(3 * n), ve(3 * n),
      ac(3 * n), st(6 * n), sn(6 * n), fg(9 * n), mm(9 * n), sv(MPMX_POD_NSVARS * n);
  std::vector<unsigned char> imm(n), stat(n);
  std::vector<uint32_t> mid(n), nsv(n);
  mpmx_pod_columns c;
  c.n = n;
  c.id = id.data();
  c.mass = mass.data();
  c.volume = vol.data();
  c.pressure = pres.data();
  c.coord = co.data();
  c.displacement = di.data();
  c.nsize = ns.data();
  c.velocity = ve.data();
  c.acceleration = ac.data();
  c.stress = st.data();
  c.strain = sn.data();
  c.defgrad = fg.data();
  c.initialise_mapping_matrix = imm.data();
  c.mapping_matrix = mm.data();
  c.status = stat.data();
  c.cell_id = cell.data();
  c.material_id = mid.data();
  c.nstate_vars = nsv.data();
  c.svars = sv.data();
  if (mpmx_pod_read(f.c_str(), &c) != 0) throw std::runtime_error(mpmx_pod_last_error());
  std::vector<double> x(Tdim * n), u(Tdim * n), v(Tdim * n), rho(n), svars(MPMX_STATE_VARS * n);
  std::vector<int> mat(n);
  std::vector<int64_t> hint(n);
  int64_t with_sv = 0, without_sv = 0;
  for (int64_t i = 0; i < n; ++i) {
    for (unsigned k = 0; k < Tdim; ++k) {
      x[i * Tdim + k] = co[3 * i + k];
      u[i * Tdim + k] = di[3 * i + k];
      v[i * Tdim + k] = ve[3 * i + k];
    }
    rho[i] = mass[i] / vol[i];  // particle.tcc:46-47
    const auto it = mat_index_.find((int64_t)mid[i]);
    if (it == mat_index_.end()) throw std::runtime_error("Material is invalid to assign to particle!");
    mat[i] = it->second;
    // a cell id beyond the mesh is none (max Index written as -1)
    hint[i] = (cell[i] >= 0 && cell[i] < (int64_t)(cells_.size() / NN)) ? cell[i] : -1;
    // the table's state variables when their count is the material's
    // (particle.tcc:160-171), else the material's initial values
    const bool mc = mat_params_[it->second * MPMX_MATERIAL_PARAMS] == 1.0;
    for (int k = 0; k < MPMX_STATE_VARS; ++k)
      svars[i * MPMX_STATE_VARS + k] = sv[MPMX_POD_NSVARS * i + k];
    if (mc) (nsv[i] == MPMX_STATE_VARS ? with_sv : without_sv) += 1;
    nsize_[id[i]] = ns[3 * i];
  }
  if (with_sv > 0 && without_sv > 0)
    throw std::runtime_error("resumed MohrCoulomb particles with and without their state variables");
  // F as stored; the table's stress and strain are the Voigt 6-vectors
  check(ctx_, mpmx_set_particles(ctx_, n, x.data(), u.data(), v.data(), st.data(), sn.data(),
                                 fg.data(), mass.data(), vol.data(), rho.data(), mat.data(),
                                 hint.data(), without_sv > 0 ? nullptr : svars.data()),
        "resumed particles");
  check(ctx_, mpmx_set_particle_ids(ctx_, n, id.data()), "resumed particle ids");
  // mapping matrices of the affine updates where the table flags them
  // (particle.tcc:111-125)
  std::vector<double> B(Tdim * Tdim * n, 0.);
  bool any_mm = false;
  for (int64_t i = 0; i < n; ++i)
    if (imm[i]) {
      any_mm = true;
      for (unsigned a = 0; a < Tdim; ++a)
        for (unsigned b = 0; b < Tdim; ++b) B[(i * Tdim + a) * Tdim + b] = mm[9 * i + 3 * a + b];
    }
  if (any_mm) check(ctx_, mpmx_set_mapping_matrix(ctx_, n, B.data()), "resumed mapping matrices");
  check(ctx_, mpmx_set_step(ctx_, step_), "resume step");
  particle_loads();
}

// <working dir><post_processing path><uuid>/<attribute><step, zero-padded
// to the digits of nsteps><ext> (IO::output_file, src/io/io.cc:94-140);
// creates the folders.
template <unsigned Tdim>
std::string MPMExplicit<Tdim>::output_file(const std::string& attribute, const std::string& ext,
                                           int64_t step) const {
  const std::string dir = io_->working_dir() + out_path_ + uuid_ + "/";
  std::string cur;
  for (char ch : dir) {
    cur += ch;
    if (ch == '/') mkdir(cur.c_str(), 0755);
  }
  const int digits = (int)std::log10((double)std::max<int64_t>(nsteps_, 1)) + 1;
  std::ostringstream name;
  name << attribute;
  name.fill('0');
  name.width(digits);
  name << step;
  name << ext;
  return dir + name.str();
}

// The particle table at output steps (MPMBase::write_outputs ->
// write_hdf5, mpm_base.tcc:577-595, 752-766; Particle::pod(),
// particle.tcc:180-305).
template <unsigned Tdim>
void MPMExplicit<Tdim>::write_outputs(int64_t step) {
  if (step % output_steps_ != 0) return;
  int64_t n = 0;
  check(ctx_, mpmx_nparticles(ctx_, &n), "particle count");
  std::vector<double> x(n * Tdim), u(n * Tdim), v(n * Tdim), s(n * 6), e(n * 6), F(n * 9),
      vol(n), m(n), sv(n * MPMX_STATE_VARS);
  std::vector<int> mat(n);
  std::vector<int64_t> cell(n), ids(n);
  check(ctx_, mpmx_get_particles(ctx_, x.data(), u.data(), v.data(), s.data(), e.data(), F.data(),
                                 m.data(), vol.data(), nullptr, mat.data(), cell.data(), nullptr,
                                 sv.data()),
        "read particles");
  check(ctx_, mpmx_get_particle_ids(ctx_, ids.data()), "particle ids");
  std::vector<double> co(3 * n, 0.), di(3 * n, 0.), ve(3 * n, 0.), ns(3 * n, 0.),
      svars(MPMX_POD_NSVARS * n, 0.);
  std::vector<uint32_t> mid(n), nsv(n, 0);
  // material ids as the deck numbers them
  std::vector<int64_t> deck_id(mat_index_.size());
  for (const auto& kv : mat_index_) deck_id[kv.second] = kv.first;
  for (int64_t i = 0; i < n; ++i) {
    for (unsigned k = 0; k < Tdim; ++k) {
      co[3 * i + k] = x[i * Tdim + k];
      di[3 * i + k] = u[i * Tdim + k];
      ve[3 * i + k] = v[i * Tdim + k];
      const auto it = nsize_.find(ids[i]);
      ns[3 * i + k] = it == nsize_.end() ? 0.0 : it->second;
    }
    mid[i] = (uint32_t)deck_id[mat[i]];
    if (mat_params_[mat[i] * MPMX_MATERIAL_PARAMS] == 1.0) {  // MohrCoulomb
      nsv[i] = MPMX_STATE_VARS;
      for (int k = 0; k < MPMX_STATE_VARS; ++k)
        svars[MPMX_POD_NSVARS * i + k] = sv[i * MPMX_STATE_VARS + k];
    }
  }
  mpmx_pod_columns c;
  std::memset(&c, 0, sizeof(c));
  c.n = n;
  c.id = ids.data();
  c.mass = m.data();
  c.volume = vol.data();
  c.coord = co.data();
  c.displacement = di.data();
  c.nsize = ns.data();
  c.velocity = ve.data();
  c.stress = s.data();
  c.strain = e.data();
  c.defgrad = F.data();
  c.cell_id = cell.data();
  c.material_id = mid.data();
  c.nstate_vars = nsv.data();
  c.svars = svars.data();
  // the affine updates' mapping matrices, flagged once they exist (after the
  // first step; Particle::pod, particle.tcc:212-218)
  std::vector<double> mm(9 * n, 0.), B(Tdim * Tdim * n);
  std::vector<unsigned char> imm(n, 0);
  const bool affine = vu_ == MPMX_APIC || vu_ == MPMX_ASFLIP || vu_ == MPMX_TPIC;
  if (affine && step > 0) {
    check(ctx_, mpmx_get_mapping_matrix(ctx_, B.data()), "mapping matrices");
    for (int64_t i = 0; i < n; ++i) {
      imm[i] = 1;
      for (unsigned a = 0; a < Tdim; ++a)
        for (unsigned b = 0; b < Tdim; ++b) mm[9 * i + 3 * a + b] = B[(i * Tdim + a) * Tdim + b];
    }
  }
  c.initialise_mapping_matrix = imm.data();
  c.mapping_matrix = mm.data();
  const std::string path = output_file("particles", ".h5", step);
  if (mpmx_pod_write(path.c_str(), &c) != 0)
    throw std::runtime_error(std::string("write ") + path + ": " + mpmx_pod_last_error());
}

template <unsigned Tdim>
void MPMExplicit<Tdim>::summary() const {
  std::printf("{\"analysis\": \"MPMExplicit%uD\", \"scheme\": \"%s\", \"velocity_update\": \"%s\", "
              "\"dt\": %.17g, \"nsteps\": %lld, \"output_steps\": %lld, \"damping\": %.17g, "
              "\"nnodes\": %lld, \"ncells\": %lld, \"node_sets\": %zu, \"cell_sets\": %zu, "
              "\"velocity_constraints\": %zu, \"acceleration_constraints\": %zu, "
              "\"math_functions\": %zu, \"nodal_forces\": %zu, \"friction_constraints\": %zu, "
              "\"nodal_euler_angles\": %zu, "
              "\"particle_velocity_constraints\": %zu, \"particles\": %lld, "
              "\"gravity\": [%.17g, %.17g, %.17g], \"ramping_time\": %.17g, \"materials\": [",
              Tdim, scheme_.c_str(),
              vu_ == MPMX_PIC ? "pic" : vu_ == MPMX_ASFLIP ? "asflip" : vu_ == MPMX_TPIC ? "tpic"
              : vu_ == MPMX_APIC ? "apic" : "flip", dt_, (long long)nsteps_,
              (long long)output_steps_, damping_, (long long)(coords_.size() / Tdim),
              (long long)(cells_.size() / NN), node_sets_.size(), cell_sets_.size(),
              vnode_.size(), anode_.size(), fn_type_.size(), cfnode_.size(), fnode_.size(),
              enode_.size(),
              pvc_particle_.size(),
              (long long)expected_particles_,
              gravity_[0],
              gravity_[1], gravity_[2], ramp_);
  for (size_t i = 0; i < mat_params_.size(); ++i)
    std::printf("%s%s%.17g%s", i % MPMX_MATERIAL_PARAMS == 0 ? (i ? ", [" : "[") : "",
                i % MPMX_MATERIAL_PARAMS == 0 ? "" : ", ", mat_params_[i],
                (i + 1) % MPMX_MATERIAL_PARAMS == 0 ? "]" : "");
  std::printf("]}\n");
}

template <unsigned Tdim>
bool MPMExplicit<Tdim>::solve() {
  read_analysis();
  initialise_materials();
  initialise_math_functions();
  initialise_mesh();
  initialise_loads();
  if (io_->dry_run()) {
    initialise_particles();  // counts only: no context
    summary();
    return true;
  }
  int scheme = MPMX_USF;
  if (scheme_ == "usl") scheme = MPMX_USL;
  else if (scheme_ == "musl") scheme = MPMX_MUSL;
  else if (scheme_ != "usf") throw std::runtime_error("mpm_scheme " + scheme_ + " is not supported");
  check(nullptr, mpmx_create(&ctx_, Tdim, scheme, vu_, blending_, io_->ordered() ? MPMX_ORDERED : 0u),
        "create solver");
  check(ctx_, mpmx_set_materials(ctx_, (int)mat_index_.size(), mat_params_.data()), "materials");
  check(ctx_, mpmx_set_mesh(ctx_, (int64_t)(coords_.size() / Tdim), coords_.data(),
                            (int64_t)(cells_.size() / NN), cells_.data(), isoparametric_ ? 1 : 0),
        "mesh");
  if (!enode_.empty())
    check(ctx_, mpmx_set_nodal_euler_angles(ctx_, (int64_t)enode_.size(), enode_.data(),
                                            eang_.data()),
          "Euler angles are not properly assigned/computed");
  if (!vnode_.empty())
    check(ctx_, mpmx_set_velocity_constraints(ctx_, (int64_t)vnode_.size(), vnode_.data(),
                                              vdir_.data(), vval_.data()),
          "velocity constraints");
  if (!fn_type_.empty())
    check(ctx_, mpmx_set_math_functions(ctx_, (int)fn_type_.size(), fn_type_.data(), fn_npts_.data(),
                                        fn_params_.data()),
          "math functions");
  if (!anode_.empty())
    check(ctx_, mpmx_set_acceleration_constraints_fn(ctx_, (int64_t)anode_.size(), anode_.data(),
                                                     adir_.data(), aval_.data(), afn_.data()),
          "acceleration constraints");
  if (!cfnode_.empty())
    check(ctx_, mpmx_set_nodal_forces(ctx_, (int64_t)cfnode_.size(), cfnode_.data(), cfdir_.data(),
                                      cfval_.data(), cffn_.data()),
          "concentrated nodal forces");
  if (!fnode_.empty())
    check(ctx_, mpmx_set_friction_constraints(ctx_, (int64_t)fnode_.size(), fnode_.data(),
                                              fdir_.data(), fsign_.data(), fmu_.data()),
          "friction constraints");
  check(ctx_, mpmx_set_gravity(ctx_, gravity_, ramp_), "gravity");
  check(ctx_, mpmx_set_locate_particles(ctx_, locate_particles_ ? 1 : 0), "locate_particles");
  if (resume_) {
    resume_particles();
  } else {
    initialise_particles();
    write_outputs(0);
  }
  // the step loop of mpm_explicit.tcc:112-190, in runs of steps between outputs
  int64_t step = step_;
  while (step < nsteps_) {
    const int64_t next = std::min(nsteps_, (step / output_steps_ + 1) * output_steps_);
    const int rc = mpmx_step(ctx_, (int)(next - step), dt_, damping_);
    if (rc == MPMX_E_UNLOCATABLE) throw std::runtime_error(mpmx_last_error(ctx_));
    check(ctx_, rc, "step");
    step = next;
    write_outputs(step);
  }
  check(ctx_, mpmx_synchronize(ctx_), "synchronize");
  return true;
}

}  // namespace mpm

// The analysis factory: the reference's keys (src/mpm.cc:26-32).
int main(int argc, char** argv) {
  try {
    auto io = std::make_shared<mpm::IO>(argc, argv);
    const std::map<std::string, std::function<bool()>> factory = {
        {"MPMExplicit2D", [&] { return mpm::MPMExplicit<2>(io).solve(); }},
        {"MPMExplicit3D", [&] { return mpm::MPMExplicit<3>(io).solve(); }},
    };
    const std::string type = io->analysis_type();
    auto it = factory.find(type);
    if (it == factory.end()) throw std::runtime_error("analysis type " + type + " is not supported");
    return it->second() ? 0 : 1;
  } catch (std::exception& e) {
    std::cerr << "MPM main: " << e.what() << std::endl;
    return 1;
  }
}
