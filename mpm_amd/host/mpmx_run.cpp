// mpmx_run: the explicit MPM analysis driver over the mpmx C ABI.
//
// Reads the reference's input deck -- the JSON file and the ASCII mesh,
// particle, constraint, stress and entity-set files it names (src/io/io.cc,
// include/io/io_mesh_ascii.tcc, include/solvers/mpm_base.tcc) -- and runs the
// analysis types MPMExplicit2D / MPMExplicit3D (mpm_explicit.tcc:24-199) with
// the time step on the GPU through include/mpmx.h. Particle states are
// written as text at the deck's output_steps (the reference writes HDF5/VTK).
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
  // objective stress rate (linear_elastic.tcc:8-17): "jaumann" for
  // LinearElastic; "green_naghdi" is passed on (the library rejects it)
  std::string rate = "none";
  if (m.contains("stress_rate") && m.at("stress_rate").is_string()) rate = m.at("stress_rate").str();
  if (type == "LinearElastic" + sfx) {
    p[0] = 0;
    p[1] = m.at("density").number();
    p[2] = m.at("youngs_modulus").number();
    p[3] = m.at("poisson_ratio").number();
    p[4] = rate == "jaumann" ? 1. : (rate == "green_naghdi" ? 2. : 0.);
    return p;
  }
  if (rate == "jaumann" || rate == "green_naghdi")
    throw std::runtime_error("objective stress rates are supported for LinearElastic only");
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
  std::map<int64_t, std::vector<int64_t>> node_sets_, cell_sets_, particle_sets_;
  std::vector<double> mat_params_;
  std::map<int64_t, int> mat_index_;
  std::vector<int64_t> vnode_, fnode_;
  std::vector<int64_t> pvc_particle_;
  std::vector<int> pvc_dir_;
  std::vector<double> pvc_val_;
  std::vector<int> vdir_, fdir_, fsign_;
  std::vector<double> vval_, fmu_;
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

  void read_analysis();
  void initialise_materials();
  void initialise_math_functions();
  void initialise_mesh();
  void initialise_particles();
  void initialise_loads();
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
    if (t == "pic") vu_ = MPMX_PIC;
    else if (t == "flip") vu_ = MPMX_FLIP;
    else throw std::runtime_error("velocity update " + t + " is not supported (flip, pic)");
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
  if (a.contains("resume") && a.at("resume").contains("resume") &&
      a.at("resume").at("resume").is_boolean() && a.at("resume").at("resume").boolean())
    throw std::runtime_error("resume from an HDF5 checkpoint is not supported");
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
  if (mp.contains("isoparametric") && mp.at("isoparametric").boolean())
    throw std::runtime_error("isoparametric cells are not supported; set \"isoparametric\": false");
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
      expected_particles_ += nc * (int64_t)std::pow(nppd, (int)Tdim);
      if (ctx_)
        check(ctx_, mpmx_generate_gauss(ctx_, nppd, cset >= 0 ? (int64_t)cells.size() : -1,
                                        cset >= 0 ? cells.data() : nullptr, mid),
              "generate gauss particles");
    } else if (type == "file") {
      const auto x = read_particles(io_->file_name(g.at("location").str()), Tdim);
      const int64_t n = (int64_t)(x.size() / Tdim);
      expected_particles_ += n;
      if (ctx_) {
        const std::vector<double> one(n, 1.0);
        const std::vector<int> mat(n, mid);
        check(ctx_, mpmx_set_particles(ctx_, n, x.data(), nullptr, nullptr, nullptr, nullptr,
                                       nullptr, one.data(), one.data(), one.data(), mat.data(),
                                       nullptr, nullptr),
              "particles from file");
      }
    } else {
      throw std::runtime_error("particle generator " + type + " is not supported (gauss, file)");
    }
  }
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
  if (!ctx_) return;
  check(ctx_, mpmx_compute_volume_mass(ctx_), "particle volume and mass");
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
    check(ctx_, mpmx_set_particle_tractions(ctx_, (int64_t)tp.size(), tp.data(), td.data(),
                                            tv.data(), tf.data()),
          "particle surface tractions");
  }
  if (!pvc_particle_.empty())
    check(ctx_, mpmx_set_particle_velocity_constraints(ctx_, (int64_t)pvc_particle_.size(),
                                                       pvc_particle_.data(), pvc_dir_.data(),
                                                       pvc_val_.data()),
          "particle velocity constraints");
  if (mp.contains("particles_stresses")) {
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

// Particle state as text: one row per particle in the reference's order.
template <unsigned Tdim>
void MPMExplicit<Tdim>::write_outputs(int64_t step) {
  int64_t n = 0;
  check(ctx_, mpmx_nparticles(ctx_, &n), "particle count");
  std::vector<double> x(n * Tdim), v(n * Tdim), s(n * 6), e(n * 6), vol(n), m(n), sv(n * 9);
  std::vector<int64_t> cell(n);
  check(ctx_, mpmx_get_particles(ctx_, x.data(), nullptr, v.data(), s.data(), e.data(), nullptr,
                                 m.data(), vol.data(), nullptr, nullptr, cell.data(), nullptr,
                                 sv.data()),
        "read particles");
  const std::string dir = io_->working_dir() + out_path_ + uuid_ + "/";
  std::string cur;
  for (char ch : dir) {
    cur += ch;
    if (ch == '/') mkdir(cur.c_str(), 0755);
  }
  const std::string path = dir + "particles" + std::to_string(step) + ".txt";
  FILE* f = std::fopen(path.c_str(), "w");
  if (!f) throw std::runtime_error("cannot write " + path);
  std::fprintf(f, "# step %lld particles %lld\n# id cell x[%u] v[%u] stress[6] strain[6] volume mass pdstrain\n",
               (long long)step, (long long)n, Tdim, Tdim);
  for (int64_t i = 0; i < n; ++i) {
    std::fprintf(f, "%lld %lld", (long long)i, (long long)cell[i]);
    for (unsigned k = 0; k < Tdim; ++k) std::fprintf(f, " %.17g", x[i * Tdim + k]);
    for (unsigned k = 0; k < Tdim; ++k) std::fprintf(f, " %.17g", v[i * Tdim + k]);
    for (int k = 0; k < 6; ++k) std::fprintf(f, " %.17g", s[i * 6 + k]);
    for (int k = 0; k < 6; ++k) std::fprintf(f, " %.17g", e[i * 6 + k]);
    std::fprintf(f, " %.17g %.17g %.17g\n", vol[i], m[i], sv[i * 9 + 8]);
  }
  std::fclose(f);
}

template <unsigned Tdim>
void MPMExplicit<Tdim>::summary() const {
  std::printf("{\"analysis\": \"MPMExplicit%uD\", \"scheme\": \"%s\", \"velocity_update\": \"%s\", "
              "\"dt\": %.17g, \"nsteps\": %lld, \"output_steps\": %lld, \"damping\": %.17g, "
              "\"nnodes\": %lld, \"ncells\": %lld, \"node_sets\": %zu, \"cell_sets\": %zu, "
              "\"velocity_constraints\": %zu, \"acceleration_constraints\": %zu, "
              "\"math_functions\": %zu, \"nodal_forces\": %zu, \"friction_constraints\": %zu, "
              "\"particle_velocity_constraints\": %zu, \"particles\": %lld, "
              "\"gravity\": [%.17g, %.17g, %.17g], \"ramping_time\": %.17g, \"materials\": [",
              Tdim, scheme_.c_str(), vu_ == MPMX_PIC ? "pic" : "flip", dt_, (long long)nsteps_,
              (long long)output_steps_, damping_, (long long)(coords_.size() / Tdim),
              (long long)(cells_.size() / NN), node_sets_.size(), cell_sets_.size(),
              vnode_.size(), anode_.size(), fn_type_.size(), cfnode_.size(), fnode_.size(),
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
                            (int64_t)(cells_.size() / NN), cells_.data(), 0),
        "mesh");
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
  initialise_particles();
  write_outputs(0);
  // the step loop of mpm_explicit.tcc:112-190, in runs of steps between outputs
  int64_t step = 0;
  while (step < nsteps_) {
    const int64_t next = std::min(nsteps_, (step / output_steps_ + 1) * output_steps_);
    const int rc = mpmx_step(ctx_, (int)(next - step), dt_, damping_);
    if (rc == MPMX_E_UNLOCATABLE) throw std::runtime_error(mpmx_last_error(ctx_));
    check(ctx_, rc, "step");
    step = next;
    if (step % output_steps_ == 0 || step == nsteps_) write_outputs(step);
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
