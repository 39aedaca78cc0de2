// Host side of the mpmx C-ABI: context, mesh preprocessing, particle
// initialisation, and the per-step launch sequence (captured once into a
// hipGraph and replayed per step).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mpmx.h"
#include "mpmx_internal.h"
#include "mpmx_adjoint.h"
#include "mpmx_kernels.h"
#include "mpmx_math.h"

using namespace mpmx;

namespace {

struct ApiError : std::runtime_error {
  int code;
  ApiError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHECK(expr)                                                                       \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      throw ApiError(MPMX_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));         \
  } while (0)

template <class T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    release();
    if (count == 0) return;
    HIPCHECK(hipMalloc((void**)&p, count * sizeof(T)));
    n = count;
  }
  void upload(const T* h, size_t count, hipStream_t st, size_t offset = 0) {
    if (count) HIPCHECK(hipMemcpyAsync(p + offset, h, count * sizeof(T), hipMemcpyHostToDevice, st));
  }
  void download(T* h, size_t count, hipStream_t st, size_t offset = 0) const {
    if (count) HIPCHECK(hipMemcpyAsync(h, p + offset, count * sizeof(T), hipMemcpyDeviceToHost, st));
  }
};

struct HostCons {
  std::map<int, double> vel, acc;  // dir -> value (first insert wins)
  std::map<int, int> accfn;        // dir -> math function of the acceleration constraint
  bool has_cf = false;             // concentrated force: per direction, one function
  double cf[3] = {0., 0., 0.};
  int cfn = -1;
  int fdir = -1, fsign = 0;
  double fmu = 0;
  bool rot = false;  // general boundary: constraints in the frame of R
  double R[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
};

struct KernelTime {
  double ms = 0;
  int64_t launches = 0;
};

}  // namespace

struct mpmx_ctx {
  int dim = 2, nn = 4;
  int scheme = USF;
  int vu = FLIP;
  double blending = 1.0;
  unsigned flags = 0;
  hipStream_t stream = nullptr;
  std::string error;

  // mesh (host copies for init-time geometry)
  int64_t nnodes = 0, ncells = 0;
  std::vector<double> node_x;  // [nnodes][dim]
  std::vector<int> cell_nodes;  // [ncells][nn]
  std::vector<double> cell_volume;
  DBuf<int> d_cell_nodes, d_nb_ptr, d_nb_idx, d_node_cells;
  DBuf<double> d_node_x, d_centroid, d_len, d_two_mean, d_dndx_c;
  // isoparametric mesh: local coordinates by transform_real_to_unit_cell
  // (cell.tcc:617-828), located by a kernel of its own each step
  bool iso = false;
  // instances (batched runs): cell ranges, cell -> instance
  int64_t ninst = 0;
  DBuf<int> d_cell_inst;
  DBuf<int64_t> d_inst_begin;

  // nodes
  DBuf<double> d_nmass, d_nmom, d_nvel, d_nacc, d_nfext, d_nfint;
  DBuf<unsigned char> d_nactive;
  DBuf<int> d_ncons;
  DBuf<NodeCons> d_table;
  std::map<int64_t, HostCons> cons;
  // math functions of time-dependent constraints and their values this step
  std::vector<MathFn> fns;
  DBuf<MathFn> d_fns;
  DBuf<double> d_fn_data, d_fnval;

  // materials
  std::vector<Material> mats;
  // per material: some particle was given strength state variables other
  // than the material's initial ones (Material::uniform then stays off)
  std::vector<char> sv_custom;
  DBuf<Material> d_mats;
  bool any_mc = false;
  bool any_rate = false;     // a material with an objective stress rate
  bool any_gn = false;       // ... with the Green-Naghdi rate
  bool any_mc_rate = false;  // a MohrCoulomb material with a stress rate

  // particle velocity constraints in the order given: (particle or -1 for
  // all, direction, value); applied per particle as a mask and values
  struct Pvc {
    int64_t particle;
    int dir;
    double value;
  };
  std::vector<Pvc> pvc;
  DBuf<unsigned char> d_pvc_mask;
  DBuf<double> d_pvc_val;
  // particle tractions (in order) and the particles' sizes pow(V0, 1/dim)
  struct Trac {
    int64_t particle;
    int dir;
    double value;
    int fn;
  };
  std::vector<Trac> trac;
  std::vector<double> psize;
  int ntr = 0;
  DBuf<int> d_tr_list, d_tr_fn;
  DBuf<double> d_tr_val, d_psize;

  // particles
  int64_t np = 0;
  // the reference's particle ids (Particle::id(): particles_.size() when the
  // particle was created, mesh.tcc:531-536, or given), in particle order
  std::vector<int64_t> ids;
  // locate_particles (mpm_base.tcc:53-54): false removes particles that leave
  // the mesh instead of failing the step (mpm_scheme.tcc:245-261)
  bool locate_all = true;
  DBuf<double> d_x, d_u, d_v, d_stress, d_strain, d_F, d_mass, d_vol, d_rho, d_xi, d_sv;
  // mapping matrices of the affine velocity updates (APIC, ASFLIP, TPIC)
  DBuf<double> d_Bmat;
  DBuf<int> d_mat, d_cell;

  // step machinery
  DBuf<int> d_skey, d_perm, d_sort_tmp, d_scan_list, d_cstart, d_cend;
  DBuf<double> d_spin;          // Particles::spin (any_rate)
  DBuf<double> d_rotF;          // Particles::rotF (any_gn)
  DBuf<double> d_plastic_trial; // Particles::plastic_trial (any_mc_rate)
  DBuf<double> d_Cmat, d_beta;  // affine transfer matrices, ASFLIP beta (affine())
  DBuf<double> d_gterm;         // Particles::gterm (ordered mode)
  DBuf<int> d_plastic_list;     // deferred Mohr-Coulomb return mappings
  DBuf<double> d_plastic_rec;
  DBuf<Status> d_status;
  Status* h_status = nullptr;  // pinned
  double g[3] = {0, 0, 0};
  double ramp = 0;
  int64_t step = 0;

  // particle order by cell: a full radix sort every `sort_every` steps (and
  // whenever the particle set changed); in between, fast mode reuses the last
  // permutation with refreshed keys
  int sort_every = 8;
  int chunk = 256;  // particles per workgroup of the LDS-staged scatters
  bool need_sort = true;
  int64_t since_sort = 0;

  // graph cache: [0] step with key refresh, [1] step with a full sort
  // [2], [3]: the same steps recording nodal quantities into the slots of
  // d_ad_node_slots (the slot index is Status::ad_slot)
  hipGraphExec_t graph_exec[4] = {nullptr, nullptr, nullptr, nullptr};
  double graph_dt = 0, graph_damping = 0;

  // adjoint: checkpoints of a recorded run (full particle state every
  // ad_every steps and at its end), the per-step records of one segment, node
  // and particle adjoint buffers
  bool ad_recorded = false;
  // the step records of the last segment were written by the recorded run
  // itself (the backward pass need not replay it)
  bool ad_last_recorded = false;
  int ad_every = 0, ad_nsteps = 0;
  double ad_dt = 0, ad_damping = 0;
  double ad_apex_ratio = 0.3;
  int64_t ad_step0 = 0, ad_np = 0;
  int64_t ad_ck_stride = 0;  // doubles per checkpoint
  DBuf<double> d_ad_ck, d_ad_slots, d_ad_node_slots, d_ad_nodes, d_ad_parts;
  DBuf<int> d_ad_ck_cell, d_ad_slot_cell, d_ad_list;
  DBuf<double> d_ad_list_buf;

  // device the context was created on, and the address there of the
  // file-scope scratch array of mpmx_step_noalloc (null: not resolved, with
  // the reason in scratch_error)
  int device = 0;
  void* scratch = nullptr;
  std::string scratch_error;

  // profiling
  bool profile = false;
  std::vector<std::string> kt_names;
  std::map<std::string, KernelTime> kt;

  ~mpmx_ctx() {
    for (auto& g : graph_exec)
      if (g) (void)hipGraphExecDestroy(g);
    if (h_status) (void)hipHostFree(h_status);
    if (stream) (void)hipStreamDestroy(stream);
  }

  // an affine velocity update: APIC, ASFLIP or TPIC
  bool affine() const { return vu == ASFLIP || vu == TPIC || vu == APIC; }

  void invalidate_graph() {
    for (auto& g : graph_exec) {
      if (g) (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
  }

  Mesh mesh() const {
    Mesh M;
    M.nnodes = nnodes;
    M.ncells = ncells;
    M.cell_nodes = d_cell_nodes.p;
    M.node_x = d_node_x.p;
    M.centroid = d_centroid.p;
    M.len = d_len.p;
    M.two_mean = d_two_mean.p;
    M.dndx_c = d_dndx_c.p;
    M.nb_ptr = d_nb_ptr.p;
    M.nb_idx = d_nb_idx.p;
    M.node_cells = d_node_cells.p;
    M.cell_inst = ninst > 0 ? d_cell_inst.p : nullptr;
    M.inst_begin = ninst > 0 ? d_inst_begin.p : nullptr;
    M.iso = iso ? 1 : 0;
    return M;
  }
  Particles parts() const {
    Particles P;
    P.np = np;
    P.x = d_x.p;
    P.u = d_u.p;
    P.v = d_v.p;
    P.stress = d_stress.p;
    P.strain = d_strain.p;
    P.F = d_F.p;
    P.mass = d_mass.p;
    P.vol = d_vol.p;
    P.rho = d_rho.p;
    P.xi = d_xi.p;
    P.sv = d_sv.p;
    P.mat = d_mat.p;
    P.cell = d_cell.p;
    P.pvc_mask = pvc.empty() ? nullptr : d_pvc_mask.p;
    P.pvc_val = pvc.empty() ? nullptr : d_pvc_val.p;
    P.spin = any_rate ? d_spin.p : nullptr;
    P.rotF = any_gn ? d_rotF.p : nullptr;
    P.plastic_trial = any_mc_rate ? d_plastic_trial.p : nullptr;
    P.gterm = (flags & MPMX_ORDERED) ? d_gterm.p : nullptr;
    P.Bmat = affine() ? d_Bmat.p : nullptr;
    P.Cmat = affine() ? d_Cmat.p : nullptr;
    P.beta = affine() ? d_beta.p : nullptr;
    P.plastic_list = any_mc ? d_plastic_list.p : nullptr;
    P.plastic_rec = any_mc ? d_plastic_rec.p : nullptr;
    P.ntr = ntr;
    P.tr_list = ntr ? d_tr_list.p : nullptr;
    P.tr_val = ntr ? d_tr_val.p : nullptr;
    P.tr_fn = ntr ? d_tr_fn.p : nullptr;
    P.psize = ntr ? d_psize.p : nullptr;
    return P;
  }
  Nodes nodes() const {
    Nodes N;
    N.mass = d_nmass.p;
    N.mom = d_nmom.p;
    N.vel = d_nvel.p;
    N.acc = d_nacc.p;
    N.fext = d_nfext.p;
    N.fint = d_nfint.p;
    N.active = d_nactive.p;
    N.cons = d_ncons.p;
    N.table = d_table.p;
    N.fnval = fns.empty() ? nullptr : d_fnval.p;
    return N;
  }
};

namespace {

// ----------------------------- cell geometry --------------------------------
// Same expressions and order as the reference's cell initialisation
// (cell.tcc:33-67, 254-277), element volume (quadrilateral_element.tcc:473-497,
// hexahedron_element.tcc:559-597) and dN/dx at the centroid.
double norm_diff(int dim, const double* a, const double* b) {
  double s = (a[0] - b[0]) * (a[0] - b[0]);
  for (int k = 1; k < dim; ++k) s = s + (a[k] - b[k]) * (a[k] - b[k]);
  return std::sqrt(s);
}

double quad_volume(const double X[8][3]) {
  const double a = norm_diff(2, X[0], X[3]);
  const double b = norm_diff(2, X[2], X[3]);
  const double c = norm_diff(2, X[1], X[2]);
  const double d = norm_diff(2, X[0], X[1]);
  const double p = norm_diff(2, X[0], X[2]);
  const double q = norm_diff(2, X[1], X[3]);
  const double t = (a * a + c * c - b * b - d * d);
  return 0.25 * std::sqrt(4 * p * p * q * q - t * t);
}

double hex_volume(const double X[8][3]) {
  auto sub = [](const double* u, const double* w, double* r) {
    for (int k = 0; k < 3; ++k) r[k] = u[k] - w[k];
  };
  auto cross = [](const double* u, const double* w, double* r) {
    r[0] = u[1] * w[2] - u[2] * w[1];
    r[1] = u[2] * w[0] - u[0] * w[2];
    r[2] = u[0] * w[1] - u[1] * w[0];
  };
  auto dot = [](const double* u, const double* w) { return (u[0] * w[0] + u[1] * w[1]) + u[2] * w[2]; };
  const double *a = X[7], *b = X[6], *c = X[2], *d = X[3], *e = X[4], *f = X[5], *g = X[1], *h = X[0];
  double bd[3], ca[3], eb[3], fa[3], de[3], ha[3], cg[3], cf[3], fg[3], hf[3], hg[3], hc[3];
  sub(b, d, bd); sub(c, a, ca); sub(e, b, eb); sub(f, a, fa); sub(d, e, de); sub(h, a, ha);
  sub(c, g, cg); sub(c, f, cf); sub(f, g, fg); sub(h, f, hf); sub(h, g, hg); sub(h, c, hc);
  double c1[3], c2[3], c3[3], c4[3], c5[3], c6[3], t[3], s[3];
  cross(bd, ca, c1); cross(eb, fa, c2); cross(de, ha, c3);
  cross(cg, cf, c4); cross(fg, hf, c5); cross(hg, hc, c6);
  sub(a, g, t);
  for (int k = 0; k < 3; ++k) s[k] = (c1[k] + c2[k]) + c3[k];
  const double v1 = (1.0 / 12) * dot(t, s);
  sub(b, g, t);
  for (int k = 0; k < 3; ++k) s[k] = c1[k] + c4[k];
  const double v2 = (1.0 / 12) * dot(t, s);
  sub(e, g, t);
  for (int k = 0; k < 3; ++k) s[k] = c2[k] + c5[k];
  const double v3 = (1.0 / 12) * dot(t, s);
  sub(d, g, t);
  for (int k = 0; k < 3; ++k) s[k] = c3[k] + c6[k];
  const double v4 = (1.0 / 12) * dot(t, s);
  return ((v1 + v2) + v3) + v4;
}

void host_grad(int dim, const double* xi, double* g /* [nn][3] */) {
  if (dim == 2) {
    g[0 * 3 + 0] = -0.25 * (1 - xi[1]); g[1 * 3 + 0] = 0.25 * (1 - xi[1]);
    g[2 * 3 + 0] = 0.25 * (1 + xi[1]);  g[3 * 3 + 0] = -0.25 * (1 + xi[1]);
    g[0 * 3 + 1] = -0.25 * (1 - xi[0]); g[1 * 3 + 1] = -0.25 * (1 + xi[0]);
    g[2 * 3 + 1] = 0.25 * (1 + xi[0]);  g[3 * 3 + 1] = 0.25 * (1 - xi[0]);
    return;
  }
  const double m0 = 1 - xi[0], p0 = 1 + xi[0], m1 = 1 - xi[1], p1 = 1 + xi[1], m2 = 1 - xi[2],
               p2 = 1 + xi[2];
  const double G[8][3] = {{-0.125 * m1 * m2, -0.125 * m0 * m2, -0.125 * m0 * m1},
                          {0.125 * m1 * m2, -0.125 * p0 * m2, -0.125 * p0 * m1},
                          {0.125 * p1 * m2, 0.125 * p0 * m2, -0.125 * p0 * p1},
                          {-0.125 * p1 * m2, 0.125 * m0 * m2, -0.125 * m0 * p1},
                          {-0.125 * m1 * p2, -0.125 * m0 * p2, 0.125 * m0 * m1},
                          {0.125 * m1 * p2, -0.125 * p0 * p2, 0.125 * p0 * m1},
                          {0.125 * p1 * p2, 0.125 * p0 * p2, 0.125 * p0 * p1},
                          {-0.125 * p1 * p2, 0.125 * m0 * p2, 0.125 * m0 * p1}};
  for (int i = 0; i < 8; ++i)
    for (int k = 0; k < 3; ++k) g[i * 3 + k] = G[i][k];
}

void host_dndx(int dim, const double* xi, const double X[8][3], double out[8][3]) {
  const int nn = dim == 2 ? 4 : 8;
  double g[24];
  host_grad(dim, xi, g);
  double J[3][3] = {{0}};
  for (int a = 0; a < dim; ++a)
    for (int b = 0; b < dim; ++b) {
      double s = g[0 * 3 + a] * X[0][b];
      for (int i = 1; i < nn; ++i) s = s + g[i * 3 + a] * X[i][b];
      J[a][b] = s;
    }
  double Ji[3][3];
  if (dim == 2) {
    const double det = J[0][0] * J[1][1] - J[1][0] * J[0][1];
    const double inv = 1.0 / det;
    Ji[0][0] = J[1][1] * inv;
    Ji[1][0] = -J[1][0] * inv;
    Ji[0][1] = -J[0][1] * inv;
    Ji[1][1] = J[0][0] * inv;
  } else {
    double c[3][3];
    c[0][0] = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    c[1][0] = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    c[2][0] = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    c[0][1] = J[2][1] * J[0][2] - J[2][2] * J[0][1];
    c[1][1] = J[2][2] * J[0][0] - J[2][0] * J[0][2];
    c[2][1] = J[2][0] * J[0][1] - J[2][1] * J[0][0];
    c[0][2] = J[0][1] * J[1][2] - J[0][2] * J[1][1];
    c[1][2] = J[0][2] * J[1][0] - J[0][0] * J[1][2];
    c[2][2] = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    const double det = (J[0][0] * c[0][0] + J[1][0] * c[0][1]) + J[2][0] * c[0][2];
    const double inv = 1.0 / det;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) Ji[i][j] = c[i][j] * inv;
  }
  for (int i = 0; i < nn; ++i)
    for (int a = 0; a < dim; ++a) {
      double s = g[i * 3 + 0] * Ji[a][0];
      for (int b = 1; b < dim; ++b) s = s + g[i * 3 + b] * Ji[a][b];
      out[i][a] = s;
    }
}

void host_shapefn(int dim, const double* xi, double* N) {
  if (dim == 2) {
    N[0] = 0.25 * (1 - xi[0]) * (1 - xi[1]);
    N[1] = 0.25 * (1 + xi[0]) * (1 - xi[1]);
    N[2] = 0.25 * (1 + xi[0]) * (1 + xi[1]);
    N[3] = 0.25 * (1 - xi[0]) * (1 + xi[1]);
    return;
  }
  N[0] = 0.125 * (1 - xi[0]) * (1 - xi[1]) * (1 - xi[2]);
  N[1] = 0.125 * (1 + xi[0]) * (1 - xi[1]) * (1 - xi[2]);
  N[2] = 0.125 * (1 + xi[0]) * (1 + xi[1]) * (1 - xi[2]);
  N[3] = 0.125 * (1 - xi[0]) * (1 + xi[1]) * (1 - xi[2]);
  N[4] = 0.125 * (1 - xi[0]) * (1 - xi[1]) * (1 + xi[2]);
  N[5] = 0.125 * (1 + xi[0]) * (1 - xi[1]) * (1 + xi[2]);
  N[6] = 0.125 * (1 + xi[0]) * (1 + xi[1]) * (1 + xi[2]);
  N[7] = 0.125 * (1 - xi[0]) * (1 + xi[1]) * (1 + xi[2]);
}

// Gauss points of the reference's quadratures, in their order
// (quadrilateral_quadrature.tcc, hexahedron_quadrature.tcc): 1 per direction:
// the centre; 2: the element's node order; 3: quadrilateral corners, edge
// midpoints (-y, +x, +y, -x) and centre, hexahedron xi0 slowest and xi2
// fastest; 4: quadrilateral xi0 slowest, hexahedron xi2 slowest, then xi0,
// then xi1. Returns the number of points.
int gauss_points(int d, int nppd, double (*qp)[3]) {
  for (int i = 0; i < 64; ++i) qp[i][0] = qp[i][1] = qp[i][2] = 0.;
  switch (nppd) {
    case 1:
      return 1;
    case 2: {
      const double g = 1. / std::sqrt(3.);
      const int n = d == 2 ? 4 : 8;
      for (int i = 0; i < n; ++i) {
        const int l = i & 3;
        qp[i][0] = (l == 0 || l == 3) ? -g : g;
        qp[i][1] = l < 2 ? -g : g;
        if (d == 3) qp[i][2] = i < 4 ? -g : g;
      }
      return n;
    }
    case 3: {
      const double a = std::sqrt(3. / 5.);
      const double c3[3] = {-a, 0., a};
      if (d == 2) {
        // corners in node order, then the edge midpoints, then the centre
        const int ix[9] = {0, 2, 2, 0, 1, 2, 1, 0, 1};
        const int iy[9] = {0, 0, 2, 2, 0, 1, 2, 1, 1};
        for (int q = 0; q < 9; ++q) {
          qp[q][0] = c3[ix[q]];
          qp[q][1] = c3[iy[q]];
        }
        return 9;
      }
      for (int q = 0; q < 27; ++q) {
        qp[q][0] = c3[q / 9];
        qp[q][1] = c3[(q / 3) % 3];
        qp[q][2] = c3[q % 3];
      }
      return 27;
    }
    case 4: {
      const double s65 = std::sqrt(6. / 5.);
      const double a = std::sqrt(3. / 7. - 2. / 7. * s65);
      const double b = std::sqrt(3. / 7. + 2. / 7. * s65);
      const double g4[4] = {-b, -a, a, b};
      const int n = d == 2 ? 16 : 64;
      for (int q = 0; q < n; ++q) {
        qp[q][0] = g4[(q / 4) % 4];
        qp[q][1] = g4[q % 4];
        if (d == 3) qp[q][2] = g4[q / 16];
      }
      return n;
    }
    default:
      throw ApiError(MPMX_E_UNSUPPORTED, "gauss generator supports nparticles_per_dir 1 to 4");
  }
}

// ----------------------------- materials ------------------------------------
Material make_material(const double* p) {
  Material m;
  std::memset(&m, 0, sizeof(m));
  m.type = (int)p[0];
  if (m.type != LE && m.type != MC) throw ApiError(MPMX_E_ARG, "unknown material type");
  m.density = p[1];
  const double E = p[2], nu = p[3];
  m.E = E;
  if (!(m.density > 0) || !(E > 0) || !(nu > -1.0 && nu < 0.5))
    throw ApiError(MPMX_E_ARG, "material density, youngs_modulus or poisson_ratio out of range");
  // objective stress rate (linear_elastic.tcc:8-17; MohrCoulomb inherits it,
  // infinitesimal_elasto_plastic.tcc:1-25): parameter 4 of LinearElastic,
  // 16 of MohrCoulomb
  m.rate = (int)p[m.type == LE ? 4 : 16];
  if (m.rate != RATE_NONE && m.rate != RATE_JAUMANN && m.rate != RATE_GREEN_NAGHDI)
    throw ApiError(MPMX_E_ARG, "unknown stress rate");
  if (m.type == LE) {
    const double K = E / (3.0 * (1. - 2. * nu));  // linear_elastic.tcc:27
    const double G = E / (2.0 * (1. + nu));       // linear_elastic.tcc:55
    m.a1 = K + (4.0 / 3.0) * G;
    m.a2 = K - (2.0 / 3.0) * G;
    m.G = G;
    return m;
  }
  const double packing = p[15] > 0. ? p[15] : 1.;     // mohr_coulomb.tcc:12-31
  const double grain_density = m.density / packing;
  m.critical_density = p[14] * grain_density;
  m.softening = p[13] != 0. ? 1 : 0;
  m.phi_peak = p[4] * M_PI / 180.;
  m.psi_peak = p[5] * M_PI / 180.;
  m.c_peak = p[6];
  m.phi_res = p[7] * M_PI / 180.;
  m.psi_res = p[8] * M_PI / 180.;
  m.c_res = p[9];
  m.pd_peak = p[10];
  m.pd_res = p[11];
  const double tcut = p[12];
  const double K = E / (3.0 * (1. - 2. * nu));  // mohr_coulomb.tcc:75-77
  m.shear = E / (2.0 * (1 + nu));
  m.a1 = K + (4.0 / 3.0) * m.shear;
  m.a2 = K - (2.0 / 3.0) * m.shear;
  m.G = m.shear;
  dmath::sincos(m.phi_peak, &m.sphi, &m.cphi);
  m.tphi = m.sphi / m.cphi;
  m.tpsi = dmath::tan(m.psi_peak);
  // initialise_state_variables (mohr_coulomb.tcc:85-111)
  m.sv0[SV_YIELD] = 0;
  m.sv0[SV_PHI] = m.phi_peak;
  m.sv0[SV_PSI] = m.psi_peak;
  m.sv0[SV_COH] = m.c_peak;
  const double apex = m.c_peak / dmath::tan(m.phi_peak);
  const double tc = (tcut < apex) ? tcut : apex;
  m.sv0[SV_TCUT] = tc > 1.0e-15 ? tc : 1.0e-15;
  // the cutoff follows the friction angle when it is the apex c / tan(phi)
  if (!(tcut < apex) && apex > 1.0e-15) {
    const double sphi = dmath::sin(m.phi_peak);
    m.dtcut_dphi = -m.c_peak / (sphi * sphi);
  }
  return m;
}

}  // namespace

// =============================== API ========================================
#define API_BEGIN(ctx)      \
  if (!(ctx)) return MPMX_E_ARG; \
  try {
#define API_END(ctx)                      \
  return MPMX_OK;                         \
  }                                       \
  catch (ApiError & e) {                  \
    (ctx)->error = e.what();              \
    return e.code;                        \
  }                                       \
  catch (std::exception & e) {            \
    (ctx)->error = e.what();              \
    return MPMX_E_ARG;                    \
  }

namespace {

void require_mesh(mpmx_ctx* c) {
  if (c->nnodes == 0) throw ApiError(MPMX_E_STATE, "mesh is not set");
}

// Inverse of the leading dim x dim block by the adjugate over the
// determinant (cofactor expansion along the first column), what
// rotation_matrix_.inverse() computes (node.tcc:388-389).
void matrix_inverse(int dim, const double (*J)[3], double (*Ji)[3]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) Ji[i][j] = i == j ? 1. : 0.;
  if (dim == 2) {
    const double det = J[0][0] * J[1][1] - J[1][0] * J[0][1];
    const double invdet = 1.0 / det;
    Ji[0][0] = J[1][1] * invdet;
    Ji[1][0] = -J[1][0] * invdet;
    Ji[0][1] = -J[0][1] * invdet;
    Ji[1][1] = J[0][0] * invdet;
    return;
  }
  double cf[3][3];
  cf[0][0] = J[1][1] * J[2][2] - J[1][2] * J[2][1];
  cf[1][0] = J[1][2] * J[2][0] - J[1][0] * J[2][2];
  cf[2][0] = J[1][0] * J[2][1] - J[1][1] * J[2][0];
  cf[0][1] = J[2][1] * J[0][2] - J[2][2] * J[0][1];
  cf[1][1] = J[2][2] * J[0][0] - J[2][0] * J[0][2];
  cf[2][1] = J[2][0] * J[0][1] - J[2][1] * J[0][0];
  cf[0][2] = J[0][1] * J[1][2] - J[0][2] * J[1][1];
  cf[1][2] = J[0][2] * J[1][0] - J[0][0] * J[1][2];
  cf[2][2] = J[0][0] * J[1][1] - J[0][1] * J[1][0];
  const double det = (J[0][0] * cf[0][0] + J[1][0] * cf[0][1]) + J[2][0] * cf[0][2];
  const double invdet = 1.0 / det;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) Ji[i][j] = cf[i][j] * invdet;
}

// mpm::geometry::rotation_matrix (src/geometry.cc:6-56) of Euler angles
// (alpha, beta) in 2D, (alpha, beta, gamma) in 3D.
void euler_rotation(int dim, const double* ang, double R[3][3]) {
  const double ca = dmath::cos(ang[0]), sa = dmath::sin(ang[0]);
  const double cb = dmath::cos(ang[1]), sb = dmath::sin(ang[1]);
  const double cacb = ca * cb, casb = ca * sb, sacb = sa * cb, sasb = sa * sb;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) R[i][j] = i == j ? 1. : 0.;
  if (dim == 2) {
    R[0][0] = cacb - sasb;
    R[0][1] = -casb - sacb;
    R[1][0] = sacb + casb;
    R[1][1] = -sasb + cacb;
    return;
  }
  const double sg = dmath::sin(ang[2]), cg = dmath::cos(ang[2]);
  R[0][0] = cacb - sasb * cg;
  R[0][1] = -casb - sacb * cg;
  R[0][2] = sa * sg;
  R[1][0] = sacb + casb * cg;
  R[1][1] = -sasb + cacb * cg;
  R[1][2] = -(ca * sg);
  R[2][0] = sb * sg;
  R[2][1] = cb * sg;
  R[2][2] = cg;
}

void upload_constraints(mpmx_ctx* c) {
  std::vector<int> idx(c->nnodes, -1);
  std::vector<NodeCons> table;
  for (auto& kv : c->cons) {
    NodeCons nc;
    std::memset(&nc, 0, sizeof(nc));
    nc.fdir = -1;
    nc.afn[0] = nc.afn[1] = nc.afn[2] = -1;
    nc.has_cf = kv.second.has_cf ? 1 : 0;
    nc.cfn = kv.second.cfn;
    for (int k = 0; k < 3; ++k) nc.cf[k] = kv.second.cf[k];
    for (auto& v : kv.second.vel) {
      nc.vmask |= 1 << v.first;
      nc.vval[v.first] = v.second;
    }
    for (auto& a : kv.second.acc) {
      nc.amask |= 1 << a.first;
      nc.aval[a.first] = a.second;
      const auto f = kv.second.accfn.find(a.first);
      if (f != kv.second.accfn.end()) nc.afn[a.first] = f->second;
    }
    if (kv.second.fdir >= 0) {
      nc.fdir = kv.second.fdir;
      nc.fsign = kv.second.fsign > 0 ? 1. : -1.;
      nc.fmu = kv.second.fmu;
    }
    nc.rot = kv.second.rot ? 1 : 0;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) nc.R[i][j] = kv.second.R[i][j];
    if (kv.second.rot) matrix_inverse(c->dim, kv.second.R, nc.Ri);
    idx[kv.first] = (int)table.size();
    table.push_back(nc);
  }
  c->d_ncons.alloc(c->nnodes);
  c->d_ncons.upload(idx.data(), idx.size(), c->stream);
  c->d_table.alloc(std::max<size_t>(table.size(), 1));
  c->d_table.upload(table.data(), table.size(), c->stream);
  HIPCHECK(hipStreamSynchronize(c->stream));
  c->invalidate_graph();
}

// Grow every particle array to hold `n` particles, keeping the first `keep`.
template <class T>
void grow(DBuf<T>& b, int64_t keep, int64_t n, int comps, hipStream_t st) {
  DBuf<T> nb;
  nb.alloc((size_t)n * comps);
  for (int k = 0; k < comps; ++k)
    if (keep > 0)
      HIPCHECK(hipMemcpyAsync(nb.p + (size_t)k * n, b.p + (size_t)k * keep, keep * sizeof(T),
                              hipMemcpyDeviceToDevice, st));
  HIPCHECK(hipStreamSynchronize(st));
  std::swap(b.p, nb.p);
  std::swap(b.n, nb.n);
}

// Upload AoS rows [n][comps] of the new particles into SoA [comps][np] at offset.
template <class T>
void put_soa(DBuf<T>& b, int64_t np, int64_t off, int64_t n, int comps, const T* rows,
             hipStream_t st) {
  std::vector<T> col(n);
  for (int k = 0; k < comps; ++k) {
    for (int64_t i = 0; i < n; ++i) col[i] = rows[i * comps + k];
    b.upload(col.data(), n, st, (size_t)k * np + off);
    HIPCHECK(hipStreamSynchronize(st));
  }
}

template <class T>
void get_soa(const DBuf<T>& b, int64_t np, int comps, T* rows, hipStream_t st) {
  std::vector<T> col(np);
  for (int k = 0; k < comps; ++k) {
    b.download(col.data(), np, st, (size_t)k * np);
    HIPCHECK(hipStreamSynchronize(st));
    for (int64_t i = 0; i < np; ++i) rows[i * comps + k] = col[i];
  }
}

// Per-particle mask and values of the particle velocity constraints, in the
// order given (a later constraint on the same particle and direction wins, as
// Mesh::apply_particle_velocity_constraints applies them in sequence,
// mesh.tcc:1454-1468).
void build_pvc(mpmx_ctx* c) {
  if (c->pvc.empty() || c->np == 0) return;
  const int64_t np = c->np;
  std::vector<unsigned char> mask(np, 0);
  std::vector<double> val((size_t)c->dim * np, 0.0);
  for (const auto& q : c->pvc) {
    const int64_t b = q.particle < 0 ? 0 : q.particle;
    const int64_t e = q.particle < 0 ? np : std::min<int64_t>(q.particle + 1, np);
    for (int64_t p = b; p < e; ++p) {
      mask[p] |= (unsigned char)(1u << q.dir);
      val[(size_t)q.dir * np + p] = q.value;
    }
  }
  c->d_pvc_mask.alloc(np);
  c->d_pvc_val.alloc((size_t)c->dim * np);
  c->d_pvc_mask.upload(mask.data(), np, c->stream);
  c->d_pvc_val.upload(val.data(), (size_t)c->dim * np, c->stream);
  HIPCHECK(hipStreamSynchronize(c->stream));
}

// Per particle and direction the last traction load that covers it
// (Particle::assign_traction overwrites per direction), the list of the
// particles with any, and the particle sizes.
void build_trac(mpmx_ctx* c) {
  c->ntr = 0;
  c->invalidate_graph();
  if (c->trac.empty() || c->np == 0) return;
  const int64_t np = c->np;
  const int d = c->dim;
  std::vector<int> fn((size_t)d * np, -2);
  std::vector<double> val((size_t)d * np, 0.0);
  for (const auto& t : c->trac) {
    const int64_t b = t.particle < 0 ? 0 : t.particle;
    const int64_t e = t.particle < 0 ? np : std::min<int64_t>(t.particle + 1, np);
    for (int64_t p = b; p < e; ++p) {
      fn[(size_t)t.dir * np + p] = t.fn;
      val[(size_t)t.dir * np + p] = t.value;
    }
  }
  std::vector<int> list;
  for (int64_t p = 0; p < np; ++p) {
    bool any = false;
    for (int k = 0; k < d; ++k) any = any || fn[(size_t)k * np + p] >= -1;
    if (any) list.push_back((int)p);
  }
  if (list.empty()) return;
  c->d_tr_list.alloc(list.size());
  c->d_tr_list.upload(list.data(), list.size(), c->stream);
  c->d_tr_fn.alloc((size_t)d * np);
  c->d_tr_fn.upload(fn.data(), (size_t)d * np, c->stream);
  c->d_tr_val.alloc((size_t)d * np);
  c->d_tr_val.upload(val.data(), (size_t)d * np, c->stream);
  c->d_psize.alloc(np);
  c->d_psize.upload(c->psize.data(), np, c->stream);
  HIPCHECK(hipStreamSynchronize(c->stream));
  c->ntr = (int)list.size();
}

// Particle size of a volume (Particle::assign_volume, particle.tcc:556-568).
double particle_size(double volume, int dim) {
  return std::pow(volume, static_cast<double>(1. / dim));
}

void alloc_step_buffers(mpmx_ctx* c) {
  const int64_t np = c->np;
  c->d_skey.alloc(std::max<int64_t>(np, 1));
  c->d_perm.alloc(std::max<int64_t>(np, 1));
  c->d_sort_tmp.alloc(radix_sort_temp_ints(std::max<int64_t>(np, 1)));
  c->d_scan_list.alloc(std::max<int64_t>(np, 1));
  if (c->any_rate) c->d_spin.alloc(3 * std::max<int64_t>(np, 1));
  if (c->any_gn) c->d_rotF.alloc((size_t)(c->dim * 3 + c->dim * c->dim) * std::max<int64_t>(np, 1));
  if (c->any_mc_rate) c->d_plastic_trial.alloc(6 * std::max<int64_t>(np, 1));
  if (c->affine()) {
    c->d_Cmat.alloc((size_t)c->dim * c->dim * std::max<int64_t>(np, 1));
    c->d_beta.alloc(std::max<int64_t>(np, 1));
  }
  if (c->any_mc) {
    c->d_plastic_list.alloc(std::max<int64_t>(np, 1));
    c->d_plastic_rec.alloc(PR_N * std::max<int64_t>(np, 1));
  }
  c->need_sort = true;
  if (c->flags & MPMX_ORDERED) {
    c->d_cstart.alloc(std::max<int64_t>(c->ncells, 1));
    c->d_cend.alloc(std::max<int64_t>(c->ncells, 1));
    c->d_gterm.alloc((size_t)c->nn * (1 + c->dim) * std::max<int64_t>(np, 1));
  }
}

void reset_status(mpmx_ctx* c) {
  Status s;
  s.step = c->step;
  s.first_bad_step = std::numeric_limits<int64_t>::max();
  s.bad_particle = -1;
  s.nscan = 0;
  s.nplastic = 0;
  s.nplastic_t = 0;
  s.ad_slot = 0;
  s.nlost = 0;
  *c->h_status = s;
  c->d_status.upload(c->h_status, 1, c->stream);
}

void read_status(mpmx_ctx* c) {
  c->d_status.download(c->h_status, 1, c->stream);
  HIPCHECK(hipStreamSynchronize(c->stream));
}

// Run the initial locate of particles [off, off+n) (hint: cell per new
// particle or -1) and fail with MPMX_E_UNLOCATABLE if any left the mesh.
void locate_new(mpmx_ctx* c, int64_t off, int64_t n, const std::vector<int>& hint) {
  Particles P = c->parts();
  // operate on the sub-range by offsetting pointers: SoA strides stay np
  DBuf<int> d_hint;
  d_hint.alloc(std::max<int64_t>(c->np, 1));
  std::vector<int> full(c->np, -1);
  for (int64_t i = 0; i < n; ++i) full[off + i] = hint[i];
  // particles before `off` keep their cells: give them their own cell as hint
  std::vector<int> cells(c->np, -1);
  if (off > 0) {
    c->d_cell.download(cells.data(), off, c->stream);
    HIPCHECK(hipStreamSynchronize(c->stream));
    for (int64_t i = 0; i < off; ++i) full[i] = cells[i];
  }
  d_hint.upload(full.data(), c->np, c->stream);
  reset_status(c);
  Mesh M = c->mesh();
  if (c->dim == 2) {
    launch_locate_init<2>(M, P, c->d_status.p, c->d_scan_list.p, d_hint.p, c->stream);
    launch_locate_scan<2>(M, P, c->d_status.p, c->d_scan_list.p, d_hint.p, 0, c->stream);
  } else {
    launch_locate_init<3>(M, P, c->d_status.p, c->d_scan_list.p, d_hint.p, c->stream);
    launch_locate_scan<3>(M, P, c->d_status.p, c->d_scan_list.p, d_hint.p, 0, c->stream);
  }
  HIPCHECK(hipGetLastError());
  read_status(c);
  if (c->h_status->first_bad_step != std::numeric_limits<int64_t>::max())
    throw ApiError(MPMX_E_UNLOCATABLE,
                   "Particle outside the mesh domain: " + std::to_string(c->h_status->bad_particle));
}

// Material::uniform from the materials and the strength state variables the
// particles were given; uploads the table when it changed (or `force`).
void refresh_uniform(mpmx_ctx* c, bool force = false) {
  bool changed = force;
  for (size_t i = 0; i < c->mats.size(); ++i) {
    Material& m = c->mats[i];
    const int u = (m.type == MC && !m.softening && !c->sv_custom[i]) ? 1 : 0;
    if (u != m.uniform) changed = true;
    m.uniform = u;
  }
  if (!changed) return;
  c->d_mats.upload(c->mats.data(), c->mats.size(), c->stream);
  HIPCHECK(hipStreamSynchronize(c->stream));
}

// Append n particles (rows given on the host); returns the offset.
int64_t append_particles(mpmx_ctx* c, int64_t n, const double* x, const double* u, const double* v,
                         const double* stress, const double* strain, const double* F,
                         const double* mass, const double* volume, const double* density,
                         const int* mat, const double* sv) {
  const int d = c->dim;
  const int64_t off = c->np, np = off + n;
  if (np > (int64_t)std::numeric_limits<int>::max() / 2)
    throw ApiError(MPMX_E_ARG, "too many particles for 32-bit particle ids");
  for (int64_t i = 0; i < n; ++i)
    if (mat[i] < 0 || mat[i] >= (int)c->mats.size()) throw ApiError(MPMX_E_ARG, "bad material id");
  hipStream_t st = c->stream;
  grow(c->d_x, off, np, d, st);
  grow(c->d_u, off, np, d, st);
  grow(c->d_v, off, np, d, st);
  grow(c->d_xi, off, np, d, st);
  grow(c->d_stress, off, np, 6, st);
  grow(c->d_strain, off, np, 6, st);
  grow(c->d_F, off, np, 9, st);
  grow(c->d_mass, off, np, 1, st);
  grow(c->d_vol, off, np, 1, st);
  grow(c->d_rho, off, np, 1, st);
  grow(c->d_mat, off, np, 1, st);
  grow(c->d_cell, off, np, 1, st);
  if (c->any_mc) grow(c->d_sv, off, np, NSV, st);
  grow(c->d_Bmat, off, np, d * d, st);
  c->np = np;
  c->ad_recorded = false;
  c->psize.resize(np);
  for (int64_t i = 0; i < n; ++i) c->psize[off + i] = particle_size(volume ? volume[i] : 1.0, d);
  c->ids.resize(np);
  for (int64_t i = 0; i < n; ++i) c->ids[off + i] = off + i;  // particles_.size() at creation
  build_pvc(c);
  build_trac(c);
  std::vector<double> zeros((size_t)n * 9, 0.0), eye((size_t)n * 9, 0.0);
  for (int64_t i = 0; i < n; ++i) eye[i * 9 + 0] = eye[i * 9 + 4] = eye[i * 9 + 8] = 1.0;
  put_soa(c->d_x, np, off, n, d, x, st);
  put_soa(c->d_u, np, off, n, d, u ? u : zeros.data(), st);
  put_soa(c->d_v, np, off, n, d, v ? v : zeros.data(), st);
  put_soa(c->d_xi, np, off, n, d, zeros.data(), st);
  put_soa(c->d_Bmat, np, off, n, d * d, zeros.data(), st);  // mapping_matrix_ starts at 0
  put_soa(c->d_stress, np, off, n, 6, stress ? stress : zeros.data(), st);
  put_soa(c->d_strain, np, off, n, 6, strain ? strain : zeros.data(), st);
  put_soa(c->d_F, np, off, n, 9, F ? F : eye.data(), st);
  put_soa(c->d_mass, np, off, n, 1, mass, st);
  put_soa(c->d_vol, np, off, n, 1, volume, st);
  put_soa(c->d_rho, np, off, n, 1, density, st);
  put_soa(c->d_mat, np, off, n, 1, mat, st);
  std::vector<int> neg(n, -1);
  put_soa(c->d_cell, np, off, n, 1, neg.data(), st);
  if (c->any_mc) {
    std::vector<double> s((size_t)n * NSV, 0.0);
    for (int64_t i = 0; i < n; ++i) {
      const Material& m = c->mats[mat[i]];
      for (int k = 0; k < NSV; ++k)
        s[i * NSV + k] = (sv && m.type == MC) ? sv[i * NSV + k] : m.sv0[k];
      if (m.type == MC)
        for (const int k : {SV_PHI, SV_PSI, SV_COH, SV_TCUT})
          if (s[i * NSV + k] != m.sv0[k]) c->sv_custom[mat[i]] = 1;
    }
    put_soa(c->d_sv, np, off, n, NSV, s.data(), st);
    refresh_uniform(c);
  }
  alloc_step_buffers(c);
  c->invalidate_graph();
  return off;
}

// Remove particles [off, np) again (after a failed initial locate).
void truncate_particles(mpmx_ctx* c, int64_t keep) {
  if (keep == c->np) return;
  hipStream_t st = c->stream;
  const int d = c->dim;
  const int64_t old = c->np;
  auto shrink = [&](auto& b, int comps) {
    using T = typename std::remove_reference<decltype(*b.p)>::type;
    DBuf<T> nb;
    nb.alloc((size_t)std::max<int64_t>(keep, 1) * comps);
    for (int k = 0; k < comps; ++k)
      if (keep > 0)
        HIPCHECK(hipMemcpyAsync(nb.p + (size_t)k * keep, b.p + (size_t)k * old, keep * sizeof(T),
                                hipMemcpyDeviceToDevice, st));
    HIPCHECK(hipStreamSynchronize(st));
    std::swap(b.p, nb.p);
    std::swap(b.n, nb.n);
  };
  shrink(c->d_x, d); shrink(c->d_u, d); shrink(c->d_v, d); shrink(c->d_xi, d);
  shrink(c->d_stress, 6); shrink(c->d_strain, 6); shrink(c->d_F, 9);
  shrink(c->d_mass, 1); shrink(c->d_vol, 1); shrink(c->d_rho, 1);
  shrink(c->d_mat, 1); shrink(c->d_cell, 1);
  if (c->any_mc) shrink(c->d_sv, NSV);
  shrink(c->d_Bmat, d * d);
  c->np = keep;
  c->ad_recorded = false;
  c->psize.resize(keep);
  c->ids.resize(keep);
  build_pvc(c);
  build_trac(c);
  alloc_step_buffers(c);
  c->invalidate_graph();
}

// ------------------------------ step sequence --------------------------------
// The device scratch of the step kernels (what they use beyond the mesh,
// node and particle state): the context's own buffers.
struct StepBuffers {
  Status* S;
  int* skey;
  int* perm;
  int* sort_tmp;
  int* scan;
  int* cstart;
  int* cend;
  int* plastic_list;
  double* plastic_rec;
  double* spin;
  double* fnval;  // the math functions' values at the step's time
  double* Cmat;   // affine velocity updates: the transfer matrices, beta
  double* beta;
  double* rotF;   // Green-Naghdi scratch; rotated trial stresses of the deferred list
  double* plastic_trial;
  double* gterm;  // ordered mode: per-particle gather terms
};

StepBuffers ctx_buffers(mpmx_ctx* c) {
  StepBuffers B;
  B.S = c->d_status.p;
  B.skey = c->d_skey.p;
  B.perm = c->d_perm.p;
  B.sort_tmp = c->d_sort_tmp.p;
  B.scan = c->d_scan_list.p;
  B.cstart = c->d_cstart.p;
  B.cend = c->d_cend.p;
  B.plastic_list = c->d_plastic_list.p;
  B.plastic_rec = c->d_plastic_rec.p;
  B.spin = c->d_spin.p;
  B.fnval = c->d_fnval.p;
  B.Cmat = c->d_Cmat.p;
  B.beta = c->d_beta.p;
  B.rotF = c->d_rotF.p;
  B.plastic_trial = c->d_plastic_trial.p;
  B.gterm = c->d_gterm.p;
  return B;
}

struct Launcher {
  mpmx_ctx* c;
  bool timed = true;  // events around each launch while profiling is enabled
  std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> evs;
  template <class F>
  void operator()(const char* name, F&& f) {
    if (!c->profile || !timed) {
      f();
      return;
    }
    hipEvent_t a, b;
    HIPCHECK(hipEventCreate(&a));
    HIPCHECK(hipEventCreate(&b));
    HIPCHECK(hipEventRecord(a, c->stream));
    f();
    HIPCHECK(hipEventRecord(b, c->stream));
    evs.push_back({name, {a, b}});
  }
  void collect() {
    if (evs.empty()) return;
    HIPCHECK(hipStreamSynchronize(c->stream));
    for (auto& e : evs) {
      float ms = 0;
      HIPCHECK(hipEventElapsedTime(&ms, e.second.first, e.second.second));
      auto it = c->kt.find(e.first);
      if (it == c->kt.end()) {
        c->kt_names.push_back(e.first);
        it = c->kt.emplace(e.first, KernelTime()).first;
      }
      it->second.ms += ms;
      it->second.launches += 1;
      (void)hipEventDestroy(e.second.first);
      (void)hipEventDestroy(e.second.second);
    }
    evs.clear();
  }
};

template <int D>
void enqueue_step(mpmx_ctx* c, const StepArgs& A, Launcher& L, bool full_sort,
                  const StepBuffers& B) {
  const Mesh M = c->mesh();
  Particles P = c->parts();
  if (P.plastic_list) {
    P.plastic_list = B.plastic_list;
    P.plastic_rec = B.plastic_rec;
  }
  if (P.spin) P.spin = B.spin;
  if (P.Cmat) {
    P.Cmat = B.Cmat;
    P.beta = B.beta;
  }
  if (P.rotF) P.rotF = B.rotF;
  if (P.plastic_trial) P.plastic_trial = B.plastic_trial;
  if (P.gterm) P.gterm = B.gterm;
  Nodes Nd = c->nodes();
  if (Nd.fnval) Nd.fnval = B.fnval;
  hipStream_t st = c->stream;
  Status* S = B.S;
  const Material* mats = c->d_mats.p;
  int* skey = B.skey;
  int* perm = B.perm;
  int* scan = B.scan;
  const bool ordered = (c->flags & MPMX_ORDERED) != 0;
  // Cartesian meshes are located inside the last particle kernel of the step;
  // isoparametric ones by a kernel of their own after it
  const bool loc = !c->iso;
  if (!c->fns.empty())
    L("time_functions", [&] {
      launch_fn_eval(c->d_fns.p, c->d_fn_data.p, (int)c->fns.size(), S, A.dt, B.fnval, st);
    });
  // the Mohr-Coulomb return mappings deferred by the stress kernels
  auto plastic = [&] {
    if (c->any_mc) L("mc_plastic", [&] { launch_plastic(P, mats, S, st); });
  };
  // affine velocity updates (APIC, ASFLIP, TPIC): the transfer matrix before
  // each particle-to-grid map, the mapping matrix before grid-to-particle
  const bool aff = c->affine();
  auto affine = [&] {
    if (aff) L("affine_matrix", [&] { launch_affine_matrix<D>(M, P, c->vu, S, st); });
  };
  auto mapping = [&] {
    if (aff) L("mapping_matrix", [&] { launch_mapping_matrix<D>(M, P, Nd, A, c->vu, S, st); });
  };
  if (full_sort || ordered)
    L("sort_by_cell", [&] {
      radix_sort_by_cell(P.cell, P.np, (int)M.ncells, skey, perm, B.sort_tmp, st);
    });
  else
    L("refresh_keys", [&] { refresh_sorted_keys(P.cell, perm, P.np, (int)M.ncells, skey, st); });
  if (ordered) {
    L("cell_ranges", [&] {
      launch_cell_ranges<D>(P.np, M.ncells, skey, B.cstart, B.cend, st);
    });
    const int* cs = B.cstart;
    const int* ce = B.cend;
    affine();
    if (c->scheme == USF) {
      L("gather_mass_momentum", [&] { launch_gather<D>(GK_MV_VEL, M, P, Nd, A, S, perm, cs, ce, st); });
      L("stress", [&] { launch_stress<D>(false, M, P, Nd, mats, A, S, scan, st); });
      plastic();
      L("gather_forces", [&] { launch_gather<D>(GK_F_ACC, M, P, Nd, A, S, perm, cs, ce, st); });
      mapping();
      L("g2p_locate", [&] { launch_g2p<D>(false, loc, M, P, Nd, mats, A, c->vu, S, scan, st); });
    } else if (c->scheme == USL) {
      L("gather_all", [&] { launch_gather<D>(GK_ALL_VEL_ACC, M, P, Nd, A, S, perm, cs, ce, st); });
      mapping();
      L("g2p_stress_locate", [&] { launch_g2p<D>(true, loc, M, P, Nd, mats, A, c->vu, S, scan, st); });
      plastic();
    } else {
      L("gather_all", [&] { launch_gather<D>(GK_ALL_VEL_ACC, M, P, Nd, A, S, perm, cs, ce, st); });
      mapping();
      L("g2p", [&] { launch_g2p<D>(false, false, M, P, Nd, mats, A, c->vu, S, scan, st); });
      affine();
      L("gather_momentum", [&] { launch_gather<D>(GK_MOM_VEL, M, P, Nd, A, S, perm, cs, ce, st); });
      L("stress_locate", [&] { launch_stress<D>(loc, M, P, Nd, mats, A, S, scan, st); });
      plastic();
    }
  } else {
    affine();
    if (c->scheme == USF) {
      L("p2g_mass_momentum", [&] { launch_p2g<D>(P2G_MV, c->chunk, M, P, Nd, A, S, skey, perm, st); });
      L("node_velocity", [&] { launch_node<D>(ND_INIT | ND_VEL, M, Nd, A, S, st); });
      L("stress_force_scatter", [&] { launch_stress_scatter<D>(c->chunk, M, P, Nd, mats, A, S, skey, perm, st); });
      if (P.ntr > 0) L("traction_scatter", [&] { launch_traction_scatter<D>(M, P, Nd, S, st); });
      L("node_acceleration", [&] { launch_node<D>(ND_ACC | ND_ZERO_ALL, M, Nd, A, S, st); });
      mapping();
      L("g2p_locate", [&] { launch_g2p<D>(false, loc, M, P, Nd, mats, A, c->vu, S, scan, st); });
    } else if (c->scheme == USL) {
      L("p2g_all", [&] { launch_p2g<D>(P2G_MV_F, c->chunk, M, P, Nd, A, S, skey, perm, st); });
      if (P.ntr > 0) L("traction_scatter", [&] { launch_traction_scatter<D>(M, P, Nd, S, st); });
      L("node_update", [&] { launch_node<D>(ND_INIT | ND_VEL | ND_ACC | ND_ZERO_ALL, M, Nd, A, S, st); });
      mapping();
      L("g2p_stress_locate", [&] { launch_g2p<D>(true, loc, M, P, Nd, mats, A, c->vu, S, scan, st); });
      plastic();
    } else {
      L("p2g_all", [&] { launch_p2g<D>(P2G_MV_F, c->chunk, M, P, Nd, A, S, skey, perm, st); });
      if (P.ntr > 0) L("traction_scatter", [&] { launch_traction_scatter<D>(M, P, Nd, S, st); });
      L("node_update", [&] { launch_node<D>(ND_INIT | ND_VEL | ND_ACC | ND_ZERO_MOM, M, Nd, A, S, st); });
      if (aff) {  // G2P, then the affine momentum map with the new positions
        mapping();
        L("g2p", [&] { launch_g2p<D>(false, false, M, P, Nd, mats, A, c->vu, S, scan, st); });
        affine();
        L("p2g_momentum", [&] { launch_p2g<D>(P2G_MOM, c->chunk, M, P, Nd, A, S, skey, perm, st); });
      } else {
        L("g2p_momentum_scatter", [&] { launch_g2p_scatter<D>(c->chunk, M, P, Nd, A, c->vu, S, skey, perm, st); });
      }
      L("node_velocity", [&] { launch_node<D>(ND_VEL | ND_ZERO_ALL, M, Nd, A, S, st); });
      L("stress_locate", [&] { launch_stress<D>(loc, M, P, Nd, mats, A, S, scan, st); });
      plastic();
    }
  }
  if (!loc) L("locate", [&] { launch_locate_step<D>(M, P, S, scan, st); });
  L("locate_scan", [&] { launch_locate_scan<D>(M, P, S, scan, nullptr, c->locate_all ? 0 : 1, st); });
  L("step_end", [&] { launch_step_end(S, st); });
}

void enqueue(mpmx_ctx* c, const StepArgs& A, Launcher& L, bool full_sort,
             const StepBuffers& B) {
  if (c->dim == 2)
    enqueue_step<2>(c, A, L, full_sort, B);
  else
    enqueue_step<3>(c, A, L, full_sort, B);
}

// Whether the next step sorts fully (ordered mode needs exact cell order
// every step); advances the counter.
bool next_step_sorts(mpmx_ctx* c) {
  const bool full = (c->flags & MPMX_ORDERED) || c->need_sort || c->since_sort >= c->sort_every;
  c->need_sort = false;
  c->since_sort = full ? 1 : c->since_sort + 1;
  return full;
}


// Queue nsteps steps on the context's stream (one graph replay each).
// ad_nodes: record each step's nodal quantities in the slots of length
// ad_stride there, step Status::ad_slot in its slot.
StepArgs step_args(const mpmx_ctx* c, double dt, double damping) {
  StepArgs A;
  A.dt = dt;
  A.damping = damping;
  A.blending = c->blending;
  for (int k = 0; k < 3; ++k) A.g[k] = c->g[k];
  A.ramp = c->ramp;
  A.ad_nodes = nullptr;
  A.ad_slot = nullptr;
  A.ad_stride = 0;
  A.ad_nslots = 0;
  return A;
}

void launch_steps(mpmx_ctx* c, int nsteps, double dt, double damping,
                  double* ad_nodes = nullptr, int64_t ad_stride = 0, int ad_nslots = 0) {
  const StepBuffers B = ctx_buffers(c);
  StepArgs A;
  A.dt = dt;
  A.damping = damping;
  A.blending = c->blending;
  for (int k = 0; k < 3; ++k) A.g[k] = c->g[k];
  A.ramp = c->ramp;
  A.ad_nodes = ad_nodes;
  A.ad_slot = ad_nodes ? &c->d_status.p->ad_slot : nullptr;
  A.ad_stride = ad_stride;
  A.ad_nslots = ad_nslots;
  Launcher L{c};
  const bool use_graph = !(c->flags & MPMX_NO_GRAPH) && !c->profile;
  const int rec = ad_nodes ? 2 : 0;
  if (use_graph) {
    if (c->graph_dt != dt || c->graph_damping != damping) {
      c->invalidate_graph();
      c->graph_dt = dt;
      c->graph_damping = damping;
    }
    for (int s = 0; s < nsteps; ++s) {
      const int kind = (next_step_sorts(c) ? 1 : 0) + rec;
      if (!c->graph_exec[kind]) {
        hipGraph_t graph;
        HIPCHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        enqueue(c, A, L, (kind & 1) == 1, B);
        HIPCHECK(hipStreamEndCapture(c->stream, &graph));
        hipError_t e = hipGraphInstantiate(&c->graph_exec[kind], graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        HIPCHECK(e);
      }
      HIPCHECK(hipGraphLaunch(c->graph_exec[kind], c->stream));
    }
  } else {
    for (int s = 0; s < nsteps; ++s) {
      enqueue(c, A, L, next_step_sorts(c), B);
      HIPCHECK(hipGetLastError());
      L.collect();
    }
  }
}

// Remove the particles whose cell is -1 (they left the mesh in remove mode),
// keeping the order of the others (Mesh::remove_particle erases from the
// particle vector with std::remove, mesh.tcc:658-665, vector.tcc:23-30).
// Particle-indexed loads follow their particles; those of removed particles
// are dropped.
void remove_lost_particles(mpmx_ctx* c) {
  const int64_t np = c->np;
  std::vector<int> cell(np);
  c->d_cell.download(cell.data(), np, c->stream);
  HIPCHECK(hipStreamSynchronize(c->stream));
  std::vector<int> keep;
  std::vector<int64_t> remap(np, -1);
  for (int64_t p = 0; p < np; ++p)
    if (cell[p] >= 0) {
      remap[p] = (int64_t)keep.size();
      keep.push_back((int)p);
    }
  const int64_t nk = (int64_t)keep.size();
  if (nk == np) return;
  hipStream_t st = c->stream;
  DBuf<int> d_idx;
  d_idx.alloc(std::max<int64_t>(nk, 1));
  d_idx.upload(keep.data(), nk, st);
  auto take = [&](auto& b, int comps) {
    using T = typename std::remove_reference<decltype(*b.p)>::type;
    DBuf<T> nb;
    nb.alloc((size_t)std::max<int64_t>(nk, 1) * comps);
    if (nk > 0) take_rows<T>(b.p, nb.p, d_idx.p, nk, np, comps, st);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(st));
    std::swap(b.p, nb.p);
    std::swap(b.n, nb.n);
  };
  const int d = c->dim;
  take(c->d_x, d); take(c->d_u, d); take(c->d_v, d); take(c->d_xi, d);
  take(c->d_stress, 6); take(c->d_strain, 6); take(c->d_F, 9);
  take(c->d_mass, 1); take(c->d_vol, 1); take(c->d_rho, 1);
  take(c->d_mat, 1); take(c->d_cell, 1);
  if (c->any_mc) take(c->d_sv, NSV);
  take(c->d_Bmat, d * d);
  std::vector<double> psize(nk);
  std::vector<int64_t> ids(nk);
  for (int64_t i = 0; i < nk; ++i) {
    psize[i] = c->psize[keep[i]];
    ids[i] = c->ids[keep[i]];
  }
  c->psize.swap(psize);
  c->ids.swap(ids);
  std::vector<mpmx_ctx::Pvc> pvc;
  for (const auto& q : c->pvc)
    if (q.particle < 0 || (q.particle < np && remap[q.particle] >= 0))
      pvc.push_back({q.particle < 0 ? q.particle : remap[q.particle], q.dir, q.value});
  c->pvc.swap(pvc);
  std::vector<mpmx_ctx::Trac> trac;
  for (const auto& t : c->trac)
    if (t.particle < 0 || (t.particle < np && remap[t.particle] >= 0))
      trac.push_back({t.particle < 0 ? t.particle : remap[t.particle], t.dir, t.value, t.fn});
  c->trac.swap(trac);
  c->np = nk;
  c->ad_recorded = false;
  build_pvc(c);
  build_trac(c);
  alloc_step_buffers(c);
  c->invalidate_graph();
}

// Wait for the queued steps; take the step counter; fail if a particle left
// the mesh (locate_particles, mpm_scheme.tcc:250-256), or, in remove mode,
// remove the particles that did (mpm_scheme.tcc:258-260).
void check_status(mpmx_ctx* c) {
  read_status(c);
  c->step = c->h_status->step;
  if (c->h_status->nlost > 0) remove_lost_particles(c);
  if (c->h_status->first_bad_step != std::numeric_limits<int64_t>::max())
    throw ApiError(MPMX_E_UNLOCATABLE,
                   "Particle(s) outside the mesh domain: " + std::to_string(c->h_status->bad_particle) +
                       " (step " + std::to_string(c->h_status->first_bad_step) + ")");
}

// ------------------------ step without allocation ----------------------------
// mpmx_step_noalloc takes the step kernels' scratch (everything StepBuffers
// holds but the status block) from one file-scope device array instead of
// the context's buffers. Each GPU holds its own copy until the process ends.
// In host code the array's name is not a device address: its address on a
// device comes from hipGetSymbolAddress with that device current, resolved
// once per device when a context is created there, and only that address is
// used.
constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_addr[kMaxDevices] = {};

// Address of the scratch array on the current device `dev` (resolved once).
hipError_t resolve_scratch(int dev, void** out) {
  *out = nullptr;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_addr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_addr[dev] = ptr;
  }
  *out = g_scratch_addr[dev];
  return hipSuccess;
}

// The step scratch laid out from `base` (256-byte aligned pieces); returns
// the bytes it spans. With base = 0 only the size is computed.
size_t carve_step_buffers(const mpmx_ctx* c, uintptr_t base, StepBuffers* B) {
  const int64_t np = std::max<int64_t>(c->np, 1);
  size_t off = 0;
  auto take = [&](size_t bytes) {
    off = (off + 255) & ~size_t(255);
    const uintptr_t p = base + off;
    off += bytes;
    return p;
  };
  B->S = c->d_status.p;
  B->skey = (int*)take(np * sizeof(int));
  B->perm = (int*)take(np * sizeof(int));
  B->sort_tmp = (int*)take(radix_sort_temp_ints(np) * sizeof(int));
  B->scan = (int*)take(np * sizeof(int));
  B->cstart = B->cend = nullptr;
  B->gterm = nullptr;
  if (c->flags & MPMX_ORDERED) {
    B->cstart = (int*)take(std::max<int64_t>(c->ncells, 1) * sizeof(int));
    B->cend = (int*)take(std::max<int64_t>(c->ncells, 1) * sizeof(int));
    B->gterm = (double*)take((size_t)c->nn * (1 + c->dim) * np * sizeof(double));
  }
  B->plastic_list = nullptr;
  B->plastic_rec = nullptr;
  if (c->any_mc) {
    B->plastic_list = (int*)take(np * sizeof(int));
    B->plastic_rec = (double*)take((size_t)PR_N * np * sizeof(double));
  }
  B->spin = c->any_rate ? (double*)take(3 * np * sizeof(double)) : nullptr;
  B->fnval = c->fns.empty() ? nullptr : (double*)take(c->fns.size() * sizeof(double));
  B->rotF = c->any_gn ? (double*)take((size_t)(c->dim * 3 + c->dim * c->dim) * np * sizeof(double))
                      : nullptr;
  B->plastic_trial = c->any_mc_rate ? (double*)take(6 * np * sizeof(double)) : nullptr;
  B->Cmat = B->beta = nullptr;
  if (c->affine()) {
    B->Cmat = (double*)take((size_t)c->dim * c->dim * np * sizeof(double));
    B->beta = (double*)take(np * sizeof(double));
  }
  return off;
}

// ------------------------------ adjoint --------------------------------------
// The particle state a checkpoint holds: every per-particle array the step
// reads or writes (mass and material are constant and not copied).
struct StateArray {
  double* p;
  int64_t comps;
};
std::vector<StateArray> ad_state(mpmx_ctx* c) {
  const int d = c->dim;
  std::vector<StateArray> v = {{c->d_x.p, d},      {c->d_u.p, d},      {c->d_v.p, d},
                               {c->d_stress.p, 6}, {c->d_strain.p, 6}, {c->d_F.p, 9},
                               {c->d_vol.p, 1},    {c->d_rho.p, 1},    {c->d_xi.p, d}};
  if (c->any_mc) v.push_back({c->d_sv.p, NSV});
  if (c->affine()) v.push_back({c->d_Bmat.p, (int64_t)d * d});
  return v;
}

void ad_checkpoint(mpmx_ctx* c, int idx, bool save) {
  double* base = c->d_ad_ck.p + (int64_t)idx * c->ad_ck_stride;
  int64_t off = 0;
  for (const auto& a : ad_state(c)) {
    const size_t bytes = (size_t)a.comps * c->np * sizeof(double);
    if (save)
      HIPCHECK(hipMemcpyAsync(base + off, a.p, bytes, hipMemcpyDeviceToDevice, c->stream));
    else
      HIPCHECK(hipMemcpyAsync(a.p, base + off, bytes, hipMemcpyDeviceToDevice, c->stream));
    off += a.comps * c->np;
  }
  int* cb = c->d_ad_ck_cell.p + (int64_t)idx * c->np;
  if (save)
    HIPCHECK(hipMemcpyAsync(cb, c->d_cell.p, c->np * sizeof(int), hipMemcpyDeviceToDevice, c->stream));
  else
    HIPCHECK(hipMemcpyAsync(c->d_cell.p, cb, c->np * sizeof(int), hipMemcpyDeviceToDevice, c->stream));
}

int64_t ad_slot_doubles(const mpmx_ctx* c) { return (2 * c->dim + 9) * c->np; }

Slot ad_slot(mpmx_ctx* c, int j) {
  const int64_t np = c->np, d = c->dim;
  double* b = c->d_ad_slots.p + (int64_t)j * ad_slot_doubles(c);
  Slot s;
  s.v = b;
  s.stress = b + d * np;
  s.vol = b + (d + 6) * np;
  s.rho = b + (d + 7) * np;
  s.xi = b + (d + 8) * np;
  s.yield = c->any_mc ? b + (2 * d + 8) * np : nullptr;
  s.cell = c->d_ad_slot_cell.p + (int64_t)j * np;
  return s;
}

void ad_save_slot(mpmx_ctx* c, int j) {
  const Slot r = ad_slot(c, j);
  SlotW w{const_cast<double*>(r.v),   const_cast<double*>(r.stress),
          const_cast<double*>(r.vol), const_cast<double*>(r.rho),
          const_cast<double*>(r.xi),  const_cast<double*>(r.yield),
          const_cast<int*>(r.cell)};
  if (c->dim == 2)
    launch_ad_save<2>(c->parts(), w, c->stream);
  else
    launch_ad_save<3>(c->parts(), w, c->stream);
  HIPCHECK(hipGetLastError());
}

// Nodal quantities recorded by the replayed step j (StepArgs::ad_nodes
// layout): mass, then momentum, force, updated velocity, acceleration, MUSL
// second momentum and its velocity.
int64_t ad_node_slot_doubles(const mpmx_ctx* c) { return (1 + 6 * c->dim) * c->nnodes; }
// The step records of one segment: K + 1 particle slots and K nodal slots.
// The recording graphs hold their addresses, so they are captured anew.
void ad_alloc_slots(mpmx_ctx* c, int K) {
  c->d_ad_slots.alloc((size_t)(K + 1) * ad_slot_doubles(c));
  c->d_ad_slot_cell.alloc((size_t)(K + 1) * c->np);
  c->d_ad_node_slots.alloc((size_t)K * ad_node_slot_doubles(c));
  for (int k = 2; k < 4; ++k)
    if (c->graph_exec[k]) {
      (void)hipGraphExecDestroy(c->graph_exec[k]);
      c->graph_exec[k] = nullptr;
    }
}

double* ad_node_slot(mpmx_ctx* c, int j) {
  return c->d_ad_node_slots.p + (int64_t)j * ad_node_slot_doubles(c);
}

AdjNodes ad_nodes(mpmx_ctx* c, int j) {
  const int64_t nn = c->nnodes, d = c->dim;
  double* r = ad_node_slot(c, j);
  AdjNodes A;
  A.m = r;
  A.P = r + nn;
  A.f = r + (1 + d) * nn;
  A.vp = r + (1 + 2 * d) * nn;
  A.a = r + (1 + 3 * d) * nn;
  A.P2 = r + (1 + 4 * d) * nn;
  A.vpp = r + (1 + 5 * d) * nn;
  // this step's pre-update velocity and the adjoints
  double* b = c->d_ad_nodes.p;
  double** vec[] = {&A.v, &A.lvp, &A.la, &A.lv, &A.lvpp, &A.lf, &A.lP, &A.lP2};
  A.lm = b;
  A.lm2 = b + nn;
  int64_t off = 2 * nn;
  for (auto* f : vec) {
    *f = b + off;
    off += d * nn;
  }
  return A;
}
int64_t ad_node_doubles(const mpmx_ctx* c) { return (2 + 8 * c->dim) * c->nnodes; }

AdjParticles ad_parts(mpmx_ctx* c) {
  const int64_t np = c->np, d = c->dim;
  double* b = c->d_ad_parts.p;
  AdjParticles L;
  L.lx = b;
  L.lv = b + d * np;
  L.ls = b + 2 * d * np;
  L.lvol = b + (2 * d + 6) * np;
  L.lxq = b + (2 * d + 7) * np;
  L.gE = b + (3 * d + 7) * np;
  L.gphi = b + (3 * d + 8) * np;
  L.gm = b + (3 * d + 9) * np;
  return L;
}
int64_t ad_part_doubles(const mpmx_ctx* c) { return (3 * c->dim + 10) * c->np; }

// What the adjoint differentiates: Cartesian cells (not isoparametric) with
// the reference's local node order, no friction constraints, no strain softening, no particle
// tractions (they depend on the particle's volume and position). Nodal
// forces and acceleration constraints, with or without math functions, are
// constants of the step and are supported.
void ad_check_supported(mpmx_ctx* c) {
  if (c->affine())
    throw ApiError(MPMX_E_UNSUPPORTED, "adjoint: APIC, ASFLIP and TPIC are not supported");
  if (c->iso)
    throw ApiError(MPMX_E_UNSUPPORTED, "adjoint: isoparametric meshes are not supported");
  if (!c->trac.empty())
    throw ApiError(MPMX_E_UNSUPPORTED, "adjoint: particle tractions are not supported");
  for (const auto& m : c->mats) {
    if (m.type == MC && m.softening)
      throw ApiError(MPMX_E_UNSUPPORTED, "adjoint: Mohr-Coulomb strain softening is not supported");
    if (m.rate != RATE_NONE)
      throw ApiError(MPMX_E_UNSUPPORTED, "adjoint: objective stress rates are not supported");
  }
  for (const auto& kv : c->cons) {
    if (kv.second.fdir >= 0)
      throw ApiError(MPMX_E_UNSUPPORTED, "adjoint: friction constraints are not supported");
    if (kv.second.rot)
      throw ApiError(MPMX_E_UNSUPPORTED, "adjoint: constraints on general boundaries are not "
                                         "supported");
  }
  const int d = c->dim, nn = c->nn;
  for (int64_t e = 0; e < c->ncells; ++e) {
    double lo[3], hi[3];
    for (int k = 0; k < d; ++k) {
      lo[k] = hi[k] = c->node_x[(int64_t)c->cell_nodes[e * nn] * d + k];
      for (int i = 1; i < nn; ++i) {
        const double x = c->node_x[(int64_t)c->cell_nodes[e * nn + i] * d + k];
        lo[k] = std::min(lo[k], x);
        hi[k] = std::max(hi[k], x);
      }
    }
    for (int i = 0; i < nn; ++i) {
      const int l = i & 3;
      const int sg[3] = {(l == 0 || l == 3) ? -1 : 1, l < 2 ? -1 : 1, i < 4 ? -1 : 1};
      for (int k = 0; k < d; ++k) {
        const double x = c->node_x[(int64_t)c->cell_nodes[e * nn + i] * d + k];
        const double want = sg[k] < 0 ? lo[k] : hi[k];
        if (std::fabs(x - want) > 1e-9 * (hi[k] - lo[k]))
          throw ApiError(MPMX_E_UNSUPPORTED,
                         "adjoint: cell " + std::to_string(e) +
                             " is not an axis-aligned rectangle/box in the element node order");
      }
    }
  }
}

void ad_backward_step(mpmx_ctx* c, int j, int64_t t) {
  AdjArgs a;
  a.dt = c->ad_dt;
  a.damping = c->ad_damping;
  a.blending = c->blending;
  a.vu = c->vu;
  a.apex_ratio = c->ad_apex_ratio;
  double gm = 1.0;
  if (c->ramp > 0.0) {
    const double r = static_cast<double>(t) * c->ad_dt / c->ramp;
    gm = r < 1.0 ? r : 1.0;
  }
  for (int k = 0; k < 3; ++k) a.g[k] = gm * c->g[k];
  const Slot s0 = ad_slot(c, j), s1 = ad_slot(c, j + 1);  // k_ad_node_fwd clears the adjoints
  AdjList Q;
  Q.cap = c->np;
  Q.count = c->d_ad_list.p;
  Q.ids = c->d_ad_list.p + 2;
  Q.apex = c->d_ad_list.p + 2 + c->np;
  Q.buf = c->d_ad_list_buf.p;
  if (c->dim == 2)
    launch_ad_backward_step<2>(c->scheme, c->mesh(), c->parts(), c->nodes(), c->d_mats.p, s0, s1,
                               ad_nodes(c, j), ad_parts(c), Q, a, c->stream);
  else
    launch_ad_backward_step<3>(c->scheme, c->mesh(), c->parts(), c->nodes(), c->d_mats.p, s0, s1,
                               ad_nodes(c, j), ad_parts(c), Q, a, c->stream);
  HIPCHECK(hipGetLastError());
}

}  // namespace

extern "C" {

int mpmx_create(mpmx_ctx** out, int dim, int scheme, int velocity_update, double blending_ratio,
                unsigned flags) {
  if (!out) return MPMX_E_ARG;
  *out = nullptr;
  if ((dim != 2 && dim != 3) || scheme < 0 || scheme > 2 || velocity_update < 0 ||
      velocity_update > 4)
    return MPMX_E_ARG;
  auto* c = new mpmx_ctx();
  c->dim = dim;
  c->nn = dim == 2 ? 4 : 8;
  c->scheme = scheme;
  c->vu = velocity_update;
  c->blending = (blending_ratio < 0. || blending_ratio > 1.) ? 1.0 : blending_ratio;
  c->flags = flags;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    c->error = "no HIP device available";
    *out = c;
    return MPMX_E_HIP;
  }
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipHostMalloc((void**)&c->h_status, sizeof(Status)) != hipSuccess) {
    c->error = "stream / pinned allocation failed";
    *out = c;
    return MPMX_E_HIP;
  }
  try {
    c->d_status.alloc(1);
  } catch (ApiError& e) {
    c->error = e.what();
    *out = c;
    return e.code;
  }
  // the no-allocation step's scratch on this device (resolved once per device)
  hipError_t se = hipGetDevice(&c->device);
  if (se == hipSuccess) se = resolve_scratch(c->device, &c->scratch);
  if (se != hipSuccess) {
    c->scratch = nullptr;
    c->scratch_error = std::string("scratch address: ") + hipGetErrorString(se);
  }
  *out = c;
  return MPMX_OK;
}

int mpmx_destroy(mpmx_ctx* c) {
  if (!c) return MPMX_E_ARG;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  delete c;
  return MPMX_OK;
}

const char* mpmx_last_error(const mpmx_ctx* c) { return c ? c->error.c_str() : "null context"; }

int mpmx_set_mesh(mpmx_ctx* c, int64_t nnodes, const double* coords, int64_t ncells,
                  const int64_t* cell_nodes, int isoparametric) {
  API_BEGIN(c)
  if (nnodes <= 0 || ncells <= 0 || !coords || !cell_nodes) throw ApiError(MPMX_E_ARG, "empty mesh");
  if (nnodes > std::numeric_limits<int>::max() / 8 || ncells > std::numeric_limits<int>::max() / 8)
    throw ApiError(MPMX_E_ARG, "mesh too large for 32-bit ids");
  if (c->np > 0) throw ApiError(MPMX_E_STATE, "mesh must be set before particles");
  const int d = c->dim, nn = c->nn;
  c->nnodes = nnodes;
  c->ncells = ncells;
  c->ninst = 0;
  c->iso = isoparametric != 0;
  c->invalidate_graph();
  c->node_x.assign(coords, coords + nnodes * d);
  c->cell_nodes.resize(ncells * nn);
  for (int64_t i = 0; i < ncells * nn; ++i) {
    if (cell_nodes[i] < 0 || cell_nodes[i] >= nnodes) throw ApiError(MPMX_E_ARG, "cell node id out of range");
    c->cell_nodes[i] = (int)cell_nodes[i];
  }
  // geometry
  std::vector<double> centroid(d * ncells), len(d * ncells), two_mean(ncells), dndx_c(ncells * nn * d);
  c->cell_volume.assign(ncells, 0.);
  static const int q_sides[4][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}};
  static const int h_sides[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                                     {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
  for (int64_t ci = 0; ci < ncells; ++ci) {
    double X[8][3] = {{0}};
    for (int i = 0; i < nn; ++i)
      for (int k = 0; k < d; ++k) X[i][k] = c->node_x[c->cell_nodes[ci * nn + i] * d + k];
    for (int k = 0; k < d; ++k) {
      double s = 0.;
      for (int i = 0; i < nn; ++i) s += X[i][k];
      centroid[k * ncells + ci] = s / nn;
    }
    const int ns = d == 2 ? 4 : 12;
    double ml = 0.;
    for (int sdx = 0; sdx < ns; ++sdx) {
      const int a = d == 2 ? q_sides[sdx][0] : h_sides[sdx][0];
      const int b = d == 2 ? q_sides[sdx][1] : h_sides[sdx][1];
      ml += norm_diff(d, X[a], X[b]);
    }
    ml /= ns;
    two_mean[ci] = ml * 2.;
    len[0 * ncells + ci] = norm_diff(d, X[0], X[1]);
    len[1 * ncells + ci] = norm_diff(d, X[1], X[2]);
    if (d == 3) len[2 * ncells + ci] = norm_diff(d, X[1], X[5]);
    c->cell_volume[ci] = d == 2 ? quad_volume(X) : hex_volume(X);
    if (!(c->cell_volume[ci] > 0))
      throw ApiError(MPMX_E_ARG, "Negative or zero volume cell, misconfigured cell: " + std::to_string(ci));
    const double zero[3] = {0, 0, 0};
    double dc[8][3];
    host_dndx(d, zero, X, dc);
    for (int i = 0; i < nn; ++i)
      for (int k = 0; k < d; ++k) dndx_c[(ci * nn + i) * d + k] = dc[i][k];
  }
  // neighbours: cells sharing a node (mesh.tcc:287-307), ascending ids
  std::vector<std::vector<int>> node_cells(nnodes);
  for (int64_t ci = 0; ci < ncells; ++ci)
    for (int i = 0; i < nn; ++i) node_cells[c->cell_nodes[ci * nn + i]].push_back((int)ci);
  std::vector<int> nb_ptr(ncells + 1, 0), nb_idx;
  std::vector<int> tmp;
  for (int64_t ci = 0; ci < ncells; ++ci) {
    tmp.clear();
    for (int i = 0; i < nn; ++i)
      for (int o : node_cells[c->cell_nodes[ci * nn + i]])
        if (o != ci) tmp.push_back(o);
    std::sort(tmp.begin(), tmp.end());
    tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
    nb_idx.insert(nb_idx.end(), tmp.begin(), tmp.end());
    nb_ptr[ci + 1] = (int)nb_idx.size();
  }
  std::vector<int> ncell_tab(nnodes * nn, -1);
  for (int64_t n = 0; n < nnodes; ++n) {
    auto& v = node_cells[n];
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    if ((int)v.size() > nn) throw ApiError(MPMX_E_ARG, "node shared by more than 2^dim cells");
    for (size_t j = 0; j < v.size(); ++j) ncell_tab[n * nn + j] = v[j];
  }
  // node coordinates SoA
  std::vector<double> nx(d * nnodes);
  for (int64_t n = 0; n < nnodes; ++n)
    for (int k = 0; k < d; ++k) nx[k * nnodes + n] = c->node_x[n * d + k];
  hipStream_t st = c->stream;
  c->d_cell_nodes.alloc(ncells * nn);
  c->d_cell_nodes.upload(c->cell_nodes.data(), ncells * nn, st);
  c->d_node_x.alloc(d * nnodes);
  c->d_node_x.upload(nx.data(), nx.size(), st);
  c->d_centroid.alloc(d * ncells);
  c->d_centroid.upload(centroid.data(), centroid.size(), st);
  c->d_len.alloc(d * ncells);
  c->d_len.upload(len.data(), len.size(), st);
  c->d_two_mean.alloc(ncells);
  c->d_two_mean.upload(two_mean.data(), two_mean.size(), st);
  c->d_dndx_c.alloc(ncells * nn * d);
  c->d_dndx_c.upload(dndx_c.data(), dndx_c.size(), st);
  c->d_nb_ptr.alloc(ncells + 1);
  c->d_nb_ptr.upload(nb_ptr.data(), nb_ptr.size(), st);
  c->d_nb_idx.alloc(std::max<size_t>(nb_idx.size(), 1));
  c->d_nb_idx.upload(nb_idx.data(), nb_idx.size(), st);
  c->d_node_cells.alloc(nnodes * nn);
  c->d_node_cells.upload(ncell_tab.data(), ncell_tab.size(), st);
  // nodes (zeroed)
  auto zero_alloc = [&](DBuf<double>& b, int64_t n) {
    b.alloc(n);
    HIPCHECK(hipMemsetAsync(b.p, 0, n * sizeof(double), st));
  };
  zero_alloc(c->d_nmass, nnodes);
  zero_alloc(c->d_nmom, d * nnodes);
  zero_alloc(c->d_nvel, d * nnodes);
  zero_alloc(c->d_nacc, d * nnodes);
  zero_alloc(c->d_nfext, d * nnodes);
  zero_alloc(c->d_nfint, d * nnodes);
  c->d_nactive.alloc(nnodes);
  HIPCHECK(hipMemsetAsync(c->d_nactive.p, 0, nnodes, st));
  HIPCHECK(hipStreamSynchronize(st));
  c->cons.clear();
  upload_constraints(c);
  alloc_step_buffers(c);
  c->invalidate_graph();
  API_END(c)
}

int mpmx_set_instances(mpmx_ctx* c, int64_t n, const int64_t* cell_begin) {
  API_BEGIN(c)
  require_mesh(c);
  if (n < 1 || !cell_begin) throw ApiError(MPMX_E_ARG, "at least one instance required");
  if (c->np > 0) throw ApiError(MPMX_E_STATE, "instances must be set before particles");
  if (cell_begin[0] != 0 || cell_begin[n] != c->ncells)
    throw ApiError(MPMX_E_ARG, "instance cell ranges must start at 0 and end at ncells");
  std::vector<int> cell_inst(c->ncells);
  std::vector<int> node_inst(c->nnodes, -1);
  for (int64_t i = 0; i < n; ++i) {
    if (cell_begin[i + 1] <= cell_begin[i])
      throw ApiError(MPMX_E_ARG, "instance " + std::to_string(i) + " has no cells");
    for (int64_t ci = cell_begin[i]; ci < cell_begin[i + 1]; ++ci) {
      cell_inst[ci] = (int)i;
      for (int q = 0; q < c->nn; ++q) {
        const int nd = c->cell_nodes[ci * c->nn + q];
        if (node_inst[nd] >= 0 && node_inst[nd] != (int)i)
          throw ApiError(MPMX_E_ARG, "node " + std::to_string(nd) + " is shared by instances " +
                                         std::to_string(node_inst[nd]) + " and " +
                                         std::to_string(i));
        node_inst[nd] = (int)i;
      }
    }
  }
  c->ninst = n;
  c->d_cell_inst.alloc(c->ncells);
  c->d_cell_inst.upload(cell_inst.data(), c->ncells, c->stream);
  c->d_inst_begin.alloc(n + 1);
  c->d_inst_begin.upload(cell_begin, n + 1, c->stream);
  HIPCHECK(hipStreamSynchronize(c->stream));
  c->invalidate_graph();
  API_END(c)
}

int mpmx_set_materials(mpmx_ctx* c, int n, const double* params) {
  API_BEGIN(c)
  if (n <= 0 || !params) throw ApiError(MPMX_E_ARG, "no materials");
  if (c->np > 0) throw ApiError(MPMX_E_STATE, "materials must be set before particles");
  c->mats.clear();
  c->sv_custom.assign(n, 0);
  c->any_mc = false;
  c->any_rate = c->any_gn = c->any_mc_rate = false;
  for (int i = 0; i < n; ++i) {
    c->mats.push_back(make_material(params + (size_t)i * MPMX_MATERIAL_PARAMS));
    if (c->mats.back().type == MC) c->any_mc = true;
    if (c->mats.back().rate != RATE_NONE) c->any_rate = true;
    if (c->mats.back().rate == RATE_GREEN_NAGHDI) c->any_gn = true;
    if (c->mats.back().type == MC && c->mats.back().rate != RATE_NONE) c->any_mc_rate = true;
  }
  c->d_mats.alloc(n);
  refresh_uniform(c, true);
  c->invalidate_graph();
  API_END(c)
}

int mpmx_set_velocity_constraints(mpmx_ctx* c, int64_t n, const int64_t* node, const int* dir,
                                  const double* value) {
  API_BEGIN(c)
  require_mesh(c);
  for (int64_t i = 0; i < n; ++i) {
    if (node[i] < 0 || node[i] >= c->nnodes || dir[i] < 0 || dir[i] >= c->dim)
      throw ApiError(MPMX_E_ARG, "velocity constraint node/dir out of range");
    c->cons[node[i]].vel.insert({dir[i], value[i]});
  }
  upload_constraints(c);
  API_END(c)
}

int mpmx_set_acceleration_constraints_fn(mpmx_ctx* c, int64_t n, const int64_t* node,
                                         const int* dir, const double* value, const int* fn) {
  API_BEGIN(c)
  require_mesh(c);
  for (int64_t i = 0; i < n; ++i) {
    if (node[i] < 0 || node[i] >= c->nnodes || dir[i] < 0 || dir[i] >= c->dim)
      throw ApiError(MPMX_E_ARG, "acceleration constraint node/dir out of range");
    if (fn && (fn[i] < -1 || fn[i] >= (int)c->fns.size()))
      throw ApiError(MPMX_E_ARG, "acceleration constraint math function index out of range");
  }
  for (int64_t i = 0; i < n; ++i) {
    auto& hc = c->cons[node[i]];
    if (hc.vel.count(dir[i]))
      throw ApiError(MPMX_E_ARG,
                     "Acceleration and velocity constraints set in the same direction for the same node");
    if (hc.acc.insert({dir[i], value[i]}).second && fn && fn[i] >= 0) hc.accfn[dir[i]] = fn[i];
  }
  upload_constraints(c);
  API_END(c)
}

int mpmx_set_nodal_forces(mpmx_ctx* c, int64_t n, const int64_t* node, const int* dir,
                          const double* force, const int* fn) {
  API_BEGIN(c)
  require_mesh(c);
  for (int64_t i = 0; i < n; ++i) {
    if (node[i] < -1 || node[i] >= c->nnodes || dir[i] < 0 || dir[i] >= c->dim)
      throw ApiError(MPMX_E_ARG, "nodal force node/dir out of range");
    if (fn && (fn[i] < -1 || fn[i] >= (int)c->fns.size()))
      throw ApiError(MPMX_E_ARG, "nodal force math function index out of range");
  }
  for (int64_t i = 0; i < n; ++i) {
    const int64_t b = node[i] < 0 ? 0 : node[i], e = node[i] < 0 ? c->nnodes : node[i] + 1;
    for (int64_t nd = b; nd < e; ++nd) {
      auto& hc = c->cons[nd];
      hc.has_cf = true;
      hc.cf[dir[i]] = force[i];
      hc.cfn = fn ? fn[i] : -1;
    }
  }
  upload_constraints(c);
  API_END(c)
}

int mpmx_set_particle_tractions(mpmx_ctx* c, int64_t n, const int64_t* particle, const int* dir,
                                const double* value, const int* fn) {
  API_BEGIN(c)
  std::vector<mpmx_ctx::Trac> t;
  for (int64_t i = 0; i < n; ++i) {
    if (particle[i] < -1 || dir[i] < 0 || dir[i] >= c->dim)
      throw ApiError(MPMX_E_ARG, "particle traction particle/dir out of range");
    if (fn && (fn[i] < -1 || fn[i] >= (int)c->fns.size()))
      throw ApiError(MPMX_E_ARG, "particle traction math function index out of range");
    t.push_back({particle[i], dir[i], value[i], fn ? fn[i] : -1});
  }
  c->trac = t;
  build_trac(c);
  API_END(c)
}

int mpmx_set_acceleration_constraints(mpmx_ctx* c, int64_t n, const int64_t* node, const int* dir,
                                      const double* value) {
  return mpmx_set_acceleration_constraints_fn(c, n, node, dir, value, nullptr);
}

static_assert(FN_LINEAR == MPMX_FN_LINEAR && FN_SIN == MPMX_FN_SIN, "math function types");

int mpmx_set_math_functions(mpmx_ctx* c, int nf, const int* type, const int* npts,
                            const double* params) {
  API_BEGIN(c)
  if (nf < 0 || nf > kMaxMathFns) throw ApiError(MPMX_E_ARG, "number of math functions out of range");
  for (auto& kv : c->cons)
    for (auto& f : kv.second.accfn)
      if (f.second >= nf)
        throw ApiError(MPMX_E_STATE, "a constraint refers to a math function being removed");
  std::vector<MathFn> fs;
  std::vector<double> data;
  size_t off = 0;
  for (int f = 0; f < nf; ++f) {
    MathFn m;
    m.type = type[f];
    m.off = (int)data.size();
    if (m.type == FN_LINEAR) {
      if (npts[f] < 1) throw ApiError(MPMX_E_ARG, "a linear math function needs points");
      m.n = npts[f];
      data.insert(data.end(), params + off, params + off + 2 * (size_t)npts[f]);
      off += 2 * (size_t)npts[f];
    } else if (m.type == FN_SIN) {
      m.n = 4;
      data.insert(data.end(), params + off, params + off + 4);
      off += 4;
    } else {
      throw ApiError(MPMX_E_ARG, "unknown math function type");
    }
    fs.push_back(m);
  }
  c->fns = fs;
  if (!fs.empty()) {
    c->d_fns.alloc(fs.size());
    c->d_fns.upload(fs.data(), fs.size(), c->stream);
    c->d_fn_data.alloc(data.size());
    c->d_fn_data.upload(data.data(), data.size(), c->stream);
    c->d_fnval.alloc(fs.size());
    HIPCHECK(hipStreamSynchronize(c->stream));
  }
  c->invalidate_graph();
  API_END(c)
}

int mpmx_set_friction_constraints(mpmx_ctx* c, int64_t n, const int64_t* node, const int* dir,
                                  const int* sign_n, const double* mu) {
  API_BEGIN(c)
  require_mesh(c);
  for (int64_t i = 0; i < n; ++i) {
    if (node[i] < 0 || node[i] >= c->nnodes || dir[i] < 0 || dir[i] >= c->dim)
      throw ApiError(MPMX_E_ARG, "friction constraint node/dir out of range");
    auto& hc = c->cons[node[i]];
    hc.fdir = dir[i];
    hc.fsign = sign_n[i];
    hc.fmu = mu[i];
  }
  upload_constraints(c);
  API_END(c)
}

int mpmx_set_particle_velocity_constraints(mpmx_ctx* c, int64_t n, const int64_t* particle,
                                           const int* dir, const double* value) {
  API_BEGIN(c)
  if (n < 0 || (n > 0 && (!particle || !dir || !value)))
    throw ApiError(MPMX_E_ARG, "particle velocity constraints: n rows of particle, dir, value");
  std::vector<mpmx_ctx::Pvc> list;
  for (int64_t i = 0; i < n; ++i) {
    if (dir[i] < 0 || dir[i] >= c->dim)
      throw ApiError(MPMX_E_ARG, "Invalid direction of velocity constraint");
    if (particle[i] < -1 || particle[i] >= c->np)
      throw ApiError(MPMX_E_ARG, "No particle set found to assign velocity constraint");
    list.push_back({particle[i], dir[i], value[i]});
  }
  c->pvc = list;
  c->d_pvc_mask.release();
  c->d_pvc_val.release();
  build_pvc(c);
  c->invalidate_graph();
  c->ad_recorded = false;
  API_END(c)
}

int mpmx_set_nodal_euler_angles(mpmx_ctx* c, int64_t n, const int64_t* node,
                                const double* angles) {
  API_BEGIN(c)
  require_mesh(c);
  if (n < 0 || (n > 0 && (!node || !angles))) throw ApiError(MPMX_E_ARG, "null arrays");
  for (int64_t i = 0; i < n; ++i) {
    if (node[i] < 0 || node[i] >= c->nnodes) throw ApiError(MPMX_E_ARG, "node id out of range");
    HostCons& h = c->cons[node[i]];
    euler_rotation(c->dim, angles + i * c->dim, h.R);
    h.rot = true;
  }
  upload_constraints(c);
  API_END(c)
}

int mpmx_set_gravity(mpmx_ctx* c, const double* g, double ramping_time) {
  API_BEGIN(c)
  if (!g) throw ApiError(MPMX_E_ARG, "null gravity");
  if (ramping_time < 0) throw ApiError(MPMX_E_ARG, "Ramping time cannot be negative");
  for (int k = 0; k < 3; ++k) c->g[k] = k < c->dim ? g[k] : 0.;
  c->ramp = ramping_time;
  c->invalidate_graph();
  API_END(c)
}

int mpmx_set_particles(mpmx_ctx* c, int64_t n, const double* x, const double* u, const double* v,
                       const double* stress, const double* strain, const double* F,
                       const double* mass, const double* volume, const double* density,
                       const int* material_id, const int64_t* cell, const double* statevars) {
  API_BEGIN(c)
  require_mesh(c);
  if (c->mats.empty()) throw ApiError(MPMX_E_STATE, "materials are not set");
  if (n <= 0 || !x || !mass || !volume || !density || !material_id)
    throw ApiError(MPMX_E_ARG, "particles need x, mass, volume, density and material_id");
  for (int64_t i = 0; i < n; ++i)
    if (!(volume[i] > 0)) throw ApiError(MPMX_E_ARG, "Particle volume cannot be negative");
  const int64_t off = append_particles(c, n, x, u, v, stress, strain, F, mass, volume, density,
                                       material_id, statevars);
  std::vector<int> hint(n, -1);
  if (cell)
    for (int64_t i = 0; i < n; ++i) hint[i] = (cell[i] >= 0 && cell[i] < c->ncells) ? (int)cell[i] : -1;
  try {
    locate_new(c, off, n, hint);
  } catch (...) {
    truncate_particles(c, off);
    throw;
  }
  API_END(c)
}

int mpmx_generate_gauss(mpmx_ctx* c, int nppd, int64_t nset, const int64_t* cset, int material_id) {
  API_BEGIN(c)
  require_mesh(c);
  if (material_id < 0 || material_id >= (int)c->mats.size()) throw ApiError(MPMX_E_ARG, "bad material id");
  const int d = c->dim, nn = c->nn;
  double qp[64][3];
  const int nq = gauss_points(d, nppd, qp);
  const int64_t ncs = nset < 0 ? c->ncells : nset;
  std::vector<double> xs;
  std::vector<int> hint;
  xs.reserve(ncs * nq * d);
  for (int64_t ic = 0; ic < ncs; ++ic) {
    const int64_t ci = nset < 0 ? ic : cset[ic];
    if (ci < 0 || ci >= c->ncells) throw ApiError(MPMX_E_ARG, "cell set id out of range");
    for (int q = 0; q < nq; ++q) {
      double N[8];
      host_shapefn(d, qp[q], N);
      for (int k = 0; k < d; ++k) {
        double s = c->node_x[c->cell_nodes[ci * nn + 0] * d + k] * N[0];
        for (int i = 1; i < nn; ++i) s = s + c->node_x[c->cell_nodes[ci * nn + i] * d + k] * N[i];
        xs.push_back(s);
      }
      hint.push_back((int)ci);
    }
  }
  const int64_t n = (int64_t)hint.size();
  if (n == 0) throw ApiError(MPMX_E_ARG, "No particles were generated!");
  const double rho = c->mats[material_id].density;
  std::vector<double> ones(n, 1.0), dens(n, rho);
  std::vector<int> mat(n, material_id);
  const int64_t off = append_particles(c, n, xs.data(), nullptr, nullptr, nullptr, nullptr, nullptr,
                                       ones.data(), ones.data(), dens.data(), mat.data(), nullptr);
  try {
    locate_new(c, off, n, hint);
  } catch (...) {
    truncate_particles(c, off);
    throw;
  }
  // compute_volume (cell volume / particles in the cell) then compute_mass
  std::vector<int> cells(c->np);
  c->d_cell.download(cells.data(), c->np, c->stream);
  HIPCHECK(hipStreamSynchronize(c->stream));
  std::vector<int64_t> count(c->ncells, 0);
  for (int64_t i = 0; i < c->np; ++i) count[cells[i]] += 1;
  std::vector<double> vol(n), mass(n);
  for (int64_t i = 0; i < n; ++i) {
    const int ci = cells[off + i];
    vol[i] = c->cell_volume[ci] / (double)count[ci];
    mass[i] = vol[i] * rho;
  }
  c->d_vol.upload(vol.data(), n, c->stream, off);
  c->d_mass.upload(mass.data(), n, c->stream, off);
  HIPCHECK(hipStreamSynchronize(c->stream));
  for (int64_t i = 0; i < n; ++i) c->psize[off + i] = particle_size(vol[i], c->dim);
  build_trac(c);
  API_END(c)
}

int mpmx_assign_velocity(mpmx_ctx* c, int64_t n, const double* v) {
  API_BEGIN(c)
  if (n != c->np || !v) throw ApiError(MPMX_E_ARG, "assign_velocity needs one row per particle");
  put_soa(c->d_v, c->np, 0, n, c->dim, v, c->stream);
  API_END(c)
}

int mpmx_compute_volume_mass(mpmx_ctx* c) {
  API_BEGIN(c)
  require_mesh(c);
  const int64_t n = c->np;
  if (n == 0) return MPMX_OK;
  std::vector<int> cell(n), mid(n);
  c->d_cell.download(cell.data(), n, c->stream);
  c->d_mat.download(mid.data(), n, c->stream);
  HIPCHECK(hipStreamSynchronize(c->stream));
  std::vector<int64_t> count(c->ncells, 0);
  for (int64_t i = 0; i < n; ++i) {
    if (cell[i] < 0 || cell[i] >= c->ncells)
      throw ApiError(MPMX_E_STATE, "particle " + std::to_string(i) + " has no cell");
    ++count[cell[i]];
  }
  std::vector<double> vol(n), mass(n), rho(n);
  for (int64_t i = 0; i < n; ++i) {
    if (mid[i] < 0 || mid[i] >= (int)c->mats.size())
      throw ApiError(MPMX_E_STATE, "particle " + std::to_string(i) + " has no material");
    vol[i] = c->cell_volume[cell[i]] / (double)count[cell[i]];
    rho[i] = c->mats[mid[i]].density;
    mass[i] = vol[i] * rho[i];
  }
  c->d_vol.upload(vol.data(), n, c->stream);
  c->d_rho.upload(rho.data(), n, c->stream);
  c->d_mass.upload(mass.data(), n, c->stream);
  HIPCHECK(hipStreamSynchronize(c->stream));
  for (int64_t i = 0; i < n; ++i) c->psize[i] = particle_size(vol[i], c->dim);
  build_trac(c);
  API_END(c)
}

int mpmx_assign_materials(mpmx_ctx* c, int64_t n, const int64_t* particle, int material_id) {
  API_BEGIN(c)
  if (n < 0 || (n > 0 && !particle)) throw ApiError(MPMX_E_ARG, "bad particle list");
  if (material_id < 0 || material_id >= (int)c->mats.size())
    throw ApiError(MPMX_E_ARG, "bad material id");
  const int64_t np = c->np;
  if (n == 0 || np == 0) return MPMX_OK;
  std::vector<int> mid(np);
  c->d_mat.download(mid.data(), np, c->stream);
  std::vector<double> sv;
  if (c->any_mc) {
    sv.resize((size_t)NSV * np);
    c->d_sv.download(sv.data(), (size_t)NSV * np, c->stream);
  }
  HIPCHECK(hipStreamSynchronize(c->stream));
  const Material& m = c->mats[material_id];
  for (int64_t i = 0; i < n; ++i) {
    const int64_t p = particle[i];
    if (p < 0 || p >= np) throw ApiError(MPMX_E_ARG, "particle index out of range");
    mid[p] = material_id;
    if (c->any_mc)
      for (int k = 0; k < NSV; ++k) sv[(size_t)k * np + p] = m.type == MC ? m.sv0[k] : 0.0;
  }
  c->d_mat.upload(mid.data(), np, c->stream);
  if (c->any_mc) c->d_sv.upload(sv.data(), (size_t)NSV * np, c->stream);
  HIPCHECK(hipStreamSynchronize(c->stream));
  c->ad_recorded = false;
  API_END(c)
}

int mpmx_assign_stresses(mpmx_ctx* c, int64_t n, const double* stress) {
  API_BEGIN(c)
  if (n != c->np || !stress)
    throw ApiError(MPMX_E_ARG, "Number of particles in mesh and initial stresses don't match");
  put_soa(c->d_stress, c->np, 0, n, 6, stress, c->stream);
  API_END(c)
}

int mpmx_nparticles(const mpmx_ctx* c, int64_t* n) {
  if (!c || !n) return MPMX_E_ARG;
  *n = c->np;
  return MPMX_OK;
}

int mpmx_get_particles(mpmx_ctx* c, double* x, double* u, double* v, double* stress, double* strain,
                       double* F, double* mass, double* volume, double* density, int* material_id,
                       int64_t* cell, double* xi, double* statevars) {
  API_BEGIN(c)
  const int64_t np = c->np;
  const int d = c->dim;
  hipStream_t st = c->stream;
  HIPCHECK(hipStreamSynchronize(st));
  if (np == 0) return MPMX_OK;
  if (x) get_soa(c->d_x, np, d, x, st);
  if (u) get_soa(c->d_u, np, d, u, st);
  if (v) get_soa(c->d_v, np, d, v, st);
  if (stress) get_soa(c->d_stress, np, 6, stress, st);
  if (strain) get_soa(c->d_strain, np, 6, strain, st);
  if (F) get_soa(c->d_F, np, 9, F, st);
  if (mass) get_soa(c->d_mass, np, 1, mass, st);
  if (volume) get_soa(c->d_vol, np, 1, volume, st);
  if (density) get_soa(c->d_rho, np, 1, density, st);
  if (material_id) get_soa(c->d_mat, np, 1, material_id, st);
  if (xi) get_soa(c->d_xi, np, d, xi, st);
  if (cell) {
    std::vector<int> cc(np);
    get_soa(c->d_cell, np, 1, cc.data(), st);
    for (int64_t i = 0; i < np; ++i) cell[i] = cc[i];
  }
  if (statevars) {
    if (c->any_mc)
      get_soa(c->d_sv, np, NSV, statevars, st);
    else
      std::fill(statevars, statevars + np * NSV, 0.0);
  }
  API_END(c)
}

int mpmx_get_nodes(mpmx_ctx* c, double* mass, double* velocity, double* acceleration, int* active) {
  API_BEGIN(c)
  require_mesh(c);
  const int64_t nn = c->nnodes;
  const int d = c->dim;
  hipStream_t st = c->stream;
  std::vector<unsigned char> act(nn);
  c->d_nactive.download(act.data(), nn, st);
  HIPCHECK(hipStreamSynchronize(st));
  if (active)
    for (int64_t i = 0; i < nn; ++i) active[i] = act[i];
  if (mass) {
    c->d_nmass.download(mass, nn, st);
    HIPCHECK(hipStreamSynchronize(st));
  }
  if (velocity) get_soa(c->d_nvel, nn, d, velocity, st);
  if (acceleration) get_soa(c->d_nacc, nn, d, acceleration, st);
  API_END(c)
}

int mpmx_get_mapping_matrix(mpmx_ctx* c, double* B) {
  API_BEGIN(c)
  if (!B) throw ApiError(MPMX_E_ARG, "null array");
  const int64_t np = c->np;
  const int dd = c->dim * c->dim;
  if (np == 0) return MPMX_OK;
  std::vector<double> soa((size_t)dd * np);
  c->d_Bmat.download(soa.data(), soa.size(), c->stream);
  HIPCHECK(hipStreamSynchronize(c->stream));
  for (int64_t p = 0; p < np; ++p)
    for (int k = 0; k < dd; ++k) B[p * dd + k] = soa[(size_t)k * np + p];
  API_END(c)
}

int mpmx_set_mapping_matrix(mpmx_ctx* c, int64_t n, const double* B) {
  API_BEGIN(c)
  if (n != c->np || (n > 0 && !B)) throw ApiError(MPMX_E_ARG, "one matrix per particle required");
  const int dd = c->dim * c->dim;
  if (n == 0) return MPMX_OK;
  std::vector<double> soa((size_t)dd * n);
  for (int64_t p = 0; p < n; ++p)
    for (int k = 0; k < dd; ++k) soa[(size_t)k * n + p] = B[p * dd + k];
  c->d_Bmat.upload(soa.data(), soa.size(), c->stream);
  HIPCHECK(hipStreamSynchronize(c->stream));
  API_END(c)
}

int mpmx_step(mpmx_ctx* c, int nsteps, double dt, double damping_factor) {
  API_BEGIN(c)
  require_mesh(c);
  if (c->mats.empty()) throw ApiError(MPMX_E_STATE, "materials are not set");
  if (nsteps < 0 || !(dt > 0)) throw ApiError(MPMX_E_ARG, "nsteps >= 0 and dt > 0 required");
  if (nsteps == 0 || c->np == 0) {
    c->step += nsteps;
    return MPMX_OK;
  }
  reset_status(c);
  launch_steps(c, nsteps, dt, damping_factor);
  check_status(c);
  API_END(c)
}

int mpmx_step_noalloc(mpmx_ctx* c, int nsteps, double dt, double damping_factor) {
  API_BEGIN(c)
  require_mesh(c);
  if (c->mats.empty()) throw ApiError(MPMX_E_STATE, "materials are not set");
  if (nsteps < 0 || !(dt > 0)) throw ApiError(MPMX_E_ARG, "nsteps >= 0 and dt > 0 required");
  if (!c->locate_all)
    throw ApiError(MPMX_E_UNSUPPORTED,
                   "mpmx_step_noalloc: removing particles that leave the mesh reallocates the "
                   "particle state; use mpmx_step with locate_particles = 0");
  if (!c->scratch) throw ApiError(MPMX_E_HIP, "mpmx_step_noalloc: " + c->scratch_error);
  if (nsteps == 0 || c->np == 0) {
    c->step += nsteps;
    return MPMX_OK;
  }
  StepBuffers B;
  const size_t need = carve_step_buffers(c, 0, &B);
  if (need > kScratchBytes)
    throw ApiError(MPMX_E_ARG, "mpmx_step_noalloc: the step needs " + std::to_string(need) +
                                   " bytes of scratch, more than its maximum of " +
                                   std::to_string(kScratchBytes) + " (use mpmx_step)");
  carve_step_buffers(c, reinterpret_cast<uintptr_t>(c->scratch), &B);
  reset_status(c);
  const StepArgs A = step_args(c, dt, damping_factor);
  Launcher L{c};
  L.timed = false;
  // the scratch holds whatever the last call on this GPU left: sort fully at
  // the first step; the context's own permutation was not advanced either
  c->need_sort = true;
  for (int s = 0; s < nsteps; ++s) enqueue(c, A, L, next_step_sorts(c), B);
  HIPCHECK(hipGetLastError());
  c->need_sort = true;
  check_status(c);
  API_END(c)
}

int mpmx_set_sort_interval(mpmx_ctx* c, int steps) {
  API_BEGIN(c)
  if (steps < 1) throw ApiError(MPMX_E_ARG, "sort interval must be >= 1");
  c->sort_every = steps;
  API_END(c)
}

int mpmx_set_locate_particles(mpmx_ctx* c, int locate_particles) {
  API_BEGIN(c)
  c->locate_all = locate_particles != 0;
  c->invalidate_graph();
  API_END(c)
}

int mpmx_get_particle_ids(mpmx_ctx* c, int64_t* ids) {
  API_BEGIN(c)
  if (!ids && c->np > 0) throw ApiError(MPMX_E_ARG, "null ids");
  std::copy(c->ids.begin(), c->ids.end(), ids);
  API_END(c)
}

int mpmx_set_particle_ids(mpmx_ctx* c, int64_t n, const int64_t* ids) {
  API_BEGIN(c)
  if (n != c->np || (!ids && n > 0)) throw ApiError(MPMX_E_ARG, "one id per particle required");
  std::copy(ids, ids + n, c->ids.begin());
  API_END(c)
}

int mpmx_set_chunk(mpmx_ctx* c, int particles_per_workgroup) {
  API_BEGIN(c)
  const int ch = particles_per_workgroup;
  if (ch != 128 && ch != 256 && ch != 512)
    throw ApiError(MPMX_E_ARG, "particles per workgroup must be 128, 256 or 512");
  if (ch != c->chunk) {
    c->chunk = ch;
    c->invalidate_graph();
  }
  API_END(c)
}

int mpmx_autotune(mpmx_ctx* c, double dt, double damping_factor, int trial_steps,
                  int* chunk_out, int* sort_every_out) {
  API_BEGIN(c)
  require_mesh(c);
  if (c->mats.empty()) throw ApiError(MPMX_E_STATE, "materials are not set");
  if (!(dt > 0) || trial_steps < 1) throw ApiError(MPMX_E_ARG, "dt > 0 and trial_steps >= 1");
  if (c->np > 0 && !(c->flags & MPMX_ORDERED)) {
    // snapshot of the particle state (the node accumulators are zero between
    // steps), restored before every trial and at the end
    int64_t per = 0;
    for (const auto& a : ad_state(c)) per += a.comps;
    DBuf<double> snap;
    DBuf<int> snap_cell;
    snap.alloc((size_t)per * c->np);
    snap_cell.alloc(c->np);
    auto copy_state = [&](bool save) {
      int64_t off = 0;
      for (const auto& a : ad_state(c)) {
        const size_t bytes = (size_t)a.comps * c->np * sizeof(double);
        if (save)
          HIPCHECK(hipMemcpyAsync(snap.p + off, a.p, bytes, hipMemcpyDeviceToDevice, c->stream));
        else
          HIPCHECK(hipMemcpyAsync(a.p, snap.p + off, bytes, hipMemcpyDeviceToDevice, c->stream));
        off += a.comps * c->np;
      }
      const size_t cb = c->np * sizeof(int);
      if (save)
        HIPCHECK(hipMemcpyAsync(snap_cell.p, c->d_cell.p, cb, hipMemcpyDeviceToDevice, c->stream));
      else
        HIPCHECK(hipMemcpyAsync(c->d_cell.p, snap_cell.p, cb, hipMemcpyDeviceToDevice, c->stream));
    };
    const int64_t step0 = c->step;
    copy_state(true);
    hipEvent_t e0, e1;
    HIPCHECK(hipEventCreate(&e0));
    HIPCHECK(hipEventCreate(&e1));
    int best_chunk = c->chunk, best_sort = c->sort_every;
    float best_ms = 0.f;
    bool first = true;
    try {
      for (const int ch : {128, 256, 512})
        for (const int se : {4, 8, 16}) {
          copy_state(false);
          c->step = step0;
          c->chunk = ch;
          c->sort_every = se;
          c->need_sort = true;
          c->invalidate_graph();
          reset_status(c);
          launch_steps(c, 1, dt, damping_factor);  // capture, first sort
          HIPCHECK(hipEventRecord(e0, c->stream));
          launch_steps(c, trial_steps, dt, damping_factor);
          HIPCHECK(hipEventRecord(e1, c->stream));
          check_status(c);
          float ms = 0.f;
          HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
          if (first || ms < best_ms) {
            best_ms = ms;
            best_chunk = ch;
            best_sort = se;
            first = false;
          }
        }
    } catch (...) {
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      copy_state(false);
      c->step = step0;
      c->need_sort = true;
      c->invalidate_graph();
      (void)hipStreamSynchronize(c->stream);
      throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    copy_state(false);
    HIPCHECK(hipStreamSynchronize(c->stream));
    c->step = step0;
    c->chunk = best_chunk;
    c->sort_every = best_sort;
    c->need_sort = true;
    c->invalidate_graph();
  }
  if (chunk_out) *chunk_out = c->chunk;
  if (sort_every_out) *sort_every_out = c->sort_every;
  API_END(c)
}

int mpmx_set_step(mpmx_ctx* c, int64_t step) {
  API_BEGIN(c)
  if (step < 0) throw ApiError(MPMX_E_ARG, "negative step");
  c->step = step;
  API_END(c)
}

int mpmx_get_step(const mpmx_ctx* c, int64_t* step) {
  if (!c || !step) return MPMX_E_ARG;
  *step = c->step;
  return MPMX_OK;
}

int mpmx_synchronize(mpmx_ctx* c) {
  API_BEGIN(c)
  HIPCHECK(hipStreamSynchronize(c->stream));
  API_END(c)
}

int mpmx_profile(mpmx_ctx* c, int enable) {
  API_BEGIN(c)
  c->profile = enable != 0;
  if (c->profile) {
    c->kt.clear();
    c->kt_names.clear();
  }
  API_END(c)
}

int mpmx_kernel_times(mpmx_ctx* c, int cap, char* names, double* total_ms, int64_t* launches,
                      int* n) {
  API_BEGIN(c)
  const int cnt = (int)c->kt_names.size();
  if (n) *n = cnt;
  for (int i = 0; i < cnt && i < cap; ++i) {
    const auto& nm = c->kt_names[i];
    if (names) {
      std::memset(names + 48 * i, 0, 48);
      std::strncpy(names + 48 * i, nm.c_str(), 47);
    }
    if (total_ms) total_ms[i] = c->kt[nm].ms;
    if (launches) launches[i] = c->kt[nm].launches;
  }
  API_END(c)
}

int mpmx_adjoint_forward(mpmx_ctx* c, int nsteps, double dt, double damping_factor,
                         int checkpoint_every) {
  API_BEGIN(c)
  require_mesh(c);
  if (c->mats.empty()) throw ApiError(MPMX_E_STATE, "materials are not set");
  if (c->np == 0) throw ApiError(MPMX_E_STATE, "no particles");
  if (nsteps < 1 || !(dt > 0) || checkpoint_every < 1)
    throw ApiError(MPMX_E_ARG, "nsteps >= 1, dt > 0 and checkpoint_every >= 1 required");
  ad_check_supported(c);
  c->ad_recorded = false;
  const int K = checkpoint_every;
  const int nseg = (nsteps + K - 1) / K;
  int64_t per = 0;
  for (const auto& a : ad_state(c)) per += a.comps;
  c->ad_ck_stride = per * c->np;
  c->d_ad_ck.alloc((size_t)(nseg + 1) * c->ad_ck_stride);
  c->d_ad_ck_cell.alloc((size_t)(nseg + 1) * c->np);
  c->ad_step0 = c->step;
  c->ad_last_recorded = false;
  ad_alloc_slots(c, K);
  reset_status(c);
  for (int s = 0; s < nseg; ++s) {
    const int ns = std::min(K, nsteps - s * K);
    ad_checkpoint(c, s, true);
    if (s < nseg - 1) {
      launch_steps(c, ns, dt, damping_factor);
      continue;
    }
    // the last segment runs recording its step records, as the backward
    // pass's replay of a segment does (slot index from 0)
    check_status(c);
    reset_status(c);
    for (int j = 0; j < ns; ++j) {
      ad_save_slot(c, j);
      launch_steps(c, 1, dt, damping_factor, c->d_ad_node_slots.p, ad_node_slot_doubles(c), K);
    }
    ad_save_slot(c, ns);
  }
  ad_checkpoint(c, nseg, true);
  check_status(c);
  c->ad_last_recorded = true;
  c->ad_every = K;
  c->ad_nsteps = nsteps;
  c->ad_dt = dt;
  c->ad_damping = damping_factor;
  c->ad_np = c->np;
  c->ad_recorded = true;
  API_END(c)
}

}  // extern "C"

namespace {
// The backward pass; `dev`: the loss gradients come in and the parameter
// gradients go out as arrays on the context's device, else on the host.
void adjoint_backward(mpmx_ctx* c, bool dev, const double* dL_dx, const double* dL_dv,
                      double* dL_dE, double* dL_drho, double* dL_dphi, double* dL_dx0,
                      double* dL_dv0) {
  if (!c->ad_recorded || c->ad_np != c->np)
    throw ApiError(MPMX_E_STATE, "no recorded run (mpmx_adjoint_forward) for these particles");
  const int K = c->ad_every;
  const int nseg = (c->ad_nsteps + K - 1) / K;
  const int64_t np = c->np;
  const int d = c->dim;
  hipStream_t st = c->stream;
  if (!c->ad_last_recorded) ad_alloc_slots(c, K);
  c->d_ad_nodes.alloc(ad_node_doubles(c));
  c->d_ad_parts.alloc(ad_part_doubles(c));
  c->d_ad_list.alloc(2 + 2 * (size_t)np);
  c->d_ad_list_buf.alloc(6 * (size_t)np);
  launch_ad_fill(c->d_ad_parts.p, ad_part_doubles(c), 0., st);
  const AdjParticles L = ad_parts(c);
  auto put_rows = [&](double* soa, int comps, const double* rows) {
    if (dev) {
      launch_rows_to_soa(rows, soa, np, comps, st);
      HIPCHECK(hipGetLastError());
      return;
    }
    std::vector<double> col(np);
    for (int k = 0; k < comps; ++k) {
      for (int64_t i = 0; i < np; ++i) col[i] = rows[i * comps + k];
      HIPCHECK(hipMemcpyAsync(soa + k * np, col.data(), np * sizeof(double),
                              hipMemcpyHostToDevice, st));
      HIPCHECK(hipStreamSynchronize(st));
    }
  };
  auto get_rows = [&](const double* soa, int comps, double* rows, double scale,
                      const std::vector<double>* per) {
    if (dev) {
      launch_soa_to_rows(soa, rows, np, comps, scale, per ? c->d_mass.p : nullptr,
                         per ? c->d_mat.p : nullptr, c->d_mats.p, st);
      HIPCHECK(hipGetLastError());
      return;
    }
    std::vector<double> col(np);
    for (int k = 0; k < comps; ++k) {
      HIPCHECK(hipMemcpyAsync(col.data(), soa + k * np, np * sizeof(double),
                              hipMemcpyDeviceToHost, st));
      HIPCHECK(hipStreamSynchronize(st));
      for (int64_t i = 0; i < np; ++i)
        rows[i * comps + k] = col[i] * scale * (per ? (*per)[i] : 1.0);
    }
  };
  if (dL_dx) put_rows(L.lx, d, dL_dx);
  if (dL_dv) put_rows(L.lv, d, dL_dv);
  // segments from the last: replay from the checkpoint recording each step,
  // then transpose the steps in reverse
  for (int s = nseg - 1; s >= 0; --s) {
    const int ns = std::min(K, c->ad_nsteps - s * K);
    if (s == nseg - 1 && c->ad_last_recorded) {
      c->ad_last_recorded = false;  // the replays below overwrite the records
    } else {
      ad_checkpoint(c, s, false);
      c->step = c->ad_step0 + (int64_t)s * K;
      c->need_sort = true;
      reset_status(c);
      for (int j = 0; j < ns; ++j) {  // reset_status set the slot index to 0
        ad_save_slot(c, j);
        launch_steps(c, 1, c->ad_dt, c->ad_damping, c->d_ad_node_slots.p,
                     ad_node_slot_doubles(c), K);
      }
      ad_save_slot(c, ns);
      check_status(c);
    }
    for (int j = ns - 1; j >= 0; --j) ad_backward_step(c, j, c->ad_step0 + (int64_t)s * K + j);
  }
  HIPCHECK(hipStreamSynchronize(st));
  // parameter gradients: E per particle; phi in degrees (the parameter
  // vector's unit); density through the mass m = V rho of the material
  std::vector<double> m_over_rho(dev ? 0 : np);  // device: taken from mass and material
  if (!dev) {
    std::vector<double> mass(np);
    std::vector<int> mat(np);
    HIPCHECK(hipMemcpyAsync(mass.data(), c->d_mass.p, np * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipMemcpyAsync(mat.data(), c->d_mat.p, np * sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    for (int64_t i = 0; i < np; ++i) m_over_rho[i] = mass[i] / c->mats[mat[i]].density;
  }
  if (dL_dE) get_rows(L.gE, 1, dL_dE, 1.0, nullptr);
  if (dL_dphi) get_rows(L.gphi, 1, dL_dphi, M_PI / 180., nullptr);
  if (dL_drho) get_rows(L.gm, 1, dL_drho, 1.0, &m_over_rho);
  if (dL_dx0) get_rows(L.lx, d, dL_dx0, 1.0, nullptr);
  if (dL_dv0) get_rows(L.lv, d, dL_dv0, 1.0, nullptr);
  // leave the context at the end of the recorded run
  ad_checkpoint(c, nseg, false);
  c->step = c->ad_step0 + c->ad_nsteps;
  c->need_sort = true;
  HIPCHECK(hipStreamSynchronize(st));
}
}  // namespace

extern "C" {

int mpmx_adjoint_backward(mpmx_ctx* c, const double* dL_dx, const double* dL_dv, double* dL_dE,
                          double* dL_drho, double* dL_dphi, double* dL_dx0, double* dL_dv0) {
  API_BEGIN(c)
  adjoint_backward(c, false, dL_dx, dL_dv, dL_dE, dL_drho, dL_dphi, dL_dx0, dL_dv0);
  API_END(c)
}

int mpmx_adjoint_backward_device(mpmx_ctx* c, const double* dL_dx, const double* dL_dv,
                                 double* dL_dE, double* dL_drho, double* dL_dphi,
                                 double* dL_dx0, double* dL_dv0) {
  API_BEGIN(c)
  adjoint_backward(c, true, dL_dx, dL_dv, dL_dE, dL_drho, dL_dphi, dL_dx0, dL_dv0);
  API_END(c)
}

int mpmx_get_particles_device(mpmx_ctx* c, double* x, double* v) {
  API_BEGIN(c)
  if (c->np == 0) return MPMX_OK;
  if (x) launch_soa_to_rows(c->d_x.p, x, c->np, c->dim, 1.0, nullptr, nullptr, nullptr, c->stream);
  if (v) launch_soa_to_rows(c->d_v.p, v, c->np, c->dim, 1.0, nullptr, nullptr, nullptr, c->stream);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(c->stream));
  API_END(c)
}


int mpmx_adjoint_set_apex_ratio(mpmx_ctx* c, double ratio) {
  API_BEGIN(c)
  if (!(ratio == ratio)) throw ApiError(MPMX_E_ARG, "apex ratio is NaN");
  c->ad_apex_ratio = ratio;
  API_END(c)
}

int mpmx_adjoint_release(mpmx_ctx* c) {
  API_BEGIN(c)
  c->ad_recorded = false;
  c->ad_last_recorded = false;
  c->d_ad_ck.release();
  c->d_ad_ck_cell.release();
  c->d_ad_slots.release();
  c->d_ad_slot_cell.release();
  c->d_ad_node_slots.release();
  for (int k = 2; k < 4; ++k)
    if (c->graph_exec[k]) {
      (void)hipGraphExecDestroy(c->graph_exec[k]);
      c->graph_exec[k] = nullptr;
    }
  c->d_ad_nodes.release();
  c->d_ad_parts.release();
  c->d_ad_list.release();
  c->d_ad_list_buf.release();
  API_END(c)
}

}  // extern "C"
