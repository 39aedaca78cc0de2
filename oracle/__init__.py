"""ctypes front-end of the oracle (CPU restatement of the reference step).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker. The product package
(mpm_amd) never imports this module.

``Oracle`` mirrors the product's solver API (same argument meaning, same
particle order) so parity tests drive both the same way. ``omp=True`` loads
the OpenMP build used as the timed CPU baseline; the default single-threaded
build sums nodal quantities in particle order and is the deterministic oracle.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}

_D = C.POINTER(C.c_double)
_I64 = C.POINTER(C.c_int64)
_I32 = C.POINTER(C.c_int)

SCHEMES = {"usf": 0, "usl": 1, "musl": 2}
VELOCITY_UPDATES = {"flip": 0, "pic": 1, "asflip": 2, "tpic": 3, "apic": 4}

# Phase ids of orc_phase (each is one reference sweep)
PH = dict(reset_nodes=0, shapefn=1, map_mass_momentum=2, compute_velocity=3,
          deformation_gradient_increment=4, strain=5, volume=6, stress=7,
          update_deformation_gradient=8, body_force=9, internal_force=10,
          acceleration_velocity=11, position=12, locate=13,
          acceleration_velocity_plain=14)
NODE_FIELDS = dict(mass=0, momentum=1, velocity=2, acceleration=3, fext=4, fint=5, status=6)
PARTICLE_FIELDS = dict(x=0, v=1, stress=2, volume=3, mass=4, density=5, u=6, strain=7, sv=8)


_BUILDS = {"portable": "liboracle.so", "platform": "liboracle_libm.so",
           "omp": "liboracle_omp.so"}


def build(force=False):
    """Compile the oracle builds with the committed Makefile."""
    if force or not all(os.path.exists(os.path.join(_HERE, f)) for f in _BUILDS.values()):
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []), check=True)


def _lib(omp=False, libm="portable"):
    """omp: OpenMP build (platform libm); else libm 'portable' or 'platform'."""
    name = _BUILDS["omp"] if omp else _BUILDS[libm]
    if name not in _LIBS:
        path = os.path.join(_HERE, name)
        if not os.path.exists(path):
            build()
        lib = C.CDLL(path)
        lib.orc_create.restype = C.c_void_p
        lib.orc_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double]
        lib.orc_destroy.argtypes = [C.c_void_p]
        lib.orc_last_error.restype = C.c_char_p
        lib.orc_last_error.argtypes = [C.c_void_p]
        for fn in ("orc_set_mesh", "orc_set_materials", "orc_set_velocity_constraints",
                   "orc_set_acceleration_constraints", "orc_set_friction_constraints",
                   "orc_set_gravity", "orc_generate_gauss", "orc_set_particles",
                   "orc_nparticles", "orc_get_particles", "orc_get_nodes", "orc_get_cells",
                   "orc_step", "orc_phase", "orc_set_node_field", "orc_set_particle_field",
                   "orc_get_particle_scratch", "orc_assign_velocity",
                   "orc_assign_stresses", "orc_set_particle_velocity_constraints",
                   "orc_set_acceleration_constraints_fn", "orc_set_math_functions",
                   "orc_set_nodal_forces", "orc_set_particle_tractions",
                   "orc_assign_materials", "orc_compute_volume_mass", "orc_set_isoparametric",
                   "orc_kat_transform_real_to_unit_cell", "orc_kat_point_in_cell",
                   "orc_set_nodal_euler_angles", "orc_get_mapping_matrix", "orc_kat_mc_rate",
                   "orc_set_mapping_matrix"):
            getattr(lib, fn).restype = C.c_int
        lib.orc_kat_cell_volume.restype = C.c_double
        lib.orc_kat_mc_yield.restype = C.c_int
        lib.orc_kat_node_update.restype = C.c_int
        lib.orc_threads.restype = C.c_int
        lib.orc_kat_math.restype = C.c_double
        lib.orc_kat_math.argtypes = [C.c_int, C.c_double]
        lib.orc_kat_math_function.restype = C.c_double
        lib.orc_kat_math_function.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_double]
        _LIBS[name] = lib
    return _LIBS[name]


def pack_math_functions(functions):
    """(type, npts, params) arrays of math functions given as ("linear",
    xvalues, fxvalues) or ("sin", x0, a, (xmin, xmax))."""
    types, npts, params = [], [], []
    for f in functions:
        if f[0] == "linear":
            x = np.asarray(f[1], dtype=np.float64)
            fx = np.asarray(f[2], dtype=np.float64)
            if x.shape != fx.shape:
                raise ValueError("xvalues and fxvalues differ in length")
            types.append(0)
            npts.append(len(x))
            params += list(x) + list(fx)
        elif f[0] == "sin":
            types.append(1)
            npts.append(4)
            params += [float(f[1]), float(f[2]), float(f[3][0]), float(f[3][1])]
        else:
            raise ValueError("unknown math function %r" % (f[0],))
    return (np.asarray(types, dtype=np.int32), np.asarray(npts, dtype=np.int32),
            np.asarray(params if params else [0.0], dtype=np.float64))


def math_function_value(function, x):
    """The oracle's value of one math function (as pack_math_functions) at x."""
    lib = _lib()
    t, n, prm = pack_math_functions([function])
    return lib.orc_kat_math_function(int(t[0]), int(n[0]), prm.ctypes.data, float(x))


def _p(a, ctype=_D):
    if a is None:
        return None
    return a.ctypes.data_as(ctype)


def _f64(a, shape=None):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


class OracleError(RuntimeError):
    pass


class Oracle:
    """One reference-semantics solver instance (particles, nodes, cells)."""

    def __init__(self, dim, scheme="usf", velocity_update="flip", blending_ratio=1.0, omp=False,
                 libm="portable"):
        self.lib = _lib(omp, libm)
        self.dim = dim
        self.nn = 4 if dim == 2 else 8
        ptr = self.lib.orc_create(dim, SCHEMES[scheme], VELOCITY_UPDATES[velocity_update],
                                  C.c_double(blending_ratio))
        if not ptr:
            raise OracleError("orc_create failed")
        self.ctx = C.c_void_p(ptr)
        self.nnodes = 0
        self.ncells = 0

    def __del__(self):
        if getattr(self, "ctx", None):
            self.lib.orc_destroy(self.ctx)
            self.ctx = None

    def _check(self, rc):
        if rc != 0:
            raise OracleError(self.lib.orc_last_error(self.ctx).decode())

    @property
    def threads(self):
        return self.lib.orc_threads()

    def set_mesh(self, coords, cells, isoparametric=False):
        coords = _f64(coords)
        cells = np.ascontiguousarray(cells, dtype=np.int64)
        self.nnodes, self.ncells = coords.shape[0], cells.shape[0]
        self._check(self.lib.orc_set_isoparametric(self.ctx, int(bool(isoparametric))))
        self._check(self.lib.orc_set_mesh(self.ctx, C.c_int64(self.nnodes), _p(coords),
                                          C.c_int64(self.ncells), _p(cells, _I64)))

    def set_materials(self, params):
        params = _f64(params).reshape(-1, 17)
        self._check(self.lib.orc_set_materials(self.ctx, params.shape[0], _p(params)))

    def _constraints(self, fn, nodes, dirs, vals, ints=False):
        nodes = np.ascontiguousarray(nodes, dtype=np.int64)
        dirs = np.ascontiguousarray(np.broadcast_to(dirs, nodes.shape), dtype=np.int32)
        if ints:
            vals = [np.ascontiguousarray(np.broadcast_to(vals[0], nodes.shape), dtype=np.int32),
                    np.ascontiguousarray(np.broadcast_to(vals[1], nodes.shape), dtype=np.float64)]
            self._check(fn(self.ctx, C.c_int64(len(nodes)), _p(nodes, _I64), _p(dirs, _I32),
                           _p(vals[0], _I32), _p(vals[1])))
        else:
            vals = np.ascontiguousarray(np.broadcast_to(vals, nodes.shape), dtype=np.float64)
            self._check(fn(self.ctx, C.c_int64(len(nodes)), _p(nodes, _I64), _p(dirs, _I32),
                           _p(vals)))

    def set_velocity_constraints(self, nodes, dirs, values):
        self._constraints(self.lib.orc_set_velocity_constraints, nodes, dirs, values)

    def set_acceleration_constraints(self, nodes, dirs, values, functions=None):
        """functions: per constraint the index of its math function (-1: none)."""
        if functions is None:
            self._constraints(self.lib.orc_set_acceleration_constraints, nodes, dirs, values)
            return
        nodes = np.ascontiguousarray(nodes, dtype=np.int64)
        dirs = np.ascontiguousarray(np.broadcast_to(dirs, nodes.shape), dtype=np.int32)
        vals = np.ascontiguousarray(np.broadcast_to(values, nodes.shape), dtype=np.float64)
        fns = np.ascontiguousarray(np.broadcast_to(functions, nodes.shape), dtype=np.int32)
        self._check(self.lib.orc_set_acceleration_constraints_fn(
            self.ctx, C.c_int64(len(nodes)), _p(nodes, _I64), _p(dirs, _I32), _p(vals),
            _p(fns, _I32)))

    def set_nodal_forces(self, nodes, dirs, forces, functions=None):
        """Concentrated nodal forces (node -1: every node)."""
        nodes = np.ascontiguousarray(nodes, dtype=np.int64)
        dirs = np.ascontiguousarray(np.broadcast_to(dirs, nodes.shape), dtype=np.int32)
        vals = np.ascontiguousarray(np.broadcast_to(forces, nodes.shape), dtype=np.float64)
        fns = np.ascontiguousarray(np.broadcast_to(-1 if functions is None else functions,
                                                   nodes.shape), dtype=np.int32)
        self._check(self.lib.orc_set_nodal_forces(self.ctx, C.c_int64(len(nodes)),
                                                  _p(nodes, _I64), _p(dirs, _I32), _p(vals),
                                                  _p(fns, _I32)))

    def set_particle_tractions(self, particles, dirs, values, functions=None):
        """Particle surface tractions in order (particle -1: every particle)."""
        particles = np.ascontiguousarray(particles, dtype=np.int64)
        dirs = np.ascontiguousarray(np.broadcast_to(dirs, particles.shape), dtype=np.int32)
        vals = np.ascontiguousarray(np.broadcast_to(values, particles.shape), dtype=np.float64)
        fns = np.ascontiguousarray(np.broadcast_to(-1 if functions is None else functions,
                                                   particles.shape), dtype=np.int32)
        self._check(self.lib.orc_set_particle_tractions(self.ctx, C.c_int64(len(particles)),
                                                        _p(particles, _I64), _p(dirs, _I32),
                                                        _p(vals), _p(fns, _I32)))

    def set_math_functions(self, functions):
        """functions: list of ("linear", xvalues, fxvalues) or
        ("sin", x0, a, (xmin, xmax))."""
        types, npts, params = pack_math_functions(functions)
        self._check(self.lib.orc_set_math_functions(
            self.ctx, C.c_int(len(types)), _p(types, _I32), _p(npts, _I32), _p(params)))

    def set_friction_constraints(self, nodes, dirs, sign_n, mu):
        self._constraints(self.lib.orc_set_friction_constraints, nodes, dirs, (sign_n, mu),
                          ints=True)

    def set_particle_velocity_constraints(self, particles, dirs, values):
        pa = np.ascontiguousarray(particles, dtype=np.int64)
        di = np.ascontiguousarray(dirs, dtype=np.int32)
        va = np.ascontiguousarray(values, dtype=np.float64)
        self._check(self.lib.orc_set_particle_velocity_constraints(
            self.ctx, C.c_int64(len(pa)), _p(pa, _I64), _p(di, _I32), _p(va)))

    def set_velocity_update(self, velocity_update, blending_ratio=1.0):
        self.lib.orc_set_velocity_update(self.ctx, VELOCITY_UPDATES[velocity_update],
                                         C.c_double(blending_ratio))

    def compute_volume_mass(self):
        self._check(self.lib.orc_compute_volume_mass(self.ctx))

    def assign_materials(self, particles, material_id):
        idx = np.ascontiguousarray(particles, dtype=np.int64)
        self._check(self.lib.orc_assign_materials(self.ctx, C.c_int64(len(idx)),
                                                  idx.ctypes.data_as(C.c_void_p),
                                                  int(material_id)))

    def set_locate_particles(self, locate_particles=True):
        self.lib.orc_set_locate_particles(self.ctx, 1 if locate_particles else 0)

    def set_nodal_euler_angles(self, nodes, angles):
        """General (inclined) boundaries: the constraints of these nodes act in
        the frame rotated by the Euler angles (rows of dim)."""
        nodes = np.ascontiguousarray(nodes, dtype=np.int64)
        ang = _f64(angles).reshape(len(nodes), -1)
        self._check(self.lib.orc_set_nodal_euler_angles(self.ctx, C.c_int64(len(nodes)),
                                                        _p(nodes, _I64), _p(ang)))

    def set_gravity(self, g, ramping_time=0.0):
        g = _f64(g)
        self._check(self.lib.orc_set_gravity(self.ctx, _p(g), C.c_double(ramping_time)))

    def generate_gauss(self, nparticles_per_dir, cell_set=None, material_id=0):
        if cell_set is None:
            self._check(self.lib.orc_generate_gauss(self.ctx, nparticles_per_dir, C.c_int64(-1),
                                                    None, material_id))
        else:
            cs = np.ascontiguousarray(cell_set, dtype=np.int64)
            self._check(self.lib.orc_generate_gauss(self.ctx, nparticles_per_dir,
                                                    C.c_int64(len(cs)), _p(cs, _I64),
                                                    material_id))

    def set_particles(self, x, mass, volume, density, material_id, v=None, u=None,
                      stress=None, strain=None, F=None, cell=None, statevars=None):
        d = self.dim
        x = _f64(x, (-1, d))
        n = x.shape[0]
        arrs = dict(u=_f64(u, (n, d)), v=_f64(v, (n, d)), stress=_f64(stress, (n, 6)),
                    strain=_f64(strain, (n, 6)), F=_f64(F, (n, 9)), mass=_f64(mass, (n,)),
                    volume=_f64(volume, (n,)), density=_f64(density, (n,)),
                    sv=_f64(statevars, (n, 9)))
        mat = np.ascontiguousarray(np.broadcast_to(material_id, (n,)), dtype=np.int32)
        cell = None if cell is None else np.ascontiguousarray(cell, dtype=np.int64)
        self._check(self.lib.orc_set_particles(
            self.ctx, C.c_int64(n), _p(x), _p(arrs["u"]), _p(arrs["v"]), _p(arrs["stress"]),
            _p(arrs["strain"]), _p(arrs["F"]), _p(arrs["mass"]), _p(arrs["volume"]),
            _p(arrs["density"]), _p(mat, _I32), _p(cell, _I64), _p(arrs["sv"])))

    def assign_velocity(self, v):
        v = _f64(v, (-1, self.dim))
        self._check(self.lib.orc_assign_velocity(self.ctx, C.c_int64(v.shape[0]), _p(v)))

    def assign_stresses(self, stress):
        stress = _f64(stress, (-1, 6))
        self._check(self.lib.orc_assign_stresses(self.ctx, C.c_int64(stress.shape[0]), _p(stress)))

    @property
    def nparticles(self):
        n = C.c_int64(0)
        self.lib.orc_nparticles(self.ctx, C.byref(n))
        return n.value

    def particles(self, fields=None):
        n, d = self.nparticles, self.dim
        out = dict(x=np.zeros((n, d)), u=np.zeros((n, d)), v=np.zeros((n, d)),
                   stress=np.zeros((n, 6)), strain=np.zeros((n, 6)), F=np.zeros((n, 9)),
                   mass=np.zeros(n), volume=np.zeros(n), density=np.zeros(n),
                   material_id=np.zeros(n, dtype=np.int32), cell=np.zeros(n, dtype=np.int64),
                   xi=np.zeros((n, d)), statevars=np.zeros((n, 9)))
        self.lib.orc_get_particles(
            self.ctx, _p(out["x"]), _p(out["u"]), _p(out["v"]), _p(out["stress"]),
            _p(out["strain"]), _p(out["F"]), _p(out["mass"]), _p(out["volume"]),
            _p(out["density"]), _p(out["material_id"], _I32), _p(out["cell"], _I64),
            _p(out["xi"]), _p(out["statevars"]))
        return out

    def nodes(self):
        n, d = self.nnodes, self.dim
        out = dict(mass=np.zeros(n), momentum=np.zeros((n, d)), velocity=np.zeros((n, d)),
                   acceleration=np.zeros((n, d)), fext=np.zeros((n, d)), fint=np.zeros((n, d)),
                   status=np.zeros(n, dtype=np.int32))
        self.lib.orc_get_nodes(self.ctx, _p(out["mass"]), _p(out["momentum"]),
                               _p(out["velocity"]), _p(out["acceleration"]), _p(out["fext"]),
                               _p(out["fint"]), _p(out["status"], _I32))
        return out

    def cells(self):
        n = self.ncells
        out = dict(volume=np.zeros(n), mean_length=np.zeros(n), centroid=np.zeros((n, self.dim)))
        self.lib.orc_get_cells(self.ctx, _p(out["volume"]), _p(out["mean_length"]),
                               _p(out["centroid"]))
        return out

    def step(self, nsteps, dt, damping_factor=0.0):
        self._check(self.lib.orc_step(self.ctx, int(nsteps), C.c_double(dt),
                                      C.c_double(damping_factor)))

    # -- replay of reference unit-test call sequences -------------------------
    def phase(self, name, dt=0.0, arg=None):
        a = None if arg is None else _f64(np.atleast_1d(arg))
        self._check(self.lib.orc_phase(self.ctx, PH[name], C.c_double(dt), _p(a)))

    def set_node_field(self, node, field, value):
        v = _f64(np.atleast_1d(value))
        self._check(self.lib.orc_set_node_field(self.ctx, C.c_int64(node), NODE_FIELDS[field],
                                                _p(v)))

    def set_particle_field(self, pid, field, value):
        v = _f64(np.atleast_1d(value))
        self._check(self.lib.orc_set_particle_field(self.ctx, C.c_int64(pid),
                                                    PARTICLE_FIELDS[field], _p(v)))

    def mapping_matrix(self):
        """The affine velocity updates' mapping matrices, (np, dim, dim)."""
        B = np.zeros((self.nparticles, self.dim, self.dim))
        self._check(self.lib.orc_get_mapping_matrix(self.ctx, _p(B)))
        return B

    def set_mapping_matrix(self, B):
        B = _f64(B)
        self._check(self.lib.orc_set_mapping_matrix(self.ctx, C.c_int64(B.shape[0]), _p(B)))

    def particle_scratch(self, pid):
        nn, d = self.nn, self.dim
        out = dict(N=np.zeros(nn), dndx=np.zeros((nn, d)), dstrain=np.zeros(6),
                   dF=np.zeros((3, 3)), dvol=np.zeros(1))
        self._check(self.lib.orc_get_particle_scratch(
            self.ctx, C.c_int64(pid), _p(out["N"]), _p(out["dndx"]), _p(out["dstrain"]),
            _p(out["dF"]), _p(out["dvol"])))
        out["dvol"] = float(out["dvol"][0])
        return out


# -- function-level known-answer hooks ----------------------------------------
def math_fn(which, x, libm="portable"):
    """Elementary function of one oracle build (see orc_kat_math)."""
    lib = _lib(libm=libm)
    names = ("sin", "cos", "tan", "acos", "pow_1p5", "pow_m1p5", "pow_m2p5")
    w = names.index(which)
    return np.array([lib.orc_kat_math(w, float(v)) for v in np.atleast_1d(x)])


def shapefn(dim, xi):
    lib = _lib()
    nn = 4 if dim == 2 else 8
    xi = _f64(xi)
    N = np.zeros(nn)
    g = np.zeros((nn, dim))
    lib.orc_kat_shapefn(dim, _p(xi), _p(N), _p(g))
    return N, g


def dn_dx(dim, xi, X):
    lib = _lib()
    xi, X = _f64(xi), _f64(X)
    out = np.zeros_like(X)
    lib.orc_kat_dn_dx(dim, _p(xi), _p(X), _p(out))
    return out


def jacobian(dim, xi, X):
    lib = _lib()
    xi, X = _f64(xi), _f64(X)
    J = np.zeros((dim, dim))
    lib.orc_kat_jacobian(dim, _p(xi), _p(X), _p(J))
    return J


def transform_real_to_unit_cell(dim, X, point):
    """Isoparametric local coordinates of `point` in the one cell with
    vertices X (nn x dim); None when the reference would throw (NaN)."""
    X, point = _f64(X), _f64(point)
    xi = np.zeros(dim)
    rc = _lib().orc_kat_transform_real_to_unit_cell(dim, _p(X), _p(point), _p(xi))
    return None if rc != 0 else xi


def point_in_cell(dim, X, point, isoparametric=True):
    """(inside, xi) of is_point_in_cell for the one cell with vertices X."""
    X, point = _f64(X), _f64(point)
    xi = np.zeros(dim)
    r = _lib().orc_kat_point_in_cell(dim, int(bool(isoparametric)), _p(X), _p(point), _p(xi))
    return bool(r), xi


def cell_volume(dim, X):
    return _lib().orc_kat_cell_volume(dim, _p(_f64(X)))


def stress_update(params, stress, dstrain, density=1000.0, statevars=None):
    """One constitutive update; returns (new stress, state variables)."""
    lib = _lib()
    params, stress, dstrain = _f64(params), _f64(stress), _f64(dstrain)
    out = np.zeros(6)
    init = statevars is None
    sv = np.zeros(9) if init else _f64(statevars).copy()
    lib.orc_kat_stress(_p(params), _p(stress), _p(dstrain), C.c_double(density), _p(sv),
                       1 if init else 0, _p(out))
    return out, sv


def le_rate_update(params, stress, dstrain, dF, F=None):
    """One LinearElastic update with the material's objective stress rate
    (params[4]: 0 none, 1 Jaumann, 2 Green-Naghdi) for a deformation gradient
    increment dF (3x3) at the step-start deformation gradient F (3x3,
    default identity)."""
    lib = _lib()
    out = np.zeros(6)
    Fp = None if F is None else _f64(np.asarray(F, dtype=np.float64).reshape(9))
    rc = lib.orc_kat_le_rate(_p(_f64(params)), _p(_f64(stress)), _p(_f64(dstrain)),
                             _p(_f64(np.asarray(dF, dtype=np.float64).reshape(9))), _p(Fp),
                             _p(out))
    if rc != 0:
        raise ValueError("not a LinearElastic material")
    return out


def mc_rate_update(params, stress, dstrain, dF, F=None, density=1000.0, statevars=None):
    """One MohrCoulomb update with the material's stress rate (params[16]);
    returns (stress, state variables)."""
    lib = _lib()
    out = np.zeros(6)
    init = statevars is None
    sv = np.zeros(9) if init else _f64(statevars).copy()
    Fp = None if F is None else _f64(np.asarray(F, dtype=np.float64).reshape(9))
    rc = lib.orc_kat_mc_rate(_p(_f64(params)), _p(_f64(stress)), _p(_f64(dstrain)),
                             _p(_f64(np.asarray(dF, dtype=np.float64).reshape(9))), _p(Fp),
                             C.c_double(density), _p(sv), 1 if init else 0, _p(out))
    if rc != 0:
        raise ValueError("not a MohrCoulomb material")
    return out, sv


def polar_rotation(F):
    R = np.zeros(9)
    _lib().orc_kat_polar_rotation(_p(_f64(np.asarray(F, dtype=np.float64).reshape(9))), _p(R))
    return R.reshape(3, 3)


def invariants(stress):
    lib = _lib()
    s = _f64(stress)
    out = np.zeros(5)
    dth = np.zeros(6)
    lib.orc_kat_invariants(_p(s), _p(out), _p(dth))
    return dict(p=out[0], q=out[1], j2=out[2], j3=out[3], lode_angle=out[4], dtheta_dsigma=dth)


def mc_initial_state(params):
    sv = np.zeros(9)
    _lib().orc_kat_mc_init_state(_p(_f64(params)), _p(sv))
    return sv


def mc_invariants(stress, statevars):
    sv = _f64(statevars).copy()
    _lib().orc_kat_mc_invariants(_p(_f64(stress)), _p(sv))
    return sv


def mc_yield(statevars):
    yf = np.zeros(2)
    t = _lib().orc_kat_mc_yield(_p(_f64(statevars)), _p(yf))
    return t, yf


def mc_df_dp(params, yield_type, statevars, stress):
    df, dp = np.zeros(6), np.zeros(6)
    dpdq, soft = C.c_double(0), C.c_double(0)
    _lib().orc_kat_mc_df_dp(_p(_f64(params)), int(yield_type), _p(_f64(statevars)),
                            _p(_f64(stress)), _p(df), _p(dp), C.byref(dpdq), C.byref(soft))
    return df, dp, dpdq.value, soft.value


def node_update(dim, mass, fext, fint, vel, acc, dt, damping=0.0, vel_constraints=None,
                acc_constraints=None, friction=None):
    """compute_acceleration_velocity_cundall on one node. friction=(dir, sign, mu)."""
    lib = _lib()
    v = _f64(vel).copy()
    a = _f64(acc).copy()
    vc = None if vel_constraints is None else _f64(vel_constraints)
    ac = None if acc_constraints is None else _f64(acc_constraints)
    fd, fs, fm = (-1, 0, 0.0) if friction is None else friction
    st = lib.orc_kat_node_update(dim, C.c_double(mass), _p(_f64(fext)), _p(_f64(fint)), _p(v),
                                 _p(a), C.c_double(dt), C.c_double(damping), _p(vc), _p(ac),
                                 int(fd), int(fs), C.c_double(fm))
    return bool(st), v, a


def node_constraints(dim, vel, acc, angles=None, velocity=(), acceleration=(), friction=None,
                     dt=0.0):
    """Constraints of one node (general boundary when `angles`, the Euler
    angles, are given): velocity [(dir, value)] -> apply_velocity_constraints,
    acceleration [(dir, value)] -> apply_acceleration_constraints, friction
    (dir, sign_n, mu) -> apply_friction_constraints(dt). Returns (vel, acc)."""
    v, a = _f64(vel).copy(), _f64(acc).copy()
    ang = None if angles is None else _f64(angles)
    vd = np.array([d for d, _ in velocity] or [0], dtype=np.int32)
    vv = np.array([x for _, x in velocity] or [0.0])
    ad = np.array([d for d, _ in acceleration] or [0], dtype=np.int32)
    av = np.array([x for _, x in acceleration] or [0.0])
    which = 0 if velocity else (1 if acceleration else 2)
    fd, fs, fm = friction if friction is not None else (0, 0, 0.0)
    _lib().orc_kat_node_constraints(dim, _p(ang), _p(v), _p(a), len(velocity), _p(vd, _I32),
                                    _p(vv), len(acceleration), _p(ad, _I32), _p(av), int(fd),
                                    int(fs), C.c_double(fm), C.c_double(dt), which)
    return v, a


def rotation_matrix(dim, angles):
    R = np.zeros((dim, dim))
    _lib().orc_kat_rotation_matrix(dim, _p(_f64(angles)), _p(R))
    return R


def friction(dim, acc, vel, dt, direction, sign_n, mu):
    a = _f64(acc).copy()
    _lib().orc_kat_friction(dim, _p(a), _p(_f64(vel)), C.c_double(dt), int(direction),
                            int(sign_n), C.c_double(mu))
    return a
