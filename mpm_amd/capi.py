"""ctypes binding of the mpmx C-ABI (include/mpmx.h).

``Solver`` is the thin host-side object over one ``mpmx_ctx``: the same
operations, argument meaning and particle order as the C-ABI, with numpy
arrays. It loads ``mpm_amd/libmpmx.so`` (or the build named by the
``MPMX_LIB`` environment variable) and raises if the library or a HIP device
is missing; there is no CPU fallback.
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("MPMX_LIB") or os.path.join(HERE, "libmpmx.so")

OK, E_ARG, E_STATE, E_HIP, E_UNLOCATABLE, E_UNSUPPORTED = 0, -1, -2, -3, -4, -5
MATERIAL_PARAMS = 17  # MPMX_MATERIAL_PARAMS
SCHEMES = {"usf": 0, "usl": 1, "musl": 2}
VELOCITY_UPDATES = {"flip": 0, "pic": 1, "asflip": 2, "tpic": 3, "apic": 4}
ORDERED = 0x1
NO_GRAPH = 0x2

# Every entry point declared in include/mpmx.h.
EXPORTS = (
    "mpmx_create", "mpmx_destroy", "mpmx_last_error", "mpmx_set_mesh", "mpmx_set_instances",
    "mpmx_set_materials",
    "mpmx_set_velocity_constraints", "mpmx_set_acceleration_constraints",
    "mpmx_set_acceleration_constraints_fn", "mpmx_set_math_functions", "mpmx_set_nodal_forces",
    "mpmx_set_particle_tractions",
    "mpmx_set_friction_constraints", "mpmx_set_particle_velocity_constraints",
    "mpmx_set_nodal_euler_angles",
    "mpmx_set_gravity", "mpmx_generate_gauss",
    "mpmx_set_particles", "mpmx_compute_volume_mass", "mpmx_assign_velocity",
    "mpmx_assign_stresses", "mpmx_assign_materials", "mpmx_nparticles", "mpmx_get_particles", "mpmx_get_nodes",
    "mpmx_get_mapping_matrix", "mpmx_set_mapping_matrix",
    "mpmx_step", "mpmx_step_noalloc", "mpmx_set_locate_particles", "mpmx_get_particle_ids", "mpmx_set_particle_ids",
    "mpmx_set_sort_interval", "mpmx_set_chunk", "mpmx_autotune", "mpmx_set_step",
    "mpmx_get_step",
    "mpmx_synchronize", "mpmx_profile", "mpmx_kernel_times",
    "mpmx_adjoint_forward", "mpmx_adjoint_backward", "mpmx_adjoint_set_apex_ratio",
    "mpmx_adjoint_release", "mpmx_adjoint_backward_device", "mpmx_get_particles_device",
)

_D = C.POINTER(C.c_double)
_I64 = C.POINTER(C.c_int64)
_I32 = C.POINTER(C.c_int)
_lib = None


class MpmxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("mpmx error %d: %s" % (code, msg))
        self.code = code


class UnlocatableParticle(MpmxError):
    pass


def load(path=LIB_PATH):
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError("mpmx HIP library not built: %s (run __graft_entry__.build())" % path)
    # torch carries its own HIP runtime: when it is in the process, its
    # device must be initialized before the library's, or torch finds no
    # device afterwards (the library then shares torch's runtime, and torch
    # tensors' device memory can be passed to the *_device entry points)
    torch = sys.modules.get("torch")
    if torch is not None and torch.cuda.is_available():
        torch.cuda.init()
    lib = C.CDLL(path)
    V = C.c_void_p
    sig = {
        "mpmx_create": [C.POINTER(V), C.c_int, C.c_int, C.c_int, C.c_double, C.c_uint],
        "mpmx_destroy": [V],
        "mpmx_last_error": [V],
        "mpmx_set_mesh": [V, C.c_int64, _D, C.c_int64, _I64, C.c_int],
        "mpmx_set_instances": [V, C.c_int64, _I64],
        "mpmx_set_materials": [V, C.c_int, _D],
        "mpmx_set_velocity_constraints": [V, C.c_int64, _I64, _I32, _D],
        "mpmx_set_acceleration_constraints": [V, C.c_int64, _I64, _I32, _D],
        "mpmx_set_acceleration_constraints_fn": [V, C.c_int64, _I64, _I32, _D, _I32],
        "mpmx_set_math_functions": [V, C.c_int, _I32, _I32, _D],
        "mpmx_set_nodal_forces": [V, C.c_int64, _I64, _I32, _D, _I32],
        "mpmx_set_particle_tractions": [V, C.c_int64, _I64, _I32, _D, _I32],
        "mpmx_set_friction_constraints": [V, C.c_int64, _I64, _I32, _I32, _D],
        "mpmx_set_nodal_euler_angles": [V, C.c_int64, _I64, _D],
        "mpmx_set_particle_velocity_constraints": [V, C.c_int64, _I64, _I32, _D],
        "mpmx_set_gravity": [V, _D, C.c_double],
        "mpmx_generate_gauss": [V, C.c_int, C.c_int64, _I64, C.c_int],
        "mpmx_set_particles": [V, C.c_int64, _D, _D, _D, _D, _D, _D, _D, _D, _D, _I32, _I64, _D],
        "mpmx_compute_volume_mass": [V],
        "mpmx_assign_velocity": [V, C.c_int64, _D],
        "mpmx_assign_stresses": [V, C.c_int64, _D],
        "mpmx_assign_materials": [V, C.c_int64, _I64, C.c_int],
        "mpmx_nparticles": [V, _I64],
        "mpmx_get_particles": [V, _D, _D, _D, _D, _D, _D, _D, _D, _D, _I32, _I64, _D, _D],
        "mpmx_get_nodes": [V, _D, _D, _D, _I32],
        "mpmx_step": [V, C.c_int, C.c_double, C.c_double],
        "mpmx_step_noalloc": [V, C.c_int, C.c_double, C.c_double],
        "mpmx_get_mapping_matrix": [V, _D],
        "mpmx_set_mapping_matrix": [V, C.c_int64, _D],
        "mpmx_set_step": [V, C.c_int64],
        "mpmx_set_locate_particles": [V, C.c_int],
        "mpmx_get_particle_ids": [V, _I64],
        "mpmx_set_particle_ids": [V, C.c_int64, _I64],
        "mpmx_set_sort_interval": [V, C.c_int],
        "mpmx_set_chunk": [V, C.c_int],
        "mpmx_autotune": [V, C.c_double, C.c_double, C.c_int, _I32, _I32],
        "mpmx_get_step": [V, _I64],
        "mpmx_synchronize": [V],
        "mpmx_profile": [V, C.c_int],
        "mpmx_kernel_times": [V, C.c_int, C.c_char_p, _D, _I64, _I32],
        "mpmx_adjoint_forward": [V, C.c_int, C.c_double, C.c_double, C.c_int],
        "mpmx_adjoint_backward": [V, _D, _D, _D, _D, _D, _D, _D],
        "mpmx_adjoint_backward_device": [V, V, V, V, V, V, V, V],
        "mpmx_get_particles_device": [V, V, V],
        "mpmx_adjoint_set_apex_ratio": [V, C.c_double],
        "mpmx_adjoint_release": [V],
    }
    for name, args in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_char_p if name == "mpmx_last_error" else C.c_int
    _lib = lib
    return lib


def _p(a, t=_D):
    return None if a is None else a.ctypes.data_as(t)


def _f64(a, shape=None):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a if shape is None else a.reshape(shape)


def pack_math_functions(functions):
    """(type, npts, params) arrays of mpmx_set_math_functions from a list of
    ("linear", xvalues, fxvalues) or ("sin", x0, a, (xmin, xmax))."""
    types, npts, params = [], [], []
    for f in functions:
        if f[0] == "linear":
            x = np.asarray(f[1], dtype=np.float64).ravel()
            fx = np.asarray(f[2], dtype=np.float64).ravel()
            if x.shape != fx.shape or len(x) == 0:
                raise ValueError("a linear function needs xvalues and fxvalues of one length")
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


class Solver:
    """One explicit-MPM solver instance on the GPU (one mpmx context)."""

    def __init__(self, dim, scheme="usf", velocity_update="flip", blending_ratio=1.0,
                 ordered=False, graph=True):
        self.lib = load()
        self.dim = dim
        self.nn = 4 if dim == 2 else 8
        flags = (ORDERED if ordered else 0) | (0 if graph else NO_GRAPH)
        ctx = C.c_void_p()
        rc = self.lib.mpmx_create(C.byref(ctx), dim, SCHEMES[scheme],
                                  VELOCITY_UPDATES[velocity_update], blending_ratio, flags)
        self.ctx = ctx
        if rc != OK:
            msg = self.lib.mpmx_last_error(ctx).decode() if ctx.value else "create failed"
            if ctx.value:
                self.lib.mpmx_destroy(ctx)
            self.ctx = None
            raise MpmxError(rc, msg)
        self.nnodes = 0
        self.ncells = 0

    def close(self):
        if getattr(self, "ctx", None):
            self.lib.mpmx_destroy(self.ctx)
            self.ctx = None

    __del__ = close

    def _check(self, rc):
        if rc != OK:
            msg = self.lib.mpmx_last_error(self.ctx).decode()
            if rc == E_UNLOCATABLE:
                raise UnlocatableParticle(rc, msg)
            raise MpmxError(rc, msg)

    # -- setup ---------------------------------------------------------------
    def set_mesh(self, coords, cells, isoparametric=False):
        coords = _f64(coords)
        cells = np.ascontiguousarray(cells, dtype=np.int64)
        self.nnodes, self.ncells = coords.shape[0], cells.shape[0]
        self._check(self.lib.mpmx_set_mesh(self.ctx, self.nnodes, _p(coords), self.ncells,
                                           _p(cells, _I64), int(bool(isoparametric))))

    def set_instances(self, cell_begin):
        """Batched run: instance i owns cells [cell_begin[i], cell_begin[i+1])."""
        cb = np.ascontiguousarray(cell_begin, dtype=np.int64)
        self._check(self.lib.mpmx_set_instances(self.ctx, len(cb) - 1, _p(cb, _I64)))

    def set_materials(self, params):
        params = _f64(params).reshape(-1, MATERIAL_PARAMS)
        self._check(self.lib.mpmx_set_materials(self.ctx, params.shape[0], _p(params)))
        self.material_params = params.copy()

    def material_types(self):
        """Per material id: 0 LinearElastic, 1 MohrCoulomb (parameter 0)."""
        return self.material_params[:, 0].astype(int)

    def _cons(self, fn, nodes, dirs, *vals):
        nodes = np.ascontiguousarray(nodes, dtype=np.int64)
        dirs = np.ascontiguousarray(np.broadcast_to(dirs, nodes.shape), dtype=np.int32)
        args = []
        for v, t in vals:
            dt = np.int32 if t is _I32 else np.float64
            a = np.ascontiguousarray(np.broadcast_to(v, nodes.shape), dtype=dt)
            args.append((a, t))
        self._check(fn(self.ctx, len(nodes), _p(nodes, _I64), _p(dirs, _I32),
                       *[_p(a, t) for a, t in args]))

    def set_velocity_constraints(self, nodes, dirs, values):
        self._cons(self.lib.mpmx_set_velocity_constraints, nodes, dirs, (values, _D))

    def set_acceleration_constraints(self, nodes, dirs, values, functions=None):
        """functions: per constraint the index of the math function scaling
        its value (-1: none)."""
        if functions is None:
            self._cons(self.lib.mpmx_set_acceleration_constraints, nodes, dirs, (values, _D))
        else:
            self._cons(self.lib.mpmx_set_acceleration_constraints_fn, nodes, dirs, (values, _D),
                       (functions, _I32))

    def set_nodal_forces(self, nodes, dirs, forces, functions=None):
        """Concentrated nodal forces (node -1: every node); functions: per
        entry the index of its math function (-1: none)."""
        fn = -1 if functions is None else functions
        self._cons(self.lib.mpmx_set_nodal_forces, nodes, dirs, (forces, _D), (fn, _I32))

    def set_particle_tractions(self, particles, dirs, values, functions=None):
        """Particle surface tractions in order (particle -1: every particle);
        functions: per load the index of its math function (-1: none)."""
        fn = -1 if functions is None else functions
        self._cons(self.lib.mpmx_set_particle_tractions, particles, dirs, (values, _D), (fn, _I32))

    def set_math_functions(self, functions):
        """Math functions of time-dependent constraints, indexed in list order:
        ("linear", xvalues, fxvalues) or ("sin", x0, a, (xmin, xmax))."""
        types, npts, params = pack_math_functions(functions)
        self._check(self.lib.mpmx_set_math_functions(self.ctx, len(types), _p(types, _I32),
                                                     _p(npts, _I32), _p(params)))

    def set_friction_constraints(self, nodes, dirs, sign_n, mu):
        self._cons(self.lib.mpmx_set_friction_constraints, nodes, dirs, (sign_n, _I32), (mu, _D))

    def set_particle_velocity_constraints(self, particles, dirs, values):
        """Velocity component dirs[i] of particle particles[i] (-1: all) is set
        to values[i] after every velocity update, in order."""
        pa = np.ascontiguousarray(particles, dtype=np.int64)
        di = np.ascontiguousarray(dirs, dtype=np.int32)
        va = np.ascontiguousarray(values, dtype=np.float64)
        self._check(self.lib.mpmx_set_particle_velocity_constraints(
            self.ctx, len(pa), _p(pa, _I64), _p(di, _I32), _p(va)))

    def set_nodal_euler_angles(self, nodes, angles):
        """General (inclined) boundaries: the constraints of these nodes act in
        the frame rotated by the Euler angles (radians, rows of dim)."""
        nodes = np.ascontiguousarray(nodes, dtype=np.int64)
        ang = _f64(angles, (len(nodes), self.dim))
        self._check(self.lib.mpmx_set_nodal_euler_angles(self.ctx, len(nodes), _p(nodes, _I64),
                                                         _p(ang)))

    def set_gravity(self, g, ramping_time=0.0):
        self._check(self.lib.mpmx_set_gravity(self.ctx, _p(_f64(g)), ramping_time))

    def generate_gauss(self, nparticles_per_dir, cell_set=None, material_id=0):
        if cell_set is None:
            self._check(self.lib.mpmx_generate_gauss(self.ctx, nparticles_per_dir, -1, None,
                                                     material_id))
        else:
            cs = np.ascontiguousarray(cell_set, dtype=np.int64)
            self._check(self.lib.mpmx_generate_gauss(self.ctx, nparticles_per_dir, len(cs),
                                                     _p(cs, _I64), material_id))

    def set_particles(self, x, mass, volume, density, material_id, v=None, u=None, stress=None,
                      strain=None, F=None, cell=None, statevars=None):
        d = self.dim
        x = _f64(x, (-1, d))
        n = x.shape[0]
        a = dict(u=_f64(u, (n, d)), v=_f64(v, (n, d)), stress=_f64(stress, (n, 6)),
                 strain=_f64(strain, (n, 6)), F=_f64(F, (n, 9)), mass=_f64(mass, (n,)),
                 volume=_f64(volume, (n,)), density=_f64(density, (n,)),
                 sv=_f64(statevars, (n, 9)))
        mat = np.ascontiguousarray(np.broadcast_to(material_id, (n,)), dtype=np.int32)
        cell = None if cell is None else np.ascontiguousarray(cell, dtype=np.int64)
        self._check(self.lib.mpmx_set_particles(
            self.ctx, n, _p(x), _p(a["u"]), _p(a["v"]), _p(a["stress"]), _p(a["strain"]),
            _p(a["F"]), _p(a["mass"]), _p(a["volume"]), _p(a["density"]), _p(mat, _I32),
            _p(cell, _I64), _p(a["sv"])))

    def compute_volume_mass(self):
        """volume = cell volume / particles in the cell, mass = volume * density."""
        self._check(self.lib.mpmx_compute_volume_mass(self.ctx))

    def assign_velocity(self, v):
        v = _f64(v, (-1, self.dim))
        self._check(self.lib.mpmx_assign_velocity(self.ctx, v.shape[0], _p(v)))

    def assign_materials(self, particles, material_id):
        """material_sets: the listed particles take material material_id and
        its initial state variables."""
        idx = np.ascontiguousarray(particles, dtype=np.int64)
        self._check(self.lib.mpmx_assign_materials(self.ctx, len(idx), _p(idx, _I64),
                                                   int(material_id)))

    def assign_stresses(self, stress):
        stress = _f64(stress, (-1, 6))
        self._check(self.lib.mpmx_assign_stresses(self.ctx, stress.shape[0], _p(stress)))

    # -- state ---------------------------------------------------------------
    @property
    def nparticles(self):
        n = C.c_int64(0)
        self._check(self.lib.mpmx_nparticles(self.ctx, C.byref(n)))
        return n.value

    def particles(self, fields=None):
        n, d = self.nparticles, self.dim
        shapes = dict(x=(n, d), u=(n, d), v=(n, d), stress=(n, 6), strain=(n, 6), F=(n, 9),
                      mass=(n,), volume=(n,), density=(n,), xi=(n, d), statevars=(n, 9))
        want = set(fields) if fields else set(shapes) | {"material_id", "cell"}
        out = {k: np.zeros(s) for k, s in shapes.items() if k in want}
        if "material_id" in want:
            out["material_id"] = np.zeros(n, dtype=np.int32)
        if "cell" in want:
            out["cell"] = np.zeros(n, dtype=np.int64)
        g = out.get
        self._check(self.lib.mpmx_get_particles(
            self.ctx, _p(g("x")), _p(g("u")), _p(g("v")), _p(g("stress")), _p(g("strain")),
            _p(g("F")), _p(g("mass")), _p(g("volume")), _p(g("density")),
            _p(g("material_id"), _I32), _p(g("cell"), _I64), _p(g("xi")), _p(g("statevars"))))
        return out

    def nodes(self):
        n, d = self.nnodes, self.dim
        out = dict(mass=np.zeros(n), velocity=np.zeros((n, d)), acceleration=np.zeros((n, d)),
                   active=np.zeros(n, dtype=np.int32))
        self._check(self.lib.mpmx_get_nodes(self.ctx, _p(out["mass"]), _p(out["velocity"]),
                                            _p(out["acceleration"]), _p(out["active"], _I32)))
        return out

    def mapping_matrix(self):
        """The affine velocity updates' mapping matrices, (np, dim, dim)."""
        B = np.zeros((self.nparticles, self.dim, self.dim))
        self._check(self.lib.mpmx_get_mapping_matrix(self.ctx, _p(B)))
        return B

    def set_mapping_matrix(self, B):
        B = _f64(B)
        self._check(self.lib.mpmx_set_mapping_matrix(self.ctx, B.shape[0], _p(B)))

    def step(self, nsteps, dt, damping_factor=0.0):
        self._check(self.lib.mpmx_step(self.ctx, int(nsteps), float(dt), float(damping_factor)))

    def step_noalloc(self, nsteps, dt, damping_factor=0.0):
        """step() with no device allocation: scratch from the library's
        fixed per-GPU array (mpmx_step_noalloc)."""
        self._check(self.lib.mpmx_step_noalloc(self.ctx, int(nsteps), float(dt),
                                               float(damping_factor)))

    @property
    def step_count(self):
        s = C.c_int64(0)
        self._check(self.lib.mpmx_get_step(self.ctx, C.byref(s)))
        return s.value

    def set_locate_particles(self, locate_particles=True):
        """True: a particle that leaves the mesh fails the step (the default);
        False: such particles are removed (the analysis key locate_particles)."""
        self._check(self.lib.mpmx_set_locate_particles(self.ctx, 1 if locate_particles else 0))

    def particle_ids(self):
        ids = np.zeros(self.nparticles, dtype=np.int64)
        self._check(self.lib.mpmx_get_particle_ids(self.ctx, _p(ids, _I64)))
        return ids

    def set_particle_ids(self, ids):
        ids = np.ascontiguousarray(ids, dtype=np.int64)
        self._check(self.lib.mpmx_set_particle_ids(self.ctx, len(ids), _p(ids, _I64)))

    def set_sort_interval(self, steps):
        """Fast mode: full particle sort by cell every `steps` steps."""
        self._check(self.lib.mpmx_set_sort_interval(self.ctx, int(steps)))

    def set_chunk(self, particles_per_workgroup):
        """Fast mode: particles per workgroup of the scatter kernels (128, 256, 512)."""
        self._check(self.lib.mpmx_set_chunk(self.ctx, int(particles_per_workgroup)))

    def autotune(self, dt, damping_factor=0.0, trial_steps=10):
        """Pick the fastest (scatter chunk, sort interval) by timing trial steps
        from the current state, which is restored. Returns the pair."""
        ch, se = C.c_int(0), C.c_int(0)
        self._check(self.lib.mpmx_autotune(self.ctx, float(dt), float(damping_factor),
                                           int(trial_steps), C.byref(ch), C.byref(se)))
        return dict(chunk=ch.value, sort_every=se.value)

    def set_step(self, step):
        self._check(self.lib.mpmx_set_step(self.ctx, int(step)))

    def synchronize(self):
        self._check(self.lib.mpmx_synchronize(self.ctx))

    def profile(self, enable=True):
        self._check(self.lib.mpmx_profile(self.ctx, 1 if enable else 0))

    def kernel_times(self):
        """{kernel name: (total ms, launches)} recorded while profiling."""
        cap = 64
        names = C.create_string_buffer(48 * cap)
        ms = np.zeros(cap)
        cnt = np.zeros(cap, dtype=np.int64)
        n = C.c_int(0)
        self._check(self.lib.mpmx_kernel_times(self.ctx, cap, names, _p(ms), _p(cnt, _I64),
                                               C.byref(n)))
        out = {}
        for i in range(min(n.value, cap)):
            nm = names.raw[48 * i:48 * i + 48].split(b"\0", 1)[0].decode()
            out[nm] = (float(ms[i]), int(cnt[i]))
        return out

    # ------------------------------- adjoint ---------------------------------
    def adjoint_forward(self, nsteps, dt, damping_factor=0.0, checkpoint_every=50):
        """Advance `nsteps` steps (as step()) recording checkpoints for
        adjoint_backward."""
        self._check(self.lib.mpmx_adjoint_forward(self.ctx, int(nsteps), float(dt),
                                                  float(damping_factor), int(checkpoint_every)))

    def adjoint_backward(self, dL_dx=None, dL_dv=None):
        """Gradients of a loss of the final positions/velocities of the
        recorded run: dict with E, rho, phi (degrees) per particle and x0, v0
        (rows of dim)."""
        n, d = self.nparticles, self.dim
        gx = _f64(dL_dx, (n, d)) if dL_dx is not None else None
        gv = _f64(dL_dv, (n, d)) if dL_dv is not None else None
        out = dict(E=np.zeros(n), rho=np.zeros(n), phi=np.zeros(n), x0=np.zeros((n, d)),
                   v0=np.zeros((n, d)))
        self._check(self.lib.mpmx_adjoint_backward(
            self.ctx, _p(gx), _p(gv), _p(out["E"]), _p(out["rho"]), _p(out["phi"]),
            _p(out["x0"]), _p(out["v0"])))
        return out

    def adjoint_backward_device(self, dL_dx=None, dL_dv=None, outputs=None):
        """adjoint_backward with device arrays: dL_dx, dL_dv and the entries
        of `outputs` (keys E, rho, phi, x0, v0) are device addresses (ints,
        e.g. torch CUDA tensors' data_ptr(), float64, rows of dim), any of
        them None. The library's stream is synchronized on return; callers
        on another stream synchronize theirs before the call."""
        o = outputs or {}
        self._check(self.lib.mpmx_adjoint_backward_device(
            self.ctx, dL_dx, dL_dv, o.get("E"), o.get("rho"), o.get("phi"), o.get("x0"),
            o.get("v0")))

    def particles_device(self, x=None, v=None):
        """Positions and velocities (rows of dim) into device arrays given by
        address (ints), either None."""
        self._check(self.lib.mpmx_get_particles_device(self.ctx, x, v))

    def adjoint_set_apex_ratio(self, ratio):
        """Mohr-Coulomb returns near the apex take constant flow directions in
        the backward pass when the plastic correction exceeds ratio * sqrt(2
        J2) of the returned stress (default 0.3); ratio <= 0: exact."""
        self._check(self.lib.mpmx_adjoint_set_apex_ratio(self.ctx, float(ratio)))

    def adjoint_release(self):
        self._check(self.lib.mpmx_adjoint_release(self.ctx))
