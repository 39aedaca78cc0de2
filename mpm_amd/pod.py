"""ctypes binding of the particle checkpoint table (include/mpmx_pod.h).

``write(filename, cols)`` / ``read(filename)`` move a dict of numpy columns
(keys as in ``mpmx_pod_columns``: id, mass, volume, pressure, coord,
displacement, nsize, velocity, acceleration, stress, strain, defgrad,
initialise_mapping_matrix, mapping_matrix, status, cell_id, material_id,
nstate_vars, svars) to and from the reference's PODParticle HDF5 table
(pod_particle.h:12-62). ``from_solver`` builds the columns of a solver's
particles as Particle::pod() does (particle.tcc:180-305).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmpmx_pod.so")
NFIELDS = 74
NSVARS = 20

_P = C.c_void_p
# (name, dtype, components per record); the order of mpmx_pod_columns
COLUMNS = (("id", np.int64, 1), ("mass", np.float64, 1), ("volume", np.float64, 1),
           ("pressure", np.float64, 1), ("coord", np.float64, 3),
           ("displacement", np.float64, 3), ("nsize", np.float64, 3),
           ("velocity", np.float64, 3), ("acceleration", np.float64, 3),
           ("stress", np.float64, 6), ("strain", np.float64, 6), ("defgrad", np.float64, 9),
           ("initialise_mapping_matrix", np.uint8, 1), ("mapping_matrix", np.float64, 9),
           ("status", np.uint8, 1), ("cell_id", np.int64, 1), ("material_id", np.uint32, 1),
           ("nstate_vars", np.uint32, 1), ("svars", np.float64, NSVARS))


class _Columns(C.Structure):
    _fields_ = [("n", C.c_int64)] + [(name, _P) for name, _, _ in COLUMNS]


_lib = None


def load(path=LIB_PATH):
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise RuntimeError("checkpoint library not built: %s (run __graft_entry__.build())"
                               % path)
        lib = C.CDLL(path)
        lib.mpmx_pod_write.argtypes = [C.c_char_p, C.POINTER(_Columns)]
        lib.mpmx_pod_read.argtypes = [C.c_char_p, C.POINTER(_Columns)]
        lib.mpmx_pod_count.argtypes = [C.c_char_p, C.POINTER(C.c_int64)]
        lib.mpmx_pod_field_names.argtypes = [C.c_char_p, C.c_char_p, C.c_int64,
                                             C.POINTER(C.c_int64)]
        lib.mpmx_pod_last_error.restype = C.c_char_p
        _lib = lib
    return _lib


def _check(rc):
    if rc != 0:
        raise RuntimeError("checkpoint table: " + load().mpmx_pod_last_error().decode())


def _pack(cols, n):
    st = _Columns()
    st.n = n
    keep = []
    for name, dt, k in COLUMNS:
        a = cols.get(name)
        if a is None:
            setattr(st, name, None)
            continue
        a = np.ascontiguousarray(a, dtype=dt).reshape(n * k)
        keep.append(a)
        setattr(st, name, a.ctypes.data)
    return st, keep


def write(filename, cols):
    n = len(cols["id"])
    st, keep = _pack(cols, n)
    _check(load().mpmx_pod_write(os.fsencode(filename), C.byref(st)))
    del keep


def count(filename):
    n = C.c_int64(0)
    _check(load().mpmx_pod_count(os.fsencode(filename), C.byref(n)))
    return n.value


def read(filename):
    n = count(filename)
    cols = {name: np.zeros((n, k) if k > 1 else n, dtype=dt) for name, dt, k in COLUMNS}
    st, keep = _pack(cols, n)
    _check(load().mpmx_pod_read(os.fsencode(filename), C.byref(st)))
    # _pack made contiguous views of the same arrays
    del keep
    return cols


def field_names(filename):
    buf = C.create_string_buffer(64 * NFIELDS)
    nf = C.c_int64(0)
    _check(load().mpmx_pod_field_names(os.fsencode(filename), buf, len(buf), C.byref(nf)))
    return [s.decode() for s in buf.raw.split(b"\0")[:nf.value]]


def from_solver(solver, nsize=None):
    """The solver's particles as PODParticle columns (Particle::pod(),
    particle.tcc:180-305): vectors padded to 3 components, no mapping matrix
    (FLIP/PIC), status true, acceleration 0 (the explicit step does not
    update it), the MohrCoulomb state variables in the material's order and
    none for LinearElastic. nsize: the natural particle size per particle
    (0 if not given, as Particle::initialise sets it)."""
    p = solver.particles()
    n, d = len(p["x"]), p["x"].shape[1]

    def pad3(a):
        out = np.zeros((n, 3))
        out[:, :d] = a
        return out

    mc = p.get("statevars") is not None
    svars = np.zeros((n, NSVARS))
    nsv = np.zeros(n, dtype=np.uint32)
    if mc:
        kinds = np.asarray(solver.material_types())[p["material_id"]]
        is_mc = kinds == 1
        svars[is_mc, :p["statevars"].shape[1]] = p["statevars"][is_mc]
        nsv[is_mc] = p["statevars"].shape[1]
    ns = np.zeros((n, 3))
    if nsize is not None:
        ns[:, :d] = np.asarray(nsize, dtype=np.float64).reshape(n, -1)[:, :d] if np.ndim(nsize) > 1 \
            else np.repeat(np.asarray(nsize, dtype=np.float64)[:, None], d, axis=1)
    return dict(id=solver.particle_ids(), mass=p["mass"], volume=p["volume"],
                pressure=np.zeros(n), coord=pad3(p["x"]), displacement=pad3(p["u"]),
                nsize=ns, velocity=pad3(p["v"]), acceleration=np.zeros((n, 3)),
                stress=p["stress"], strain=p["strain"], defgrad=p["F"].reshape(n, 9),
                initialise_mapping_matrix=np.zeros(n, dtype=np.uint8),
                mapping_matrix=np.zeros((n, 9)), status=np.ones(n, dtype=np.uint8),
                cell_id=p["cell"].astype(np.int64), material_id=p["material_id"].astype(np.uint32),
                nstate_vars=nsv, svars=svars)
