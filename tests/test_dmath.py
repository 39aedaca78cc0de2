"""The deterministic FP64 libm (mpm_amd/csrc/mpmx_math.h) against the platform
libm, and the effect of that choice on the oracle's Mohr-Coulomb step.

The reference calls the platform libm; the oracle's default build and the GPU
kernels call mpmx_math.h so they agree bit for bit. These tests bound the
difference that substitution makes: each function within 2 ulp of the
platform result on the argument ranges the return mapping uses, and a full
Mohr-Coulomb step within 1e-12 (normwise) of the platform-libm oracle.
"""
import numpy as np
import pytest

import oracle
from mpm_amd import workloads


def _ulps(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    sp = np.spacing(np.maximum(np.abs(a), np.abs(b)))
    return np.abs(a - b) / sp


CASES = [
    ("sin", np.sin, np.linspace(-10.0, 10.0, 20001)),
    ("cos", np.cos, np.linspace(-10.0, 10.0, 20001)),
    ("tan", np.tan, np.linspace(0.0, 1.5, 5001)),
    ("acos", np.arccos, np.linspace(-1.0, 1.0, 40001)),
    ("pow_1p5", lambda x: np.power(x, 1.5), np.geomspace(1e-12, 1e12, 5001)),
    ("pow_m1p5", lambda x: np.power(x, -1.5), np.geomspace(1e-12, 1e12, 5001)),
    ("pow_m2p5", lambda x: np.power(x, -2.5), np.geomspace(1e-12, 1e12, 5001)),
]


@pytest.mark.parametrize("name,ref,xs", CASES, ids=[c[0] for c in CASES])
def test_portable_math_within_2ulp(name, ref, xs):
    got = oracle.math_fn(name, xs)
    want = ref(xs)
    # sin/cos near their zeros: compare in absolute ulps of 1 (argument-reduction floor)
    u = _ulps(got, want)
    if name in ("sin", "cos"):
        u = np.minimum(u, np.abs(got - want) / np.spacing(1.0))
    assert np.max(u) <= 2.0, (name, float(np.max(u)), xs[np.argmax(u)])


def test_portable_math_special_points():
    assert oracle.math_fn("acos", [1.0])[0] == 0.0
    assert abs(oracle.math_fn("acos", [-1.0])[0] - np.pi) <= np.spacing(np.pi)
    assert oracle.math_fn("sin", [0.0])[0] == 0.0
    assert oracle.math_fn("cos", [0.0])[0] == 1.0


def test_libm_choice_changes_a_mohr_coulomb_step_below_1e12():
    w = workloads.collapse_3d(nx=8, ny=4, nz=4, block=(4, 4, 4))
    a = w.setup(oracle.Oracle(3, w.scheme, libm="portable"))
    a.step(30, w.dt)
    pa = a.particles()
    worst = 0.0
    for s in range(5):
        b = oracle.Oracle(3, w.scheme, libm="platform")
        b.set_mesh(w.coords, w.cells)
        b.set_materials(w.materials)
        b.set_velocity_constraints(w.vel_nodes, w.vel_dirs, w.vel_values)
        b.set_friction_constraints(w.fric_nodes, w.fric_dirs, w.fric_sign, w.fric_mu)
        b.set_gravity(w.gravity)
        b.set_particles(x=pa["x"], mass=pa["mass"], volume=pa["volume"], density=pa["density"],
                        material_id=pa["material_id"], v=pa["v"], u=pa["u"],
                        stress=pa["stress"], strain=pa["strain"], F=pa["F"], cell=pa["cell"],
                        statevars=pa["statevars"])
        a.step(1, w.dt)
        b.step(1, w.dt)
        pa, pb = a.particles(), b.particles()
        for f in ("x", "v", "stress", "strain"):
            d = np.max(np.abs(pa[f] - pb[f])) / np.max(np.abs(pb[f]))
            worst = max(worst, d)
    assert worst <= 1e-12, worst
