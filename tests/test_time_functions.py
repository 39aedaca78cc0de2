"""Math functions of time-dependent boundary conditions on the oracle:
LinearFunction (src/functions/linear_function.cc:19-41) and SinFunction
(src/functions/sin_function.cc:17-20) against their closed forms, and the
acceleration constraints they scale (acceleration_constraint.h:37-43,
Mesh::update_nodal_acceleration_constraints at time step * dt,
mpm_scheme.tcc:219), concentrated nodal forces (node.tcc:103-111,
mpm_scheme.tcc:156-160) and particle tractions (particle.tcc:962-990) on
small oracle runs. CPU only."""
import numpy as np
import pytest

import oracle
from mpm_amd import workloads
from mpm_amd.capi import pack_math_functions


def test_linear_function_table():
    x, fx = [0.0, 1.0, 3.0], [0.0, 2.0, -1.0]
    f = ("linear", x, fx)
    # clamped outside the table, linear inside; the knots themselves exact
    assert oracle.math_function_value(f, -1.0) == 0.0
    assert oracle.math_function_value(f, 4.0) == -1.0
    for xi, want in [(0.0, 0.0), (0.5, 1.0), (1.0, 2.0), (2.0, 0.5), (3.0, -1.0)]:
        assert oracle.math_function_value(f, xi) == want, xi
    grid = np.linspace(-0.5, 3.5, 401)
    got = np.array([oracle.math_function_value(f, v) for v in grid])
    np.testing.assert_allclose(got, np.interp(grid, x, fx), rtol=0, atol=1e-15)


def test_linear_function_single_point_and_gap():
    assert oracle.math_function_value(("linear", [2.0], [7.0]), 1.0) == 7.0
    assert oracle.math_function_value(("linear", [2.0], [7.0]), 3.0) == 7.0


def test_sin_function_range():
    a, x0 = 2 * np.pi, 0.1
    f = ("sin", x0, a, (0.0, 1.0))
    for v in (-0.01, 1.01):
        assert oracle.math_function_value(f, v) == 0.0
    grid = np.linspace(0.0, 1.0, 101)
    got = np.array([oracle.math_function_value(f, v) for v in grid])
    np.testing.assert_allclose(got, np.sin(a * (grid - x0)), rtol=0, atol=2e-16)


def test_pack_rejects_bad_tables():
    with pytest.raises(ValueError):
        pack_math_functions([("linear", [0.0, 1.0], [1.0])])
    with pytest.raises(ValueError):
        pack_math_functions([("cubic", [0.0], [1.0])])


def test_oracle_constraint_follows_function_of_step_time():
    """After each step the constrained nodes' accelerations are
    value * f(step * dt) with the step's own index (not step + 1)."""
    w = workloads.column_2d(nx=4, ny=6)
    xmax = w.coords[:, 0].max()
    row = np.nonzero((np.abs(w.coords[:, 1] - 0.15) < 1e-9) & (w.coords[:, 0] > 1e-9)
                     & (w.coords[:, 0] < xmax - 1e-9))[0]
    assert len(row) >= 2
    dt = w.dt
    fns = [("linear", [0.0, 4 * dt], [0.0, 1.0]), ("sin", 0.0, np.pi / (8 * dt), (dt, 1.0))]
    fn = np.arange(len(row)) % 2
    o = w.setup(oracle.Oracle(w.dim, w.scheme))
    o.set_math_functions(fns)
    o.set_acceleration_constraints(row, 0, 0.2, fn)
    for step in range(6):
        o.step(1, dt)
        t = step * dt
        want = np.where(fn == 0, 0.2 * min(step / 4.0, 1.0),
                        0.2 * (np.sin(np.pi / (8 * dt) * t) if t >= dt else 0.0))
        got = o.nodes()["acceleration"][row, 0]
        np.testing.assert_allclose(got, want, rtol=1e-14, atol=1e-15)


def test_oracle_concentrated_force_adds_to_external_force():
    """One step with and without a Linear-scaled concentrated force on the
    nodes of a row: the external force differs by f(step * dt) * F on those
    nodes and not elsewhere; the function's value is taken at the step's
    own time."""
    w = workloads.column_2d(nx=4, ny=6)
    row = np.nonzero(np.abs(w.coords[:, 1] - 0.15) < 1e-9)[0]
    dt = w.dt
    fns = [("linear", [0.0, 4 * dt], [0.0, 1.0])]
    for nsteps in (1, 3):
        out = []
        for force in (False, True):
            o = w.setup(oracle.Oracle(w.dim, w.scheme))
            o.set_math_functions(fns)
            if force:
                o.set_nodal_forces(row, 1, -7.0, 0)
            o.step(nsteps, dt)
            out.append(o.nodes()["fext"])
        d = out[1] - out[0]
        t = (nsteps - 1) * dt
        if nsteps == 1:  # identical states before the forced step
            np.testing.assert_allclose(d[row, 1], -7.0 * (t / (4 * dt)), rtol=0, atol=1e-12)
            mask = np.ones(len(d), dtype=bool)
            mask[row] = False
            assert np.all(d[mask] == 0.0) and np.all(d[row, 0] == 0.0)
        else:
            assert np.all(np.abs(d[row, 1]) > 0)


def test_oracle_traction_total_force():
    """The nodal shape functions of a particle sum to one, so one step with
    and without tractions differs in total external force by the sum of the
    particles' traction vectors: value * f(t) * V / V0^(1/dim) (V the
    particle's volume, V0 the volume it was created with)."""
    w = workloads.column_2d(nx=4, ny=6)
    dt = w.dt
    fns = [("linear", [0.0, 4 * dt], [0.0, 2.0])]
    out = []
    for load in (False, True):
        o = w.setup(oracle.Oracle(w.dim, w.scheme))
        o.set_math_functions(fns)
        v0 = o.particles(["volume"])["volume"].copy()
        if load:
            o.set_particle_tractions([2, 5, 2], [1, 0, 1], [-10.0, 4.0, -20.0], [0, -1, 0])
        o.step(2, dt)
        out.append((o.nodes()["fext"], o.particles(["volume"])["volume"]))
    d = out[1][0].sum(axis=0) - out[0][0].sum(axis=0)
    vol = out[1][1]  # volumes of the second step's force computation (USL: step start)
    t = 1 * dt       # time of the second step
    f = 2.0 * t / (4 * dt)
    size = np.sqrt(v0)
    # particle 2: the later load (-20) wins in y; particle 5: constant 4 in x
    want = np.array([4.0 * vol[5] / size[5], -20.0 * f * vol[2] / size[2]])
    np.testing.assert_allclose(d, want, rtol=1e-12)
