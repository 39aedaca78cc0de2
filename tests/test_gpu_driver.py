"""The C++ analysis driver end to end on the GPU: the reference-format deck
run by mpm_amd/mpmx_run (ordered mode) gives the same particle states, bit
for bit, as the same problem set up through the Python binding."""
import os

import numpy as np
import pytest

from decks import DRIVER, read_output, run_driver, write_deck
from mpm_amd import workloads
from test_driver import LE, MC3

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(DRIVER), reason="driver not built")]


def _python_run(w, steps):
    from mpm_amd.capi import Solver
    s = w.setup(Solver(w.dim, w.scheme, ordered=True))
    s.step(steps, w.dt)
    return s.particles()


def _compare(out, p, dim):
    assert np.array_equal(out["cell"], p["cell"])
    for f in ("x", "v", "stress", "strain", "volume", "mass"):
        assert np.array_equal(out[f], p[f]), f


def test_driver_2d_column_matches_binding(tmp_path):
    w = workloads.column_2d(nx=6, ny=10)
    d = str(tmp_path)
    write_deck(w, d, nsteps=30, output_steps=10, material_json=LE)
    run_driver(d, "--ordered")
    for step in (0, 10, 20, 30):
        assert os.path.exists(os.path.join(d, "results", "deck", "particles%d.txt" % step))
    out = read_output(os.path.join(d, "results", "deck", "particles30.txt"), 2)
    _compare(out, _python_run(w, 30), 2)


def test_driver_3d_mohr_coulomb_with_stresses_matches_binding(tmp_path):
    from mpm_amd.capi import Solver
    w = workloads.collapse_3d(nx=8, ny=4, nz=4, block=(4, 4, 4))
    w.velocity_seed = None  # initial velocities reach the reference only through HDF5
    s0 = w.setup(Solver(3, w.scheme, ordered=True))
    stresses = s0.particles(["stress"])["stress"]
    d = str(tmp_path)
    write_deck(w, d, nsteps=20, output_steps=20, material_json=MC3, stresses=stresses)
    run_driver(d, "--ordered")
    out = read_output(os.path.join(d, "results", "deck", "particles20.txt"), 3)
    s0.step(20, w.dt)
    p = s0.particles()
    _compare(out, p, 3)
    assert np.array_equal(out["pdstrain"], p["statevars"][:, 8])


def test_driver_particle_velocity_constraints_match_binding(tmp_path):
    """particles_velocity_constraints with particle sets and pset_id -1."""
    w = workloads.column_2d(nx=6, ny=10)
    pcs = [(-1, 0, 0.001), ([3, 7, 11], 1, -0.01), ([7], 1, 0.002)]
    d = str(tmp_path)
    write_deck(w, d, nsteps=20, output_steps=20, material_json=LE, particle_constraints=pcs)
    run_driver(d, "--ordered")
    out = read_output(os.path.join(d, "results", "deck", "particles20.txt"), 2)
    from mpm_amd.capi import Solver
    s = w.setup(Solver(2, w.scheme, ordered=True))
    s.set_particle_velocity_constraints([-1, 3, 7, 11, 7], [0, 1, 1, 1, 1],
                                        [0.001, -0.01, -0.01, -0.01, 0.002])
    s.step(20, w.dt)
    p = s.particles()
    _compare(out, p, 2)
    assert p["v"][7, 1] == 0.002 and p["v"][3, 1] == -0.01


def test_driver_time_function_acceleration_constraints_match_binding(tmp_path):
    """acceleration_constraints with Linear (inline and from a file) and Sin
    math functions."""
    w = workloads.column_2d(nx=6, ny=10)
    xmax = w.coords[:, 0].max()
    row = np.nonzero((np.abs(w.coords[:, 1] - 0.25) < 1e-9) & (w.coords[:, 0] > 1e-9)
                     & (w.coords[:, 0] < xmax - 1e-9))[0]
    dt = w.dt
    fns = [{"id": 2, "type": "Linear", "xvalues": [0.0, 10 * dt], "fxvalues": [0.0, 1.0]},
           {"id": 4, "type": "Linear", "file": "table.csv", "xvalues": [0.0, 5 * dt, 15 * dt],
            "fxvalues": [1.0, -0.5, 0.25]},
           {"id": 8, "type": "Sin", "x0": 2 * dt, "a": 2 * np.pi / (12 * dt),
            "xrange": [3 * dt, 1.0]}]
    groups = [row[0::4], row[1::4], row[2::4], row[3::4]]
    acs = [(groups[0], 0, 0.05, 2), (groups[1], 0, -0.03, 4), (groups[2], 0, 0.02, 8),
           (groups[3], 0, 0.01, None)]
    d = str(tmp_path)
    write_deck(w, d, nsteps=20, output_steps=20, material_json=LE, math_functions=fns,
               acceleration_constraints=acs)
    run_driver(d, "--ordered")
    out = read_output(os.path.join(d, "results", "deck", "particles20.txt"), 2)
    from mpm_amd.capi import Solver
    s = w.setup(Solver(2, w.scheme, ordered=True))
    s.set_math_functions([("linear", [0.0, 10 * dt], [0.0, 1.0]),
                          ("linear", [0.0, 5 * dt, 15 * dt], [1.0, -0.5, 0.25]),
                          ("sin", 2 * dt, 2 * np.pi / (12 * dt), (3 * dt, 1.0))])
    nodes = np.concatenate(groups)
    vals = np.concatenate([np.full(len(g), v) for g, (_, _, v, _) in zip(groups, acs)])
    fn = np.concatenate([np.full(len(g), i if i < 3 else -1) for i, g in enumerate(groups)])
    s.set_acceleration_constraints(nodes, 0, vals, fn)
    s.step(20, w.dt)
    _compare(out, s.particles(), 2)


def test_driver_concentrated_nodal_forces_match_binding(tmp_path):
    """concentrated_nodal_forces from node sets, with and without a math
    function."""
    w = workloads.column_2d(nx=6, ny=10)
    top = np.nonzero(np.abs(w.coords[:, 1] - 0.5) < 1e-9)[0]
    dt = w.dt
    fns = [{"id": 3, "type": "Linear", "xvalues": [0.0, 10 * dt], "fxvalues": [0.0, 1.0]}]
    d = str(tmp_path)
    write_deck(w, d, nsteps=20, output_steps=20, material_json=LE, math_functions=fns,
               nodal_forces=[(top[:4], 1, -30.0, 3), (top[4:6], 0, 10.0, None)])
    run_driver(d, "--ordered")
    out = read_output(os.path.join(d, "results", "deck", "particles20.txt"), 2)
    from mpm_amd.capi import Solver
    s = w.setup(Solver(2, w.scheme, ordered=True))
    s.set_math_functions([("linear", [0.0, 10 * dt], [0.0, 1.0])])
    s.set_nodal_forces(top[:4], 1, -30.0, 0)
    s.set_nodal_forces(top[4:6], 0, 10.0, -1)
    s.step(20, w.dt)
    _compare(out, s.particles(), 2)


def test_driver_particle_tractions_match_binding(tmp_path):
    """particle_surface_traction from particle sets and -1, with and without
    a math function."""
    w = workloads.column_2d(nx=6, ny=10)
    dt = w.dt
    fns = [{"id": 6, "type": "Linear", "xvalues": [0.0, 10 * dt], "fxvalues": [0.0, 1.0]}]
    d = str(tmp_path)
    write_deck(w, d, nsteps=20, output_steps=20, material_json=LE, math_functions=fns,
               tractions=[([10, 11, 12, 13], 1, -200.0, 6), (-1, 0, 5.0, None)])
    run_driver(d, "--ordered")
    out = read_output(os.path.join(d, "results", "deck", "particles20.txt"), 2)
    from mpm_amd.capi import Solver
    s = w.setup(Solver(2, w.scheme, ordered=True))
    s.set_math_functions([("linear", [0.0, 10 * dt], [0.0, 1.0])])
    s.set_particle_tractions([10, 11, 12, 13, -1], [1, 1, 1, 1, 0],
                             [-200.0, -200.0, -200.0, -200.0, 5.0], [0, 0, 0, 0, -1])
    s.step(20, w.dt)
    _compare(out, s.particles(), 2)
