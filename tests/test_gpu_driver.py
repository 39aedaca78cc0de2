"""The C++ analysis driver end to end on the GPU: a reference-format deck run
by mpm_amd/mpmx_run (ordered mode) against the oracle set up with the same
problem, bit for bit, read from the particle tables (the reference's
PODParticle HDF5 format) the driver writes at the output steps; checkpoint
resume from those tables; the deck the reference's own solver tests write
(tests/io/write_mesh_particles.cc)."""
import os
import shutil

import numpy as np
import pytest

import oracle
from decks import DRIVER, output_path, read_output, run_driver, write_deck, write_fixture_deck
from mpm_amd import materials as mat
from mpm_amd import workloads
from test_driver import LE, MC3

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(DRIVER), reason="driver not built")]

FIELDS = ("x", "u", "v", "stress", "strain", "F", "volume", "mass")


def _oracle_run(w, steps, extra=None):
    o = w.setup(oracle.Oracle(w.dim, w.scheme))
    if extra:
        extra(o)
    o.step(steps, w.dt, w.damping)
    return o.particles()


def _compare(out, p, dim, mc=False):
    assert np.array_equal(out["cell"], p["cell"])
    for f in FIELDS:
        a = out[f]
        b = p[f].reshape(a.shape)
        bad = np.argwhere(a != b)
        assert bad.size == 0, "%s differs at %s" % (f, bad[:3].tolist())
    if mc:
        assert np.array_equal(out["statevars"], p["statevars"])


def test_driver_2d_column_matches_oracle(tmp_path):
    w = workloads.column_2d(nx=6, ny=10)
    d = str(tmp_path)
    write_deck(w, d, nsteps=30, output_steps=10, material_json=LE)
    run_driver(d, "--ordered")
    for step in (0, 10, 20, 30):
        assert os.path.exists(output_path(d, step, 30)), step
    out = read_output(output_path(d, 30, 30), 2)
    assert list(out["id"]) == list(range(len(out["id"])))
    assert np.all(out["material_id"] == 3) and np.all(out["nstate_vars"] == 0)
    np.testing.assert_array_equal(out["nsize"], 1.0)  # 2 / sqrt(4 particles per cell)
    _compare(out, _oracle_run(w, 30), 2)


def test_driver_nodal_euler_angles_match_oracle(tmp_path):
    """A deck whose base nodes lie on a general boundary: the mesh's
    "nodal_euler_angles" file tilts their frame, the velocity constraint
    holds the local normal velocity and friction acts in that frame
    (node.tcc:385-401, 540-583); the block slides down the tilted base."""
    w = workloads.column_2d(nx=6, ny=10)
    base = np.nonzero(np.abs(w.coords[:, 1]) <= 1e-9)[0].astype(np.int64)
    keep = ~np.isin(w.vel_nodes, base)
    w.vel_nodes = np.concatenate([w.vel_nodes[keep], base])
    w.vel_dirs = np.concatenate([w.vel_dirs[keep], np.ones(len(base))]).astype(np.int32)
    w.vel_values = np.concatenate([np.asarray(w.vel_values)[keep], np.zeros(len(base))])
    w.fric_nodes = base
    w.fric_dirs = np.ones(len(base), dtype=np.int32)
    w.fric_sign = np.full(len(base), -1, dtype=np.int32)
    w.fric_mu = np.full(len(base), 0.3)
    ang = np.tile(np.radians([15.0, 10.0]), (len(base), 1))
    d = str(tmp_path)
    write_deck(w, d, nsteps=30, output_steps=30, material_json=LE, euler_angles=(base, ang))
    run_driver(d, "--ordered")
    out = read_output(output_path(d, 30, 30), 2)
    p = _oracle_run(w, 30, extra=lambda o: o.set_nodal_euler_angles(base, ang))
    _compare(out, p, 2)
    assert np.max(np.abs(out["x"][:, 0] - w.setup(oracle.Oracle(2, w.scheme)).particles()["x"][:, 0])) > 0


def test_driver_isoparametric_default_matches_oracle(tmp_path):
    """A deck without the "isoparametric" key on a mesh of skewed quads: the
    driver takes the reference's default (isoparametric, mpm_base.tcc:768-783)
    and matches the oracle's isoparametric run bit for bit."""
    w = workloads.column_2d(nx=6, ny=10)
    X = w.coords.copy()
    inner = (X[:, 0] > 1e-9) & (X[:, 0] < X[:, 0].max() - 1e-9) & (X[:, 1] > 1e-9) & \
        (X[:, 1] < X[:, 1].max() - 1e-9)
    X[inner, 0] += 0.012 * np.sin(7.0 * X[inner, 1])
    X[inner, 1] += 0.009 * np.cos(5.0 * X[inner, 0])
    w.coords = X
    w.isoparametric = True
    d = str(tmp_path)
    write_deck(w, d, nsteps=30, output_steps=30, material_json=LE)
    run_driver(d, "--ordered")
    out = read_output(output_path(d, 30, 30), 2)
    _compare(out, _oracle_run(w, 30), 2)


def test_driver_3d_mohr_coulomb_with_stresses_matches_oracle(tmp_path):
    from mpm_amd.capi import Solver
    w = workloads.collapse_3d(nx=8, ny=4, nz=4, block=(4, 4, 4))
    w.velocity_seed = None  # initial velocities reach the reference only through a checkpoint
    s0 = w.setup(Solver(3, w.scheme, ordered=True))
    stresses = s0.particles(["stress"])["stress"]
    d = str(tmp_path)
    write_deck(w, d, nsteps=20, output_steps=20, material_json=MC3, stresses=stresses)
    run_driver(d, "--ordered")
    out = read_output(output_path(d, 20, 20), 3)
    assert np.all(out["nstate_vars"] == 9)
    _compare(out, _oracle_run(w, 20), 3, mc=True)


def test_driver_particle_velocity_constraints_match_oracle(tmp_path):
    """particles_velocity_constraints with particle sets and pset_id -1."""
    w = workloads.column_2d(nx=6, ny=10)
    pcs = [(-1, 0, 0.001), ([3, 7, 11], 1, -0.01), ([7], 1, 0.002)]
    d = str(tmp_path)
    write_deck(w, d, nsteps=20, output_steps=20, material_json=LE, particle_constraints=pcs)
    run_driver(d, "--ordered")
    out = read_output(output_path(d, 20, 20), 2)

    def extra(o):
        o.set_particle_velocity_constraints([-1, 3, 7, 11, 7], [0, 1, 1, 1, 1],
                                            [0.001, -0.01, -0.01, -0.01, 0.002])
    p = _oracle_run(w, 20, extra)
    _compare(out, p, 2)
    assert p["v"][7, 1] == 0.002 and p["v"][3, 1] == -0.01


def test_driver_time_function_acceleration_constraints_match_oracle(tmp_path):
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
    out = read_output(output_path(d, 20, 20), 2)

    def extra(o):
        o.set_math_functions([("linear", [0.0, 10 * dt], [0.0, 1.0]),
                              ("linear", [0.0, 5 * dt, 15 * dt], [1.0, -0.5, 0.25]),
                              ("sin", 2 * dt, 2 * np.pi / (12 * dt), (3 * dt, 1.0))])
        nodes = np.concatenate(groups)
        vals = np.concatenate([np.full(len(g), v) for g, (_, _, v, _) in zip(groups, acs)])
        fn = np.concatenate([np.full(len(g), i if i < 3 else -1) for i, g in enumerate(groups)])
        o.set_acceleration_constraints(nodes, np.zeros(len(nodes), dtype=np.int32), vals, fn)
    _compare(out, _oracle_run(w, 20, extra), 2)


def test_driver_concentrated_nodal_forces_match_oracle(tmp_path):
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
    out = read_output(output_path(d, 20, 20), 2)

    def extra(o):
        o.set_math_functions([("linear", [0.0, 10 * dt], [0.0, 1.0])])
        o.set_nodal_forces(np.concatenate([top[:4], top[4:6]]),
                           np.array([1] * 4 + [0] * 2, dtype=np.int32),
                           np.array([-30.0] * 4 + [10.0] * 2),
                           np.array([0] * 4 + [-1] * 2, dtype=np.int32))
    _compare(out, _oracle_run(w, 20, extra), 2)


def test_driver_particle_tractions_match_oracle(tmp_path):
    """particle_surface_traction from particle sets and -1, with and without
    a math function."""
    w = workloads.column_2d(nx=6, ny=10)
    dt = w.dt
    fns = [{"id": 6, "type": "Linear", "xvalues": [0.0, 10 * dt], "fxvalues": [0.0, 1.0]}]
    d = str(tmp_path)
    write_deck(w, d, nsteps=20, output_steps=20, material_json=LE, math_functions=fns,
               tractions=[([10, 11, 12, 13], 1, -200.0, 6), (-1, 0, 5.0, None)])
    run_driver(d, "--ordered")
    out = read_output(output_path(d, 20, 20), 2)

    def extra(o):
        o.set_math_functions([("linear", [0.0, 10 * dt], [0.0, 1.0])])
        o.set_particle_tractions([10, 11, 12, 13, -1], [1, 1, 1, 1, 0],
                                 [-200.0, -200.0, -200.0, -200.0, 5.0], [0, 0, 0, 0, -1])
    _compare(out, _oracle_run(w, 20, extra), 2)


def _fixture_oracle(deck, dim, scheme, steps):
    """The fixture deck set up in the oracle as the reference's
    initialise_particles would: particles from the file (material 0),
    material_sets (particle set 2 -> material 1), volume and mass, loads."""
    import json
    d = deck
    nodes = np.array(d["_nodes"], dtype=float)
    o = oracle.Oracle(dim, scheme, velocity_update="pic")
    o.set_mesh(nodes, np.array(d["_cells"]))
    ms = d["materials"]
    o.set_materials([mat.linear_elastic(m["youngs_modulus"], m["poisson_ratio"], m["density"])
                     for m in ms])
    cons = np.array(d["_cons"])
    o.set_velocity_constraints(cons[:, 0].astype(np.int64), cons[:, 1].astype(np.int32),
                               cons[:, 2])
    o.set_math_functions([("linear", [0.0, 0.5, 1.0], [0.0, 1.0, 1.0])])
    o.set_nodal_forces([-1], [1], [10.5], [0])
    o.set_gravity(d["external_loading_conditions"]["gravity"])
    x = np.array(d["_particles"])
    n = len(x)
    o.set_particles(x=x, mass=np.ones(n), volume=np.ones(n), density=np.ones(n), material_id=0)
    o.assign_materials(np.arange(n), 1)  # set 2 = ids 0..15; absent ids are skipped
    o.compute_volume_mass()
    o.set_particle_tractions([-1], [1], [10.5], [0])
    o.step(steps, 0.001, 0.02)
    _ = json
    return o.particles()


@pytest.mark.parametrize("dim,scheme", [(2, "usf"), (3, "musl"), (2, "usl")])
def test_driver_reference_fixture_deck_matches_oracle(tmp_path, dim, scheme):
    d = str(tmp_path)
    deck = write_fixture_deck(d, dim, scheme)
    run_driver(d, "--ordered")
    for step in (0, 5, 10):
        assert os.path.exists(output_path(d, step, 10, uuid="fixture")), step
    out = read_output(output_path(d, 10, 10, uuid="fixture"), dim)
    assert np.all(out["material_id"] == 1)  # the material set applied
    # the deck's geometry, as written, for the oracle
    from decks import write_fixture_deck as _w  # noqa: F401
    import json
    with open(os.path.join(d, "mpm.json")) as f:
        assert json.load(f)["material_sets"][0]["pset_id"] == 2
    rows = [r.split() for r in open(os.path.join(d, "mesh-%dd.txt" % dim)) if "!" not in r]
    nn, nc = int(rows[0][0]), int(rows[0][1])
    deck["_nodes"] = [[float(v) for v in r] for r in rows[1:1 + nn]]
    deck["_cells"] = [[int(v) for v in r] for r in rows[1 + nn:1 + nn + nc]]
    deck["_cons"] = [[float(v) for v in r.split()] for r in
                     open(os.path.join(d, "velocity-constraints.txt"))]
    prow = [r.split() for r in open(os.path.join(d, "particles-%dd.txt" % dim))]
    deck["_particles"] = [[float(v) for v in r] for r in prow[1:]]
    p = _fixture_oracle(deck, dim, scheme, 10)
    _compare(out, p, dim)


def test_driver_apic_deck_matches_oracle_and_resumes(tmp_path):
    """"velocity_update": "apic" (VelocityUpdateType, src/mpm.cc:19-24): the
    driver's tables match the oracle's APIC run bit for bit, mapping matrices
    included (written from the first step on, flagged as the reference's
    Particle::pod does); resuming from the table of step 10 carries the
    mapping matrices and continues bit for bit."""
    import json
    w = workloads.column_2d(nx=6, ny=10)
    w.scheme = "musl"
    d = str(tmp_path)
    write_deck(w, d, nsteps=30, output_steps=10, material_json=LE, velocity_update="apic")
    run_driver(d, "--ordered")
    o = w.setup(oracle.Oracle(2, "musl", velocity_update="apic"))
    o.step(30, w.dt)
    out = read_output(output_path(d, 30, 30), 2)
    _compare(out, o.particles(), 2)
    assert np.all(out["initialise_mapping_matrix"] == 1)
    assert np.array_equal(out["mapping_matrix"], o.mapping_matrix())
    assert np.all(read_output(output_path(d, 0, 30), 2)["initialise_mapping_matrix"] == 0)
    first = read_output(output_path(d, 30, 30), 2)
    with open(os.path.join(d, "mpm.json")) as f:
        deck = json.load(f)
    deck["analysis"]["resume"] = {"resume": True, "uuid": "deck", "step": 10}
    with open(os.path.join(d, "mpm.json"), "w") as f:
        json.dump(deck, f)
    os.remove(output_path(d, 20, 30))
    os.remove(output_path(d, 30, 30))
    run_driver(d, "--ordered")
    again = read_output(output_path(d, 30, 30), 2)
    for k in first:
        assert np.array_equal(first[k], again[k]), k


def test_driver_resume_continues_bitwise(tmp_path):
    """A Mohr-Coulomb collapse run for 30 steps with tables every 10, then
    resumed from the table of step 10 ("resume": {"resume": true, "uuid",
    "step": 10}): the resumed run writes the tables of steps 20 and 30 equal,
    field by field and bit for bit, to the uninterrupted run's (the state
    the step carries is in the table; the density, which only the
    separated-state check reads, restarts as mass / volume as in the
    reference)."""
    import json
    from mpm_amd.capi import Solver
    w = workloads.collapse_3d(nx=8, ny=4, nz=4, block=(4, 4, 4))
    w.velocity_seed = None
    s0 = w.setup(Solver(3, w.scheme, ordered=True))
    stresses = s0.particles(["stress"])["stress"]
    d = str(tmp_path)
    write_deck(w, d, nsteps=30, output_steps=10, material_json=MC3, stresses=stresses)
    run_driver(d, "--ordered")
    first = {s: read_output(output_path(d, s, 30), 3) for s in (10, 20, 30)}
    yields = np.bincount(first[30]["statevars"][:, 0].astype(int), minlength=3)
    assert yields[1] + yields[2] > 0
    keep = os.path.join(d, "first")
    shutil.copytree(os.path.join(d, "results"), keep)
    with open(os.path.join(d, "mpm.json")) as f:
        deck = json.load(f)
    deck["analysis"]["resume"] = {"resume": True, "uuid": "deck", "step": 10}
    with open(os.path.join(d, "mpm.json"), "w") as f:
        json.dump(deck, f)
    os.remove(output_path(d, 20, 30))
    os.remove(output_path(d, 30, 30))
    run_driver(d, "--ordered")
    for s in (20, 30):
        again = read_output(output_path(d, s, 30), 3)
        for k in first[s]:
            assert np.array_equal(first[s][k], again[k]), (s, k)
