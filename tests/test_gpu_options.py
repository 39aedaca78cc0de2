"""Step options beyond the benchmark decks, GPU (ordered mode) against the
oracle bit for bit: Cundall damping, PIC and blended FLIP velocity updates,
acceleration constraints (constant and scaled by math functions of time),
concentrated nodal forces, particle surface tractions,
prescribed non-zero velocities, particle velocity constraints, gravity
ramping, one particle per cell, 2D Mohr-Coulomb under every scheme (with
softening, and with separated particles); and fast mode on the same decks
within the per-step tolerance."""
import numpy as np
import pytest

import oracle
from mpm_amd import materials as mat
from mpm_amd import workloads

pytestmark = pytest.mark.gpu

FIELDS = ("x", "u", "v", "stress", "strain", "F", "volume", "density", "xi", "statevars")


def _solver(w, ordered, **kw):
    from mpm_amd.capi import Solver
    return Solver(w.dim, kw.get("scheme", w.scheme), velocity_update=kw.get("vu", "flip"),
                  blending_ratio=kw.get("blend", 1.0), ordered=ordered)


def _oracle(w, **kw):
    return oracle.Oracle(w.dim, kw.get("scheme", w.scheme), velocity_update=kw.get("vu", "flip"),
                         blending_ratio=kw.get("blend", 1.0))


def _extra(obj, acc=None, ramp=None, pvc=None, fns=None):
    if fns is not None:
        obj.set_math_functions(fns)
    if acc is not None:
        obj.set_acceleration_constraints(*acc)
    if ramp is not None:
        obj.set_gravity(ramp[0], ramp[1])
    if pvc is not None:
        obj.set_particle_velocity_constraints(*pvc)
    return obj


def _run(w, steps, damping=0.0, acc=None, ramp=None, pvc=None, ordered=True, fns=None, **kw):
    g = _extra(w.setup(_solver(w, ordered, **kw)), acc, ramp, pvc, fns)
    o = _extra(w.setup(_oracle(w, **kw)), acc, ramp, pvc, fns)
    for _ in range(steps // 10):
        g.step(10, w.dt, damping)
        o.step(10, w.dt, damping)
    return g.particles(), o.particles()


def _check(pg, po, exact, tol=1e-9):
    assert np.array_equal(pg["cell"], po["cell"])
    for f in FIELDS:
        if exact:
            assert np.array_equal(pg[f], po[f]), f
        else:
            scale = max(float(np.max(np.abs(po[f]))), 1e-300)
            assert float(np.max(np.abs(pg[f] - po[f]))) / scale <= tol, f


def _column_mc(nx=6, ny=10):
    w = workloads.column_2d(nx=nx, ny=ny)
    w.materials = mat.mohr_coulomb(1e6, 0.3, 1800.0, friction=30.0, dilation=5.0,
                                   cohesion=2000.0)[None, :]
    return w


@pytest.mark.parametrize("ordered", [True, False])
def test_cundall_damping(ordered):
    w = workloads.column_2d(nx=6, ny=10)
    pg, po = _run(w, 50, damping=0.05, ordered=ordered)
    _check(pg, po, exact=ordered)


@pytest.mark.parametrize("vu,blend", [("pic", 1.0), ("flip", 0.7)])
def test_pic_and_blended_flip(vu, blend):
    w = workloads.column_2d(nx=6, ny=10)
    pg, po = _run(w, 40, vu=vu, blend=blend)
    _check(pg, po, exact=True)


def test_acceleration_constraints_and_prescribed_velocity():
    w = workloads.column_2d(nx=6, ny=10)
    top = np.nonzero(np.abs(w.coords[:, 1] - w.coords[:, 1].max()) < 1e-9)[0]
    w.vel_nodes = np.concatenate([w.vel_nodes, top])
    w.vel_dirs = np.concatenate([w.vel_dirs, np.ones(len(top), dtype=np.int32)])
    w.vel_values = np.concatenate([w.vel_values, np.full(len(top), -0.01)])
    xmax = w.coords[:, 0].max()
    mid = np.nonzero((np.abs(w.coords[:, 1] - 0.25) < 1e-9) & (w.coords[:, 0] > 1e-9)
                     & (w.coords[:, 0] < xmax - 1e-9))[0]
    assert len(mid) > 0
    acc = (mid, np.zeros(len(mid), dtype=np.int32), np.full(len(mid), 0.05))
    pg, po = _run(w, 40, acc=acc)
    _check(pg, po, exact=True)


@pytest.mark.parametrize("scheme,ordered", [("usl", True), ("musl", True), ("usf", True),
                                            ("musl", False)])
def test_time_function_acceleration_constraints(scheme, ordered):
    """Acceleration constraints scaled by math functions of time: a Linear
    ramp that starts, rises and holds within the run, a Sin that switches
    on part-way, and a constant constraint beside them; the constrained
    nodes' accelerations follow value * f(step * dt)."""
    w = workloads.column_2d(nx=6, ny=10)
    xmax = w.coords[:, 0].max()
    row = np.nonzero((np.abs(w.coords[:, 1] - 0.25) < 1e-9) & (w.coords[:, 0] > 1e-9)
                     & (w.coords[:, 0] < xmax - 1e-9))[0]
    assert len(row) >= 3
    fns = [("linear", [0.0, 10 * w.dt, 30 * w.dt], [0.0, 1.0, 0.25]),
           ("sin", 5 * w.dt, 2 * np.pi / (20 * w.dt), (8 * w.dt, 1.0))]
    fn = np.arange(len(row)) % 3 - 1  # -1 (none), 0 (linear), 1 (sin)
    acc = (row, np.zeros(len(row), dtype=np.int32), np.full(len(row), 0.05), fn)
    g = _extra(w.setup(_solver(w, ordered, scheme=scheme)), acc=acc, fns=fns)
    o = _extra(w.setup(_oracle(w, scheme=scheme)), acc=acc, fns=fns)
    for _ in range(4):
        g.step(10, w.dt)
        o.step(10, w.dt)
    _check(g.particles(), o.particles(), exact=ordered, tol=1e-9)
    ng, no = g.nodes(), o.nodes()
    active = no["status"][row] != 0  # nodes of the last step (the GPU clears its flags)
    assert np.all(active)
    t = 39 * w.dt  # time of the last step
    f = np.where(fn == 0, np.interp(t, fns[0][1], fns[0][2]),
                 np.where(fn == 1, np.sin(fns[1][2] * (t - fns[1][1])), 1.0))
    a = ng["acceleration"][row, 0]
    np.testing.assert_allclose(a, 0.05 * f, rtol=1e-13, atol=1e-15)
    if ordered:
        assert np.array_equal(a, no["acceleration"][row, 0])


@pytest.mark.parametrize("scheme,ordered", [("usl", True), ("musl", True), ("usf", True),
                                            ("usl", False)])
def test_concentrated_nodal_forces(scheme, ordered):
    """Concentrated nodal forces: a Linear-ramped downward load on the top
    row of the column, a constant sideways load on one node, then a second
    assignment on one loaded node that replaces its function (the node's
    function is that of its last assignment)."""
    w = workloads.column_2d(nx=6, ny=10)
    top = np.nonzero((np.abs(w.coords[:, 1] - 0.5) < 1e-9) & (w.coords[:, 0] < 0.3 + 1e-9))[0]
    assert len(top) >= 3
    fns = [("linear", [0.0, 20 * w.dt], [0.0, 1.0]), ("sin", 0.0, 2 * np.pi / (40 * w.dt), (0.0, 1.0))]

    def setup(obj):
        w.setup(obj)
        obj.set_math_functions(fns)
        obj.set_nodal_forces(top, 1, -50.0, 0)
        obj.set_nodal_forces([top[1]], [0], [20.0], [-1])
        obj.set_nodal_forces([top[2]], [0], [5.0], [1])
        return obj

    g = setup(_solver(w, ordered, scheme=scheme))
    o = setup(_oracle(w, scheme=scheme))
    for _ in range(4):
        g.step(10, w.dt)
        o.step(10, w.dt)
    _check(g.particles(), o.particles(), exact=ordered)


@pytest.mark.parametrize("scheme,ordered", [("usl", True), ("musl", True), ("usf", True),
                                            ("usl", False)])
def test_particle_tractions(scheme, ordered):
    """Particle surface tractions: a Linear-ramped downward traction on the
    particles of the top cell row, a constant sideways one on every
    particle that a later load replaces for three particles, and a Sin load
    on one particle in the other direction."""
    w = workloads.column_2d(nx=6, ny=10)
    fns = [("linear", [0.0, 20 * w.dt], [0.0, 1.0]), ("sin", 0.0, 2 * np.pi / (40 * w.dt), (0.0, 1.0))]

    def setup(obj):
        w.setup(obj)
        x = obj.particles(["x"])["x"]
        top = np.nonzero(x[:, 1] > x[:, 1].max() - 0.05)[0]
        obj.set_math_functions(fns)
        ids = np.concatenate([top, [-1], [3, 4, 5], [7]])
        dirs = np.concatenate([np.ones(len(top), dtype=np.int32), [0], [0, 0, 0], [1]])
        vals = np.concatenate([np.full(len(top), -400.0), [25.0], [-60.0, -60.0, -60.0], [90.0]])
        fn = np.concatenate([np.zeros(len(top), dtype=np.int32), [-1], [-1, -1, -1], [1]])
        obj.set_particle_tractions(ids, dirs, vals, fn)
        return obj

    g = setup(_solver(w, ordered, scheme=scheme))
    o = setup(_oracle(w, scheme=scheme))
    for _ in range(4):
        g.step(10, w.dt)
        o.step(10, w.dt)
    _check(g.particles(), o.particles(), exact=ordered)


@pytest.mark.parametrize("ordered", [True, False])
def test_loads_3d(ordered):
    """Hex8: particle tractions (Linear-ramped, on the top particles; one
    constant on every particle) and concentrated nodal forces (Sin-scaled)
    under MUSL."""
    w = workloads.collapse_3d(nx=8, ny=4, nz=4, block=(4, 4, 4))
    w.materials = mat.linear_elastic(5e6, 0.3, 1800.0)[None, :]
    fns = [("linear", [0.0, 20 * w.dt], [0.0, 1.0]), ("sin", 0.0, 2 * np.pi / (40 * w.dt), (0.0, 1.0))]
    top_nodes = np.nonzero((np.abs(w.coords[:, 2] - 0.04) < 1e-9) & (w.coords[:, 0] < 0.04 + 1e-9))[0]
    assert len(top_nodes) > 0

    def setup(obj):
        w.setup(obj)
        x = obj.particles(["x"])["x"]
        top = np.nonzero(x[:, 2] > x[:, 2].max() - 0.005)[0]
        obj.set_math_functions(fns)
        obj.set_particle_tractions(np.concatenate([top, [-1]]),
                                   np.concatenate([np.full(len(top), 2), [0]]).astype(np.int32),
                                   np.concatenate([np.full(len(top), -300.0), [15.0]]),
                                   np.concatenate([np.zeros(len(top)), [-1]]).astype(np.int32))
        obj.set_nodal_forces(top_nodes, 1, 4.0, 1)
        return obj

    g = setup(_solver(w, ordered))
    o = setup(_oracle(w))
    for _ in range(3):
        g.step(10, w.dt)
        o.step(10, w.dt)
    _check(g.particles(), o.particles(), exact=ordered)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("ordered", [True, False])
@pytest.mark.parametrize("scheme", ["usf", "usl", "musl"])
def test_jaumann_stress_rate(dim, ordered, scheme):
    """LinearElastic with the Jaumann rate (linear_elastic.tcc:96-124) on a
    soft block under gravity with random initial velocities, so that the
    particles spin; every scheme, so that each kernel build with the
    rotation runs (USF: the stress-then-scatter kernel in fast mode; USL:
    G2P with the stress update; MUSL: the stress kernel with locate)."""
    if dim == 2:
        w = workloads.column_2d(nx=6, ny=10)
    else:
        w = workloads.collapse_3d(nx=8, ny=4, nz=4, block=(4, 4, 4))
    w.velocity_seed, w.velocity_amplitude = 7, 0.05
    w.materials = mat.linear_elastic(2e5, 0.3, 1800.0, stress_rate="jaumann")[None, :]
    pg, po = _run(w, 40, ordered=ordered, scheme=scheme)
    _check(pg, po, exact=ordered)
    # the rotation matters: the rate-free update differs
    w.materials = mat.linear_elastic(2e5, 0.3, 1800.0)[None, :]
    pn, _ = _run(w, 40, ordered=True, scheme=scheme)
    assert not np.array_equal(pn["stress"], po["stress"])


def test_stress_rate_validation():
    from mpm_amd.capi import MpmxError, E_ARG, Solver
    s = Solver(2, "usf")
    bad = mat.linear_elastic(1e6, 0.3, 1800.0)
    bad[4] = 3
    with pytest.raises(MpmxError) as e:
        s.set_materials(bad[None, :])
    assert e.value.code == E_ARG
    s.close()


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("ordered", [True, False])
@pytest.mark.parametrize("scheme", ["usf", "usl", "musl"])
def test_green_naghdi_stress_rate(dim, ordered, scheme):
    """LinearElastic with the Green-Naghdi rate (linear_elastic.tcc:129-164):
    a spinning block, every scheme, against the oracle."""
    if dim == 2:
        w = workloads.column_2d(nx=6, ny=10)
    else:
        w = workloads.collapse_3d(nx=8, ny=4, nz=4, block=(4, 4, 4))
    w.velocity_seed, w.velocity_amplitude = 7, 0.05
    w.materials = mat.linear_elastic(2e5, 0.3, 1800.0, stress_rate="green_naghdi")[None, :]
    pg, po = _run(w, 40, ordered=ordered, scheme=scheme)
    _check(pg, po, exact=ordered)
    w.materials = mat.linear_elastic(2e5, 0.3, 1800.0, stress_rate="jaumann")[None, :]
    pj, _ = _run(w, 40, ordered=True, scheme=scheme)
    assert not np.array_equal(pj["stress"], po["stress"])


@pytest.mark.parametrize("rate", ["jaumann", "green_naghdi"])
@pytest.mark.parametrize("scheme", ["usf", "usl", "musl"])
def test_mohr_coulomb_stress_rates(rate, scheme):
    """MohrCoulomb with an objective stress rate (compute_trial_stress,
    infinitesimal_elasto_plastic.tcc:1-25): the rotated trial stress feeds the
    return mapping, also when it is deferred; ordered mode bitwise against the
    oracle on the collapse deck, with a rate-free MohrCoulomb and a
    LinearElastic material beside it."""
    w = workloads.collapse_3d(nx=8, ny=4, nz=4, block=(4, 4, 4))
    w.velocity_seed, w.velocity_amplitude = 7, 0.05
    mc = mat.mohr_coulomb(5e6, 0.3, 1800.0, friction=30.0, dilation=0.0, cohesion=0.0,
                          stress_rate=rate)
    w.materials = np.stack([mc, mat.mohr_coulomb(5e6, 0.3, 1800.0, friction=30.0, dilation=0.0,
                                                 cohesion=0.0),
                            mat.linear_elastic(5e6, 0.3, 1800.0, stress_rate=rate)])
    n = w.nparticles
    ids = np.arange(n)

    def assign(o):
        o.assign_materials(ids[ids % 3 == 1], 1)
        o.assign_materials(ids[ids % 3 == 2], 2)
        return o

    g = assign(w.setup(_solver(w, True, scheme=scheme)))
    o = assign(w.setup(_oracle(w, scheme=scheme)))
    for _ in range(4):
        g.step(10, w.dt)
        o.step(10, w.dt)
    pg, po = g.particles(), o.particles()
    _check(pg, po, exact=True)
    y = po["statevars"][ids % 3 == 0, 0]
    assert np.sum((y == 1) | (y == 2)) > 0, "no rate particle yields"


def test_gravity_ramp():
    w = workloads.column_2d(nx=6, ny=10)
    pg, po = _run(w, 40, ramp=(w.gravity, 20 * w.dt))
    _check(pg, po, exact=True)


def test_one_particle_per_cell():
    w = workloads.column_2d(nx=6, ny=10)
    w.nppd = 1
    pg, po = _run(w, 30)
    assert len(pg["x"]) == 60
    _check(pg, po, exact=True)


@pytest.mark.parametrize("scheme", ["usf", "usl", "musl"])
def test_mohr_coulomb_2d(scheme):
    w = _column_mc()
    pg, po = _run(w, 60, scheme=scheme)
    _check(pg, po, exact=True)
    assert np.any(po["statevars"][:, 0] > 0), "no plastic state reached"


def _collapse_mc(**kw):
    w = workloads.collapse_2d(8, 8, 0.01, 1e6, 30.0, 1800.0)
    w.materials = mat.mohr_coulomb(1e6, 0.3, 1800.0, friction=30.0, dilation=0.0, cohesion=0.0,
                                   **kw)[None, :]
    return w


@pytest.mark.parametrize("scheme,ordered", [("usl", True), ("musl", True), ("usf", True)])
def test_mohr_coulomb_softening(scheme, ordered):
    """Linear softening of friction and cohesion with the plastic deviatoric
    strain (mohr_coulomb.tcc:360-384): the strength state variables change
    in the stress kernel, the deferred return mapping reads them back.
    Ordered mode only: over 100 steps of a cohesionless collapse, fast
    mode's atomic-order rounding moves particles across yield thresholds."""
    w = _collapse_mc(residual_friction=22.0, residual_cohesion=0.0, peak_pdstrain=1e-6,
                     residual_pdstrain=1e-3, softening=True)
    pg, po = _run(w, 100, scheme=scheme, ordered=ordered)
    _check(pg, po, exact=ordered)
    sv = po["statevars"]
    assert np.any(sv[:, 0] == 1) and np.any((sv[:, 8] > 1e-6) & (sv[:, 8] < 1e-3))
    assert np.any(sv[:, 1] < np.radians(30.0) - 1e-9), "no softened friction angle"


@pytest.mark.parametrize("scheme", ["usl", "musl"])
def test_mohr_coulomb_separated_state(scheme):
    """Particles whose density falls to the critical packing density are
    separated: zero stress, yield state 3, invariants left as they were
    (mohr_coulomb.tcc:391-396)."""
    w = _collapse_mc(packing_fraction=0.6, packing_fraction_minimum=0.6 * (1 - 3e-4))
    pg, po = _run(w, 60, scheme=scheme)
    _check(pg, po, exact=True)
    sep = po["statevars"][:, 0] == 3
    assert np.any(sep) and np.all(po["stress"][sep] == 0.0)


@pytest.mark.parametrize("scheme", ["usf", "usl", "musl"])
def test_particle_velocity_constraints(scheme):
    """Every particle's x velocity fixed, particle 3's y velocity set twice
    (the later constraint wins), particle 7's y velocity fixed."""
    w = workloads.column_2d(nx=6, ny=10)
    pvc = ([-1, 3, 7, 3], [0, 1, 1, 1], [0.001, -0.02, 0.005, -0.01])
    pg, po = _run(w, 40, pvc=pvc, scheme=scheme)
    _check(pg, po, exact=True)
    assert np.all(pg["v"][:, 0] == 0.001)
    assert pg["v"][3, 1] == -0.01 and pg["v"][7, 1] == 0.005


@pytest.mark.parametrize("dim,nppd", [(2, 1), (2, 3), (2, 4), (3, 3), (3, 4)])
def test_gauss_generator_counts(dim, nppd):
    """1 to 4 particles per direction: positions, volumes, masses and cells
    as the oracle generates them, and 10 ordered steps bit for bit."""
    if dim == 2:
        w = workloads.column_2d(nx=4, ny=6)
    else:
        w = workloads.collapse_3d(nx=6, ny=3, nz=4, block=(3, 3, 3))
        w.velocity_seed = None
        w.materials = mat.linear_elastic(5e6, 0.3, 1800.0)[None, :]
    w.nppd = nppd
    pg, po = _run(w, 10, scheme=w.scheme)
    assert len(pg["x"]) == len(w.cell_set if w.cell_set is not None else w.cells) * nppd ** dim
    _check(pg, po, exact=True)


def _inclined_base(w, angles):
    """The deck's base nodes on a general boundary: only the normal velocity
    (local direction dim-1) is held at 0, with Coulomb friction mu = 0.3 in
    the local frame; the base is tilted by the Euler angles."""
    d = w.dim
    base = np.nonzero(np.abs(w.coords[:, d - 1]) <= 1e-9)[0].astype(np.int64)
    keep = ~np.isin(w.vel_nodes, base)
    w.vel_nodes = np.concatenate([w.vel_nodes[keep], base])
    w.vel_dirs = np.concatenate([w.vel_dirs[keep],
                                 np.full(len(base), d - 1)]).astype(np.int32)
    w.vel_values = np.concatenate([np.asarray(w.vel_values)[keep], np.zeros(len(base))])
    w.fric_nodes = base
    w.fric_dirs = np.full(len(base), d - 1, dtype=np.int32)
    w.fric_sign = np.full(len(base), -1, dtype=np.int32)
    w.fric_mu = np.full(len(base), 0.3)
    ang = np.tile(np.radians(angles), (len(base), 1))
    return base, ang


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("ordered", [True, False])
def test_general_boundary_constraints(dim, ordered):
    """Velocity and friction constraints of the base nodes in a tilted frame
    (node.tcc:385-401, 540-583, 602-655): the block slides down the inclined
    base; GPU against the oracle (MUSL, Mohr-Coulomb in 3D)."""
    if dim == 2:
        w = workloads.column_2d(nx=6, ny=10)
        angles = (15.0, 10.0)
    else:
        w = workloads.collapse_3d(nx=8, ny=4, nz=4, block=(4, 4, 4))
        angles = (10.0, 8.0, 6.0)
    base, ang = _inclined_base(w, angles)
    o = w.setup(_oracle(w, scheme="musl"))
    o.set_nodal_euler_angles(base, ang)
    if ordered:
        g = w.setup(_solver(w, True, scheme="musl"))
        g.set_nodal_euler_angles(base, ang)
        for _ in range(4):
            g.step(10, w.dt)
            o.step(10, w.dt)
        pg, po = g.particles(), o.particles()
        _check(pg, po, exact=True)
        done = 40
    else:
        # fast mode, each step from the oracle's state: friction's stick/slip
        # tests and the yield surface amplify atomic-order rounding over steps
        for _ in range(8):
            o.step(5, w.dt)
            po = o.particles()
            from mpm_amd.capi import Solver
            g = Solver(w.dim, "musl", ordered=False)
            g.set_mesh(w.coords, w.cells)
            g.set_materials(w.materials)
            g.set_velocity_constraints(w.vel_nodes, w.vel_dirs, w.vel_values)
            g.set_friction_constraints(w.fric_nodes, w.fric_dirs, w.fric_sign, w.fric_mu)
            g.set_nodal_euler_angles(base, ang)
            g.set_gravity(w.gravity)
            g.set_particles(x=po["x"], mass=po["mass"], volume=po["volume"],
                            density=po["density"], material_id=po["material_id"], v=po["v"],
                            u=po["u"], stress=po["stress"], strain=po["strain"], F=po["F"],
                            cell=po["cell"], statevars=po["statevars"])
            g.step(1, w.dt)
            o.step(1, w.dt)
            _check(g.particles(), o.particles(), exact=False, tol=1e-12)
            g.close()
        po = o.particles()
        done = 48
    # the tilt matters: the Cartesian base gives other velocities
    c = w.setup(_solver(w, True, scheme="musl"))
    c.step(done, w.dt)
    assert not np.array_equal(c.particles()["v"], po["v"])


@pytest.mark.parametrize("vu", ["apic", "asflip", "tpic"])
@pytest.mark.parametrize("scheme", ["usf", "usl", "musl"])
@pytest.mark.parametrize("dim", [2, 3])
def test_affine_velocity_updates(vu, scheme, dim):
    """APIC, ASFLIP (blend 0.9) and TPIC (particle.tcc:661-724, 1073-1140):
    ordered mode bit for bit against the oracle in every particle field and
    the mapping matrices, under every scheme; fast mode to rounding."""
    if dim == 2:
        w = workloads.column_2d(nx=6, ny=10)
    else:  # elastic, so that fast mode's rounding is not amplified by yielding
        w = workloads.collapse_3d(nx=8, ny=4, nz=4, block=(4, 4, 4))
        w.materials = mat.linear_elastic(5e6, 0.3, 1800.0)[None, :]
    w.velocity_seed, w.velocity_amplitude = 3, 0.05
    blend = 0.9 if vu == "asflip" else 1.0
    kw = dict(scheme=scheme, vu=vu, blend=blend)
    g = w.setup(_solver(w, True, **kw))
    f = w.setup(_solver(w, False, **kw))
    o = w.setup(_oracle(w, **kw))
    for _ in range(3):
        g.step(10, w.dt)
        f.step(10, w.dt)
        o.step(10, w.dt)
    po = o.particles()
    _check(g.particles(), po, exact=True)
    Bo = o.mapping_matrix()
    assert np.array_equal(g.mapping_matrix(), Bo)
    assert np.max(np.abs(Bo)) > 0.0
    _check(f.particles(), po, exact=False, tol=1e-9)
    scale = float(np.max(np.abs(Bo)))
    assert float(np.max(np.abs(f.mapping_matrix() - Bo))) <= 1e-9 * scale
