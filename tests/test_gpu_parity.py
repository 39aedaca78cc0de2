"""GPU parity: the HIP step (through the C-ABI) against the oracle.

Ordered mode sums nodal quantities in particle order, as the single-threaded
reference does; with LinearElastic it must reproduce the oracle bit for bit
(np.array_equal on every particle field). Mohr-Coulomb calls sin/cos/tan/acos/
pow, whose device and host library implementations may differ in the last
bit, so its check is normwise relative <= 1e-12 per step. The default fast
mode adds per-cell partial sums with FP64 atomics in arbitrary order; its
check is normwise relative <= 1e-12 per step (per field: max|a-b| / max|b|)
and bit-exact cell ids.
"""
import numpy as np
import pytest

import oracle
from mpm_amd import workloads
from mpm_amd import materials as mat
from mpm_amd.mesh import structured_mesh

pytestmark = pytest.mark.gpu

FIELDS = ("x", "u", "v", "stress", "strain", "F", "volume", "density", "xi")
STEP_TOL = 1e-12


def _solver(dim, scheme, ordered):
    from mpm_amd.capi import Solver
    return Solver(dim, scheme, ordered=ordered)


def _pair(w, ordered, scheme=None):
    scheme = scheme or w.scheme
    g = w.setup(_solver(w.dim, scheme, ordered))
    o = w.setup(oracle.Oracle(w.dim, scheme))
    return g, o


def _relerr(a, b):
    scale = float(np.max(np.abs(b))) if b.size else 0.0
    if scale == 0.0:
        return float(np.max(np.abs(a))) if a.size else 0.0
    return float(np.max(np.abs(a - b))) / scale


def _compare(g, o, exact, tol=STEP_TOL, what=""):
    pg, po = g.particles(), o.particles()
    assert np.array_equal(pg["cell"], po["cell"]), what + ": cell ids differ"
    errs = {}
    for f in FIELDS + ("statevars",):
        if exact:
            bad = np.argwhere(pg[f] != po[f])
            assert bad.size == 0, "%s: %s differs bitwise at %s (gpu %r oracle %r)" % (
                what, f, bad[:3].tolist(), pg[f][tuple(bad[0])], po[f][tuple(bad[0])])
        errs[f] = _relerr(pg[f], po[f])
        if not exact:
            assert errs[f] <= tol, "%s: %s rel err %.3e > %.1e" % (what, f, errs[f], tol)
    return errs


def _column(nx=8, ny=16):
    return workloads.column_2d(nx=nx, ny=ny)


def _collapse(le=False):
    w = workloads.collapse_3d(nx=8, ny=4, nz=4, block=(4, 4, 4))
    if le:
        w.materials = mat.linear_elastic(5e6, 0.3, 1800.0)[None, :]
    return w


@pytest.mark.parametrize("scheme", ["usf", "usl", "musl"])
def test_ordered_2d_linear_elastic_bitwise(scheme):
    w = _column()
    g, o = _pair(w, ordered=True, scheme=scheme)
    _compare(g, o, exact=True, what="init")
    for s in range(20):
        g.step(1, w.dt)
        o.step(1, w.dt)
        _compare(g, o, exact=True, what="%s step %d" % (scheme, s))
    g.step(80, w.dt)
    o.step(80, w.dt)
    _compare(g, o, exact=True, what="%s after 100 steps" % scheme)


@pytest.mark.parametrize("scheme", ["usf", "usl", "musl"])
def test_ordered_3d_linear_elastic_bitwise(scheme):
    w = _collapse(le=True)
    g, o = _pair(w, ordered=True, scheme=scheme)
    _compare(g, o, exact=True, what="init")
    for s in range(10):
        g.step(5, w.dt)
        o.step(5, w.dt)
        _compare(g, o, exact=True, what="%s step %d" % (scheme, 5 * (s + 1)))


@pytest.mark.parametrize("scheme", ["musl", "usl", "usf"])
def test_ordered_3d_mohr_coulomb_bitwise(scheme):
    """Granular collapse deck (K0 start, friction base): every field and MC
    state variable bit-identical to the oracle over 120 steps."""
    w = _collapse()
    g, o = _pair(w, ordered=True, scheme=scheme)
    for s in range(12):
        g.step(10, w.dt)
        o.step(10, w.dt)
        _compare(g, o, exact=True, what="MC %s step %d" % (scheme, 10 * (s + 1)))
        assert np.array_equal(g.particles()["statevars"], o.particles()["statevars"])
    yields = np.bincount(o.particles()["statevars"][:, 0].astype(int), minlength=3)
    assert yields[1] + yields[2] > 0, "deck never yields: return mapping not exercised"


def test_fast_3d_mohr_coulomb_per_step():
    """Fast mode, each step from the oracle's state: 1e-12 per step."""
    w = _collapse()
    o = w.setup(oracle.Oracle(3, w.scheme))
    worst = 0.0
    for s in range(40):
        g = _gpu_from(o, w, w.scheme, False)
        g.step(1, w.dt)
        o.step(1, w.dt)
        e = _compare(g, o, exact=False, what="MC step %d" % s)
        worst = max(worst, max(e.values()))
        g.close()
    print("fast MC worst per-step field rel err over 40 steps: %.3e" % worst)


def test_fast_3d_odd_grid_per_step():
    """Fast mode on a 7 x 5 x 6 grid (no power-of-two dimension: cell and
    node strides that are not shifts), each step from the oracle's state, to
    1e-12."""
    from mpm_amd import workloads
    w = workloads.collapse_3d(nx=7, ny=5, nz=6, block=(3, 3, 3))
    o = w.setup(oracle.Oracle(3, w.scheme))
    for s in range(10):
        g = _gpu_from(o, w, w.scheme, False)
        g.step(1, w.dt)
        o.step(1, w.dt)
        _compare(g, o, exact=False, what="odd grid step %d" % s)
        g.close()


def _gpu_from(o, w, scheme, ordered=False):
    """A fresh GPU solver holding exactly the oracle's particle state."""
    from mpm_amd.capi import Solver
    po = o.particles()
    n = Solver(w.dim, scheme, ordered=ordered)
    n.set_mesh(w.coords, w.cells)
    n.set_materials(w.materials)
    if len(w.vel_nodes):
        n.set_velocity_constraints(w.vel_nodes, w.vel_dirs, w.vel_values)
    if len(w.fric_nodes):
        n.set_friction_constraints(w.fric_nodes, w.fric_dirs, w.fric_sign, w.fric_mu)
    n.set_gravity(w.gravity)
    n.set_particles(x=po["x"], mass=po["mass"], volume=po["volume"], density=po["density"],
                    material_id=po["material_id"], v=po["v"], u=po["u"], stress=po["stress"],
                    strain=po["strain"], F=po["F"], cell=po["cell"], statevars=po["statevars"])
    return n


@pytest.mark.parametrize("scheme", ["usf", "usl", "musl"])
def test_fast_2d_per_step(scheme):
    """Each step from the oracle's state: one step's atomic-order error only."""
    w = _column()
    o = w.setup(oracle.Oracle(2, scheme))
    o.step(5, w.dt)
    for s in range(10):
        g = _gpu_from(o, w, scheme)
        g.step(1, w.dt)
        o.step(1, w.dt)
        _compare(g, o, exact=False, what="fast %s step %d" % (scheme, s))
        g.close()


@pytest.mark.parametrize("scheme", ["usf", "musl"])
def test_fast_2d_full_run(scheme):
    """100 steps in fast mode stay within 1e-9 of the oracle."""
    w = _column()
    g, o = _pair(w, ordered=False, scheme=scheme)
    g.step(100, w.dt)
    o.step(100, w.dt)
    e = _compare(g, o, exact=False, tol=1e-9, what="fast %s 100 steps" % scheme)
    print("fast %s 100-step errors: %s" % (scheme, {k: "%.2e" % v for k, v in e.items()}))


def test_elastic_column_full_config():
    """The 2D elastic column as specified: quad4 32x64 cells of 0.05 m, 8192
    particles, USF, dt = 1e-4, all 1000 steps. Ordered mode is bitwise equal
    to the oracle at the end; fast mode is within 1e-9 (normwise relative,
    per field) and has the same cell ids."""
    w = workloads.column_2d()
    o = w.setup(oracle.Oracle(2, "usf"))
    g = w.setup(_solver(2, "usf", True))
    f = w.setup(_solver(2, "usf", False))
    for _ in range(4):
        o.step(w.nsteps // 4, w.dt)
        g.step(w.nsteps // 4, w.dt)
        f.step(w.nsteps // 4, w.dt)
    assert o.nparticles == 8192 and g.step_count == 1000
    _compare(g, o, exact=True, what="ordered column 1000 steps")
    e = _compare(f, o, exact=False, tol=1e-9, what="fast column 1000 steps")
    print("fast column 1000-step errors: %s" % {k: "%.2e" % v for k, v in e.items()})


@pytest.mark.parametrize("dim", [2, 3])
def test_cell_crossing_gpu(kats, dim):
    c = kats["cell_crossing_%dd" % dim]
    from mpm_amd.capi import Solver
    for ordered in (True, False):
        g = Solver(dim, "usf", ordered=ordered)
        g.set_mesh(np.array(c["nodes"], dtype=float), np.array(c["cells"]))
        m = c["material"]
        g.set_materials([mat.linear_elastic(m["youngs_modulus"], m["poisson_ratio"],
                                            m["density"])])
        n = len(c["particles"])
        vol = 1.0 / n
        g.set_particles(x=c["particles"], mass=[vol * m["density"]] * n, volume=[vol] * n,
                        density=[m["density"]] * n, material_id=0)
        assert list(g.particles()["cell"]) == c["cells_before"]
        vc = np.array(c["velocity_constraints"])
        g.set_velocity_constraints(vc[:, 0].astype(int), vc[:, 1].astype(int), vc[:, 2])
        g.step(1, c["dt"])
        cells = g.particles()["cell"]
        assert list(np.bincount(cells, minlength=2)) == c["cell_counts_after"]


def test_unlocatable_particle_raises():
    from mpm_amd.capi import Solver, UnlocatableParticle
    coords, cells = structured_mesh((0.0, 0.0), (1.0, 1.0), (2, 1))
    g = Solver(2, "usf", ordered=False)
    g.set_mesh(coords, cells)
    g.set_materials([mat.linear_elastic(1e7, 0.3, 1000.0)])
    g.set_particles(x=[[1.5, 0.5]], mass=[1000.0], volume=[1.0], density=[1000.0],
                    material_id=0, v=[[3000.0, 0.0]])
    with pytest.raises(UnlocatableParticle):
        g.step(1, 0.001)
    with pytest.raises(UnlocatableParticle):
        g.set_particles(x=[[5.0, 0.5]], mass=[1.0], volume=[1.0], density=[1.0], material_id=0)


def test_graph_replay_matches_direct_launches():
    w = _column()
    from mpm_amd.capi import Solver
    a = w.setup(Solver(2, "usf", ordered=False, graph=True))
    b = w.setup(Solver(2, "usf", ordered=False, graph=False))
    for _ in range(5):
        a.step(10, w.dt)
        b.step(10, w.dt)
        pa, pb = a.particles(), b.particles()
        assert np.array_equal(pa["cell"], pb["cell"])
        for f in ("x", "v", "stress"):
            assert _relerr(pa[f], pb[f]) <= 1e-12


def test_profiled_steps_record_every_kernel_and_match():
    """Profiling launches each kernel between HIP events (no graph): every
    kernel of the sequence is timed, and the ordered-mode state equals an
    unprofiled run's bit for bit."""
    w = _collapse()
    from mpm_amd.capi import Solver
    a = w.setup(Solver(3, "musl", ordered=True))
    b = w.setup(Solver(3, "musl", ordered=True))
    a.profile(True)
    a.step(4, w.dt)
    kt = a.kernel_times()
    a.profile(False)
    b.step(4, w.dt)
    for k in ("sort_by_cell", "gather_all", "g2p", "gather_momentum", "stress_locate",
              "mc_plastic", "locate_scan", "step_end"):
        assert k in kt and kt[k][1] == 4 and kt[k][0] > 0.0, (k, kt)
    pa, pb = a.particles(), b.particles()
    for f in FIELDS + ("statevars", "cell"):
        assert np.array_equal(pa[f], pb[f]), f


def test_sort_interval_changes_only_summation_order():
    """Reusing an earlier sort's order (particles that changed cell form their
    own runs) changes only the atomic summation order: a run sorted every step
    and one sorted every 16 steps agree within 1e-10 and bin identically."""
    from mpm_amd.capi import Solver, MpmxError
    w = _column()
    a = w.setup(Solver(2, w.scheme, ordered=False))
    b = w.setup(Solver(2, w.scheme, ordered=False))
    a.set_sort_interval(1)
    b.set_sort_interval(16)
    with pytest.raises(MpmxError):
        b.set_sort_interval(0)
    for _ in range(5):
        a.step(20, w.dt)
        b.step(20, w.dt)
        pa, pb = a.particles(), b.particles()
        assert np.array_equal(pa["cell"], pb["cell"])
        for f in FIELDS:
            assert _relerr(pa[f], pb[f]) <= 1e-10, f


def _lost_pair(ordered, sort_every=None, locate=True):
    from mpm_amd.capi import Solver
    import lost_particle_deck as deck
    g = deck.setup(Solver(2, "usf", ordered=ordered))
    o = deck.setup(oracle.Oracle(2, "usf"))
    if sort_every:
        g.set_sort_interval(sort_every)
    g.set_locate_particles(locate)
    o.set_locate_particles(locate)
    return g, o, deck.DT


def _oracle_failing_step(o, dt, limit=200):
    for s in range(limit):
        try:
            o.step(1, dt)
        except oracle.OracleError:
            return s
    raise AssertionError("particle never left the mesh")


@pytest.mark.parametrize("ordered,sort_every", [(True, None), (False, 4), (False, 3)])
def test_unlocatable_stops_at_the_failing_step(ordered, sort_every):
    """A particle leaves the mesh between full sorts while the others keep
    moving: the step raises, the state returned is that of the end of the
    failing step (later queued steps do nothing) and every other particle
    matches the oracle's state at its throw."""
    from mpm_amd.capi import UnlocatableParticle
    g, o, dt = _lost_pair(ordered, sort_every)
    k = _oracle_failing_step(o, dt)
    assert k > 4 and (sort_every is None or k % sort_every != 0), k
    with pytest.raises(UnlocatableParticle) as e:
        g.step(k + 25, dt)
    assert "step %d" % k in str(e.value)
    assert g.step_count == k + 1
    pg, po = g.particles(), o.particles()
    assert pg["cell"][0] == -1
    assert np.array_equal(pg["cell"][1:], po["cell"][1:])
    for f in FIELDS:
        a, b = pg[f][1:], po[f][1:]
        if ordered:
            assert np.array_equal(a, b), f
        else:
            assert _relerr(a, b) <= 1e-10, (f, _relerr(a, b))


@pytest.mark.parametrize("ordered", [True, False])
def test_locate_particles_false_removes_the_lost_particle(ordered):
    """locate_particles: false: the particle is removed at its failing step
    and the run goes on; the others keep their order, ids and the particle
    velocity constraint of the last particle follow the particles, and the
    state matches the oracle's removal."""
    g, o, dt = _lost_pair(ordered, None if ordered else 4, locate=False)
    n0 = g.nparticles
    ids0 = g.particle_ids()
    assert list(ids0) == list(range(n0))
    g.step(60, dt)
    o.step(60, dt)
    assert g.nparticles == n0 - 1 == o.nparticles
    assert list(g.particle_ids()) == list(range(1, n0))
    pg, po = g.particles(), o.particles()
    assert np.array_equal(pg["cell"], po["cell"])
    assert pg["v"][-1, 0] == 1e-3
    for f in FIELDS:
        if ordered:
            assert np.array_equal(pg[f], po[f]), f
        else:
            assert _relerr(pg[f], po[f]) <= 1e-10, (f, _relerr(pg[f], po[f]))
