"""Isoparametric meshes on the GPU against the oracle (through the C ABI).

Cells whose interior nodes are moved off the structured grid: particles are
located by the inverse element map (transform_real_to_unit_cell,
cell.tcc:617-828) instead of the Cartesian formula. In 2D every located
point is answered by the closed-form quad4 inverse; in 3D the warped hexes
need the affine guess and the Newton-Raphson iteration. Ordered mode must
match the oracle bit for bit (cell ids, local coordinates and every particle
field), fast mode to the atomic-order rounding.
"""
import numpy as np
import pytest

import oracle
from mpm_amd import workloads
from mpm_amd import materials as mat

pytestmark = pytest.mark.gpu

FIELDS = ("x", "u", "v", "stress", "strain", "F", "volume", "density", "xi", "statevars")


def _warp(w, amp):
    """Move interior nodes by amp * h * (smooth bumps), boundary nodes kept."""
    X = w.coords.copy()
    lo, hi = X.min(0), X.max(0)
    L = hi - lo
    h = np.min(np.abs(np.diff(np.unique(X[:, 0]))))
    u = (X - lo) / L
    bump = np.prod(np.sin(np.pi * u), axis=1)
    for k in range(X.shape[1]):
        X[:, k] += amp * h * bump * np.cos((k + 1) * np.pi * u[:, (k + 1) % X.shape[1]])
    w.coords = X
    w.isoparametric = True
    return w


def _column_2d():
    """A LinearElastic block of 4 x 6 quad4 cells of 0.25 m in a 10 x 8 cell
    mesh, thrown at up to 3 m/s so that particles cross cells. Cells of this
    size keep the reference's iteration (whose matrix is that of the unit
    cell) to a few hundred steps where the closed form has no root."""
    w = workloads.collapse_2d(4, 6, 0.25, 1e6, 30.0, 1800.0, dt=1e-3)
    w.materials = mat.linear_elastic(1e6, 0.3, 1800.0)[None, :]
    return _warp(w, 0.35)


def _collapse_3d():
    """The Mohr-Coulomb collapse deck on hex8 cells of 0.5 m (8 x 4 x 6, a
    4 x 4 x 4 block), warped so that the cells are not parallelepipeds."""
    w = workloads.collapse_3d(nx=8, ny=4, nz=6, h=0.5, block=(4, 4, 4), dt=1e-3)
    return _warp(w, 0.25)


def _compare(pg, po, exact, tol=1e-9, what=""):
    assert np.array_equal(pg["cell"], po["cell"]), what + ": cells differ"
    for f in FIELDS:
        if exact:
            bad = np.argwhere(pg[f] != po[f])
            assert bad.size == 0, "%s: %s differs at %s (%r vs %r)" % (
                what, f, bad[:3].tolist(), pg[f][tuple(bad[0])], po[f][tuple(bad[0])])
        else:
            scale = max(float(np.max(np.abs(po[f]))), 1e-300)
            err = float(np.max(np.abs(pg[f] - po[f]))) / scale
            assert err <= tol, "%s: %s rel err %.3e" % (what, f, err)


def _setup(w, solver):
    """The deck, then a rigid spin of the block about its centre (4 rad/s)
    plus a drift (1 m/s along x), so that particles cross skewed cell faces."""
    w.setup(solver)
    x = solver.particles(["x"])["x"]
    r = x - x.mean(0)
    v = np.zeros_like(x)
    v[:, 0] = 1.0 - 4.0 * r[:, 1]
    v[:, 1] = 4.0 * r[:, 0]
    solver.assign_velocity(v)
    return solver


def _pair(w, scheme, ordered):
    from mpm_amd.capi import Solver
    g = _setup(w, Solver(w.dim, scheme, ordered=ordered))
    o = _setup(w, oracle.Oracle(w.dim, scheme))
    return g, o


@pytest.mark.parametrize("scheme", ["usf", "usl", "musl"])
def test_isoparametric_2d_ordered_bitwise(scheme):
    w = _column_2d()
    g, o = _pair(w, scheme, True)
    _compare(g.particles(), o.particles(), True, what="placement")
    c0 = o.particles(["cell"])["cell"].copy()
    for s in range(6):
        g.step(10, w.dt)
        o.step(10, w.dt)
        _compare(g.particles(), o.particles(), True, what="%s step %d" % (scheme, 10 * (s + 1)))
    assert np.any(o.particles(["cell"])["cell"] != c0), "no particle changed cell"


@pytest.mark.parametrize("scheme", ["usf", "musl"])
def test_isoparametric_2d_fast(scheme):
    w = _column_2d()
    g, o = _pair(w, scheme, False)
    g.step(40, w.dt)
    o.step(40, w.dt)
    _compare(g.particles(), o.particles(), False, what="fast %s" % scheme)


def test_isoparametric_3d_mohr_coulomb_ordered():
    w = _collapse_3d()
    g, o = _pair(w, "musl", True)
    _compare(g.particles(), o.particles(), True, what="placement")
    c0 = o.particles(["cell"])["cell"].copy()
    for s in range(4):
        g.step(10, w.dt)
        o.step(10, w.dt)
        _compare(g.particles(), o.particles(), True, what="3D MUSL step %d" % (10 * (s + 1)))
    assert np.any(o.particles(["cell"])["cell"] != c0), "no particle changed cell"


def _gpu_from(o, w, scheme):
    """A fast-mode GPU solver holding the oracle's particle state (cells as
    hints; the local coordinates are located again)."""
    from mpm_amd.capi import Solver
    po = o.particles()
    g = Solver(w.dim, scheme, ordered=False)
    g.set_mesh(w.coords, w.cells, isoparametric=True)
    g.set_materials(w.materials)
    g.set_velocity_constraints(w.vel_nodes, w.vel_dirs, w.vel_values)
    if len(w.fric_nodes):
        g.set_friction_constraints(w.fric_nodes, w.fric_dirs, w.fric_sign, w.fric_mu)
    g.set_gravity(w.gravity)
    g.set_particles(x=po["x"], mass=po["mass"], volume=po["volume"], density=po["density"],
                    material_id=po["material_id"], v=po["v"], u=po["u"], stress=po["stress"],
                    strain=po["strain"], F=po["F"], cell=po["cell"], statevars=po["statevars"])
    return g


def test_isoparametric_3d_mohr_coulomb_fast_per_step():
    """Fast mode on the warped hexes, each step from the oracle's state
    (Mohr-Coulomb amplifies the atomic-order rounding over several steps)."""
    w = _collapse_3d()
    o = _setup(w, oracle.Oracle(3, "musl"))
    o.step(10, w.dt)
    for s in range(20):
        g = _gpu_from(o, w, "musl")
        _compare(g.particles(), o.particles(), True, what="relocated step %d" % s)
        g.step(1, w.dt)
        o.step(1, w.dt)
        _compare(g.particles(), o.particles(), False, tol=1e-12, what="fast step %d" % s)
        g.close()


def test_isoparametric_cells_hold_their_points():
    """After placement on the warped meshes, each particle's position is the
    element map of its local coordinates in its cell: N(xi) X = x."""
    from mpm_amd.capi import Solver
    for w in (_column_2d(), _collapse_3d()):
        g = _setup(w, Solver(w.dim, w.scheme, ordered=True))
        g.step(20, w.dt)
        p = g.particles(["x", "xi", "cell"])
        for i in range(0, len(p["x"]), 7):
            X = w.coords[w.cells[p["cell"][i]]]
            N, _ = oracle.shapefn(w.dim, p["xi"][i])
            assert np.max(np.abs(N @ X - p["x"][i])) <= 1e-9 * (X.max() - X.min() + 1)


def test_isoparametric_adjoint_refused():
    from mpm_amd.capi import Solver, MpmxError, E_UNSUPPORTED
    w = _column_2d()
    g = _setup(w, Solver(2, "usl"))
    with pytest.raises(MpmxError) as e:
        g.adjoint_forward(2, w.dt)
    assert e.value.code == E_UNSUPPORTED
