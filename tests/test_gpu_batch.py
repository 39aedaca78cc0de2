"""Batched runs (mpmx_set_instances) on the GPU.

A batch of independent instances must advance each instance exactly as a
solver holding that instance alone: in ordered mode (particle-order nodal
sums) the batch reproduces the single-instance runs bit for bit; in fast mode
the nodal sums of one instance never mix with another's, so the two agree to
the atomic-order rounding. The full-scan fallback of locate must stay inside
the particle's instance even where another instance's mesh overlaps in space.
"""
import numpy as np
import pytest

from mpm_amd import workloads
from mpm_amd import materials as mat
from mpm_amd.mesh import structured_mesh

pytestmark = pytest.mark.gpu

FIELDS = ("x", "u", "v", "stress", "strain", "F", "volume", "density", "xi", "statevars")


def _small_batch(n=4, elastic=False):
    b = workloads.batch_2d(ninst=n, seed=7, nmin=200, nmax=3000)
    if elastic:  # the same instances with LinearElastic (same E, nu, density)
        for w in b.instances:
            m = w.materials[0]
            w.materials = mat.linear_elastic(m[2], m[3], m[1])[None, :]
        b.materials = np.concatenate([w.materials for w in b.instances])
    return b


def _run_solo(w, steps, ordered):
    from mpm_amd.capi import Solver
    s = w.setup(Solver(w.dim, w.scheme, ordered=ordered))
    s.step(steps, w.dt)
    return s.particles()


@pytest.mark.parametrize("ordered", [True, False])
def test_batch_equals_instances_run_alone(ordered):
    """Ordered mode: the Mohr-Coulomb batch bit for bit. Fast mode: the
    elastic batch to 1e-10 after 40 steps (with Mohr-Coulomb, rounding-level
    differences of the atomic sums can move a particle across a yield-state
    threshold within 40 steps, and the comparison would measure that)."""
    from mpm_amd.capi import Solver
    b = _small_batch(elastic=not ordered)
    g = b.setup(Solver(b.dim, b.scheme, ordered=ordered))
    steps = 40
    g.step(steps, b.dt)
    pb = g.particles()
    off = 0
    for i, w in enumerate(b.instances):
        ps = _run_solo(w, steps, ordered)
        n = w.nparticles
        assert np.array_equal(pb["cell"][off:off + n] - b.cell_begin[i], ps["cell"]), i
        for f in FIELDS:
            a, e = pb[f][off:off + n], ps[f]
            if ordered:
                assert np.array_equal(a, e), (i, f)
            else:
                scale = max(float(np.max(np.abs(e))), 1e-300)
                assert float(np.max(np.abs(a - e))) / scale <= 1e-10, (i, f)
        off += n


def test_instances_must_not_share_nodes():
    from mpm_amd.capi import Solver, MpmxError
    coords, cells = structured_mesh((0.0, 0.0), (1.0, 1.0), (2, 1))
    s = Solver(2, "usf")
    s.set_mesh(coords, cells)
    with pytest.raises(MpmxError):
        s.set_instances([0, 1, 2])  # the two cells share two nodes


def test_full_scan_stays_in_the_instance():
    """Instance 0 covers [0,1]^2, instance 1 covers [0,2]x[0,1] at the same
    place. A particle of instance 0 thrown past x = 1 is outside its own mesh:
    the step must fail, not relocate it into instance 1."""
    from mpm_amd.capi import Solver, UnlocatableParticle
    c0, e0 = structured_mesh((0.0, 0.0), (1.0, 1.0), (1, 1))
    c1, e1 = structured_mesh((0.0, 0.0), (1.0, 1.0), (2, 1))
    coords = np.concatenate([c0, c1])
    cells = np.concatenate([e0, e1 + len(c0)])
    s = Solver(2, "usf")
    s.set_mesh(coords, cells)
    s.set_instances([0, 1, 3])
    s.set_materials([mat.linear_elastic(1e7, 0.3, 1000.0)])
    s.set_particles(x=[[0.5, 0.5]], mass=[1000.0], volume=[1.0], density=[1000.0],
                    material_id=0, v=[[1000.0, 0.0]], cell=[0])
    with pytest.raises(UnlocatableParticle):
        s.step(1, 0.001)
    # the same particle placed with a hint in instance 1 moves within it
    t = Solver(2, "usf")
    t.set_mesh(coords, cells)
    t.set_instances([0, 1, 3])
    t.set_materials([mat.linear_elastic(1e7, 0.3, 1000.0)])
    t.set_particles(x=[[0.5, 0.5]], mass=[1000.0], volume=[1.0], density=[1000.0],
                    material_id=0, cell=[1])
    assert list(t.particles()["cell"]) == [1]


class _NoInstances:
    """The oracle on the batch's union mesh: it has no notion of instances and
    needs none, since no node is shared between them and no particle of these
    decks leaves its cell's neighbourhood (the full-scan fallback, the only
    instance-aware part of the step, is never reached: checked below)."""

    def __init__(self, o):
        self.o = o

    def set_instances(self, cell_begin):
        pass

    def __getattr__(self, name):
        return getattr(self.o, name)


def test_batch_ordered_equals_oracle_on_the_union_mesh():
    """The batched Mohr-Coulomb USL run (union mesh, per-instance materials and
    K0 starts) in ordered mode against the oracle run of the same union mesh:
    bit for bit in every particle field after 60 steps."""
    import oracle
    from mpm_amd.capi import Solver
    b = _small_batch(n=5)
    g = b.setup(Solver(b.dim, b.scheme, ordered=True))
    o = b.setup(_NoInstances(oracle.Oracle(b.dim, b.scheme)))
    for _ in range(3):
        g.step(20, b.dt)
        o.step(20, b.dt)
        pg, po = g.particles(), o.particles()
        assert np.array_equal(pg["cell"], po["cell"])
        for f in FIELDS:
            assert np.array_equal(pg[f], po[f]), f
    # every particle stayed in its own instance's cells
    own = np.repeat(np.arange(len(b.instances)), [w.nparticles for w in b.instances])
    inst = np.searchsorted(b.cell_begin, pg["cell"], side="right") - 1
    assert np.array_equal(inst, own)


def test_full_batch_equals_instances_run_alone():
    """The full 256-instance batch (4.6M particles), ordered mode, 10 steps:
    every instance bit for bit equal to a solver holding it alone."""
    from mpm_amd.capi import Solver
    b = workloads.batch_2d()
    g = b.setup(Solver(b.dim, b.scheme, ordered=True))
    steps = 10
    g.step(steps, b.dt)
    pb = g.particles(["x", "v", "stress", "cell", "statevars"])
    off = 0
    for i, w in enumerate(b.instances):
        s = w.setup(Solver(w.dim, w.scheme, ordered=True))
        s.step(steps, w.dt)
        ps = s.particles(["x", "v", "stress", "cell", "statevars"])
        s.close()
        n = w.nparticles
        assert np.array_equal(pb["cell"][off:off + n] - b.cell_begin[i], ps["cell"]), i
        for f in ("x", "v", "stress", "statevars"):
            assert np.array_equal(pb[f][off:off + n], ps[f]), (i, f)
        off += n
    assert off == g.nparticles
