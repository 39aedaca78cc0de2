"""Adjoint of a batched run (mpmx_set_instances) against central differences
of the oracle: the adjoint inversion workload's shape at a small size, four
column instances of 8x8 quad4 cells on their own meshes (LinearElastic and
cohesionless MohrCoulomb alternating, K0 starts, USL), 20 and 60 steps
recorded with a checkpoint every 10.

The loss is sum(wv * v_T) over every particle (velocities carry the
parameters' effect at full relative precision; positions bury it under
their own rounding); the parameters are each instance's Young's modulus,
density and (MohrCoulomb instances) friction angle, perturbed relatively
along a random per-instance direction: all instances over 20 steps, the
LinearElastic ones (E, density) and the MohrCoulomb ones (friction angle)
over 60, where the cohesionless instances' yield-state kinks leave the
other quotients without a converged step. The
oracle runs the union mesh without instances (no node is shared between
them, so the instances do not interact; the batch tests check that the
batched step equals the oracle's on such a mesh). Central differences are
taken at the largest step whose quotient agrees to 3e-7 with that of a ten
times smaller one, and the adjoint (exact derivatives: apex ratio 0) must
agree to 1e-6 relative."""
import numpy as np
import pytest

from mpm_amd import workloads

pytestmark = pytest.mark.gpu

COLUMN = {"E": 2, "rho": 1, "phi": 4}  # material parameter columns


class _NoInstances:
    """The oracle on the batch's union mesh (it has no instances)."""

    def __init__(self, o):
        self.o = o

    def set_instances(self, cell_begin):
        pass

    def __getattr__(self, name):
        return getattr(self.o, name)


def _batch(nsteps):
    return workloads.adjoint_2d(ninst=4, columns=8, nsteps=nsteps, checkpoint_every=10)


def _oracle_loss(b, materials, wv):
    import oracle
    saved = b.materials
    b.materials = materials
    try:
        o = b.setup(_NoInstances(oracle.Oracle(b.dim, b.scheme)))
    finally:
        b.materials = saved
    o.step(b.nsteps, b.dt)
    return float(np.sum(wv * o.particles(["v"])["v"]))


def _fd(b, col, d, h, wv):
    out = []
    for sign in (1.0, -1.0):
        m = b.materials.copy()
        m[:, col] *= 1.0 + sign * h * d
        out.append(_oracle_loss(b, m, wv))
    return (out[0] - out[1]) / (2 * h)


def _fd_converged(b, col, d, wv, h=1e-2, agree=3e-7, levels=7):
    prev = _fd(b, col, d, h, wv)
    for _ in range(levels):
        h /= 10.0
        cur = _fd(b, col, d, h, wv)
        if abs(prev - cur) <= agree * abs(prev):
            return prev
        prev = cur
    raise AssertionError("finite differences did not converge")


_RUNS = {}


def _gradients(nsteps):
    """The batch of `nsteps` steps, the loss weights and the adjoint's
    gradients (one recorded run per length, shared by the cases)."""
    if nsteps not in _RUNS:
        from mpm_amd.capi import Solver
        b = _batch(nsteps)
        s = b.setup(Solver(b.dim, b.scheme))
        s.adjoint_set_apex_ratio(0.0)
        s.adjoint_forward(b.nsteps, b.dt, 0.0, checkpoint_every=b.checkpoint_every)
        n = s.nparticles
        wv = np.random.default_rng(21).normal(size=(n, b.dim))
        g = s.adjoint_backward(np.zeros((n, b.dim)), wv)
        s.close()
        _RUNS[nsteps] = (b, wv, g)
    return _RUNS[nsteps]


@pytest.mark.parametrize("param,which,nsteps", [("E", "all", 20), ("rho", "all", 20),
                                                ("phi", "mc", 20), ("E", "le", 60),
                                                ("rho", "le", 60), ("phi", "mc", 60)])
def test_batched_adjoint_matches_oracle_differences(param, which, nsteps):
    b, wv, g = _gradients(nsteps)
    nin = len(b.instances)
    bounds = np.cumsum([0] + [w.nparticles for w in b.instances])
    per_instance = np.array([g[param][bounds[i]:bounds[i + 1]].sum() for i in range(nin)])
    d = np.random.default_rng(13).normal(size=nin)
    if which == "le":
        d[1::2] = 0.0
    if which == "mc":  # the LinearElastic instances have no friction angle
        d[0::2] = 0.0
    if param == "phi":
        assert np.all(g["phi"][bounds[0]:bounds[1]] == 0.0)
    col = COLUMN[param]
    ad = float(np.sum(per_instance * b.materials[:, col] * d))
    fd = _fd_converged(b, col, d, wv)
    assert abs(ad - fd) <= 1e-6 * abs(fd), (param, which, nsteps, fd, ad)
