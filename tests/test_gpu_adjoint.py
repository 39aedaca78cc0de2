"""Adjoint of the explicit step (mpmx_adjoint_forward / _backward) against
central finite differences of the oracle's forward run, FP64.

Each case builds one small problem (one material per particle) on the GPU
and on the oracle, records the run on the GPU, takes the gradient of the
loss sum(wx * x_T) + sum(wv * v_T) with respect to per-particle E, density,
friction angle, initial positions and velocities, and compares its
directional derivative along a random direction with the oracle's central
difference along the same direction. The difference is taken at the
largest step whose quotient agrees to 3e-7 with that of a ten times smaller
step (Problem.fd_converged): below the kinks of the run (yield-state,
return-mapping iteration-count or cell changes between the two runs) and
above rounding noise. The bar is 1e-6 relative."""
import numpy as np
import pytest
# torch in the process before the first solver: mpm_amd.capi.load then
# initializes torch's device ahead of the library's, which the torch
# binding's CUDA-tensor test needs (test modules are imported before any
# test runs)
import torch  # noqa: F401

from adjoint_problem import Problem

pytestmark = pytest.mark.gpu

# largest central-difference steps tried: relative for E and rho, absolute
# otherwise
H = dict(E=1e-2, rho=1e-2, phi=1e-1, v0=1e-3, x0=1e-4)
TOL = 1e-6

CASES = {
    "usl_le": dict(scheme="usl", material="le"),
    "usf_le": dict(scheme="usf", material="le"),
    "musl_le": dict(scheme="musl", material="le"),
    "usl_le_pic": dict(scheme="usl", material="le", velocity_update="pic"),
    "usl_le_options": dict(scheme="usl", material="le", blending=0.8, damping=0.05,
                           ramp=0.002, acc_constraint=True),
    "usl_mc": dict(scheme="usl", material="mc", geostatic=True),
    "musl_le_timed_loads": dict(scheme="musl", material="le", nodal_forces=True,
                                acc_constraint=True),
    "musl_le_particle_constraints": dict(scheme="musl", material="le", particle_constraints=True),
    "usf_mc_particle_constraints": dict(scheme="usf", material="mc", geostatic=True,
                                        particle_constraints=True),
    "usf_mc": dict(scheme="usf", material="mc", geostatic=True),
    "usl_mc_apex_cutoff": dict(scheme="usl", material="mc", geostatic=True,
                               tension_cutoff=1e5),
    "musl_mix": dict(scheme="musl", material="mix", geostatic=True),
    "musl_mc_3d": dict(dim=3, scheme="musl", material="mc", ncells=(4, 3, 6), block=(2, 2, 3),
                       geostatic=True, nsteps=30),
    "usf_le_3d": dict(dim=3, scheme="usf", material="le", ncells=(4, 3, 6), block=(2, 2, 3),
                      nsteps=30),
}


def _gpu_gradients(P, ordered=False, every=7, apex_ratio=0.0):
    """Adjoint gradients of the problem's loss; exact derivatives unless an
    apex ratio is given."""
    from mpm_amd.capi import Solver
    s = Solver(P.dim, P.scheme, velocity_update=P.vu, blending_ratio=P.blending,
               ordered=ordered)
    P.setup(s, P.params())
    s.adjoint_set_apex_ratio(apex_ratio)
    s.adjoint_forward(P.nsteps, P.dt, P.damping, checkpoint_every=every)
    final = s.particles(["x", "v"])
    g = s.adjoint_backward(P.wx, P.wv)
    after = s.particles(["x", "v"])
    s.close()
    return g, final, after


@pytest.mark.parametrize("case", list(CASES))
def test_adjoint_matches_finite_differences(case):
    P = Problem(**CASES[case])
    g, final, after = _gpu_gradients(P)
    # the backward pass leaves the state at the end of the recorded run
    for f in ("x", "v"):
        assert np.array_equal(final[f], after[f])
    rng = np.random.default_rng(11)
    for name in ("E", "rho", "phi", "v0", "x0"):
        d = rng.normal(size=P.params()[name].shape)
        ad = P.predicted(g, name, d)
        if name == "phi" and P.material == "le":
            assert P.fd(name, d, H[name]) == 0.0 and ad == 0.0
            continue
        fd = P.fd_converged(name, d, H[name])
        assert abs(ad - fd) <= TOL * abs(fd), (name, fd, ad)


def test_adjoint_ordered_mode_and_segment_lengths():
    """Ordered mode records the oracle's trajectory bit for bit; the gradient
    does not depend on how the run is cut into checkpoint segments."""
    P = Problem(scheme="usl", material="mix", geostatic=True, nsteps=24)
    g1, final, _ = _gpu_gradients(P, ordered=True, every=1)
    o_x = P.oracle_loss(P.params())
    assert P.loss_of(final) == o_x
    for every in (5, 24, 100):
        g2, _, _ = _gpu_gradients(P, ordered=True, every=every)
        for k in ("E", "rho", "phi", "x0", "v0"):
            scale = np.max(np.abs(g1[k])) + 1e-300
            assert np.max(np.abs(g1[k] - g2[k])) <= 1e-12 * scale, (every, k)


def test_adjoint_individual_particles():
    """Per-particle gradients (not only directional sums) for a few
    particles, including the friction angle of yielding ones."""
    P = Problem(scheme="usl", material="mc", geostatic=True)
    g, _, _ = _gpu_gradients(P)
    picks = np.argsort(-np.abs(g["phi"]))[:3]
    for p in picks:
        for name in ("E", "phi"):
            d = np.zeros(P.np)
            d[p] = 1.0
            fd = P.fd_converged(name, d, H[name], agree=1e-6)
            ad = P.predicted(g, name, d)
            assert abs(ad - fd) <= 1e-5 * abs(fd), (p, name, fd, ad)


def test_adjoint_unsupported_options():
    from mpm_amd.capi import MpmxError, E_UNSUPPORTED, E_STATE, Solver
    from mpm_amd import materials as mat
    P = Problem(scheme="usl", material="le", nsteps=5)
    s = Solver(2, "usl")
    P.setup(s, P.params())
    with pytest.raises(MpmxError) as e:
        s.adjoint_backward(P.wx)
    assert e.value.code == E_STATE
    base = np.nonzero(np.abs(P.coords[:, 1]) < 1e-12)[0]
    s.set_friction_constraints(base, np.ones(len(base), dtype=np.int32),
                               -np.ones(len(base), dtype=np.int32), np.full(len(base), 0.3))
    with pytest.raises(MpmxError) as e:
        s.adjoint_forward(5, P.dt)
    assert e.value.code == E_UNSUPPORTED
    s.close()
    s = Solver(2, "usl")
    P.setup(s, P.params())
    s.set_particle_tractions([0], [1], [-10.0])
    with pytest.raises(MpmxError) as e:
        s.adjoint_forward(5, P.dt)
    assert e.value.code == E_UNSUPPORTED
    s.close()
    s = Solver(2, "usl")
    s.set_mesh(P.coords, P.cells)
    s.set_materials(mat.linear_elastic(1e6, 0.3, 1800.0, stress_rate="jaumann")[None, :])
    s.set_particles(P.x0, mass=P.volume * 1800.0, volume=P.volume,
                    density=np.full(P.np, 1800.0), material_id=0)
    with pytest.raises(MpmxError) as e:
        s.adjoint_forward(5, P.dt)
    assert e.value.code == E_UNSUPPORTED
    s.close()
    s = Solver(2, "usl")
    prm = P.params()
    s.set_mesh(P.coords, P.cells)
    m = mat.mohr_coulomb(1e6, 0.3, 1800.0, friction=30.0, dilation=0.0, cohesion=1000.0,
                         residual_friction=20.0, peak_pdstrain=0.0, residual_pdstrain=0.1,
                         softening=True)
    s.set_materials(m[None, :])
    s.set_particles(prm["x0"], mass=P.volume * 1800.0, volume=P.volume,
                    density=np.full(P.np, 1800.0), material_id=0)
    with pytest.raises(MpmxError) as e:
        s.adjoint_forward(5, P.dt)
    assert e.value.code == E_UNSUPPORTED
    s.close()


def test_torch_autograd_function():
    """mpm_amd.adjoint.trajectory: torch.autograd through the C-ABI adjoint
    gives the gradients of mpmx_adjoint_backward for the same loss."""
    from mpm_amd.adjoint import trajectory
    P = Problem(scheme="usl", material="mix", geostatic=True)

    def setup(s, E, rho, phi):
        P.setup(s, dict(E=E, rho=rho, phi=phi, x0=P.x0, v0=P.v0))

    E = torch.tensor(P.E, requires_grad=True)
    rho = torch.tensor(P.rho, requires_grad=True)
    phi = torch.tensor(P.phi, requires_grad=True)
    x, v = trajectory(setup, E, rho, phi, P.nsteps, P.dt, checkpoint_every=7, dim=2,
                      scheme="usl")
    loss = (torch.from_numpy(P.wx) * x).sum() + (torch.from_numpy(P.wv) * v).sum()
    loss.backward()
    assert abs(loss.item() - P.oracle_loss(P.params())) <= 1e-12 * abs(loss.item())
    g, _, _ = _gpu_gradients(P)
    for name, t in (("E", E), ("rho", rho), ("phi", phi)):
        ref = g[name]
        assert np.max(np.abs(t.grad.numpy() - ref)) <= 1e-9 * np.max(np.abs(ref)), name
    # CUDA tensors: final state and gradients stay on the device
    # (mpmx_get_particles_device, mpmx_adjoint_backward_device); another
    # fast-mode run, so equal to the atomic-order rounding
    dev = torch.device("cuda:0")
    Ed, rd, pd = (torch.tensor(a, device=dev, requires_grad=True) for a in (P.E, P.rho, P.phi))
    xd, vd = trajectory(setup, Ed, rd, pd, P.nsteps, P.dt, checkpoint_every=7, dim=2,
                        scheme="usl")
    assert xd.device.type == "cuda" and vd.device.type == "cuda"
    for a, b in ((xd, x), (vd, v)):
        b = b.detach().numpy()
        assert np.max(np.abs(a.detach().cpu().numpy() - b)) <= 1e-12 * np.max(np.abs(b))
    lossd = ((torch.from_numpy(P.wx).to(dev) * xd).sum() +
             (torch.from_numpy(P.wv).to(dev) * vd).sum())
    lossd.backward()
    for td, t in ((Ed, E), (rd, rho), (pd, phi)):
        assert td.grad.device.type == "cuda"
        ref = t.grad.numpy()
        assert np.max(np.abs(td.grad.cpu().numpy() - ref)) <= 1e-9 * np.max(np.abs(ref))


def test_apex_ratio_bounds_long_collapse_gradients():
    """A cohesionless column collapse over 500 steps: the exact derivative
    of the return mapping near the apex (free-surface particles at low
    stress) makes the gradient overflow, the default apex treatment keeps it
    of order one."""
    from mpm_amd.capi import Solver
    from mpm_amd import workloads
    mc = workloads.adjoint_2d(ninst=2, columns=16).instances[1]

    def grad(w, ratio):
        s = w.setup(Solver(w.dim, w.scheme, ordered=True))
        s.adjoint_set_apex_ratio(ratio)
        s.adjoint_forward(500, w.dt, 0.0, 50)
        x = s.particles(["x"])["x"]
        g = s.adjoint_backward(np.ones_like(x))
        s.close()
        return g

    g_exact = grad(mc, 0.0)
    g_reg = grad(mc, 0.3)
    assert not np.all(np.abs(g_exact["x0"]) < 1e6)
    assert np.all(np.isfinite(g_reg["x0"])) and np.max(np.abs(g_reg["x0"])) < 10.0
