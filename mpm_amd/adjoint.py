"""torch.autograd binding of the adjoint of the explicit step.

``trajectory(setup, E, rho, phi, nsteps, dt, ...)`` runs an explicit MPM
simulation whose particles have the given per-particle Young's modulus,
density and friction angle (degrees) and returns the final positions and
velocities as float64 tensors; its backward is mpmx_adjoint_backward, so
any scalar loss of (x_T, v_T) can be differentiated with respect to E, rho
and phi by torch.autograd. ``setup(solver, E, rho, phi)`` configures a fresh
mpm_amd.capi.Solver (mesh, constraints, gravity, materials, particles) from
numpy arrays of the parameters; the run itself is mpmx_adjoint_forward,
with checkpoints every ``checkpoint_every`` steps.

The GPU solver is the only path: there is no CPU fallback.
"""
import torch

from .capi import Solver


class _Trajectory(torch.autograd.Function):
    @staticmethod
    def forward(ctx, E, rho, phi, setup, nsteps, dt, damping, checkpoint_every, solver_kw):
        arrs = [t.detach().to("cpu", torch.float64).numpy() for t in (E, rho, phi)]
        solver = Solver(**solver_kw)
        setup(solver, *arrs)
        solver.adjoint_forward(nsteps, dt, damping, checkpoint_every)
        p = solver.particles(["x", "v"])
        ctx.solver = solver
        ctx.device = E.device
        x = torch.from_numpy(p["x"]).to(E.device)
        v = torch.from_numpy(p["v"]).to(E.device)
        return x, v

    @staticmethod
    def backward(ctx, gx, gv):
        solver = ctx.solver
        gx = None if gx is None else gx.detach().to("cpu", torch.float64).numpy()
        gv = None if gv is None else gv.detach().to("cpu", torch.float64).numpy()
        g = solver.adjoint_backward(gx, gv)
        solver.close()
        ctx.solver = None
        dev = ctx.device
        out = [torch.from_numpy(g[k]).to(dev) for k in ("E", "rho", "phi")]
        return out[0], out[1], out[2], None, None, None, None, None, None


def trajectory(setup, E, rho, phi, nsteps, dt, damping_factor=0.0, checkpoint_every=50,
               **solver_kw):
    """Final (x, v) of an explicit run, differentiable in E, rho, phi.

    solver_kw: arguments of mpm_amd.capi.Solver (dim, scheme, velocity_update,
    blending_ratio, ordered)."""
    return _Trajectory.apply(E, rho, phi, setup, int(nsteps), float(dt), float(damping_factor),
                             int(checkpoint_every), dict(solver_kw))

