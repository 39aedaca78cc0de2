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

The GPU solver is the only path: there is no CPU fallback. With CUDA
tensors the final state and the gradients stay on the device
(mpmx_get_particles_device, mpmx_adjoint_backward_device); the parameters
go through the host once, to build the material table of the run. torch
brings its own HIP runtime: import torch (this module does) before the
first mpm_amd.capi.Solver of the process, so that the library shares it.
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
        ctx.solver = solver
        ctx.device = E.device
        if E.device.type == "cuda":
            n, d = solver.nparticles, solver.dim
            x = torch.empty((n, d), dtype=torch.float64, device=E.device)
            v = torch.empty((n, d), dtype=torch.float64, device=E.device)
            torch.cuda.synchronize(E.device)
            solver.particles_device(x.data_ptr(), v.data_ptr())
            return x, v
        p = solver.particles(["x", "v"])
        return torch.from_numpy(p["x"]), torch.from_numpy(p["v"])

    @staticmethod
    def backward(ctx, gx, gv):
        solver = ctx.solver
        dev = ctx.device
        if dev.type == "cuda":
            n = solver.nparticles
            gx = None if gx is None else gx.detach().to(dev, torch.float64).contiguous()
            gv = None if gv is None else gv.detach().to(dev, torch.float64).contiguous()
            out = [torch.empty(n, dtype=torch.float64, device=dev) for _ in range(3)]
            torch.cuda.synchronize(dev)  # the loss gradients are written on torch's stream
            solver.adjoint_backward_device(
                None if gx is None else gx.data_ptr(), None if gv is None else gv.data_ptr(),
                dict(E=out[0].data_ptr(), rho=out[1].data_ptr(), phi=out[2].data_ptr()))
            solver.close()
            ctx.solver = None
            return out[0], out[1], out[2], None, None, None, None, None, None
        gx = None if gx is None else gx.detach().to("cpu", torch.float64).numpy()
        gv = None if gv is None else gv.detach().to("cpu", torch.float64).numpy()
        g = solver.adjoint_backward(gx, gv)
        solver.close()
        ctx.solver = None
        out = [torch.from_numpy(g[k]) for k in ("E", "rho", "phi")]
        return out[0], out[1], out[2], None, None, None, None, None, None


def trajectory(setup, E, rho, phi, nsteps, dt, damping_factor=0.0, checkpoint_every=50,
               **solver_kw):
    """Final (x, v) of an explicit run, differentiable in E, rho, phi.

    solver_kw: arguments of mpm_amd.capi.Solver (dim, scheme, velocity_update,
    blending_ratio, ordered)."""
    return _Trajectory.apply(E, rho, phi, setup, int(nsteps), float(dt), float(damping_factor),
                             int(checkpoint_every), dict(solver_kw))

