"""Small problems for the adjoint tests: a block of particles on a Cartesian
mesh with one material per particle, so that Young's modulus, density and
friction angle can be perturbed particle by particle, and a linear loss of
the final positions and velocities. The same problem is built on the GPU
solver (mpm_amd.capi.Solver) and on the oracle, whose central finite
differences are the reference for the adjoint's gradients."""
import numpy as np

import oracle
from mpm_amd import materials as mat
from mpm_amd.mesh import structured_mesh


class Problem:
    def __init__(self, dim=2, scheme="usl", material="le", ncells=(6, 10), block=(3, 6),
                 h=0.05, nsteps=40, dt=1e-4, seed=3, damping=0.0, velocity_update="flip",
                 blending=1.0, ramp=0.0, acc_constraint=False, vel_amplitude=0.05,
                 geostatic=False, tension_cutoff=0.0, particle_constraints=False,
                 nodal_forces=False):
        rng = np.random.default_rng(seed)
        self.dim, self.scheme, self.nsteps, self.dt = dim, scheme, nsteps, dt
        self.damping, self.vu, self.blending, self.ramp = damping, velocity_update, blending, ramp
        self.coords, self.cells = structured_mesh((0.0,) * dim, (h,) * dim, ncells)
        nc = np.asarray(ncells)
        ci = np.arange(len(self.cells))
        idx = [(ci // int(np.prod(nc[:k]))) % nc[k] for k in range(dim)]
        inside = np.ones(len(ci), dtype=bool)
        for k in range(dim):
            inside &= idx[k] < block[k]
        self.cell_set = ci[inside]
        # constraints: fixed base, roller at x = 0
        base = np.nonzero(np.abs(self.coords[:, dim - 1]) < 1e-12)[0]
        left = np.nonzero(np.abs(self.coords[:, 0]) < 1e-12)[0]
        vn = [base] * dim + [left]
        vd = [np.full(len(base), k) for k in range(dim)] + [np.zeros(len(left))]
        self.vel = (np.concatenate(vn), np.concatenate(vd).astype(np.int32),
                    np.zeros(sum(len(a) for a in vn)))
        # concentrated forces on the nodes of a row inside the block, ramped by
        # a Linear math function, and a Sin-scaled acceleration constraint
        self.forces = None
        if nodal_forces:
            row = np.nonzero((np.abs(self.coords[:, dim - 1] - 2 * h) < 1e-12)
                             & (self.coords[:, 0] > 1e-12) & (self.coords[:, 0] < h * block[0]))[0]
            self.forces = (row, np.full(len(row), dim - 1, dtype=np.int32),
                           np.full(len(row), -30.0), np.zeros(len(row), dtype=np.int32))
            self.functions = [("linear", [0.0, 10 * dt], [0.0, 1.0]),
                              ("sin", 0.0, np.pi / (20 * dt), (0.0, 1.0))]
        self.acc = None
        if acc_constraint:
            top = nc[dim - 1] * h
            row = np.nonzero((np.abs(self.coords[:, dim - 1] - 0.5 * top) < 1e-12)
                             & (self.coords[:, 0] > 1e-12))[0]
            self.acc = (row, np.zeros(len(row), dtype=np.int32), np.full(len(row), 0.2))
        self.gravity = np.zeros(dim)
        self.gravity[dim - 1] = -9.81
        # particles at the Gauss points (positions, volumes, cells from the oracle)
        o = oracle.Oracle(dim, scheme)
        o.set_mesh(self.coords, self.cells)
        o.set_materials(mat.linear_elastic(1e6, 0.3, 1800.0)[None, :])
        o.generate_gauss(2, self.cell_set, 0)
        p = o.particles()
        self.np = len(p["x"])
        self.x0 = p["x"].copy()
        self.volume = p["volume"].copy()
        self.cell = p["cell"].copy()
        self.v0 = vel_amplitude * rng.uniform(-1, 1, (self.np, dim))
        self.material = material
        self.tension_cutoff = tension_cutoff
        # particle velocity constraints: x velocity of every fifth particle,
        # y velocity of particle 1
        self.pvc = None
        if particle_constraints:
            ids = np.arange(0, self.np, 5)
            self.pvc = (np.concatenate([ids, [1]]),
                        np.concatenate([np.zeros(len(ids), dtype=np.int32), [1]]),
                        np.concatenate([np.full(len(ids), 0.01), [-0.02]]))
        if material == "le":
            kind = np.zeros(self.np, dtype=int)
        elif material == "mc":
            kind = np.ones(self.np, dtype=int)
        else:  # half and half
            kind = np.arange(self.np) % 2
        self.kind = kind
        self.E = 1e6 * (1 + 0.2 * rng.uniform(-1, 1, self.np))
        self.rho = 1800.0 * (1 + 0.1 * rng.uniform(-1, 1, self.np))
        self.phi = 30.0 + 4.0 * rng.uniform(-1, 1, self.np)
        self.stress0 = np.zeros((self.np, 6))
        if geostatic:
            top = block[dim - 1] * h
            depth = top - self.x0[:, dim - 1]
            sv = -self.rho * 9.81 * depth
            k0 = 1 - np.sin(np.radians(self.phi))
            for k in range(3):
                self.stress0[:, k] = sv if k == dim - 1 else k0 * sv
        self.wx = rng.normal(size=(self.np, dim))
        self.wv = rng.normal(size=(self.np, dim)) * 0.1

    def params(self):
        return dict(E=self.E.copy(), rho=self.rho.copy(), phi=self.phi.copy(),
                    x0=self.x0.copy(), v0=self.v0.copy())

    def materials(self, E, rho, phi):
        out = []
        for p in range(self.np):
            if self.kind[p] == 0:
                out.append(mat.linear_elastic(E[p], 0.3, rho[p]))
            else:
                out.append(mat.mohr_coulomb(E[p], 0.3, rho[p], friction=phi[p], dilation=0.0,
                                            cohesion=500.0,
                                            tension_cutoff=self.tension_cutoff))
        return np.array(out)

    def setup(self, obj, prm):
        obj.set_mesh(self.coords, self.cells)
        obj.set_materials(self.materials(prm["E"], prm["rho"], prm["phi"]))
        obj.set_velocity_constraints(*self.vel)
        if self.forces is not None:
            obj.set_math_functions(self.functions)
            obj.set_nodal_forces(*self.forces)
        if self.acc is not None:
            if self.forces is not None:
                obj.set_acceleration_constraints(*self.acc, np.ones(len(self.acc[0]), dtype=np.int32))
            else:
                obj.set_acceleration_constraints(*self.acc)
        obj.set_gravity(self.gravity, self.ramp)
        obj.set_particles(prm["x0"], mass=self.volume * prm["rho"], volume=self.volume,
                          density=prm["rho"], material_id=np.arange(self.np), v=prm["v0"],
                          stress=self.stress0, cell=self.cell)
        if self.pvc is not None:
            obj.set_particle_velocity_constraints(*self.pvc)
        return obj

    def loss_of(self, p):
        return float(np.sum(self.wx * p["x"]) + np.sum(self.wv * p["v"]))

    def oracle_loss(self, prm):
        o = oracle.Oracle(self.dim, self.scheme, velocity_update=self.vu,
                          blending_ratio=self.blending)
        self.setup(o, prm)
        o.step(self.nsteps, self.dt, self.damping)
        return self.loss_of(o.particles(["x", "v"]))

    def fd(self, name, direction, h):
        """Central difference of the oracle loss along `direction` of parameter
        `name` (relative steps for E, rho; absolute for phi, x0, v0)."""
        base = self.params()
        out = []
        for s in (1.0, -1.0):
            prm = self.params()
            if name in ("E", "rho"):
                prm[name] = base[name] * (1 + s * h * direction)
            else:
                prm[name] = base[name] + s * h * direction
            out.append(self.oracle_loss(prm))
        return (out[0] - out[1]) / (2 * h)

    def fd_converged(self, name, direction, h, agree=3e-7, levels=7):
        """The central difference at the largest of h, h/10, h/100, ... whose
        value agrees with the next smaller step's to `agree` relative: a
        step below any kink of the run (yield state, return-mapping
        iteration count, cell) and above rounding noise. Raises if no pair
        agrees."""
        prev = self.fd(name, direction, h)
        for _ in range(levels):
            h /= 10.0
            cur = self.fd(name, direction, h)
            if abs(prev - cur) <= agree * abs(prev):
                return prev
            prev = cur
        raise AssertionError("finite differences of %s did not converge" % name)

    def predicted(self, grads, name, direction):
        """The adjoint's directional derivative matching fd()."""
        g = grads[name]
        if name in ("E", "rho"):
            return float(np.sum(g * self.params()[name] * direction))
        return float(np.sum(g * direction))
