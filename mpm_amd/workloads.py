"""Synthetic input decks of the benchmark configurations.

Each builder returns a ``Workload``: the mesh, materials, constraints,
gravity, particle generation and run parameters of one configuration. The
same object drives the GPU ``Solver`` and any object with the same setup
methods, so every consumer starts from identical state.

* ``column_2d``: plain 2D elastic column under gravity (quad4 32x64 cells of
  0.05 m, 4 particles per cell, LinearElastic E=1e6 Pa, nu=0.3, rho=1800,
  fixed bottom, roller sides, USF, dt=1e-4).
* ``collapse_3d``: 3D granular column collapse (hex8 128x64x64 cells of
  0.01 m, 8 particles per cell in a 0.64 m cube block = 2,097,152
  particles, MohrCoulomb phi=30, psi=0, c=0, E=5e6, nu=0.3, rho=1800,
  frictional base mu=0.3, roller walls, MUSL, dt=5e-5, initial velocities
  uniform in +-1e-4 m/s from seed 1). The block starts from a K0 = 1 - sin(phi)
  geostatic stress: with c = 0 the reference's return mapping divides by
  dP/dsigma = 0 at zero stress (mohr_coulomb.tcc:453-463) and returns NaN, so
  a stress-free cohesionless start cannot run in the reference either.

Sizes are parameters so tests run the same decks at small scale.
"""
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import materials as mat
from .mesh import structured_mesh

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(seed, n):
    """n 64-bit outputs of SplitMix64 (Steele, Lea & Flood 2014) from `seed`."""
    with np.errstate(over="ignore"):
        z = (np.uint64(seed) + np.arange(1, n + 1, dtype=np.uint64) *
             np.uint64(0x9E3779B97F4A7C15))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed, n, lo, hi):
    """n doubles uniform in [lo, hi) from the top 53 bits of SplitMix64."""
    u = (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)
    return lo + (hi - lo) * u


@dataclass
class Workload:
    name: str
    dim: int
    scheme: str
    coords: np.ndarray
    cells: np.ndarray
    materials: np.ndarray
    gravity: np.ndarray
    dt: float
    nsteps: int
    nppd: int = 2
    cell_set: Optional[np.ndarray] = None
    vel_nodes: np.ndarray = field(default_factory=lambda: np.zeros(0, dtype=np.int64))
    vel_dirs: np.ndarray = field(default_factory=lambda: np.zeros(0, dtype=np.int32))
    vel_values: np.ndarray = field(default_factory=lambda: np.zeros(0))
    fric_nodes: np.ndarray = field(default_factory=lambda: np.zeros(0, dtype=np.int64))
    fric_dirs: np.ndarray = field(default_factory=lambda: np.zeros(0, dtype=np.int32))
    fric_sign: np.ndarray = field(default_factory=lambda: np.zeros(0, dtype=np.int32))
    fric_mu: np.ndarray = field(default_factory=lambda: np.zeros(0))
    velocity_seed: Optional[int] = None
    velocity_amplitude: float = 0.0
    # K0 geostatic initial stress: (K0, unit weight rho*|g|, free-surface
    # height: a number, or a function of the particle positions)
    geostatic: Optional[tuple] = None
    damping: float = 0.0
    info: dict = field(default_factory=dict)
    # Steps the deck can run before a particle leaves the mesh (None: all
    # nsteps). The reference would stop there too (mpm_scheme.tcc:250-256).
    runnable_steps: Optional[int] = None
    # the mesh's "isoparametric" key: locate by the inverse element map
    isoparametric: bool = False

    def setup(self, solver):
        """Apply this deck to a solver (mesh, materials, constraints, particles)."""
        solver.set_mesh(self.coords, self.cells, isoparametric=self.isoparametric)
        solver.set_materials(self.materials)
        if len(self.vel_nodes):
            solver.set_velocity_constraints(self.vel_nodes, self.vel_dirs, self.vel_values)
        if len(self.fric_nodes):
            solver.set_friction_constraints(self.fric_nodes, self.fric_dirs, self.fric_sign,
                                            self.fric_mu)
        solver.set_gravity(self.gravity)
        solver.generate_gauss(self.nppd, self.cell_set, 0)
        if self.geostatic is not None:
            k0, gamma, top = self.geostatic
            x = solver.particles(["x"])["x"]
            z = x[:, self.dim - 1]
            if callable(top):  # free-surface height as a function of the positions
                top = top(x)
            sv = -gamma * (top - z)
            stress = np.zeros((len(z), 6))
            stress[:, self.dim - 1] = sv
            for k in range(self.dim - 1):
                stress[:, k] = k0 * sv
            if self.dim == 2:
                stress[:, 2] = k0 * sv
            solver.assign_stresses(stress)
        if self.velocity_seed is not None:
            n = solver.nparticles
            a = self.velocity_amplitude
            solver.assign_velocity(uniform(self.velocity_seed, n * self.dim, -a, a)
                                   .reshape(n, self.dim))
        return solver

    @property
    def nparticles(self):
        ncell = len(self.cells) if self.cell_set is None else len(self.cell_set)
        return ncell * self.nppd ** self.dim


def _node_sets(coords, axis, value):
    return np.nonzero(np.abs(coords[:, axis] - value) <= 1e-9)[0].astype(np.int64)


def column_2d(nx=32, ny=64, h=0.05, nsteps=1000, dt=1e-4):
    coords, cells = structured_mesh((0.0, 0.0), (h, h), (nx, ny))
    bottom = _node_sets(coords, 1, 0.0)
    left = _node_sets(coords, 0, 0.0)
    right = _node_sets(coords, 0, nx * h)
    vn = np.concatenate([bottom, bottom, left, right])
    vd = np.concatenate([np.zeros(len(bottom)), np.ones(len(bottom)),
                         np.zeros(len(left)), np.zeros(len(right))]).astype(np.int32)
    return Workload(
        name="column_2d", dim=2, scheme="usf", coords=coords, cells=cells,
        materials=mat.linear_elastic(1e6, 0.3, 1800.0)[None, :],
        gravity=np.array([0.0, -9.81]), dt=dt, nsteps=nsteps, nppd=2,
        vel_nodes=vn, vel_dirs=vd, vel_values=np.zeros(len(vn)),
        info=dict(mesh="quad4 %dx%d @ %g m" % (nx, ny, h), material="LinearElastic",
                  particles_per_cell=4))


def collapse_3d(nx=128, ny=64, nz=64, h=0.01, block=(64, 64, 64), nsteps=2000, dt=5e-5,
                seed=1):
    coords, cells = structured_mesh((0.0, 0.0, 0.0), (h, h, h), (nx, ny, nz))
    ci = np.arange(len(cells), dtype=np.int64)
    i = ci % nx
    j = (ci // nx) % ny
    k = ci // (nx * ny)
    cset = ci[(i < block[0]) & (j < block[1]) & (k < block[2])]
    base = _node_sets(coords, 2, 0.0)
    x0, x1 = _node_sets(coords, 0, 0.0), _node_sets(coords, 0, nx * h)
    y0, y1 = _node_sets(coords, 1, 0.0), _node_sets(coords, 1, ny * h)
    vn = np.concatenate([base, x0, x1, y0, y1])
    vd = np.concatenate([np.full(len(base), 2), np.zeros(len(x0)), np.zeros(len(x1)),
                         np.ones(len(y0)), np.ones(len(y1))]).astype(np.int32)
    return Workload(
        name="collapse_3d", dim=3, scheme="musl", coords=coords, cells=cells,
        materials=mat.mohr_coulomb(5e6, 0.3, 1800.0, friction=30.0, dilation=0.0,
                                   cohesion=0.0)[None, :],
        gravity=np.array([0.0, 0.0, -9.81]), dt=dt, nsteps=nsteps, nppd=2, cell_set=cset,
        vel_nodes=vn, vel_dirs=vd, vel_values=np.zeros(len(vn)),
        fric_nodes=base, fric_dirs=np.full(len(base), 2, dtype=np.int32),
        fric_sign=np.full(len(base), -1, dtype=np.int32), fric_mu=np.full(len(base), 0.3),
        velocity_seed=seed, velocity_amplitude=1e-4,
        geostatic=(1.0 - np.sin(np.radians(30.0)), 1800.0 * 9.81, block[2] * h),
        # Undamped FLIP/MUSL: the toe's velocities grow from ~0.25 m/s at step
        # 200 to ~2.4 m/s at step 800 and without bound after ~900; a particle
        # leaves the mesh near step 1000, in ordered mode (the reference's
        # arithmetic, bitwise) as in fast mode.
        runnable_steps=900 if nsteps >= 900 else None,
        info=dict(mesh="hex8 %dx%dx%d @ %g m" % (nx, ny, nz, h),
                  block_cells=list(block), material="MohrCoulomb phi=30 psi=0 c=0",
                  particles_per_cell=8,
                  initial_stress="K0 geostatic: sigma_zz = -rho g (H - z), "
                                 "sigma_xx = sigma_yy = (1 - sin phi) sigma_zz",
                  initial_velocity="uniform in +-1e-4 m/s per component, SplitMix64 seed %d"
                                   % seed))


def collapse_2d(width, height, h, youngs, friction, density, dt=1e-5, nsteps=1500,
                name="collapse_2d"):
    """2D granular column collapse: a block of width x height quad4 cells of
    size h (4 particles per cell) on a mesh 2*width+2 cells wide and height+2
    high, MohrCoulomb (nu=0.3, psi=0, c=0) from a K0 = 1 - sin(phi) start,
    fixed base, roller side walls, USL."""
    nx, ny = 2 * width + 2, height + 2
    coords, cells = structured_mesh((0.0, 0.0), (h, h), (nx, ny))
    ci = np.arange(len(cells), dtype=np.int64)
    cset = ci[(ci % nx < width) & (ci // nx < height)]
    base = _node_sets(coords, 1, 0.0)
    x0, x1 = _node_sets(coords, 0, 0.0), _node_sets(coords, 0, nx * h)
    vn = np.concatenate([base, base, x0, x1])
    vd = np.concatenate([np.zeros(len(base)), np.ones(len(base)), np.zeros(len(x0)),
                         np.zeros(len(x1))]).astype(np.int32)
    return Workload(
        name=name, dim=2, scheme="usl", coords=coords, cells=cells,
        materials=mat.mohr_coulomb(youngs, 0.3, density, friction=friction, dilation=0.0,
                                   cohesion=0.0)[None, :],
        gravity=np.array([0.0, -9.81]), dt=dt, nsteps=nsteps, nppd=2, cell_set=cset,
        vel_nodes=vn, vel_dirs=vd, vel_values=np.zeros(len(vn)),
        geostatic=(1.0 - np.sin(np.radians(friction)), density * 9.81, height * h),
        info=dict(mesh="quad4 %dx%d @ %g m" % (nx, ny, h), block_cells=[width, height]))


class Batch:
    """Independent instances run as one batched solver (mpmx_set_instances):
    the union of their meshes with node and cell ids offset per instance,
    material i for instance i, each instance's constraints and K0 start."""

    def __init__(self, name, instances):
        self.name = name
        self.instances = instances
        w0 = instances[0]
        self.dim, self.scheme, self.dt = w0.dim, w0.scheme, w0.dt
        self.nsteps, self.runnable_steps = w0.nsteps, None
        self.gravity = w0.gravity
        for w in instances:
            assert (w.dim, w.scheme, w.dt) == (self.dim, self.scheme, self.dt)
        node_off = np.cumsum([0] + [len(w.coords) for w in instances])
        cell_off = np.cumsum([0] + [len(w.cells) for w in instances])
        self.node_begin, self.cell_begin = node_off, cell_off
        self.coords = np.concatenate([w.coords for w in instances])
        self.cells = np.concatenate([w.cells + node_off[i] for i, w in enumerate(instances)])
        self.materials = np.concatenate([w.materials for w in instances])
        self.info = dict(instances=len(instances), mesh="union of quad4 meshes @ %g m" %
                         (instances[0].coords[1, 0] - instances[0].coords[0, 0]),
                         material="MohrCoulomb per instance", particles_per_cell=4)

    @property
    def nparticles(self):
        return sum(w.nparticles for w in self.instances)

    def setup(self, solver):
        solver.set_mesh(self.coords, self.cells)
        solver.set_instances(self.cell_begin)
        solver.set_materials(self.materials)
        nodes, dirs, vals = [], [], []
        for i, w in enumerate(self.instances):
            nodes.append(w.vel_nodes + self.node_begin[i])
            dirs.append(w.vel_dirs)
            vals.append(w.vel_values)
        solver.set_velocity_constraints(np.concatenate(nodes), np.concatenate(dirs),
                                        np.concatenate(vals))
        solver.set_gravity(self.gravity)
        for i, w in enumerate(self.instances):
            solver.generate_gauss(w.nppd, w.cell_set + self.cell_begin[i], i)
        x = solver.particles(["x"])["x"]
        stress = np.zeros((len(x), 6))
        off = 0
        for w in self.instances:
            n = w.nparticles
            k0, gamma, top = w.geostatic
            sv = -gamma * (top - x[off:off + n, self.dim - 1])
            stress[off:off + n, self.dim - 1] = sv
            for k in range(self.dim - 1):
                stress[off:off + n, k] = k0 * sv
            if self.dim == 2:
                stress[off:off + n, 2] = k0 * sv
            off += n
        solver.assign_stresses(stress)
        return solver


def batch_2d(ninst=256, seed=4, h=0.01, nmin=2000, nmax=64000, dt=1e-5, nsteps=1500):
    """Batched 2D collapses: per instance a particle count log-uniform in
    [nmin, nmax] (4 per cell), block aspect ratio (height/width) uniform in
    [0.5, 4], E log-uniform in [1e5, 1e7] Pa, phi uniform in [25, 40] deg, rho
    uniform in [1500, 2200] kg/m^3; all draws from SplitMix64(seed)."""
    u = uniform(seed, 5 * ninst, 0.0, 1.0).reshape(ninst, 5)
    inst = []
    for i in range(ninst):
        npart = np.exp(np.log(nmin) + u[i, 0] * (np.log(nmax) - np.log(nmin)))
        aspect = 0.5 + 3.5 * u[i, 1]
        ncell = max(1, int(round(npart / 4)))
        width = max(1, int(round(np.sqrt(ncell / aspect))))
        height = max(1, int(round(ncell / width)))
        youngs = np.exp(np.log(1e5) + u[i, 2] * (np.log(1e7) - np.log(1e5)))
        phi = 25.0 + 15.0 * u[i, 3]
        rho = 1500.0 + 700.0 * u[i, 4]
        inst.append(collapse_2d(width, height, h, youngs, phi, rho, dt=dt, nsteps=nsteps,
                                name="batch_2d[%d]" % i))
    return Batch("batch_2d", inst)


def adjoint_2d(ninst=64, seed=5, h=0.01, columns=64, dt=1e-5, nsteps=500, checkpoint_every=50,
               perturb=0.1):
    """Adjoint inversion batch: `ninst` 2D column collapses, each a column of
    columns x columns quad4 cells (4 particles per cell: 16384 particles for
    64) on a mesh twice as wide (runout space) plus walls, USL, dt=1e-5,
    500 steps recorded with a checkpoint every 50. Instances alternate
    LinearElastic (even) and MohrCoulomb (odd, phi uniform in [28, 38] deg,
    c=0, psi=0); E log-uniform in [5e5, 5e6] Pa and rho uniform in
    [1600, 2000] kg/m^3 per instance, from SplitMix64(seed). The target
    trajectory comes from the same batch with every instance's E, rho and
    phi scaled by factors uniform in 1 +- `perturb` (SplitMix64(seed + 1))."""
    u = uniform(seed, 3 * ninst, 0.0, 1.0).reshape(ninst, 3)
    inst = []
    for i in range(ninst):
        youngs = np.exp(np.log(5e5) + u[i, 0] * (np.log(5e6) - np.log(5e5)))
        rho = 1600.0 + 400.0 * u[i, 1]
        phi = 28.0 + 10.0 * u[i, 2]
        w = collapse_2d(columns, columns, h, youngs, phi, rho, dt=dt, nsteps=nsteps,
                        name="adjoint_2d[%d]" % i)
        if i % 2 == 0:
            nu = 0.3
            w.materials = mat.linear_elastic(youngs, nu, rho)[None, :]
            w.geostatic = (nu / (1.0 - nu), rho * 9.81, columns * h)
        inst.append(w)
    b = Batch("adjoint_2d", inst)
    b.nsteps, b.checkpoint_every = nsteps, checkpoint_every
    f = uniform(seed + 1, 3 * ninst, 1.0 - perturb, 1.0 + perturb).reshape(ninst, 3)
    target = b.materials.copy()
    target[:, 2] *= f[:, 0]          # E
    target[:, 1] *= f[:, 1]          # rho
    target[1::2, 4] *= f[1::2, 2]    # phi of the MohrCoulomb instances
    b.target_materials = target
    b.info = dict(instances=ninst, mesh="union of quad4 %dx%d meshes @ %g m" %
                  (2 * columns + 2, columns + 2, h), particles_per_cell=4,
                  material="LinearElastic (even) / MohrCoulomb (odd) per instance",
                  nsteps=nsteps, checkpoint_every=checkpoint_every)
    return b


def wedge_3d(nx=256, ny=128, nz=128, h=0.02, crest=1.28, nsteps=1000, dt=2e-5):
    """Large 3D slope runout: hex8 nx x ny x nz cells of h (5.12 x 2.56 x 2.56
    m for the defaults), 8 particles per cell in the cells whose centroid lies
    under a surface that is flat at the mesh top up to x = crest and falls at
    45 degrees beyond (half of the cells: 2,097,152 cells, 16,777,216
    particles), MohrCoulomb phi=35 psi=0 c=2 kPa E=2e7 nu=0.3 rho=2000, MUSL,
    dt=2e-5; fixed base, roller side walls, K0 = 1 - sin(phi) start under the
    local surface height."""
    coords, cells = structured_mesh((0.0, 0.0, 0.0), (h, h, h), (nx, ny, nz))
    H = nz * h

    def surface(x):
        return np.minimum(H, H - (x[:, 0] - crest))

    ci = np.arange(len(cells), dtype=np.int64)
    xc = np.stack([(ci % nx + 0.5) * h, ((ci // nx) % ny + 0.5) * h, (ci // (nx * ny) + 0.5) * h],
                  axis=1)
    cset = ci[xc[:, 2] <= surface(xc)]
    base = _node_sets(coords, 2, 0.0)
    x0, x1 = _node_sets(coords, 0, 0.0), _node_sets(coords, 0, nx * h)
    y0, y1 = _node_sets(coords, 1, 0.0), _node_sets(coords, 1, ny * h)
    vn = np.concatenate([base, base, base, x0, x1, y0, y1])
    vd = np.concatenate([np.zeros(len(base)), np.ones(len(base)), np.full(len(base), 2),
                         np.zeros(len(x0)), np.zeros(len(x1)), np.ones(len(y0)),
                         np.ones(len(y1))]).astype(np.int32)
    phi = 35.0
    return Workload(
        name="wedge_3d", dim=3, scheme="musl", coords=coords, cells=cells,
        materials=mat.mohr_coulomb(2e7, 0.3, 2000.0, friction=phi, dilation=0.0,
                                   cohesion=2000.0)[None, :],
        gravity=np.array([0.0, 0.0, -9.81]), dt=dt, nsteps=nsteps, nppd=2, cell_set=cset,
        vel_nodes=vn, vel_dirs=vd, vel_values=np.zeros(len(vn)),
        geostatic=(1.0 - np.sin(np.radians(phi)), 2000.0 * 9.81, surface),
        info=dict(mesh="hex8 %dx%dx%d @ %g m" % (nx, ny, nz, h),
                  material="MohrCoulomb phi=35 psi=0 c=2 kPa E=2e7",
                  particles_per_cell=8, slope="45 deg from x = %g m" % crest))
