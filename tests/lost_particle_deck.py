"""A 2D deck in which one particle leaves the mesh mid-run.

A LinearElastic block sits on a fixed base in cells x in [0, 4), y in [0, 2)
of a 16 x 4 mesh of 1 m cells. Particle 0, alone in cell 8 (between the
block's cells 0-3 and 16-19 in cell order), starts at (8.5, 0.5) moving down
at 1 m/s under gravity and leaves through the bottom of the mesh after a few
tens of steps. The last block particle carries a particle velocity
constraint so that a removal that shifted the indices wrongly would move it.
"""
import numpy as np

from mpm_amd import materials as mat
from mpm_amd.mesh import structured_mesh

DT = 0.01


def setup(solver, pvc=True):
    coords, cells = structured_mesh((0.0, 0.0), (1.0, 1.0), (16, 4))
    solver.set_mesh(coords, cells)
    solver.set_materials([mat.linear_elastic(1e6, 0.3, 1000.0)])
    base = np.nonzero((np.abs(coords[:, 1]) < 1e-9) & (coords[:, 0] < 4.0 + 1e-9))[0]
    nodes = np.concatenate([base, base]).astype(np.int64)
    dirs = np.concatenate([np.zeros(len(base)), np.ones(len(base))]).astype(np.int32)
    solver.set_velocity_constraints(nodes, dirs, np.zeros(len(nodes)))
    solver.set_gravity([0.0, -9.81])
    solver.set_particles(x=[[8.5, 0.5]], mass=[1000.0], volume=[1.0], density=[1000.0],
                         material_id=0, v=[[0.0, -1.0]])
    block = np.array([r * 16 + c for r in range(2) for c in range(4)], dtype=np.int64)
    solver.generate_gauss(2, block, 0)
    n = solver.nparticles
    if pvc:
        solver.set_particle_velocity_constraints([n - 1], [0], [1e-3])
    return solver
