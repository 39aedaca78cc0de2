"""Pins of the oracle's affine velocity updates (APIC, TPIC, ASFLIP).

The reference has no known-answer tests for them, so the checks come from
what the transfers are built to do (particle.tcc:661-724, 1073-1140,
1662-1730):

* APIC and TPIC carry an affine velocity field through a grid-to-particle
  and particle-to-grid round trip exactly: with nodal velocities
  v_i = v0 + L x_i, G2P gives v_p = v0 + L x_p and mapping matrix
  B = L D_p (APIC, D_p = sum_i N_i (x_i - x_p)(x_i - x_p)^T) or B = L (TPIC,
  the velocity gradient), and P2G with v_p + B D_p^-1 (x_i - x_p) (APIC) or
  v_p + B (x_i - x_p) (TPIC) gives back v0 + L x_i at every node with mass;
  plain PIC averages the field and does not (the check discriminates);
* ASFLIP moves the particle with beta * FLIP + (1 - beta) * PIC velocity,
  beta = 0 next to a velocity-constrained node or where the step's
  deformation gradient increment compresses (det < 1), else 1.

A transposed mapping matrix, a missing shape-tensor inverse, a wrong branch
vector sign or a wrong beta rule fails here.
"""
import numpy as np
import pytest

import oracle
from mpm_amd import materials
from mpm_amd.mesh import structured_mesh


def _grid(dim, vu, blend=1.0, n=3, h=0.5):
    coords, cells = structured_mesh((0.0,) * dim, (h,) * dim, (n,) * dim)
    o = oracle.Oracle(dim, "usf", velocity_update=vu, blending_ratio=blend)
    o.set_mesh(coords, cells)
    o.set_materials(materials.linear_elastic(1e6, 0.3, 1000.0)[None, :])
    o.generate_gauss(2, None, 0)
    o.phase("reset_nodes")
    o.phase("shapefn")
    return o, coords, cells


def _set_affine_nodes(o, coords, v0, L):
    for i, x in enumerate(coords):
        o.set_node_field(i, "velocity", v0 + L @ x)
        o.set_node_field(i, "acceleration", np.zeros(len(x)))


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("vu", ["apic", "tpic"])
def test_affine_field_round_trip(dim, vu):
    rng = np.random.default_rng(7 + dim)
    v0 = rng.uniform(-1, 1, dim)
    L = rng.uniform(-2, 2, (dim, dim))
    o, coords, cells = _grid(dim, vu)
    _set_affine_nodes(o, coords, v0, L)
    o.phase("position", dt=0.0)  # G2P without moving the particles
    p = o.particles(["x", "v", "cell"])
    B = o.mapping_matrix()
    np.testing.assert_allclose(p["v"], v0 + p["x"] @ L.T, rtol=0, atol=1e-13)
    for q in range(0, len(p["x"]), 5):
        if vu == "tpic":
            want = L
        else:
            N = o.particle_scratch(q)["N"]
            r = coords[cells[p["cell"][q]]] - p["x"][q]
            want = L @ np.einsum("i,ia,ib->ab", N, r, r)
        np.testing.assert_allclose(B[q], want, rtol=0, atol=1e-13)
    o.phase("reset_nodes")
    o.phase("map_mass_momentum")
    o.phase("compute_velocity")
    nd = o.nodes()
    has = nd["mass"] > 0
    assert has.sum() == len(coords)
    np.testing.assert_allclose(nd["velocity"][has], v0 + coords[has] @ L.T, rtol=0, atol=1e-12)


def test_pic_does_not_keep_the_affine_field():
    o, coords, cells = _grid(2, "pic")
    v0, L = np.array([0.1, -0.2]), np.array([[0.5, 1.0], [-1.5, 0.25]])
    _set_affine_nodes(o, coords, v0, L)
    o.phase("position", dt=0.0)
    o.phase("reset_nodes")
    o.phase("map_mass_momentum")
    o.phase("compute_velocity")
    err = np.max(np.abs(o.nodes()["velocity"] - (v0 + coords @ L.T)))
    assert err > 1e-3


def _one_cell_asflip(L, blend, constrained=False, dt=0.01):
    coords, cells = structured_mesh((0.0, 0.0), (1.0, 1.0), (1, 1))
    o = oracle.Oracle(2, "usf", velocity_update="asflip", blending_ratio=blend)
    o.set_mesh(coords, cells)
    o.set_materials(materials.linear_elastic(1e6, 0.3, 1000.0)[None, :])
    if constrained:
        o.set_velocity_constraints([0], [1], [0.0])
    o.set_particles(x=[[0.3, 0.6]], mass=[1.0], volume=[1.0], density=[1.0], material_id=0,
                    v=[[0.7, -0.4]], cell=[0])
    o.phase("reset_nodes")
    o.phase("shapefn")
    acc = np.array([[1.0, 2.0], [-3.0, 0.5], [0.25, -1.0], [2.0, 2.0]])
    for i, x in enumerate(coords):
        o.set_node_field(i, "velocity", np.array([0.05, -0.02]) + L @ x)
        o.set_node_field(i, "acceleration", acc[i])
    N = o.particle_scratch(0)["N"]  # in the cell's local node order
    loc = cells[0]
    nv = N @ (np.array([0.05, -0.02]) + coords[loc] @ L.T)
    na = N @ acc[loc]
    x0 = np.array([0.3, 0.6])
    o.phase("position", dt=dt)
    p = o.particles(["x", "v", "u"])
    flip = np.array([0.7, -0.4]) + na * dt
    return p, x0, flip, nv, dt


@pytest.mark.parametrize("blend", [1.0, 0.7])
def test_asflip_beta_rules(blend):
    expand, compress = 0.5 * np.eye(2), -0.5 * np.eye(2)
    # away from constraints, expanding: beta = 1
    p, x0, flip, nv, dt = _one_cell_asflip(expand, blend)
    sep = blend * flip + (1 - blend) * nv
    np.testing.assert_allclose(p["v"][0], blend * flip + (1 - blend) * nv, rtol=0, atol=1e-15)
    np.testing.assert_allclose(p["x"][0], x0 + sep * dt, rtol=0, atol=1e-15)
    np.testing.assert_allclose(p["u"][0], sep * dt, rtol=0, atol=1e-15)
    # compressing: beta = 0, the particle moves with the grid velocity
    p, x0, flip, nv, dt = _one_cell_asflip(compress, blend)
    np.testing.assert_allclose(p["x"][0], x0 + nv * dt, rtol=0, atol=1e-15)
    np.testing.assert_allclose(p["v"][0], blend * flip + (1 - blend) * nv, rtol=0, atol=1e-15)
    # a velocity-constrained node of the cell: beta = 0
    p, x0, flip, nv, dt = _one_cell_asflip(expand, blend, constrained=True)
    np.testing.assert_allclose(p["x"][0], x0 + nv * dt, rtol=0, atol=1e-15)
