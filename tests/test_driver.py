"""The C++ analysis driver (mpm_amd/mpmx_run) reading the reference's input
deck. On the CPU: --dry-run reads and checks the deck and prints what it
found, compared here with the deck's own content."""
import json
import os

import numpy as np
import pytest

from decks import DRIVER, run_driver, write_deck
from mpm_amd import materials as mat
from mpm_amd import workloads

pytestmark = pytest.mark.skipif(not os.path.exists(DRIVER), reason="driver not built")

LE = {"id": 3, "type": "LinearElastic2D", "density": 1800.0, "youngs_modulus": 1e6,
      "poisson_ratio": 0.3}
MC3 = {"id": 3, "type": "MohrCoulomb3D", "density": 1800.0, "youngs_modulus": 5e6,
       "poisson_ratio": 0.3, "friction": 30.0, "dilation": 0.0, "cohesion": 0.0,
       "residual_friction": 30.0, "residual_dilation": 0.0, "residual_cohesion": 0.0,
       "peak_pdstrain": 0.0, "residual_pdstrain": 0.0, "tension_cutoff": 0.0,
       "softening": False}


@pytest.mark.parametrize("constraints_as", ["sets", "file"])
def test_dry_run_reads_the_deck(tmp_path, constraints_as):
    w = workloads.column_2d(nx=6, ny=10)
    write_deck(w, str(tmp_path), nsteps=20, output_steps=10, constraints_as=constraints_as,
               material_json=LE)
    s = json.loads(run_driver(str(tmp_path), "--dry-run").stdout)
    assert s["analysis"] == "MPMExplicit2D" and s["scheme"] == "usf"
    assert s["velocity_update"] == "flip"
    assert (s["nnodes"], s["ncells"]) == (len(w.coords), len(w.cells))
    assert s["velocity_constraints"] == len(w.vel_nodes)
    assert s["particles"] == len(w.cells) * 4
    assert s["dt"] == w.dt and s["nsteps"] == 20 and s["output_steps"] == 10
    assert s["gravity"][:2] == list(w.gravity)
    np.testing.assert_array_equal(s["materials"][0], mat.from_json(LE, 2))
    assert s["particle_velocity_constraints"] == 0


def test_dry_run_particle_velocity_constraints(tmp_path):
    w = workloads.column_2d(nx=6, ny=10)
    write_deck(w, str(tmp_path), nsteps=5, output_steps=5, material_json=LE,
               particle_constraints=[(-1, 0, 0.0), ([1, 2, 3], 1, 0.5)])
    s = json.loads(run_driver(str(tmp_path), "--dry-run").stdout)
    assert s["particle_velocity_constraints"] == 4


def test_dry_run_acceleration_constraints_and_math_functions(tmp_path):
    w = workloads.column_2d(nx=6, ny=10)
    fns = [{"id": 5, "type": "Linear", "xvalues": [0.0, 1e-3], "fxvalues": [0.0, 1.0]},
           {"id": 7, "type": "Linear", "file": "ramp.csv", "xvalues": [0.0, 2e-3, 4e-3],
            "fxvalues": [1.0, 0.5, 0.0]},
           {"id": 9, "type": "Sin", "x0": 0.0, "a": 100.0, "xrange": [0.0, 1.0]}]
    write_deck(w, str(tmp_path), nsteps=5, output_steps=5, material_json=LE,
               math_functions=fns,
               acceleration_constraints=[([40, 41], 0, 0.1, 5), ([42], 0, 0.2, None),
                                         ([43, 44, 45], 0, -0.1, 9)])
    s = json.loads(run_driver(str(tmp_path), "--dry-run").stdout)
    assert s["acceleration_constraints"] == 6 and s["math_functions"] == 3
    deck = json.load(open(os.path.join(str(tmp_path), "mpm.json")))
    deck["mesh"]["boundary_conditions"]["acceleration_constraints"][0]["math_function_id"] = 6
    json.dump(deck, open(os.path.join(str(tmp_path), "mpm.json"), "w"))
    r = run_driver(str(tmp_path), "--dry-run", check=False)
    assert r.returncode != 0 and "math function" in r.stderr
    d2 = os.path.join(str(tmp_path), "forces")
    write_deck(w, d2, nsteps=5, output_steps=5, material_json=LE, math_functions=fns[:1],
               nodal_forces=[([50, 51, 52], 1, -2.0, 5), (-1, 0, 0.5, None)])
    s = json.loads(run_driver(d2, "--dry-run").stdout)
    assert s["nodal_forces"] == 4 and s["math_functions"] == 1


def test_dry_run_mohr_coulomb_3d_with_friction(tmp_path):
    w = workloads.collapse_3d(nx=8, ny=4, nz=4, block=(4, 4, 4))
    write_deck(w, str(tmp_path), nsteps=5, output_steps=5, material_json=MC3)
    s = json.loads(run_driver(str(tmp_path), "--dry-run").stdout)
    assert s["analysis"] == "MPMExplicit3D" and s["scheme"] == "musl"
    assert s["friction_constraints"] == len(w.fric_nodes)
    assert s["particles"] == len(w.cell_set) * 8
    np.testing.assert_array_equal(s["materials"][0], mat.from_json(MC3, 3))


def test_dry_run_nodal_euler_angles(tmp_path):
    """The mesh's "nodal_euler_angles" file (mpm_base.tcc:1041-1063): one
    row per node, a node's first row counting (io_mesh_ascii.tcc:351-391)."""
    w = workloads.column_2d(nx=4, ny=4)
    base = np.nonzero(np.abs(w.coords[:, 1]) <= 1e-9)[0]
    ang = np.tile(np.radians([15.0, 10.0]), (len(base), 1))
    write_deck(w, str(tmp_path), nsteps=1, output_steps=1, material_json=LE,
               euler_angles=(base, ang))
    with open(os.path.join(str(tmp_path), "euler-angles.txt"), "a") as f:
        f.write("%d 0.5 0.5\n" % base[0])  # a second row for a node: ignored
    s = json.loads(run_driver(str(tmp_path), "--dry-run").stdout)
    assert s["nodal_euler_angles"] == len(base)


def test_deck_errors_are_reported(tmp_path):
    w = workloads.column_2d(nx=2, ny=2)
    bad = dict(LE)
    del bad["youngs_modulus"]
    write_deck(w, str(tmp_path), nsteps=1, output_steps=1, material_json=bad)
    r = run_driver(str(tmp_path), "--dry-run", check=False)
    assert r.returncode == 1 and "youngs_modulus" in r.stderr
    with open(os.path.join(str(tmp_path), "mpm.json"), "w") as f:
        f.write('{"analysis": {"type": "MPMExplicit2D", "dt": 1e-3,}')
    r = run_driver(str(tmp_path), "--dry-run", check=False)
    assert r.returncode == 1 and "JSON parse error" in r.stderr


def test_dry_run_stress_rate(tmp_path):
    """"stress_rate" (linear_elastic.tcc:8-17, inherited by MohrCoulomb):
    "jaumann" and "green_naghdi" map to 1 and 2, in parameter [4] for
    LinearElastic and [16] for MohrCoulomb; any other name is no rate."""
    import json
    w = workloads.column_2d(nx=2, ny=2)
    mc = dict(LE, type="MohrCoulomb2D", friction=30.0, dilation=0.0, cohesion=1000.0,
              residual_friction=30.0, residual_dilation=0.0, residual_cohesion=1000.0,
              peak_pdstrain=0.0, residual_pdstrain=0.0, tension_cutoff=0.0,
              softening=False)
    for m, idx in ((LE, 4), (mc, 16)):
        for rate, want in (("jaumann", 1.0), ("green_naghdi", 2.0), ("other", 0.0)):
            write_deck(w, str(tmp_path), nsteps=1, output_steps=1,
                       material_json=dict(m, stress_rate=rate))
            r = run_driver(str(tmp_path), "--dry-run", check=False)
            assert r.returncode == 0, r.stderr
            params = json.loads(r.stdout.strip().splitlines()[-1])["materials"][0]
            assert len(params) == 17 and params[idx] == want, (m["type"], rate, params)
