"""Write a workload as the reference's input deck: mpm.json plus the ASCII
mesh, entity-set, constraint and particle-stress files it names (formats of
include/io/io_mesh_ascii.tcc and src/io/io.cc)."""
import json
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "mpm_amd", "mpmx_run")


def write_deck(w, d, nsteps, output_steps, constraints_as="sets", uuid="deck",
               material_json=None, stresses=None, particle_constraints=None,
               acceleration_constraints=None, math_functions=None, nodal_forces=None,
               tractions=None, velocity_update=False, euler_angles=None):
    """w: a Workload; d: directory. Velocity constraints go to node sets
    (``constraints_as="sets"``) or to a constraints file (``"file"``).
    particle_constraints: [(particle ids or -1, dir, velocity)], written as
    particle sets and particles_velocity_constraints.
    acceleration_constraints: [(node ids, dir, acceleration, math function id
    or None)], written as node sets; math_functions: the deck's
    "math_functions" entries (a Linear one with "file" gets its table
    written to that file from "xvalues"/"fxvalues").
    nodal_forces: [(node ids or -1, dir, force, math function id or None)],
    written as node sets and concentrated_nodal_forces.
    tractions: [(particle ids or -1, dir, traction, math function id or
    None)], written as particle sets and particle_surface_traction.
    euler_angles: (node ids, angles[n, dim] in radians), written to the
    "nodal_euler_angles" file of the mesh's boundary conditions."""
    os.makedirs(d, exist_ok=True)
    dim = w.dim
    with open(os.path.join(d, "mesh.txt"), "w") as f:
        f.write("! mesh\n%d %d\n" % (len(w.coords), len(w.cells)))
        for x in w.coords:
            f.write(" ".join("%.17g" % v for v in x) + "\n")
        for c in w.cells:
            f.write(" ".join(str(int(n)) for n in c) + "\n")
    cset = list(range(len(w.cells))) if w.cell_set is None else [int(c) for c in w.cell_set]
    sets = {"cell_sets": [{"id": 1, "set": cset}], "node_sets": []}
    bc = {}
    if len(w.vel_nodes):
        if constraints_as == "file":
            with open(os.path.join(d, "velocity-constraints.txt"), "w") as f:
                f.write("# node dir velocity\n")
                for n, k, v in zip(w.vel_nodes, w.vel_dirs, w.vel_values):
                    f.write("%d %d %.17g\n" % (n, k, v))
            bc["velocity_constraints"] = [{"file": "velocity-constraints.txt"}]
        else:
            vc = []
            for k in sorted(set(int(x) for x in w.vel_dirs)):
                for val in sorted(set(float(v) for v in np.asarray(w.vel_values)[w.vel_dirs == k])):
                    sel = (w.vel_dirs == k) & (np.asarray(w.vel_values) == val)
                    sid = 10 + len(sets["node_sets"])
                    sets["node_sets"].append({"id": sid, "set": [int(n) for n in w.vel_nodes[sel]]})
                    vc.append({"nset_id": sid, "dir": k, "velocity": val})
            bc["velocity_constraints"] = vc
    if euler_angles is not None:
        with open(os.path.join(d, "euler-angles.txt"), "w") as f:
            f.write("! node alpha beta (gamma)\n")
            for n, a in zip(*euler_angles):
                f.write("%d %s\n" % (n, " ".join("%.17g" % v for v in a)))
        bc["nodal_euler_angles"] = "euler-angles.txt"
    if len(w.fric_nodes):
        fc = []
        sid = 100
        sets["node_sets"].append({"id": sid, "set": [int(n) for n in w.fric_nodes]})
        fc.append({"nset_id": sid, "dir": int(w.fric_dirs[0]), "sign_n": int(w.fric_sign[0]),
                   "friction": float(w.fric_mu[0])})
        bc["friction_constraints"] = fc
    if acceleration_constraints:
        ac = []
        for ids, k, val, fid in acceleration_constraints:
            sid = 300 + len(sets["node_sets"])
            sets["node_sets"].append({"id": sid, "set": [int(n) for n in ids]})
            c = {"nset_id": sid, "dir": int(k), "acceleration": float(val)}
            if fid is not None:
                c["math_function_id"] = int(fid)
            ac.append(c)
        bc["acceleration_constraints"] = ac
    nf = []
    for ids, k, val, fid in nodal_forces or []:
        if isinstance(ids, int) and ids == -1:
            sid = -1
        else:
            sid = 500 + len(sets["node_sets"])
            sets["node_sets"].append({"id": sid, "set": [int(n) for n in ids]})
        c = {"nset_id": sid, "dir": int(k), "force": float(val)}
        if fid is not None:
            c["math_function_id"] = int(fid)
        nf.append(c)
    fns = []
    for f in math_functions or []:
        f = dict(f)
        if f.get("type") == "Linear" and "file" in f:
            with open(os.path.join(d, f["file"]), "w") as fh:
                for x, fx in zip(f.pop("xvalues"), f.pop("fxvalues")):
                    fh.write("%.17g,%.17g\n" % (x, fx))
        fns.append(f)
    tr = []
    for ids, k, val, fid in tractions or []:
        sets.setdefault("particle_sets", [])
        if isinstance(ids, int) and ids == -1:
            sid = -1
        else:
            sid = 700 + len(sets["particle_sets"])
            sets["particle_sets"].append({"id": sid, "set": [int(i) for i in ids]})
        c = {"pset_id": sid, "dir": int(k), "traction": float(val)}
        if fid is not None:
            c["math_function_id"] = int(fid)
        tr.append(c)
    if particle_constraints:
        sets.setdefault("particle_sets", [])
        pc = []
        for ids, k, val in particle_constraints:
            if isinstance(ids, int) and ids == -1:
                pc.append({"pset_id": -1, "dir": int(k), "velocity": float(val)})
                continue
            sid = 200 + len(sets["particle_sets"])
            sets["particle_sets"].append({"id": sid, "set": [int(i) for i in ids]})
            pc.append({"pset_id": sid, "dir": int(k), "velocity": float(val)})
        bc["particles_velocity_constraints"] = pc
    with open(os.path.join(d, "entity_sets.json"), "w") as f:
        json.dump(sets, f)
    mesh = {"mesh": "mesh.txt", "entity_sets": "entity_sets.json", "io_type": "Ascii%dD" % dim,
            "isoparametric": False, "check_duplicates": True,
            "node_type": "N%dD" % dim, "cell_type": "ED2Q4" if dim == 2 else "ED3H8",
            "boundary_conditions": bc}
    if w.isoparametric:  # the key left out: the reference's default is isoparametric
        del mesh["isoparametric"]
    if stresses is not None:
        with open(os.path.join(d, "stresses.txt"), "w") as f:
            f.write("%d\n" % len(stresses))
            for s in stresses:
                f.write(" ".join("%.17g" % v for v in s) + "\n")
        mesh["particles_stresses"] = {"type": "file", "location": "stresses.txt"}
    deck = {
        "title": "mpmx deck",
        "mesh": mesh,
        "particles": [{"generator": {"type": "gauss", "material_id": 3, "cset_id": 1,
                                     "pset_id": 0, "nparticles_per_dir": w.nppd,
                                     "particle_type": "P%dD" % dim}}],
        "materials": [material_json],
        "external_loading_conditions": {"gravity": [float(g) for g in w.gravity]},
        "analysis": {"type": "MPMExplicit%dD" % dim, "mpm_scheme": w.scheme,
                     "velocity_update": velocity_update, "locate_particles": True, "dt": w.dt,
                     "nsteps": nsteps, "uuid": uuid,
                     "damping": {"type": "Cundall", "damping_factor": w.damping},
                     "resume": {"resume": False}},
        "post_processing": {"path": "results/", "output_steps": output_steps},
    }
    if fns:
        deck["math_functions"] = fns
    if nf:
        deck["external_loading_conditions"]["concentrated_nodal_forces"] = nf
    if tr:
        deck["external_loading_conditions"]["particle_surface_traction"] = tr
    with open(os.path.join(d, "mpm.json"), "w") as f:
        json.dump(deck, f, indent=1)
    return os.path.join(d, "mpm.json")


def run_driver(d, *extra, check=True):
    r = subprocess.run([DRIVER, "-f", d, "-i", "mpm.json"] + list(extra),
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    if check and r.returncode != 0:
        raise RuntimeError("mpmx_run failed: " + r.stderr)
    return r


def output_path(d, step, nsteps, uuid="deck", attribute="particles", ext=".h5"):
    """<d>/results/<uuid>/<attribute><step zero-padded to the digits of
    nsteps><ext> (IO::output_file, src/io/io.cc:94-140)."""
    import math
    digits = int(math.log10(max(nsteps, 1))) + 1
    return os.path.join(d, "results", uuid, "%s%0*d%s" % (attribute, digits, step, ext))


def read_output(path, dim):
    """A particle table written by the driver (the reference's PODParticle
    HDF5 table) as arrays in its record order: id, cell, x, u, v, stress,
    strain, F, volume, mass, statevars (the first 9 state variables),
    material_id, nsize."""
    from mpm_amd import pod
    c = pod.read(path)
    return {"id": c["id"], "cell": c["cell_id"], "x": c["coord"][:, :dim],
            "u": c["displacement"][:, :dim], "v": c["velocity"][:, :dim],
            "stress": c["stress"], "strain": c["strain"], "F": c["defgrad"],
            "volume": c["volume"], "mass": c["mass"], "statevars": c["svars"][:, :9],
            "pdstrain": c["svars"][:, 8], "material_id": c["material_id"],
            "nstate_vars": c["nstate_vars"], "nsize": c["nsize"][:, 0],
            "mapping_matrix": c["mapping_matrix"].reshape(-1, 3, 3)[:, :dim, :dim],
            "initialise_mapping_matrix": c["initialise_mapping_matrix"]}


def write_fixture_deck(d, dim, scheme="usf", nsteps=10, output_steps=5, resume=None,
                       uuid="fixture"):
    """The deck the reference's solver tests run (tests/io/write_mesh_particles.cc:
    write_json, write_entity_set, write_mesh_2d/3d, write_particles_2d/3d):
    two quad4/hex8 cells of 0.5 m, 8 (2D) or 16 (3D) particles from a file,
    two LinearElastic materials with the second given to particle set 2 by
    "material_sets", velocity constraints from a file, a Linear math function
    scaling a particle traction on every particle and a concentrated force on
    every node, Cundall damping 0.02, velocity_update true (PIC), dt 1e-3.
    Its rotation_forces entry is left out: this build rejects them.
    resume: None or the step to resume from."""
    os.makedirs(d, exist_ok=True)
    if dim == 2:
        nodes = [(0., 0.), (0.5, 0.), (0.5, 0.5), (0., 0.5), (1.0, 0.), (1.0, 0.5)]
        cells = [(0, 1, 2, 3), (1, 4, 5, 2)]
        parts = [(0.125, 0.125), (0.375, 0.125), (0.375, 0.375), (0.125, 0.375),
                 (0.625, 0.125), (0.875, 0.125), (0.875, 0.375), (0.625, 0.375)]
        cons = [(0, 1, 0.0), (1, 1, 0.0), (4, 1, 0.0)]
        gravity = [0.0, -9.81]
        node_set = [4, 5]
    else:
        nodes = [(0., 0., 0.), (0.5, 0., 0.), (0.5, 0.5, 0.), (0., 0.5, 0.), (0., 0., 0.5),
                 (0.5, 0., 0.5), (0.5, 0.5, 0.5), (0., 0.5, 0.5), (1.0, 0., 0.), (1.0, 0.5, 0.),
                 (1.0, 0., 0.5), (1.0, 0.5, 0.5)]
        cells = [(0, 1, 2, 3, 4, 5, 6, 7), (1, 8, 9, 2, 5, 10, 11, 6)]
        parts = [(0.125, 0.125, 0.125), (0.375, 0.125, 0.125), (0.375, 0.375, 0.125),
                 (0.125, 0.375, 0.125), (0.125, 0.125, 0.375), (0.375, 0.125, 0.375),
                 (0.375, 0.375, 0.375), (0.125, 0.375, 0.375), (0.625, 0.125, 0.125),
                 (0.875, 0.125, 0.125), (0.875, 0.375, 0.125), (0.625, 0.375, 0.125),
                 (0.675, 0.125, 0.375), (0.875, 0.125, 0.375), (0.875, 0.375, 0.375),
                 (0.675, 0.375, 0.375)]
        cons = [(n, 2, 0.0) for n in (0, 1, 2, 3, 8, 9)]
        gravity = [0.0, 0.0, -9.81]
        node_set = [8, 9, 10, 11]
    tag = "%dd" % dim
    with open(os.path.join(d, "mesh-%s.txt" % tag), "w") as f:
        f.write("! elementShape hexahedron\n! elementNumPoints 8\n")
        f.write("%d\t%d\n" % (len(nodes), len(cells)))
        for x in nodes:
            f.write("\t".join("%g" % v for v in x) + "\t\n")
        for c in cells:
            f.write("\t".join(str(n) for n in c) + "\t\n")
    with open(os.path.join(d, "velocity-constraints.txt"), "w") as f:
        for n, k, v in cons:
            f.write("%d\t%d\t%g\n" % (n, k, v))
    with open(os.path.join(d, "particles-%s.txt" % tag), "w") as f:
        f.write("%d\n" % len(parts))
        for x in parts:
            f.write("\t".join("%g" % v for v in x) + "\t\n")
    with open(os.path.join(d, "entity_sets_0.json"), "w") as f:
        json.dump({"particle_sets": [{"id": 2, "set": list(range(16))}],
                   "node_sets": [{"id": 1, "set": node_set}]}, f, indent=2)
    mtype = "LinearElastic%dD" % dim
    deck = {
        "title": "Example JSON Input for MPM",
        "mesh": {"mesh": "mesh-%s.txt" % tag, "entity_sets": "entity_sets_0.json",
                 "io_type": "Ascii%dD" % dim, "check_duplicates": True, "isoparametric": False,
                 "node_type": "N%dD" % dim,
                 "boundary_conditions": {"velocity_constraints": [
                     {"file": "velocity-constraints.txt"}]},
                 "cell_type": "ED2Q4" if dim == 2 else "ED3H8"},
        "particles": [{"generator": {"type": "file", "material_id": 0, "pset_id": 0,
                                     "io_type": "Ascii%dD" % dim,
                                     "particle_type": "P%dD" % dim, "check_duplicates": True,
                                     "location": "particles-%s.txt" % tag}}],
        "materials": [{"id": 0, "type": mtype, "density": 1000., "youngs_modulus": 1.0E+8,
                       "poisson_ratio": 0.495},
                      {"id": 1, "type": mtype, "density": 2300., "youngs_modulus": 1.5E+6,
                       "poisson_ratio": 0.25}],
        "material_sets": [{"material_id": 1, "phase_id": 0, "pset_id": 2}],
        "external_loading_conditions": {
            "gravity": gravity,
            "particle_surface_traction": [{"math_function_id": 0, "pset_id": -1, "dir": 1,
                                           "traction": 10.5}],
            "concentrated_nodal_forces": [{"math_function_id": 0, "nset_id": -1, "dir": 1,
                                           "force": 10.5}]},
        "math_functions": [{"id": 0, "type": "Linear", "xvalues": [0.0, 0.5, 1.0],
                            "fxvalues": [0.0, 1.0, 1.0]}],
        "analysis": {"type": "MPMExplicit%dD" % dim, "mpm_scheme": scheme,
                     "locate_particles": True, "velocity_update": True, "dt": 0.001,
                     "uuid": uuid, "nsteps": nsteps,
                     "resume": {"resume": resume is not None, "uuid": uuid,
                                "step": resume if resume is not None else 5},
                     "damping": {"type": "Cundall", "damping_factor": 0.02},
                     "newmark": {"beta": 0.25, "gamma": 0.5}},
        "post_processing": {"path": "results/", "vtk": ["stresses", "strains", "velocities"],
                            "vtk_statevars": [{"phase_id": 0, "statevars": ["pdstrain"]}],
                            "output_steps": output_steps},
    }
    with open(os.path.join(d, "mpm.json"), "w") as f:
        json.dump(deck, f, indent=2)
    return deck
