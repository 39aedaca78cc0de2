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
               tractions=None):
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
    None)], written as particle sets and particle_surface_traction."""
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
                     "velocity_update": False, "locate_particles": True, "dt": w.dt,
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


def read_output(path, dim):
    """Rows of a particles<step>.txt: id cell x v stress strain volume mass pdstrain."""
    a = np.loadtxt(path, comments="#", ndmin=2)
    k = 2
    out = {"cell": a[:, 1].astype(np.int64)}
    for name, n in (("x", dim), ("v", dim), ("stress", 6), ("strain", 6), ("volume", 1),
                    ("mass", 1), ("pdstrain", 1)):
        out[name] = a[:, k:k + n] if n > 1 else a[:, k]
        k += n
    return out
