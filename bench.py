"""Benchmark: explicit MPM step throughput on one MI355X.

Workload (default): the 3D granular column collapse — hex8 128x64x64 cells of
0.01 m, 2,097,152 particles (8 per cell in a 0.64 m cube block), MohrCoulomb
phi=30 psi=0 c=0, MUSL, FLIP, dt=5e-5, frictional base. A "step" is one full
explicit step (binning, P2G, node update, G2P, stress, locate) over all
particles, with the particle state resident in HBM.

The timed window starts after a fixed pre-roll of simulated steps
(--preroll, default 300 on the collapse: by step 200 about two thirds of the
particles yield, so the Mohr-Coulomb return mappings are in full swing) plus
--warmup steps, whatever --steps and --warmup are; `config.window` names the
simulated steps timed. The run from the initial state is also timed, and its
end state compared with an ordered-mode run of the same deck (`full_run`):
all of the workload's steps, or the steps it can run before a particle leaves
the mesh (the undamped collapse deck: 900 of its 2000, see
`Workload.runnable_steps`).

Prints one JSON line: value = particle-updates/s = particles * steps / time.
`roofline` is the dominant kernel's algorithmic HBM bytes per launch over its
average launch time (HIP events on the library's stream); `step_roofline` is
the whole step's compulsory bytes over the step time. `cpu_baseline` times
the oracle's OpenMP build (the reference algorithm restated in C++) on the
host's cores over a bounded number of steps of the same workload.
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "particle-updates/s (FP64 explicit step) + batched & adjoint steps/s at ≤1e-9 rel"

# Algorithmic bytes (compulsory HBM traffic of a faithful step, FP64, int32
# cell id): per particle and per active node, split by the kernel that owns
# each field; MohrCoulomb, fast mode.
PER_PARTICLE_MC = {
    (3, "musl"): {
        "p2g_all": 8,                 # mass read
        "g2p_momentum_scatter": 144,  # x, u, v read + write
        "stress_locate": 424,         # stress, strain, F, volume, density r/w, MC state, cell r/w
                                      # (with mc_plastic, which completes the yielding ones)
    },
    (2, "usl"): {
        "p2g_all": 8,
        "g2p_stress_locate": 360,     # x, u, v, stress(4), strain(3), V, rho, F(4) r/w, MC state, cell
    },
}
PER_NODE = {
    (3, "musl"): {
        "p2g_all": 80,        # mass, momentum, f_ext, f_int written
        "node_update": 128,   # those read, velocity + acceleration written
        "g2p_momentum_scatter": 48,  # velocity + acceleration read
    },
    (2, "usl"): {"p2g_all": 56, "node_update": 88, "g2p_stress_locate": 32},
}
STEP_BYTES_PER_PARTICLE = {(3, "mc"): 576, (3, "le"): 528, (2, "mc"): 368, (2, "le"): 320}
STEP_BYTES_PER_NODE = {3: 256, 2: 176}


def summarize_trace(path, first, count, scale=None):
    """Per-kernel statistics of the dispatches in a rocprofv3 CSV (kernel
    trace, or one counter of a --pmc pass) that belong to simulated steps
    first+1 .. first+count of the run: those after the first-th k_step_end
    dispatch up to the (first+count)-th. Steps are counted from the first
    solver the command creates, so run it without the autotuner (--chunk and
    --sort-every fixed), whose trial steps would shift the count. Returns
    markdown rows."""
    import csv
    rows = list(csv.DictReader(open(path)))
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    ends = 0
    sel = []
    for r in rows:
        name = r["Kernel_Name"]
        if first <= ends < first + count:
            sel.append(r)
        if name.startswith("mpmx::k_step_end"):
            ends += 1
    if ends < first + count:
        raise SystemExit("only %d steps in %s" % (ends, path))
    per = {}
    for r in sel:
        name = r["Kernel_Name"].split("(")[0].replace("void ", "")
        if scale is None:
            v = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
        else:
            v = float(r["Counter_Value"]) * 1024.0 * scale
        per.setdefault(name, []).append(v)
    out = []
    if scale is None:
        total = sum(sum(v) for v in per.values())
        out.append("| kernel | calls | avg us | min us | max us | % |")
        out.append("|---|---|---|---|---|---|")
        for k, v in sorted(per.items(), key=lambda kv: -sum(kv[1])):
            out.append("| `%s` | %d | %.1f | %.1f | %.1f | %.1f |" % (
                k, len(v), sum(v) / len(v), min(v), max(v), 100 * sum(v) / total))
        out.append("")
        out.append("kernel time per step: %.4f ms" % (total / count / 1e3))
    else:
        out.append("| kernel | dispatches | median bytes | median MB |")
        out.append("|---|---|---|---|")
        for k, v in sorted(per.items(), key=lambda kv: -sorted(kv[1])[len(kv[1]) // 2]):
            med = sorted(v)[len(v) // 2]
            out.append("| `%s` | %d | %.0f | %.1f |" % (k, len(v), med, med / 1e6))
    return "\n".join(out)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def rel_dev(a, b):
    scale = float(np.max(np.abs(b))) if b.size else 0.0
    return float(np.max(np.abs(a - b))) / scale if scale > 0 else float(np.max(np.abs(a)))


def full_run(w, args, tuned):
    """The workload's whole run from its initial state, timed (graph
    replays, fast mode), and its end state against an ordered-mode run of the
    same deck (nodal sums in particle order, bitwise equal to the oracle)."""
    from mpm_amd.capi import Solver
    s = w.setup(Solver(w.dim, w.scheme, ordered=args.ordered))
    if tuned:
        s.set_chunk(tuned["chunk"])
        s.set_sort_interval(tuned["sort_every"])
    s.synchronize()
    nsteps = w.runnable_steps or w.nsteps
    t0 = time.perf_counter()
    s.step(nsteps, w.dt)
    s.synchronize()
    wall = time.perf_counter() - t0
    out = {"steps": nsteps, "workload_steps": w.nsteps, "wall_s": wall,
           "ms_per_step": 1e3 * wall / nsteps,
           "particle_updates_per_s": s.nparticles * nsteps / wall}
    if nsteps < w.nsteps:
        out["why_shorter"] = ("a particle leaves the mesh near step 1000 in ordered mode (the "
                              "reference's arithmetic) as in fast mode: undamped FLIP/MUSL "
                              "velocities at the toe grow without bound after ~900 steps")
    s.close()
    if args.no_accuracy or args.ordered:
        return out
    # Accuracy: a fast-mode run and an ordered-mode run (particle-order nodal
    # sums, bitwise equal to the oracle) of the same deck side by side,
    # compared at a few steps along the run.
    f = w.setup(Solver(w.dim, w.scheme, ordered=False))
    if tuned:
        f.set_chunk(tuned["chunk"])
        f.set_sort_interval(tuned["sort_every"])
    o = w.setup(Solver(w.dim, w.scheme, ordered=True))
    marks = sorted({m for m in (1, 10, 50, 100, 200, 400, 600, nsteps) if m <= nsteps})
    done, ordered_wall, along = 0, 0.0, []
    for m in marks:
        f.step(m - done, w.dt)
        t0 = time.perf_counter()
        o.step(m - done, w.dt)
        o.synchronize()
        ordered_wall += time.perf_counter() - t0
        done = m
        pf = f.particles(["x", "v", "stress", "cell", "statevars"])
        po = o.particles(["x", "v", "stress", "cell", "statevars"])
        dev = {"step": m}
        dev.update({k: rel_dev(pf[k], po[k]) for k in ("x", "v", "stress")})
        dev["cells_differing"] = int(np.sum(pf["cell"] != po["cell"]))
        if pf.get("statevars") is not None and po.get("statevars") is not None:
            dev["yield_states_differing"] = int(np.sum(pf["statevars"][:, 0] !=
                                                       po["statevars"][:, 0]))
        along.append(dev)
    f.close()
    o.close()
    out["vs_ordered"] = dict(what="fast mode against ordered mode (particle-order nodal sums, "
                                  "bitwise equal to the oracle) at steps along the run, "
                                  "max |a - b| / max |b| per field", along=along)
    out["ordered_ms_per_step"] = 1e3 * ordered_wall / nsteps
    return out


def active_nodes(w, cells_of_particles):
    occ = np.unique(cells_of_particles[cells_of_particles >= 0])
    return int(np.unique(w.cells[occ].ravel()).size)


def cpu_baseline(w, budget_s):
    """Oracle OpenMP build on the host: steps/s over a bounded window."""
    import oracle
    o = w.setup(oracle.Oracle(w.dim, w.scheme, omp=True))
    t0 = time.perf_counter()
    o.step(1, w.dt)
    one = time.perf_counter() - t0
    n = max(1, min(200, int(budget_s / max(one, 1e-6))))
    t0 = time.perf_counter()
    o.step(n, w.dt)
    el = time.perf_counter() - t0
    return dict(value=o.nparticles * n / el, unit="particle-updates/s", cores=o.threads,
                kind="port", cpu_model=cpu_model(),
                sample="%d steps (after 1 warm-up step) of the same %d-particle workload, "
                       "oracle OpenMP build, %d threads, %.1f s" % (n, o.nparticles, o.threads, el))


def cpu_baseline_batch(b, budget_s):
    """Oracle OpenMP build, instances one after another (the reference runs one
    mesh per analysis): particle-updates/s over a bounded window."""
    import oracle
    done, el, threads, ninst = 0, 0.0, 1, 0
    for w in b.instances:
        o = w.setup(oracle.Oracle(w.dim, w.scheme, omp=True))
        o.step(1, w.dt)
        t0 = time.perf_counter()
        o.step(10, w.dt)
        el += time.perf_counter() - t0
        done += o.nparticles * 10
        threads = o.threads
        ninst += 1
        if el > budget_s:
            break
    return dict(value=done / el, unit="particle-updates/s", cores=threads, kind="port",
                cpu_model=cpu_model(),
                sample="10 steps (after 1 warm-up step) of each of the first %d instances, "
                       "one at a time, oracle OpenMP build, %d threads, %.1f s"
                       % (ninst, threads, el))


def sequential_gpu(b, steps, warmup=5):
    """Each instance alone on the GPU (its own solver), one after another:
    the sum over instances of the per-step time."""
    from mpm_amd.capi import Solver
    total = 0.0
    for w in b.instances:
        s = w.setup(Solver(w.dim, w.scheme))
        s.step(warmup, w.dt)
        s.synchronize()
        t0 = time.perf_counter()
        s.step(steps, w.dt)
        s.synchronize()
        total += (time.perf_counter() - t0) / steps
        s.close()
    return total


def bench_adjoint(b, args):
    """Adjoint inversion batch: the plain forward run, the recorded forward
    run (checkpoints) and the backward pass of the squared distance to a
    target trajectory, each timed whole; then central differences of that
    loss (GPU forward runs, h relative per instance parameter) along random
    per-instance directions of E, rho and phi against the adjoint's
    directional derivatives."""
    from mpm_amd.capi import Solver
    nsteps = args.steps or b.nsteps
    every = b.checkpoint_every

    def fresh(materials):
        saved = b.materials
        b.materials = materials
        s = b.setup(Solver(b.dim, b.scheme, ordered=args.ordered))
        b.materials = saved
        s.synchronize()
        return s

    def final_x(materials, ordered=None):
        saved = b.materials
        b.materials = materials
        s = b.setup(Solver(b.dim, b.scheme,
                           ordered=args.ordered if ordered is None else ordered))
        b.materials = saved
        s.step(nsteps, b.dt)
        x = s.particles(["x"])["x"]
        s.close()
        return x

    xt = final_x(b.target_materials)
    s = fresh(b.materials)
    npart = s.nparticles
    t0 = time.perf_counter()
    s.step(nsteps, b.dt)
    s.synchronize()
    fwd = time.perf_counter() - t0
    s.close()
    s = fresh(b.materials)
    t0 = time.perf_counter()
    s.adjoint_forward(nsteps, b.dt, 0.0, every)
    s.synchronize()
    rec = time.perf_counter() - t0
    x = s.particles(["x"])["x"]
    loss = float(np.sum((x - xt) ** 2))
    t0 = time.perf_counter()
    g = s.adjoint_backward(2.0 * (x - xt))
    bwd = time.perf_counter() - t0
    s.adjoint_release()
    s.close()
    # directional checks: central differences of the loss along random
    # per-instance directions, from ordered-mode (deterministic) forward runs
    # and summed over the perturbed instances only (instances are
    # independent, so the other terms of the loss do not depend on the
    # perturbed parameters); the difference of the two losses is summed as
    # sum((x+ - x-) (x+ + x- - 2 x*)), free of the cancellation of two
    # nearly equal sums of squares. LinearElastic instances (smooth map) at
    # h = 1e-4 ... 1e-6: the quotient's error falls as h^2 down to a floor
    # set by the rounding of the two perturbed trajectories, which grows as
    # 1/h below it; the reported agreement is that at the step whose quotient
    # agrees best with the next smaller step's. Mohr-Coulomb instances at h =
    # 1e-3 and 1e-6 (their run crosses yield-state and iteration-count kinks,
    # so the quotient does not converge as h shrinks).
    nin = len(b.instances)
    bounds = np.cumsum([0] + [w.nparticles for w in b.instances])
    checks = {}
    xt_o = None if args.no_check else final_x(b.target_materials, ordered=True)
    for label, col, sel, hs in (("E_le", 2, slice(0, None, 2), (1e-4, 1e-5, 1e-6)),
                                ("rho_le", 1, slice(0, None, 2), (1e-4, 1e-5, 1e-6)),
                                ("E_mc", 2, slice(1, None, 2), (1e-3, 1e-6)),
                                ("rho_mc", 1, slice(1, None, 2), (1e-3, 1e-6)),
                                ("phi_mc", 4, slice(1, None, 2), (1e-3, 1e-6))):
        if args.no_check:
            break
        rng = np.random.default_rng(12)
        d = np.zeros(nin)
        d[sel] = rng.normal(size=len(range(nin)[sel]))
        name = label.split("_")[0]
        per_inst = np.array([g[name][bounds[i]:bounds[i + 1]].sum() for i in range(nin)])
        ad = float(np.sum(per_inst * b.materials[:, col] * d))
        entry = {"adjoint": ad}
        mask = np.zeros(npart, dtype=bool)
        for i in range(nin)[sel]:
            mask[bounds[i]:bounds[i + 1]] = True
        fds = []
        for h in hs:
            xs = []
            for sg in (1.0, -1.0):
                m = b.materials.copy()
                m[:, col] *= 1.0 + sg * h * d
                xs.append(final_x(m, ordered=True)[mask])
            fd = float(np.sum((xs[0] - xs[1]) * (xs[0] + xs[1] - 2.0 * xt_o[mask]))) / (2 * h)
            fds.append(fd)
            entry["fd_h%g" % h] = fd
            entry["rel_h%g" % h] = abs(ad - fd) / max(abs(fd), 1e-300)
        if label.endswith("_le"):
            k = int(np.argmin([abs(fds[i] - fds[i + 1]) / abs(fds[i])
                               for i in range(len(fds) - 1)]))
            entry["converged_h"] = hs[k]
            entry["rel_converged"] = entry["rel_h%g" % hs[k]]
        checks[label] = entry
    return {
        "metric": METRIC,
        "value": npart * nsteps / bwd,
        "unit": "particle-updates/s (adjoint backward)",
        "n_gpus": 1,
        "steps": nsteps,
        "warmup": 0,
        "ms_per_step": 1e3 * bwd / nsteps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": b.name, "particles": npart, "scheme": b.scheme.upper(),
                   "mode": "ordered" if args.ordered else "fast", **b.info, "dt": b.dt},
        "adjoint": {
            "forward_ms_per_step": 1e3 * fwd / nsteps,
            "recorded_forward_ms_per_step": 1e3 * rec / nsteps,
            "backward_ms_per_step": 1e3 * bwd / nsteps,
            "backward_over_forward": bwd / fwd,
            "loss": loss,
            "fd_check": dict(directions="per-instance parameter scalings, normal(seed 12)",
                             **checks),
        },
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default 200; adjoint_2d: its 500-step run)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--preroll", type=int, default=None,
                    help="simulated steps run before the warm-up (default: 300 on the "
                         "collapse, inside its plastic phase; 0 otherwise)")
    ap.add_argument("--no-full-run", action="store_true",
                    help="skip the whole-run timing and its ordered-mode comparison")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="time the whole run but skip the ordered-mode comparison")
    ap.add_argument("--workload", default="collapse_3d",
                    choices=["collapse_3d", "wedge_3d", "column_2d", "batch_2d", "adjoint_2d"])
    ap.add_argument("--sequential-steps", type=int, default=20,
                    help="batch_2d: steps per instance of the one-instance-at-a-time comparison")
    ap.add_argument("--ordered", action="store_true", help="particle-order nodal sums")
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of timed CPU work")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch the timed steps' kernels one by one instead of replaying "
                         "the captured step (kernel traces without graph launches)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default scatter chunk and sort interval")
    ap.add_argument("--chunk", type=int, default=None,
                    help="scatter chunk (128, 256, 512) instead of autotuning")
    ap.add_argument("--sort-every", type=int, default=None,
                    help="steps between cell sorts instead of autotuning")
    ap.add_argument("--no-check", action="store_true",
                    help="adjoint_2d: skip the finite-difference checks")
    ap.add_argument("--summarize", metavar="CSV", default=None,
                    help="print per-kernel statistics of the timed window's dispatches from a "
                         "rocprofv3 kernel-trace CSV (or a --pmc counter CSV with --scale) and "
                         "exit; the window is --preroll + --warmup steps in, --steps long")
    ap.add_argument("--scale", type=float, default=None,
                    help="with --summarize of a counter CSV: bytes per KiB counted / 1024")
    args = ap.parse_args()
    if args.summarize:
        pre = 300 if args.preroll is None and args.workload == "collapse_3d" else (args.preroll or 0)
        print(summarize_trace(args.summarize, pre + args.warmup, args.steps or 200, args.scale))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: the explicit step is built and measured on a "
              "single MI355X (multi-GPU sharding is a design, not part of this build)",
              file=sys.stderr)
        sys.exit(2)

    from mpm_amd import workloads
    from mpm_amd.capi import Solver

    w = getattr(workloads, args.workload)()
    if args.preroll is None:
        args.preroll = 300 if args.workload == "collapse_3d" else 0
    if args.workload == "adjoint_2d":
        print(json.dumps(bench_adjoint(w, args)))
        return
    if args.steps is None:
        args.steps = 200
    t0 = time.perf_counter()
    s = w.setup(Solver(w.dim, w.scheme, ordered=args.ordered, graph=not args.no_graph))
    setup_s = time.perf_counter() - t0
    npart = s.nparticles
    # scatter chunk and sort interval from timed trial steps (state restored)
    if args.chunk or args.sort_every:
        tuned = dict(chunk=args.chunk or 256, sort_every=args.sort_every or 8, fixed=True)
        s.set_chunk(tuned["chunk"])
        s.set_sort_interval(tuned["sort_every"])
    elif args.no_autotune:
        tuned = None
    else:
        tuned = s.autotune(w.dt, trial_steps=10)
    s.step(args.preroll + args.warmup, w.dt)
    s.synchronize()
    t0 = time.perf_counter()
    s.step(args.steps, w.dt)
    s.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    value = npart * args.steps / elapsed

    # Per-kernel times: HIP events around each launch on the library's stream,
    # over the same simulated steps as the timed region (a second solver from
    # the same deck: warm-up steps, then the timed steps launched one kernel at
    # a time between event pairs).
    cells = s.particles(["cell"])["cell"]
    nact = active_nodes(w, cells)
    s.close()
    s = w.setup(Solver(w.dim, w.scheme, ordered=args.ordered))
    if tuned:
        s.set_chunk(tuned["chunk"])
        s.set_sort_interval(tuned["sort_every"])
    s.step(args.preroll + args.warmup, w.dt)
    s.profile(True)
    s.step(args.steps, w.dt)
    kt = s.kernel_times()
    s.profile(False)
    profiled_ms_per_step = sum(v[0] for v in kt.values()) / args.steps
    kernels = {k: dict(avg_ms=v[0] / v[1], launches=v[1]) for k, v in kt.items()}
    dom = max(kt, key=lambda k: kt[k][0])
    mc = int(w.materials[0][0]) == 1
    per_p = PER_PARTICLE_MC.get((w.dim, w.scheme), {}) if mc else {}
    per_n = PER_NODE.get((w.dim, w.scheme), {}) if mc else {}
    alg = per_p.get(dom, 0) * npart + per_n.get(dom, 0) * nact
    # The Mohr-Coulomb update runs in two kernels: the stress kernel (kinematics,
    # elastic predictor, elastic particles) and mc_plastic (the return mappings it
    # defers). Their algorithmic bytes are those of the one constitutive update,
    # so the roofline row is the pair: the sum of their average durations.
    parts = [dom] + (["mc_plastic"] if dom in per_p and "mc_plastic" in kernels else [])
    dom_ms = sum(kernels[k]["avg_ms"] for k in parts)
    achieved = alg / (dom_ms * 1e-3) / 1e9 if alg else None
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic_%s.json" % args.workload)
    if os.path.exists(tfile):
        with open(tfile) as f:
            tr = json.load(f)
        if all(k in tr for k in parts):
            traffic = sum(tr[k] for k in parts)
    step_bytes = (STEP_BYTES_PER_PARTICLE[(w.dim, "mc" if mc else "le")] * npart +
                  STEP_BYTES_PER_NODE[w.dim] * nact)
    step_gbs = step_bytes / (ms_per_step * 1e-3) / 1e9

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": w.name, "particles": npart, "scheme": w.scheme.upper(),
                   "mode": "ordered" if args.ordered else "fast", **w.info,
                   "active_nodes": nact, "dt": w.dt,
                   "window": {"first_step": args.preroll + args.warmup, "steps": args.steps,
                              "preroll": args.preroll, "run_steps": w.nsteps,
                              "runnable_steps": w.runnable_steps or w.nsteps}},
        "roofline": {"bound": "hbm", "kernel": "+".join(parts), "achieved": achieved,
                     "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
                     "traffic": traffic, "alg_bytes_per_launch": alg, "avg_ms": dom_ms},
        "step_roofline": {"bytes_per_step": step_bytes, "achieved": step_gbs,
                          "frac": step_gbs / HBM_PEAK_GBS, "unit": "GB/s"},
        "kernels": kernels,
        "kernels_sum_ms_per_step": profiled_ms_per_step,
        "autotune": tuned,
        "setup_s": setup_s,
    }
    if hasattr(w, "instances"):
        seq = sequential_gpu(w, args.sequential_steps)
        out["sequential"] = {
            "what": "the same %d instances, each alone in its own solver, one after another"
                    % len(w.instances),
            "ms_per_step": 1e3 * seq, "value": npart / seq,
            "batched_speedup": seq / (ms_per_step * 1e-3),
            "steps_per_instance": args.sequential_steps}
    if not args.no_full_run and not hasattr(w, "instances"):
        out["full_run"] = full_run(w, args, tuned)
        if "ordered_ms_per_step" in out["full_run"]:
            # fast mode agrees with the reference's arithmetic to 1e-12 per
            # step but not to 1e-9 over a chaotic Mohr-Coulomb run (vs_ordered);
            # ordered mode reproduces the run bit for bit
            oms = out["full_run"]["ordered_ms_per_step"]
            out["bitwise_mode"] = {
                "what": "ordered mode (particle-order nodal sums, bitwise equal to the "
                        "oracle's run) over the whole runnable run",
                "ms_per_step": oms, "value": npart / (oms * 1e-3)}
    if not args.no_cpu:
        out["cpu_baseline"] = (cpu_baseline_batch(w, args.cpu_budget) if hasattr(w, "instances")
                               else cpu_baseline(w, args.cpu_budget))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
