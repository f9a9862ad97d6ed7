"""Throughput of the D2Q9 MRT lattice update on one MI355X.

Headline (``value``): GLUPS of the FP32 fused MRT step on the largest single-GPU
workload, the 32768 x 16384 domain (536.9 M cells) with Smagorinsky + Van Driest LES
(c_s = 0.12, tau = 0.505, u = 0.1) and 128 seeded obstacles (64 circles r 20-60 with
exact-circle Mei links, 64 plates 155 x 2 at angle of attack in [-15, 15] deg on a
staircase with simple bounce-back), Zou/He inlet, outflow copy, periodic y -- the
reference's semantics (src/lbm.hpp:293-582).  Inputs are resident in HBM; the timed
region is exactly --steps steps after --warmup untimed ones, launched through the
library's hipGraphs of --graph-steps steps (100, the library's default; a shorter run
as plain launches).
This workload as specified (impulsive start at u = 0.1, tau = 0.505) is not stable:
every storage form, FP64 included, produces its first non-finite density between steps
350 and 400, at the surface of an r = 59 circle, with Mei or with Bouzidi links.  The
headline therefore runs warmup + steps + 120 steps of roofline timing, 230 by default,
and reports the NaN flag.

``roofline`` is for the dominant kernel of the headline: algorithmic bytes per launch =
(N_fluid*18*s + N_cells*1 + N_les*2) per step covered by the launch (s = bytes per stored
population, the u16 wall distance with LES), divided by that kernel's average launch
duration measured here with HIP events on the library's stream (torch events recorded on
the stream the library launches on): the difference between replays of hipGraphs of 40
and of 20 steps, which differ by exactly 10 two-step launches (or 20 one-step launches),
against 8 TB/s.  A two-step launch of this lattice is the bulk-tile kernel with the edge
tiles (the bulk kernel's edge form) and the other tiles (obstacle surfaces, corners) running
beside it on side streams; the launch time covers all three.  ``traffic`` is the HBM bytes per launch read from the committed
rocprofv3 PMC summary of the same workload and launch shape (null when the autotuner
picks a shape the summary does not hold).

``forms`` (same JSON line): the 16-bit storage forms on the same domain (stochastic
rounding, the form that holds Cd/St within 1% of FP64, then round-to-nearest), FP32 with
one step per launch on it, the 6400 x 2400 LES cylinder in FP32 / FP16S / BF16S, the
800 x 300 cylinder (cache-resident: its two population buffers, 35 MB, fit in the 256 MB
Infinity Cache), the 256-lattice batched ensemble against the same lattices stepped one
launch each, and the thermal kernel.

``cpu_baseline`` times the FP64 CPU restatement of the step (oracle/, with the LES wall
distance computed once, since the reference recomputes it serially every step) on a
labelled 8192 x 4096 sub-domain of the headline workload with the same obstacle density,
on this host's cores (OMP_NUM_THREADS as set) and on one thread, in child processes; the
reference itself (oracle/_ref) is timed beside it on the same sub-domain without LES.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "lbm-2d_amd"))

HBM_PEAK_GBS = 8000.0
LARGE_PROFILE = os.path.join(ROOT, "profiles", "step_kernel_large32768x16384_les_f32.md")


def log(msg):
    print(f"# {msg}", file=sys.stderr, flush=True)


def cylinder_lattice(nx, ny, cx, cy, r, tau, u, storage, stream, les=False, cs=0.1, perturb=True, spec=False):
    """spec: the configs' boundary wording (equilibrium inlet, free-slip top and bottom)
    instead of the reference's (Zou/He inlet, periodic y)."""
    import lbm2d as G
    d = G.desc_for_case(G.CYLINDER, nx, ny, tau=tau, u_in=u, les=les, cs=cs)
    if spec:
        d.inlet, d.y_edge = G.INLET_EQUILIBRIUM, G.EDGE_FREE_SLIP
    ens = G.Ensemble([d], storage=storage, stream=stream)
    mask = G.circle_mask(nx, ny, cx, cy, r)
    ens.set_obstacle(0, mask)
    ens.set_bb_circle(0, cx, cy, r, G.BB_MEI)
    if perturb:
        f = G.equilibrium(np.ones((ny, nx)), np.full((ny, nx), u), np.zeros((ny, nx)))
        # seed-42 transverse perturbation of the cylinder driver (src/main.cpp:68-83)
        G.perturb_seed42(f, mask, cx + 5, cx + 60)
        ens.upload(0, f)
    else:
        ens.init_equilibrium(0, 1.0, u, 0.0)
    return ens, mask


def algorithmic_bytes(info, les_bytes=2):
    """Bytes per step: 18 population moves per fluid node, the flag byte of every node,
    the wall distance of every node of an LES lattice."""
    s = info.bytes_per_pop
    return info.fluid_cells * 18 * s + info.cells * 1 + info.les_cells * les_bytes


def time_steps(ens, steps, warmup, torch):
    """Seconds for exactly `steps` steps after `warmup` untimed ones, bracketed by
    synchronisations of the library's stream and the device."""
    if warmup:
        ens.step(warmup)
    ens.step(0)  # commits and builds the step graphs outside the timed region
    ens.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ens.step(steps)
    ens.sync()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def graph_step_ms(ens, torch, replays=20, graph_steps=100):
    """Average device time per step: HIP events on the library's stream around each
    replay of a hipGraph of graph_steps steps."""
    ens.set_launch(graph_steps=graph_steps)
    stream = torch.cuda.current_stream()
    ens.step(0)
    ens.step(graph_steps)
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(replays + 1)]
    evs[0].record(stream)
    for i in range(replays):
        ens.step(graph_steps)
        evs[i + 1].record(stream)
    torch.cuda.synchronize()
    return float(np.mean([evs[i].elapsed_time(evs[i + 1]) for i in range(replays)])) / graph_steps


def launches(n, two_step):
    """Step-kernel launches of a batch of n steps (two-step launches end with one-step ones)."""
    if not two_step or n < 2:
        return n
    pairs = (n - 1) // 2
    return pairs + (n - 2 * pairs)


def kernel_launch_ms(ens, torch, two_step, short=20, long=40, replays=1):
    """Average duration of one launch of the dominant step kernel: HIP events on the
    library's stream around replays of hipGraphs of `short` and of `long` steps; their
    difference is (long - short) / 2 two-step launches (or long - short one-step
    launches) and nothing else of note (the longer batch's force flush covers more
    launches: ~1 us).  Steps taken: 2 (1 + replays) (short + long)."""
    per = {}
    for g in (short, long):
        per[g] = graph_step_ms(ens, torch, replays=replays, graph_steps=g) * g
    extra = launches(long, two_step) - launches(short, two_step)
    return (per[long] - per[short]) / extra, per


def bench_form(nx, ny, r, tau, storage, steps, warmup, torch, stream, les=False, cs=0.1, launch="", sr=False,
               spec=False):
    ens, mask = cylinder_lattice(nx, ny, nx // 4, ny // 2, r, tau, 0.1, storage, stream,
                                 les=les, cs=cs, perturb=False, spec=spec)
    if sr:
        ens.set_stochastic_rounding(True, 1)
    if launch:
        shape = tuple(int(v) for v in launch.split(","))
        ens.set_launch(*shape)
    else:
        shape, _ = ens.autotune()
    info = ens.info()
    t = time_steps(ens, steps, warmup, torch)
    kms = graph_step_ms(ens, torch, replays=5)
    b = algorithmic_bytes(info)
    out = {
        "lattice": f"{nx}x{ny}", "storage": storage, "les": les, "rounding": "stochastic" if sr else "nearest-even",
        "boundaries": "equilibrium inlet, free-slip y" if spec else "Zou/He inlet, periodic y",
        "GLUPS": info.cells * steps / t / 1e9,
        "ms_per_step": 1e3 * t / steps,
        "kernel_ms_per_step": kms,
        "bytes_per_step": b,
        "hbm_frac": b / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "launch": list(shape),
        "nan": ens.nan_flag(0),
    }
    ens.close()
    return out


def bench_thermal(nx, ny, storage, steps, warmup, torch, stream, launch=""):
    """Heated cylinder with the thermal double distribution (src/heated_cylinder.cpp at
    Re = 100, Pr = 0.71, T_wall = 1.5, buoyancy on: beta = 0.001, g = -1e-5): circle
    r = ny/10 at (nx/4, ny/2), f and g at equilibrium with u = 0.1, T = 1; the driver's
    wall loop (solids at the wall temperature) as device conditions."""
    import lbm2d as G
    r = ny // 10
    tau = 0.5 + 3.0 * 0.1 * 2 * r / 100.0
    alpha = (tau - 0.5) / 3.0 / 0.71
    ens, mask = cylinder_lattice(nx, ny, nx // 4, ny // 2, r, tau, 0.1, storage, stream, perturb=False)
    ens.set_thermal(0, 1.0 / (0.5 + 3.0 * alpha), 0.001, -1e-5, 1.0)
    w = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
    eu = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1], float) * 0.1
    g_fluid = w * (1.0 + 3.0 * eu + 4.5 * eu * eu - 1.5 * 0.01)
    g = np.empty((ny, nx, 9))
    g[...] = g_fluid
    g[mask != 0] = w * 1.5
    ens.upload_thermal(0, g)
    del g
    ens.set_thermal_rules(0, mask, [(1.5, G.THERMAL_EQ_REST)])
    if launch:
        ens.set_launch(*(int(v) for v in launch.split(",")))
    info = ens.info()
    t = time_steps(ens, steps, warmup, torch)
    kms = graph_step_ms(ens, torch, replays=5)
    s = info.bytes_per_pop
    b = info.fluid_cells * 36 * s + info.cells * 2  # f and g planes, flag and condition bytes
    out = {
        "lattice": f"{nx}x{ny}", "storage": storage, "tau": tau,
        "GLUPS": info.cells * steps / t / 1e9,
        "ms_per_step": 1e3 * t / steps,
        "kernel_ms_per_step": kms,
        "bytes_per_step": b,
        "hbm_frac": b / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS,
    }
    ens.close()
    return out


def bench_ensemble(steps, warmup, torch, stream, n=256, launch=""):
    """The seeded ensemble (workloads.ensemble_descs): one launch per step over every
    lattice (or one per two steps), hipGraphs of 100 steps, against the same lattices
    stepped one launch each.  launch: bx,rows,nodes,options instead of the autotuner's."""
    import workloads as W
    lats = W.ensemble_descs(n=n)
    ens = W.build_ensemble(lats, "f32", stream)
    if launch:
        shape = tuple(int(v) for v in launch.split(","))
        ens.set_launch(*shape)
    else:
        shape, _ = ens.autotune()
    ens.set_launch(graph_steps=100)
    info = ens.info()
    t = time_steps(ens, steps, warmup, torch)
    kms = graph_step_ms(ens, torch, replays=3)
    # the same lattices, one launch per lattice per step, no graph
    plain_steps = max(2, steps // 20)
    for l in range(n):
        ens.step_lattice(l, 1)
    ens.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(plain_steps):
        for l in range(n):
            ens.step_lattice(l, 1)
    ens.sync()
    torch.cuda.synchronize()
    tp = time.perf_counter() - t0
    b = algorithmic_bytes(info)
    glups = info.cells * steps / t / 1e9
    glups_plain = info.cells * plain_steps / tp / 1e9
    nans = [l for l in range(n) if ens.nan_flag(l)]
    # per-lattice Cd = 2 Fx / (D u^2), D the largest diameter, mean over the last 100
    # steps run: early-time values of flows started from rest about the obstacles
    cds = []
    for l, (d, _, circles) in enumerate(lats):
        fxl, _ = ens.forces(l)
        dmax = 2 * max(c[2] for c in circles)
        cds.append(float(2.0 * fxl[-100:, 0].mean() / (dmax * 0.01)))
    out = {
        "lattices": n, "cells": info.cells, "storage": "f32",
        "les_lattices": int(sum(1 for d, _, _ in lats if d.les)),
        "GLUPS": glups, "ms_per_step": 1e3 * t / steps,
        "kernel_ms_per_step": kms,
        "hbm_frac": b / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "bytes_per_step": b,
        "plain_per_lattice_GLUPS": glups_plain,
        "speedup_vs_plain": glups / glups_plain,
        "launch": list(shape),
        "nan_lattices": nans,
        "steps_run": ens.steps_done(0),
        "cd_per_lattice": [float(f"{c:.4g}") for c in cds],
    }
    ens.close()
    return out


def bench_large(geom, form, steps, warmup, torch, stream, launch="", graph_steps=100, roofline=False):
    """One storage form of the large LES domain: set-up, autotune (or --launch), exactly
    `steps` timed steps after `warmup`; with roofline, the dominant kernel's launch
    duration from graph replays."""
    import workloads as W
    mask, links = geom
    storage = form.split("_")[0]
    t0 = time.perf_counter()
    e = W.large_lattice(mask, links, storage, stream, sr=form.endswith("_sr"))
    e.sync()
    t_setup = time.perf_counter() - t0
    if launch:
        shape = tuple(int(v) for v in launch.split(","))
        e.set_launch(*shape)
    else:
        shape, _ = e.autotune()
    e.set_launch(graph_steps=graph_steps)
    info = e.info()
    m0 = e.mass(0)
    log(f"32768x16384 {form}: set up in {t_setup:.2f} s, launch {tuple(shape)}")
    t = time_steps(e, steps, warmup, torch)
    b = algorithmic_bytes(info)
    out = {"storage": storage, "rounding": "stochastic" if form.endswith("_sr") else "nearest-even",
           "GLUPS": info.cells * steps / t / 1e9, "ms_per_step": 1e3 * t / steps, "seconds": t,
           "bytes_per_step": b, "launch": list(shape), "cells": info.cells, "fluid_cells": info.fluid_cells,
           "les_cells": info.les_cells, "setup_s": t_setup,
           "bulk_tiles": info.bulk_tiles, "edge_tiles": info.edge_tiles, "other_tiles": info.other_tiles,
           "empty_tiles": info.empty_tiles}
    two = bool(shape[3] & 4)
    if roofline:
        lms, per = kernel_launch_ms(e, torch, two)
        out["launch_ms"] = lms
        out["graph_ms"] = {str(k): v for k, v in per.items()}
        out["steps_per_launch"] = 2 if two else 1
        out["hbm_frac"] = b * out["steps_per_launch"] / (lms * 1e-3) / 1e9 / HBM_PEAK_GBS
    else:
        kms = graph_step_ms(e, torch, replays=2, graph_steps=20)
        out["kernel_ms_per_step"] = kms
        out["hbm_frac"] = b / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS
    out["nan"] = e.nan_flag(0)
    out["fp32_mass_drift"] = e.mass(0) / m0 - 1.0
    out["steps_run"] = e.steps_done(0)
    e.close()
    log(f"32768x16384 {form}: {out['GLUPS']:.1f} GLUPS")
    return out


def bench_large_one_step(geom, torch, stream, shapes=((64, 1, 1, 1), (256, 1, 1, 3), (64, 1, 4, 1), (128, 1, 2, 1))):
    """FP32 on the large LES domain with one step per launch: per-step device time of
    hipGraphs of 20 steps for a few one-step shapes; the fastest is reported."""
    import workloads as W
    mask, links = geom
    e = W.large_lattice(mask, links, "f32", stream)
    info = e.info()
    b = algorithmic_bytes(info)
    per = {}
    for shape in shapes:
        e.set_launch(*shape, graph_steps=20)
        per[shape] = graph_step_ms(e, torch, replays=2, graph_steps=20)
    best = min(per, key=per.get)
    kms = per[best]
    out = {"storage": "f32", "steps_per_launch": 1, "launch": list(best), "kernel_ms_per_step": kms,
           "GLUPS": info.cells / (kms * 1e-3) / 1e9, "hbm_frac": b / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS,
           "shapes_ms": {",".join(str(v) for v in k): v for k, v in per.items()}}
    e.close()
    log(f"32768x16384 f32 one step per launch: {out['GLUPS']:.1f} GLUPS")
    return out


CPU_SUB = dict(nx=8192, ny=4096, n_circles=4, n_plates=4)


def cpu_baseline(seconds=10.0):
    """The FP64 CPU step on a 8192 x 4096 sub-domain of the headline workload: 4 circles
    and 4 plates placed as large_domain() places 64 + 64 on the full domain (the same
    obstacle density), tau = 0.505, u = 0.1, Zou/He inlet, outflow, periodic y.  The
    reference keeps one interpolated-bounce-back geometry, so the first circle is it.
    Timed: the oracle with LES (wall distance computed once) on OMP_NUM_THREADS threads
    and on one thread, and the reference itself without LES on OMP_NUM_THREADS threads,
    each for about `seconds` of steps after one untimed step, in a child process."""
    code = f"""
import json, os, sys, time
sys.path.insert(0, {ROOT!r}); sys.path.insert(0, {os.path.join(ROOT, 'lbm-2d_amd')!r})
import numpy as np
import oracle as O
import workloads as W
kind = sys.argv[1]
nx, ny, nc, npl = {CPU_SUB['nx']}, {CPU_SUB['ny']}, {CPU_SUB['n_circles']}, {CPU_SUB['n_plates']}
mask, links, shapes = W.large_domain(42, nx, ny, nc, npl, with_shapes=True)
_, cx0, cy0, r0 = shapes[0]  # the reference's single bounce-back geometry
if kind == 'reference':
    s = O.RefLattice(nx, ny, O.CYLINDER, les=False, fast=True)
else:
    s = O.Oracle(nx, ny, O.CYLINDER, les=True, cs=0.12, faithful=False)
s.set_obstacle(mask)
s.place_cylinder(cx0, cy0, r0)
s.init_equilibrium(1.0, 0.1, 0.0)
tau, u = 0.505, 0.1
s.step(tau, u, 1)
n, total = 0, 0.0
while total < {seconds}:
    t0 = time.perf_counter(); s.step(tau, u, 1); total += time.perf_counter() - t0; n += 1
print(json.dumps({{'steps': n, 'seconds': total, 'threads': int(os.environ['OMP_NUM_THREADS']),
                  'glups': nx * ny * n / total / 1e9}}))
"""
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    runs = {}
    for name, kind, nthr in (("port", "port", threads), ("port_1thread", "port", 1),
                             ("reference_no_les", "reference", threads)):
        import oracle as O
        if kind == "reference" and not O.ref_available(fast=True):
            continue
        env = dict(os.environ)
        env["OMP_NUM_THREADS"] = str(nthr)
        env["OMP_PROC_BIND"] = "close"
        env["OMP_PLACES"] = "cores"
        out = subprocess.run([sys.executable, "-c", code, kind], capture_output=True, text=True, env=env,
                             timeout=600)
        if out.returncode != 0:
            runs[name] = {"error": out.stderr[-400:]}
            continue
        runs[name] = json.loads(out.stdout.strip().splitlines()[-1])
        log(f"cpu baseline {name}: {runs[name]['glups'] * 1e3:.1f} MLUPS on {nthr} threads")
    model = ""
    try:
        for line in subprocess.run(["lscpu"], capture_output=True, text=True).stdout.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except OSError:
        pass
    p = runs.get("port", {})
    if "glups" not in p:
        return {"error": p.get("error", "no run"), "runs": runs}
    sub = f"{CPU_SUB['nx']}x{CPU_SUB['ny']}"
    return {"value": p["glups"], "unit": "GLUPS", "cores": p["threads"], "kind": "port",
            "sample": f"{sub} sub-domain of the 32768x16384 LES workload ({CPU_SUB['n_circles']} circles + "
                      f"{CPU_SUB['n_plates']} plates, same obstacle density), FP64 AoS restatement of the "
                      f"step with the LES wall distance computed once (the reference recomputes it "
                      f"serially every step), {p['steps']} steps in {p['seconds']:.1f} s on "
                      f"{p['threads']} OpenMP threads",
            "one_thread_GLUPS": runs.get("port_1thread", {}).get("glups"),
            "reference_no_les": runs.get("reference_no_les"),
            "nproc": len(os.sched_getaffinity(0)), "cpu_count": os.cpu_count(), "cpu_model": model}


def profiled_traffic(launch, path=LARGE_PROFILE):
    """HBM bytes per launch of the headline kernel in this launch shape from the committed
    rocprofv3 PMC summary (FETCH_SIZE and WRITE_SIZE scaled by a plain-copy calibration);
    None when the summary holds no figure for the shape."""
    key = "traffic_bytes_per_launch[" + ",".join(str(int(v)) for v in launch) + "]:"
    try:
        for line in open(path):
            if line.strip().startswith(key):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def run_form(name, args, torch, stream, geom_cache):
    """One extra form by name (profiling runs use these alone)."""
    if name.startswith("large_"):
        if "geom" not in geom_cache:
            import workloads as W
            geom_cache["geom"] = W.large_domain()
        form = name[len("large_"):]
        return bench_large(geom_cache["geom"], form, args.steps, args.warmup, torch, stream,
                           launch=args.launch, graph_steps=args.graph_steps)
    if name == "batched_256":
        return bench_ensemble(max(args.steps, 200), 20, torch, stream, launch=args.launch)
    if name.startswith("thermal_"):
        storage, size = name.split("_")[1:3]
        nx, ny = (int(v) for v in size.split("x"))
        return bench_thermal(nx, ny, storage, 200, 20, torch, stream, launch=args.launch)
    if name == "f32_800x300":
        return bench_form(800, 300, 30, 0.68, "f32", 2000, 100, torch, stream, launch=args.launch)
    storage = name.split("_")[0]
    return bench_form(6400, 2400, 240, 0.644, storage, 200, 20, torch, stream, les=True,
                      launch=args.launch, sr="_sr_" in name)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100, help="timed steps of the headline workload")
    ap.add_argument("--warmup", type=int, default=10, help="untimed steps before them")
    ap.add_argument("--launch", default="", help="fix the step-kernel launch: bx,rows,nodes,options "
                                                   "(default: autotune)")
    ap.add_argument("--graph-steps", type=int, default=100,
                    help="steps per hipGraph (even; 100 is the library's default: an odd batch of two-step "
                         "launches ends with a one-step launch)")
    ap.add_argument("--no-forms", action="store_true", help="headline workload only")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--form", default="", help="run one workload only and print it (profiling): "
                                                "large_<f32|f16s|bf16s|f16s_sr|bf16s_sr> (takes --steps, "
                                                "--warmup, --launch), f32_6400x2400_les, f16s_6400x2400_les, "
                                                "f16s_sr_6400x2400_les, bf16s_6400x2400_les, f32_800x300 "
                                                "(these take --launch), batched_256, thermal_<f32|f64>_<nx>x<ny>")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB device to device (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported by this build "
              "(the lattice update is measured on a single MI355X).", file=sys.stderr)
        sys.exit(2)
    if args.steps < 1 or args.warmup < 0 or args.graph_steps < 2 or args.graph_steps % 2:
        print("--steps >= 1, --warmup >= 0 and an even --graph-steps >= 2 are required", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        import torch
        a = torch.ones(256 * 1024 * 1024, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": a.numel() * 4, "copy_bytes_written": b.numel() * 4}))
        return

    if args.warmup + args.steps + 120 > 340 and not args.form:
        log(f"warning: {args.warmup + args.steps + 120} steps of the 32768x16384 workload: it diverges "
            "after about 350 steps (FP64 too); the rate is unaffected, the NaN flag will be set")
    import torch
    import workloads as W
    torch.cuda.init()
    stream = torch.cuda.Stream()
    geom_cache = {}
    if args.form:
        with torch.cuda.stream(stream):
            out = run_form(args.form, args, torch, stream.cuda_stream, geom_cache)
        print(json.dumps({args.form: out}))
        return

    with torch.cuda.stream(stream):
        t0 = time.perf_counter()
        geom_cache["geom"] = W.large_domain()
        log(f"32768x16384 geometry ({len(geom_cache['geom'][1])} links) in {time.perf_counter() - t0:.1f} s")
        head = bench_large(geom_cache["geom"], "f32", args.steps, args.warmup, torch, stream.cuda_stream,
                           launch=args.launch, graph_steps=args.graph_steps, roofline=True)
        forms = {}
        if not args.no_forms:
            # the 16-bit forms with stochastic rounding (their accuracy form), then with
            # round-to-nearest; FP32 one step per launch, the best of a few shapes
            for form in ("f16s_sr", "bf16s_sr", "f16s"):
                forms[f"large_{form}"] = bench_large(geom_cache["geom"], form, 100, 10, torch, stream.cuda_stream)
            forms["large_f32_one_step"] = bench_large_one_step(geom_cache["geom"], torch, stream.cuda_stream)
            forms["fp16s_sr_over_fp32_large"] = forms["large_f16s_sr"]["GLUPS"] / head["GLUPS"]
            forms["fp16s_over_fp32_large"] = forms["large_f16s"]["GLUPS"] / head["GLUPS"]
            forms["fp16s_sr_over_fp32_one_step_large"] = (forms["large_f16s_sr"]["GLUPS"]
                                                          / forms["large_f32_one_step"]["GLUPS"])
            geom_cache.clear()
            for name in ("f32_6400x2400_les", "f16s_sr_6400x2400_les", "f16s_6400x2400_les", "bf16s_sr_6400x2400_les"):
                forms[name] = run_form(name, args, torch, stream.cuda_stream, geom_cache)
                log(f"{name}: {forms[name]['GLUPS']:.1f} GLUPS")
            forms["fp16s_sr_over_fp32_6400x2400"] = (forms["f16s_sr_6400x2400_les"]["GLUPS"]
                                                     / forms["f32_6400x2400_les"]["GLUPS"])
            forms["fp16s_over_fp32_6400x2400"] = forms["f16s_6400x2400_les"]["GLUPS"] / forms["f32_6400x2400_les"]["GLUPS"]
            forms["f32_800x300_cache_resident"] = bench_form(800, 300, 30, 0.68, "f32", 2000, 100, torch,
                                                             stream.cuda_stream)
            forms["f32_800x300_spec_boundaries"] = bench_form(800, 300, 30, 0.68, "f32", 2000, 100, torch,
                                                              stream.cuda_stream, spec=True)
            forms["batched_256"] = bench_ensemble(1000, 100, torch, stream.cuda_stream)
            log(f"batched ensemble: {forms['batched_256']['GLUPS']:.1f} GLUPS")
            forms["thermal_f32_6400x2400"] = bench_thermal(6400, 2400, "f32", 200, 20, torch, stream.cuda_stream)
            log("thermal done")

    cpu = None if args.no_cpu else cpu_baseline(args.cpu_seconds)
    two = head["steps_per_launch"] == 2
    bytes_launch = head["bytes_per_step"] * head["steps_per_launch"]
    achieved = bytes_launch / (head["launch_ms"] * 1e-3) / 1e9
    line = {
        "metric": "GLUPS + % HBM peak, D2Q9 MRT step: FP32 plain vs FP16-store vs batched",
        "value": head["GLUPS"],
        "unit": "GLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "32768x16384 domain, D2Q9 MRT + Smagorinsky/Van Driest LES (c_s=0.12), tau=0.505, "
                        "u=0.1, 128 seeded obstacles (64 circles r 20-60 with exact-circle Mei links, 64 plates "
                        "155x2 at AoA in [-15,15] deg, staircase bounce-back), Zou/He inlet, outflow copy, "
                        "periodic y, FP32 storage (f-w) and compute, uniform equilibrium start, "
                        f"hipGraphs of {args.graph_steps} steps",
            "cells": head["cells"], "fluid_cells": head["fluid_cells"], "launch": head["launch"],
            "setup_s": head["setup_s"], "nan": head["nan"], "steps_run": head["steps_run"],
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": profiled_traffic(head["launch"]),
            "kernel": (f"step2_bulk_kernel<1, {head['launch'][1]}, false, false, false> with its edge form "
                       f"(..., true) on the edge tiles and step2_kernel<1, {min(head['launch'][1], 16)}, false> "
                       f"on the other tiles" if two and head["bulk_tiles"] else
                       f"step2_kernel<1, {head['launch'][1]}, true>" if two else
                       "step_kernel<1, {}, {}, {}, true>".format(*head["launch"][:3])),
            "bulk_tiles": head["bulk_tiles"], "edge_tiles": head["edge_tiles"],
            "other_tiles": head["other_tiles"], "empty_tiles": head["empty_tiles"],
            "steps_per_launch": head["steps_per_launch"],
            "launch_ms": head["launch_ms"],
            "bytes_per_launch": bytes_launch,
        },
        "cpu_baseline": cpu,
        "forms": forms,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
