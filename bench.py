"""Throughput of the D2Q9 MRT lattice update on one MI355X.

Headline (``value``): GLUPS of the FP32 fused MRT step on the cylinder-wake
lattice, 800x300, circle r=30 at (200,150), Re=100 (tau=0.68), u=0.1, seed-42
perturbation, Mei interpolated bounce-back, Zou/He inlet, outflow copy,
periodic y -- the reference's semantics (src/main.cpp, src/lbm.hpp).  Inputs are
resident in HBM; the timed region is K steps replayed from hipGraphs of 100.

Extra forms on the same JSON line (``forms``): FP32 and FP16-shifted storage on
the 6400x2400 LES cylinder, the 256-lattice batched ensemble against the same
lattices stepped one launch each, and the 32768x16384 domain when
``--large`` is given.

``roofline`` is for the dominant kernel (the fused step kernel of the headline
workload): algorithmic bytes per launch = N_fluid*18*s + N_cells*1 (s = bytes per
stored population; + 2 B per node of the u16 wall distance with LES), divided by the
average launch duration measured here with HIP events on the library's stream
around hipGraph replays, against 8 TB/s.  ``traffic`` is the HBM bytes per launch
read from the committed rocprofv3 PMC summary of the same workload.

``cpu_baseline`` times the reference itself (oracle/_ref/libref_fast.so, built
from the reference sources) or, when that is absent, the FP64 oracle, on this
host's cores on a bounded sample of the same workload, in a child process.
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


def cylinder_lattice(nx, ny, cx, cy, r, tau, u, storage, stream, les=False, cs=0.1, perturb=True):
    import lbm2d as G
    d = G.desc_for_case(G.CYLINDER, nx, ny, tau=tau, u_in=u, les=les, cs=cs)
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


def algorithmic_bytes(info, les_bytes):
    s = info.bytes_per_pop
    return info.fluid_cells * 18 * s + info.cells * 1 + info.les_cells * les_bytes


def time_steps(ens, steps, warmup, torch, graph_steps=100):
    # the first call of graph_steps or more steps captures the hipGraph: keep it untimed
    ens.step(max(warmup, graph_steps))
    ens.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ens.step(steps)
    ens.sync()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def graph_step_ms(ens, torch, replays=20, graph_steps=100):
    """Average duration of one step-kernel launch: HIP events on the library's stream
    around each replay of a hipGraph of graph_steps launches, whose kernels run back to
    back (the per-step counter update is one 1-thread kernel per replay)."""
    stream = torch.cuda.current_stream()
    ens.step(graph_steps)
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(replays + 1)]
    evs[0].record(stream)
    for i in range(replays):
        ens.step(graph_steps)
        evs[i + 1].record(stream)
    torch.cuda.synchronize()
    return float(np.mean([evs[i].elapsed_time(evs[i + 1]) for i in range(replays)])) / graph_steps


def bench_form(nx, ny, r, tau, storage, steps, warmup, torch, stream, les=False, cs=0.1, launch="", sr=False):
    ens, mask = cylinder_lattice(nx, ny, nx // 4, ny // 2, r, tau, 0.1, storage, stream,
                                 les=les, cs=cs, perturb=False)
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
    b = algorithmic_bytes(info, 2)
    m0 = ens.mass(0)
    out = {
        "lattice": f"{nx}x{ny}", "storage": storage, "les": les, "rounding": "stochastic" if sr else "nearest-even",
        "GLUPS": info.cells * steps / t / 1e9,
        "ms_per_step": 1e3 * t / steps,
        "kernel_ms": kms,
        "bytes_per_step": b,
        "hbm_frac": b / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "launch": list(shape),
        "mass": m0,
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
    T = ens.temperature(0, ny // 2, ny // 2 + 1)
    out = {
        "lattice": f"{nx}x{ny}", "storage": storage, "tau": tau,
        "GLUPS": info.cells * steps / t / 1e9,
        "ms_per_step": 1e3 * t / steps,
        "kernel_ms": kms,
        "bytes_per_step": b,
        "hbm_frac": b / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "T_centre_row_max": float(T.max()),
        "launch_options": info.launch_options,
    }
    ens.close()
    return out


def ensemble_descs(seed=42, n=256):
    """Batched-ensemble lattices: NX~U[256,1024], NY~U[96,384], 1-4 circles r~U[8,40],
    centres uniform with a 2r margin, non-overlapping; Re log-uniform in [50,400];
    tau = 0.5 + 3*0.1*D/Re with D the largest diameter.  numpy Generator(PCG64(seed))."""
    import lbm2d as G
    rng = np.random.default_rng(seed)
    lats = []
    for _ in range(n):
        nx = int(rng.integers(256, 1025))
        ny = int(rng.integers(96, 385))
        k = int(rng.integers(1, 5))
        circles = []
        tries = 0
        while len(circles) < k and tries < 200:
            tries += 1
            r = int(rng.integers(8, 41))
            if 4 * r >= ny - 2 or 4 * r >= nx - 2:
                continue
            cx = int(rng.integers(2 * r, nx - 2 * r))
            cy = int(rng.integers(2 * r, ny - 2 * r))
            if all((cx - a) ** 2 + (cy - b) ** 2 > (r + c + 2) ** 2 for a, b, c in circles):
                circles.append((cx, cy, r))
        re = float(np.exp(rng.uniform(np.log(50), np.log(400))))
        dmax = 2 * max(c[2] for c in circles)
        tau = 0.5 + 3 * 0.1 * dmax / re
        d = G.desc_for_case(G.CYLINDER, nx, ny, tau=tau, u_in=0.1)
        mask = np.zeros((ny, nx), np.uint8)
        for cx, cy, r in circles:
            mask |= G.circle_mask(nx, ny, cx, cy, r)
        lats.append((d, mask, circles))
    return lats


def ensemble_links(mask, circles, nx, ny):
    """Per-obstacle exact-circle q for every link (nearest circle owns the link)."""
    import lbm2d as G
    links = []
    ys, xs = np.nonzero(mask == 0)
    solid = mask.astype(bool)
    for i in range(1, 9):
        tx = xs + G.CXV[i]
        ty = (ys + G.CYV[i]) % ny
        ok = (tx >= 0) & (tx < nx)
        sel = np.nonzero(ok)[0]
        hit = sel[solid[ty[sel], tx[sel]]]
        for k in hit:
            x, y = int(xs[k]), int(ys[k])
            sx, sy = x + G.CXV[i], y + G.CYV[i]
            c = min(circles, key=lambda c: (sx - c[0]) ** 2 + (sy - c[1]) ** 2)
            q = _q_circle(c[0], c[1], c[2], x, y, i)
            links.append((y * nx + x, i, G.BB_MEI, 0, q, 0.0, 0.0))
    return links


def _q_circle(ccx, ccy, r, xf, yf, i):
    import lbm2d as G
    dx, dy = xf - ccx, yf - ccy
    ex, ey = float(G.CXV[i]), float(G.CYV[i])
    a = ex * ex + ey * ey
    b = 2.0 * (dx * ex + dy * ey)
    c = dx * dx + dy * dy - r * r
    disc = b * b - 4.0 * a * c
    if disc < 0.0:
        return 1.0
    sd = np.sqrt(disc)
    t1, t2 = (-b - sd) / (2 * a), (-b + sd) / (2 * a)
    if 0.0 < t1 <= 1.0:
        return float(t1)
    if 0.0 < t2 <= 1.0:
        return float(t2)
    return 1.0


def bench_ensemble(steps, warmup, torch, stream, n=256):
    import lbm2d as G
    lats = ensemble_descs(n=n)
    ens = G.Ensemble([d for d, _, _ in lats], storage="f32", stream=stream)
    for l, (d, mask, circles) in enumerate(lats):
        ens.set_obstacle(l, mask)
        ens.set_links(l, ensemble_links(mask, circles, d.nx, d.ny))
        f = G.equilibrium(np.ones((d.ny, d.nx)), np.full((d.ny, d.nx), 0.1), np.zeros((d.ny, d.nx)))
        ens.upload(l, f)
    shape, _ = ens.autotune()
    info = ens.info()
    t = time_steps(ens, steps, warmup, torch)
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
    b = algorithmic_bytes(info, 0)
    glups = info.cells * steps / t / 1e9
    glups_plain = info.cells * plain_steps / tp / 1e9
    # per-lattice Cd = 2 Fx / (D u^2), D the largest diameter, mean over the last 100 steps
    # run (the flows are still starting up: these are early-time values)
    cds = []
    for l, (d, _, circles) in enumerate(lats):
        fxl, _ = ens.forces(l)
        dmax = 2 * max(c[2] for c in circles)
        cds.append(float(2.0 * fxl[-100:, 0].mean() / (dmax * 0.01)))
    out = {
        "lattices": n, "cells": info.cells, "storage": "f32",
        "GLUPS": glups, "ms_per_step": 1e3 * t / steps,
        "hbm_frac": b / (t / steps) / 1e9 / HBM_PEAK_GBS,
        "bytes_per_step": b,
        "plain_per_lattice_GLUPS": glups_plain,
        "speedup_vs_plain": glups / glups_plain,
        "launch": list(shape),
        "cd_per_lattice": [float(f"{c:.4g}") for c in cds],
    }
    ens.close()
    return out


def large_domain(seed=42, nx=32768, ny=16384, n_circles=64, n_plates=64):
    """Obstacles of the 32768x16384 domain: n_circles circles r~U[20,60] (exact-circle
    Mei links, force slots 1..n_circles) and n_plates flat plates 155x2 at angle of attack
    ~U[-15,15] deg (polygon staircase, simple bounce-back, force slot 0).  Centres uniform
    in x in [nx/16, 15nx/16) and in y clear of the periodic seam, bounding circles at least
    8 cells apart; numpy Generator(PCG64(seed)), circles first.  Returns the mask and the
    explicit link list."""
    import lbm2d as G
    rng = np.random.default_rng(seed)
    mask = np.zeros((ny, nx), np.uint8)
    placed, links = [], []
    half_diag = float(np.hypot(77.5, 1.0))
    for k in range(n_circles + n_plates):
        circle = k < n_circles
        while True:
            if circle:
                r = int(rng.integers(20, 61))
                br = float(r)
            else:
                aoa = float(rng.uniform(-15.0, 15.0))
                br = half_diag
            cx = int(rng.integers(nx // 16, nx - nx // 16))
            cy = int(rng.integers(int(br) + 4, ny - int(br) - 4))
            if all((cx - a) ** 2 + (cy - b) ** 2 > (br + c + 8) ** 2 for a, b, c in placed):
                break
        placed.append((cx, cy, br))
        x0, y0 = cx - int(br) - 2, cy - int(br) - 2
        w = h = 2 * int(br) + 5
        if circle:
            yy, xx = np.mgrid[y0:y0 + h, x0:x0 + w]
            dx, dy = (xx - cx).astype(float), (yy - cy).astype(float)
            mask[y0:y0 + h, x0:x0 + w] |= (np.sqrt(dx * dx + dy * dy) < r).astype(np.uint8)
            sub = mask[y0:y0 + h, x0:x0 + w]
            ys, xs = np.nonzero(sub[1:-1, 1:-1] == 0)
            ys, xs = ys + 1, xs + 1
            for i in range(1, 9):
                sx, sy = (xs + G.CXV[i] + x0 - cx).astype(float), (ys + G.CYV[i] + y0 - cy).astype(float)
                hit = (sub[ys + G.CYV[i], xs + G.CXV[i]] != 0) & (np.sqrt(sx * sx + sy * sy) < r)
                for yl, xl in zip(ys[hit], xs[hit]):
                    x, y = x0 + int(xl), y0 + int(yl)
                    links.append((y * nx + x, i, G.BB_MEI, 1 + k, _q_circle(cx, cy, r, x, y, i), 0.0, 0.0))
        else:
            xy = G.plate_polygon(float(cx), float(cy), 155.0, aoa, 2.0)
            mask[y0:y0 + h, x0:x0 + w] |= G.polygon_mask(w, h, xy - np.array([x0, y0], float))
    return mask, links


def bench_large(steps, warmup, torch, stream, horizon=100, seed=42):
    """The 32768x16384 LES domain (tau=0.505, u=0.1, cs=0.12, periodic y): FP64 / FP32 /
    FP16S / BF16S lattices from the same geometry and a uniform equilibrium start; after
    `horizon` steps the velocity field of each reduced form against FP32 and against the
    FP64 form (RMS and max of |u - u_ref| over fluid nodes, in units of u_in); then
    autotuned throughput of each reduced form."""
    import lbm2d as G
    nx, ny, tau, u, cs = 32768, 16384, 0.505, 0.1, 0.12
    t0 = time.perf_counter()
    mask, links = large_domain(seed, nx, ny)
    t_geom = time.perf_counter() - t0
    n_slots = 1 + max(k[3] for k in links)
    lat = {}
    for form in ("f64", "f32", "f16s", "bf16s", "f16s_sr", "bf16s_sr"):
        storage = form.split("_")[0]
        t1 = time.perf_counter()
        d = G.desc_for_case(G.CYLINDER, nx, ny, tau=tau, u_in=u, les=True, cs=cs)
        e = G.Ensemble([d], storage=storage, stream=stream)
        if form.endswith("_sr"):
            e.set_stochastic_rounding(True, 1)
        e.set_obstacle(0, mask)
        e.set_force_slots(0, n_slots)
        e.set_links(0, links)
        e.init_equilibrium(0, 1.0, u, 0.0)
        if storage == "f64":
            m0 = e.mass(0)
        e.step(horizon)
        e.sync()
        lat[form] = e
        print(f"# 32768x16384 {form}: built and stepped {horizon} in {time.perf_counter() - t1:.1f} s",
              file=sys.stderr, flush=True)
    out = {"cells": nx * ny, "obstacles": "64 circles r 20-60 (Mei links) + 64 plates 155x2 (staircase)",
           "seed": seed, "links": len(links), "host_geometry_s": t_geom, "horizon": horizon}
    strip = 1024
    fluid = mask == 0
    out["fp64_mass_drift"] = lat["f64"].mass(0) / m0 - 1.0
    pairs = [(s, r) for r in ("f64", "f32") for s in ("f32", "f16s", "bf16s", "f16s_sr", "bf16s_sr") if s != r]
    acc = {p: [0.0, 0.0, 0] for p in pairs}
    for y0 in range(0, ny, strip):
        y1 = min(ny, y0 + strip)
        fl = fluid[y0:y1]
        vel = {k: e.macros(0, y0, y1)[1:] for k, e in lat.items()}
        for s, r in pairs:
            d2 = ((vel[s][0] - vel[r][0]) ** 2 + (vel[s][1] - vel[r][1]) ** 2)[fl]
            a = acc[(s, r)]
            a[0] += float(d2.sum())
            a[1] = max(a[1], float(np.sqrt(d2.max())) if d2.size else 0.0)
            a[2] += int(d2.size)
        if y0 % 4096 == 0:
            print(f"# 32768x16384 velocity errors: rows {y1}/{ny}", file=sys.stderr, flush=True)
    for (s, r), (sq, mx, n) in acc.items():
        out[f"vel_err_{s}_vs_{r}"] = {"rms": float(np.sqrt(sq / n)) / u, "max": mx / u}
    lat.pop("f64").close()
    for s in ("f32", "f16s", "bf16s", "f16s_sr", "bf16s_sr"):
        e = lat.pop(s)
        shape, _ = e.autotune()
        info = e.info()
        t = time_steps(e, steps, warmup, torch)
        kms = graph_step_ms(e, torch, replays=2, graph_steps=20)
        b = algorithmic_bytes(info, 2)
        out[s] = {"GLUPS": info.cells * steps / t / 1e9, "ms_per_step": 1e3 * t / steps, "kernel_ms": kms,
                  "bytes_per_step": b, "hbm_frac": b / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                  "launch": list(shape), "fluid_cells": info.fluid_cells, "nan": e.nan_flag(0)}
        e.close()
        print(f"# 32768x16384 {s}: {out[s]['GLUPS']:.1f} GLUPS", file=sys.stderr, flush=True)
    return out


def cpu_baseline(seconds=12.0):
    """Child-process timing of the reference (or the oracle) on the cylinder workload."""
    code = f"""
import json, os, sys, time
sys.path.insert(0, {ROOT!r})
import oracle as O
threads = int(os.environ.get('OMP_NUM_THREADS', os.cpu_count()))
kind = 'reference' if O.ref_available(fast=True) else 'port'
nx, ny, tau, u = 800, 300, 0.68, 0.1
if kind == 'reference':
    s = O.RefLattice(nx, ny, O.CYLINDER, fast=True)
else:
    s = O.Oracle(nx, ny, O.CYLINDER, faithful=True)
s.place_cylinder(200, 150, 30)
s.init_equilibrium(1.0, u, 0.0)
s.perturb_seed42(205, 260)
s.step(tau, u, 2)
n, total = 0, 0.0
chunk = 20
while total < {seconds}:
    t0 = time.perf_counter(); s.step(tau, u, chunk); total += time.perf_counter() - t0; n += chunk
print(json.dumps({{'kind': kind, 'steps': n, 'seconds': total, 'threads': threads,
                  'mlups': nx * ny * n / total / 1e6}}))
"""
    env = dict(os.environ)
    env.setdefault("OMP_NUM_THREADS", str(os.cpu_count()))
    env["OMP_PROC_BIND"] = "close"
    env["OMP_PLACES"] = "cores"
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env,
                         timeout=300)
    if out.returncode != 0:
        return {"error": out.stderr[-400:]}
    r = json.loads(out.stdout.strip().splitlines()[-1])
    return {"value": r["mlups"] / 1e3, "unit": "GLUPS", "cores": r["threads"], "kind": r["kind"],
            "sample": f"800x300 cylinder MRT, {r['steps']} steps in {r['seconds']:.1f} s, "
                      f"FP64 AoS, OpenMP {r['threads']} threads"}


def profiled_traffic(path=os.path.join(ROOT, "profiles", "step_kernel_cylinder800x300_f32.md")):
    """HBM bytes per launch of the headline kernel from the committed rocprofv3 PMC
    summary (FETCH_SIZE and WRITE_SIZE scaled by a plain-copy calibration)."""
    try:
        for line in open(path):
            if line.strip().startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50000)
    ap.add_argument("--warmup", type=int, default=1000)
    ap.add_argument("--launch", default="", help="fix the step-kernel launch: bx,rows,nodes,options "
                                                   "(default: autotune)")
    ap.add_argument("--large", action="store_true", help="also time the 32768x16384 domain")
    ap.add_argument("--no-forms", action="store_true", help="headline workload only")
    ap.add_argument("--graph-steps", type=int, default=500,
                    help="launches per hipGraph of the headline run (forces still recorded every step)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--form", default="", help="run one extra form only and print it (profiling): "
                                                "f32_6400x2400_les, f16s_6400x2400_les, f16s_sr_6400x2400_les, "
                                                "bf16s_6400x2400_les "
                                                "(6400x2400 forms take --launch), "
                                                "batched_256, large_32768x16384, "
                                                "thermal_<f32|f64>_<nx>x<ny>")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB device to device (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        a = torch.ones(256 * 1024 * 1024, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": a.numel() * 4, "copy_bytes_written": b.numel() * 4}))
        return
    if args.form:
        import torch
        torch.cuda.init()
        stream = torch.cuda.Stream()
        with torch.cuda.stream(stream):
            if args.form == "batched_256":
                out = bench_ensemble(200, 20, torch, stream.cuda_stream)
            elif args.form == "large_32768x16384":
                out = bench_large(100, 10, torch, stream.cuda_stream)
            elif args.form.startswith("thermal_"):
                storage, size = args.form.split("_")[1:3]
                nx, ny = (int(v) for v in size.split("x"))
                out = bench_thermal(nx, ny, storage, 100, 10, torch, stream.cuda_stream, launch=args.launch)
            else:
                storage = args.form.split("_")[0]
                out = bench_form(6400, 2400, 240, 0.644, storage, 100, 10, torch, stream.cuda_stream, les=True,
                                 launch=args.launch, sr="_sr_" in args.form)
        print(json.dumps({args.form: out}))
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported by this build "
              "(the lattice update is measured on a single MI355X).", file=sys.stderr)
        sys.exit(2)
    gs = args.graph_steps
    if args.steps % gs != 0:
        args.steps = max(gs, args.steps // gs * gs)

    import torch
    import lbm2d as G
    torch.cuda.init()
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        ens, mask = cylinder_lattice(800, 300, 200, 150, 30, 0.68, 0.1, "f32", stream.cuda_stream)
        if args.launch:
            shape = tuple(int(v) for v in args.launch.split(","))
            ens.set_launch(*shape)
        else:
            shape, _ = ens.autotune()
        ens.set_launch(graph_steps=gs)
        info = ens.info()
        print(f"# headline lattice ready, launch {shape}", file=sys.stderr, flush=True)
        t = time_steps(ens, args.steps, args.warmup, torch, gs)
        glups = info.cells * args.steps / t / 1e9
        kms = graph_step_ms(ens, torch, replays=max(2, 2000 // gs), graph_steps=gs)
        bytes_launch = algorithmic_bytes(info, 0)
        fx, fy = ens.forces(0)
        cd_last = 2.0 * fx[-1, 0] / (60 * 0.01)
        ens.close()
        print(f"# headline done: {glups:.1f} GLUPS", file=sys.stderr, flush=True)

        forms = {}
        if not args.no_forms:
            big_steps = 200
            for storage in ("f32", "f16s", "bf16s"):
                forms[f"{storage}_6400x2400_les"] = bench_form(6400, 2400, 240, 0.644, storage, big_steps, 20,
                                                               torch, stream.cuda_stream, les=True)
                print(f"# {storage} 6400x2400 done", file=sys.stderr, flush=True)
            forms["f16s_sr_6400x2400_les"] = bench_form(6400, 2400, 240, 0.644, "f16s", big_steps, 20, torch,
                                                        stream.cuda_stream, les=True, sr=True)
            forms["fp16s_over_fp32"] = forms["f16s_6400x2400_les"]["GLUPS"] / forms["f32_6400x2400_les"]["GLUPS"]
            # FP32 restricted to one step per launch (the autotuner may pick two)
            forms["f32_6400x2400_les_one_step"] = bench_form(6400, 2400, 240, 0.644, "f32", big_steps, 20, torch,
                                                             stream.cuda_stream, les=True, launch="256,1,1,0")
            forms["fp16s_over_fp32_one_step"] = (forms["f16s_6400x2400_les"]["GLUPS"] /
                                                 forms["f32_6400x2400_les_one_step"]["GLUPS"])
            forms["batched_256"] = bench_ensemble(1000, 100, torch, stream.cuda_stream)
            print("# batched ensemble done", file=sys.stderr, flush=True)
            forms["thermal_f32_800x300"] = bench_thermal(800, 300, "f32", 2000, 100, torch, stream.cuda_stream)
            forms["thermal_f32_6400x2400"] = bench_thermal(6400, 2400, "f32", 200, 20, torch, stream.cuda_stream)
            print("# thermal done", file=sys.stderr, flush=True)
            if args.large:
                forms["large_32768x16384_les"] = bench_large(100, 10, torch, stream.cuda_stream)

    cpu = None if args.no_cpu else cpu_baseline()
    achieved = bytes_launch / (kms * 1e-3) / 1e9
    line = {
        "metric": "GLUPS + % HBM peak, D2Q9 MRT step: FP32 plain vs FP16-store vs batched",
        "value": glups,
        "unit": "GLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "cylinder wake 800x300, D2Q9 MRT, FP32 storage (f-w) and compute, "
                        "r=30 at (200,150), Re=100 tau=0.68 u=0.1, Mei interpolated bounce-back, "
                        "Zou/He inlet, outflow copy, periodic y, seed-42 perturbation, "
                        f"hipGraph of {gs} steps",
            "cells": info.cells, "fluid_cells": info.fluid_cells, "launch": list(shape),
            "cd_last_step": cd_last,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": profiled_traffic(),
            "kernel_ms": kms,
            "bytes_per_launch": bytes_launch,
        },
        "cpu_baseline": cpu,
        "forms": forms,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
