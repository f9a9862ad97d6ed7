"""Writes the golden fixtures in this directory from the reference solver itself.

Runs only where oracle/_ref/libref_check.so was built from the reference sources
(make -C oracle).  For every small case of tests/cases.py it records, from the
reference's execute_time_step:
  - the total force (sum of fx_body, fy_body in node order) after every step,
  - the FP64 mass over fluid nodes after the last step,
  - every 4th row of the final populations and per-direction column sums and
    sums of squares of all of them;
for every thermal case, per-step forces and total heat and the final f and g rows;
plus the cylinder-wake lattice at full size (800x300, the driver's (200,151)
circle and the (200,150) one) and the 256x256 cavity over a short horizon.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "tests"), os.path.join(ROOT, "lbm-2d_amd")]

import oracle as O  # noqa: E402
from cases import CASES, THERMAL_CASES, build_oracle, build_thermal_oracle, thermal_step  # noqa: E402


def record(obj, steps, tau, u):
    fx, fy = [], []
    for _ in range(steps):
        obj.step(tau, u)
        a, b = obj.total_force()
        fx.append(a)
        fy.append(b)
    f = obj.f.copy()
    fluid = obj.obstacle == 0
    mass = float(f[fluid].sum())
    return {
        "fx": np.array(fx), "fy": np.array(fy), "mass": np.array(mass),
        "rows": f[::4].copy(),
        "colsum": f.sum(axis=(0, 1)), "colsq": (f * f).sum(axis=(0, 1)),
    }


def record_thermal(obj, tc):
    """Per-step total force and heat (sum of g over every node) with the driver's wall
    loop after each step; every 4th row of the final f and g and their column sums."""
    fx, fy, heat = [], [], []
    for _ in range(tc.steps):
        thermal_step(obj, tc)
        a, b = obj.total_force()
        fx.append(a)
        fy.append(b)
        heat.append(float(np.cumsum(obj.g.ravel())[-1]))
    f, g = obj.f.copy(), obj.g.copy()
    return {"fx": np.array(fx), "fy": np.array(fy), "heat": np.array(heat),
            "rows": f[::4].copy(), "grows": g[::4].copy(),
            "colsum": f.sum(axis=(0, 1)), "gcolsum": g.sum(axis=(0, 1))}


def main():
    if not O.ref_available():
        sys.exit("oracle/_ref/libref_check.so is missing: run `make -C oracle` where the reference is present")
    for name, c in sorted(CASES.items()):
        r = build_oracle(c, cls=O.RefLattice)
        rec = record(r, c.steps, c.tau, c.u)
        np.savez_compressed(os.path.join(HERE, f"case_{name}.npz"), steps=c.steps, **rec)
        print(name, c.steps, rec["fx"][-1], rec["mass"])
    for name, tc in sorted(THERMAL_CASES.items()):
        r = build_thermal_oracle(tc, cls=O.RefLattice)
        rec = record_thermal(r, tc)
        np.savez_compressed(os.path.join(HERE, f"thermal_{name}.npz"), steps=tc.steps, **rec)
        print("thermal", name, tc.steps, rec["heat"][-1])
    # full-size cylinder wake (configs' 800x300, tau = 0.68), 20 steps, both centres
    for tag, cy in (("driver", 151), ("spec", 150)):
        r = O.RefLattice(800, 300, O.CYLINDER)
        r.place_cylinder(200, cy, 30)
        r.init_equilibrium(1.0, 0.1, 0.0)
        r.perturb_seed42(205, 260)
        rec = record(r, 20, 0.68, 0.1)
        rec["rows"] = rec["rows"][::8, ::4]
        np.savez_compressed(os.path.join(HERE, f"cylinder800x300_{tag}.npz"), steps=20, cy=cy, **rec)
        print(tag, "Cd step0", 2 * rec["fx"][0] / (60 * 0.01))
    # 256x256 cavity, tau = 1.268, 200 steps
    r = O.RefLattice(256, 256, O.CAVITY)
    r.place_walls()
    r.init_equilibrium(1.0, 0.0, 0.0)
    rec = record(r, 200, 1.268, 0.1)
    rec["rows"] = rec["rows"][::4, ::4]
    np.savez_compressed(os.path.join(HERE, "cavity256.npz"), steps=200, **rec)
    print("cavity mass", rec["mass"])


if __name__ == "__main__":
    main()
