"""Writes tests/golden/cylinder800x300_re100_forces.npz: the per-step total force of
the reference solver (oracle/_ref, FP64) on the cylinder-wake lattice of the configs
(800x300, circle r=30 at (200,150), tau=0.68, u=0.1, seed-42 perturbation) for 50,000
steps.  Used by the GPU accuracy tests for Cd / Cl / Strouhal of the reduced-precision
forms.  Runs where oracle/_ref was built; takes several minutes.

    python tests/golden/make_long_forces.py [steps]
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
import oracle as O  # noqa: E402


def main():
    steps = int(sys.argv[1]) if len(sys.argv) > 1 else 50000
    if not O.ref_available():
        sys.exit("oracle/_ref is missing: run `make -C oracle` where the reference is present")
    # the -ffp-contract=off build: bit-identical to the FP64 oracle
    r = O.RefLattice(800, 300, O.CYLINDER)
    r.place_cylinder(200, 150, 30)
    r.init_equilibrium(1.0, 0.1, 0.0)
    r.perturb_seed42(205, 260)
    fx = np.zeros(steps)
    fy = np.zeros(steps)
    t0 = time.time()
    for s in range(steps):
        r.step(0.68, 0.1)
        fx[s], fy[s] = r.total_force()
        if s % 5000 == 0:
            print(s, fx[s], fy[s], f"{time.time() - t0:.0f}s", flush=True)
    f = r.f
    np.savez_compressed(os.path.join(HERE, "cylinder800x300_re100_forces.npz"), fx=fx, fy=fy, steps=steps,
                        final_sum=f.sum(), final_sumsq=(f * f).sum(),
                        final_rows=f[::20, ::20].copy())


if __name__ == "__main__":
    main()
