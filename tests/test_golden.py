"""The FP64 oracle against golden fixtures written by the reference solver
(tests/golden/make_golden.py): bit-for-bit populations, forces and mass."""
import os

import numpy as np
import pytest

import oracle as O
from cases import CASES, build_oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _record(obj, steps, tau, u):
    fx, fy = [], []
    for _ in range(steps):
        obj.step(tau, u)
        a, b = obj.total_force()
        fx.append(a)
        fy.append(b)
    f = obj.f
    return np.array(fx), np.array(fy), f.copy(), float(f[obj.obstacle == 0].sum())


def _check(g, fx, fy, f, mass, rows):
    np.testing.assert_array_equal(fx, g["fx"])
    np.testing.assert_array_equal(fy, g["fy"])
    np.testing.assert_array_equal(rows, g["rows"])
    np.testing.assert_array_equal(f.sum(axis=(0, 1)), g["colsum"])
    np.testing.assert_array_equal((f * f).sum(axis=(0, 1)), g["colsq"])
    assert mass == float(g["mass"])


@pytest.mark.parametrize("name", sorted(CASES))
def test_case_matches_reference_fixture(name):
    c = CASES[name]
    g = np.load(os.path.join(GOLD, f"case_{name}.npz"))
    assert int(g["steps"]) == c.steps
    fx, fy, f, mass = _record(build_oracle(c), c.steps, c.tau, c.u)
    _check(g, fx, fy, f, mass, f[::4])


@pytest.mark.parametrize("tag,cy", [("driver", 151), ("spec", 150)])
def test_cylinder_800x300(tag, cy):
    g = np.load(os.path.join(GOLD, f"cylinder800x300_{tag}.npz"))
    o = O.Oracle(800, 300, O.CYLINDER)
    o.place_cylinder(200, cy, 30)
    o.init_equilibrium(1.0, 0.1, 0.0)
    o.perturb_seed42(205, 260)
    fx, fy, f, mass = _record(o, int(g["steps"]), 0.68, 0.1)
    _check(g, fx, fy, f, mass, f[::4][::8, ::4])
    # the cylinder driver's first-step drag coefficient, 2 Fx / (D u^2)
    if tag == "driver":
        assert abs(2 * fx[0] / (60 * 0.01) - 23.877556) < 1e-6


def test_cavity_256():
    g = np.load(os.path.join(GOLD, "cavity256.npz"))
    o = O.Oracle(256, 256, O.CAVITY)
    o.place_walls()
    o.init_equilibrium(1.0, 0.0, 0.0)
    fx, fy, f, mass = _record(o, int(g["steps"]), 1.268, 0.1)
    _check(g, fx, fy, f, mass, f[::4][::4, ::4])
    # the lid injects no mass: 64770 fluid nodes at rho = 1
    assert abs(mass - 64770.0) < 1e-8
