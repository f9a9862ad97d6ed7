"""The reference solver interface (lbm-2d_amd/include/lbm.hpp) on the GPU: a case
driver written against it (examples/cylinder_wake.cpp) against the FP64 oracle."""
import json
import os
import subprocess

import numpy as np
import pytest

import lbm2d as G
import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "lbm-2d_amd", "examples", "bin", "cylinder_wake")


def _driver(nx, ny, cx, cy, r, tau, steps, storage, mode):
    out = subprocess.run([EXE, str(nx), str(ny), str(cx), str(cy), str(r), str(tau), str(steps), storage, mode],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    lines = [json.loads(l) for l in out.stdout.splitlines()]
    return lines[:-1], lines[-1]


def _oracle(nx, ny, cx, cy, r, tau, steps):
    o = O.Oracle(nx, ny, O.CYLINDER)
    o.place_cylinder(cx, cy, r)
    o.init_equilibrium(1.0, 0.1, 0.0)
    o.perturb_seed42(cx + 5, min(nx, cx + 60))
    fx = []
    for _ in range(steps):
        o.step(tau, 0.1)
        fx.append(o.total_force())
    f = o.f
    return np.array(fx), float(f.sum()), float((f * f).sum())


@pytest.mark.parametrize("mode", ["step", "run"])
def test_cylinder_driver_f64_matches_oracle(mode):
    args = (96, 48, 24, 25, 6, 0.56, 40)
    per_step, chk = _driver(*args, "f64", mode)
    fx, s, sq = _oracle(*args)
    assert len(per_step) == args[-1]
    got = np.array([[p["fx"], p["fy"]] for p in per_step])
    np.testing.assert_allclose(got, fx, rtol=0, atol=1e-10)
    assert abs(chk["sum"] - s) < 1e-9 and abs(chk["sumsq"] - sq) < 1e-9
    # the driver's C++ force analysis (analysis.hpp) on its own history = lbm2d's on the oracle's
    cd, cl = G.coefficients(fx[:, 0], fx[:, 1], 12.0, 0.1)
    # (forces agree to 1e-10, i.e. coefficients to 2e-9)
    assert chk["cd_mean"] == pytest.approx(G.late_stats(cd, cl)[0], rel=0, abs=2e-9)
    assert chk["cl_amp"] == pytest.approx(G.late_stats(cd, cl)[1], rel=0, abs=2e-9)


def test_cylinder_driver_f32_close_to_oracle():
    args = (96, 48, 24, 25, 6, 0.56, 40)
    per_step, chk = _driver(*args, "f32", "run")
    fx, s, sq = _oracle(*args)
    got = np.array([[p["fx"], p["fy"]] for p in per_step])
    assert np.abs(got - fx).max() < 1e-4 * (np.abs(fx).max() + 1e-3)


NC = os.path.join(ROOT, "lbm-2d_amd", "examples", "bin", "natural_convection")


def _convection_oracle(n, ra, steps):
    import dataclasses
    from cases import THERMAL_CASES, build_thermal_oracle, thermal_step
    tau, beta = 0.8, 0.001
    nu = (tau - 0.5) / 3.0
    alpha = nu / 0.71
    base = THERMAL_CASES["convection"]
    assert base.base.nx == n and base.omega_k == 1.0 / (0.5 + 3.0 * alpha)
    tc = dataclasses.replace(base, g_buoyancy=-(ra * nu * alpha / (1.0 * n ** 3)) / beta, steps=steps)
    o = build_thermal_oracle(tc)
    for _ in range(steps):
        thermal_step(o, tc)
    t = o.g.sum(axis=-1)
    nu_avg = float(np.mean(-(t[1:n - 1, 2] - t[1:n - 1, 1]) * n / 1.0))
    return o, nu_avg


@pytest.mark.parametrize("mode", ["step", "run"])
def test_natural_convection_driver_matches_oracle(mode):
    n, ra, steps = 32, 1e5, 120
    out = subprocess.run([NC, str(n), str(ra), str(steps), "f64", mode], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    got = json.loads(out.stdout.splitlines()[-1])
    o, nu_avg = _convection_oracle(n, ra, steps)
    print(f"\n[natural convection {mode}] Nu {got['nu']:.12f} (oracle {nu_avg:.12f})")
    assert abs(got["nu"] - nu_avg) <= 1e-9 * max(1.0, abs(nu_avg))
    assert abs(got["fsum"] - o.f.sum()) <= 1e-9 * abs(o.f.sum())
    assert abs(got["gsum"] - o.g.sum()) <= 1e-9 * abs(o.g.sum())
    assert abs(got["gsumsq"] - (o.g * o.g).sum()) <= 1e-9 * (o.g * o.g).sum()
