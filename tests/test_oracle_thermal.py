"""The FP64 oracle's thermal double distribution (src/lbm.hpp:356-397, 489-528;
src/thermal.hpp) against the reference and against what the mathematics fixes.

1. Bit-for-bit trajectories of f, g and per-node forces against the reference itself
   (oracle/_ref; skipped where it was not built) and against the golden fixtures it wrote
   (tests/golden/thermal_*.npz), with the drivers' wall loops after every step.
2. Closed forms: the moments of the thermal equilibrium; the collision relaxes g to it
   (omega_k = 1) and keeps T; a closed domain keeps its total heat; the Boussinesq
   forcing adds exactly F = -beta (T - T_ref) g to each node's momentum (and 3 F.u to
   its mass: the reference's term has no -u.F part); the streaming
   wraps y even where the momentum bounces, bounces g back on open x edges and solids.
"""
import os

import numpy as np
import pytest

import oracle as O
from cases import THERMAL_CASES, build_thermal_oracle, thermal_eq, thermal_step

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CX = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1])
CY = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1])
W = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("name", sorted(THERMAL_CASES))
def test_thermal_trajectory_bitwise_vs_reference(name):
    tc = THERMAL_CASES[name]
    o = build_thermal_oracle(tc)
    r = build_thermal_oracle(tc, cls=O.RefLattice)
    for s in range(tc.steps):
        thermal_step(o, tc)
        thermal_step(r, tc)
        assert np.array_equal(o.f, r.f), f"{name}: f differs at step {s}"
        assert np.array_equal(o.g, r.g), f"{name}: g differs at step {s}"
    assert np.array_equal(o.fx_body, r.fx_body) and np.array_equal(o.fy_body, r.fy_body)


@pytest.mark.parametrize("name", sorted(THERMAL_CASES))
def test_thermal_case_matches_reference_fixture(name):
    tc = THERMAL_CASES[name]
    z = np.load(os.path.join(GOLD, f"thermal_{name}.npz"))
    assert int(z["steps"]) == tc.steps
    o = build_thermal_oracle(tc)
    fx, fy, heat = [], [], []
    for _ in range(tc.steps):
        thermal_step(o, tc)
        a, b = o.total_force()
        fx.append(a)
        fy.append(b)
        heat.append(float(np.cumsum(o.g.ravel())[-1]))
    np.testing.assert_array_equal(fx, z["fx"])
    np.testing.assert_array_equal(fy, z["fy"])
    np.testing.assert_array_equal(heat, z["heat"])
    np.testing.assert_array_equal(o.f[::4], z["rows"])
    np.testing.assert_array_equal(o.g[::4], z["grows"])
    np.testing.assert_array_equal(o.f.sum(axis=(0, 1)), z["colsum"])
    np.testing.assert_array_equal(o.g.sum(axis=(0, 1)), z["gcolsum"])


def test_thermal_equilibrium_moments():
    rng = np.random.default_rng(3)
    for _ in range(100):
        T = rng.uniform(0.3, 2.0)
        u, v = rng.uniform(-0.15, 0.15, 2)
        g = np.array([O.thermal_equilibrium(i, T, u, v) for i in range(9)])
        assert g.sum() == pytest.approx(T, rel=1e-15, abs=0)
        assert (g * CX).sum() == pytest.approx(T * u, abs=1e-15)
        assert (g * CY).sum() == pytest.approx(T * v, abs=1e-15)
        # second moments of the D2Q9 equilibrium: T (c_s^2 delta + u u)
        assert (g * CX * CX).sum() == pytest.approx(T * (1 / 3 + u * u), rel=1e-14)
        assert (g * CX * CY).sum() == pytest.approx(T * u * v, abs=1e-15)
        np.testing.assert_allclose(g, thermal_eq(T, u, v)[...], rtol=1e-15)


def _uniform(nx, ny, case, f9, g9):
    o = O.Oracle(nx, ny, case)
    o.set_f(np.broadcast_to(f9, (ny, nx, 9)))
    o.set_g(np.broadcast_to(g9, (ny, nx, 9)))
    return o


def test_collision_relaxes_to_equilibrium_and_keeps_T():
    # uniform periodic lattice: streaming is the identity, so one step is the collision
    rng = np.random.default_rng(4)
    f = np.array([O.equilibrium(i, 1.0, 0.05, -0.03) for i in range(9)])
    g = rng.uniform(0.02, 0.5, 9)
    T = g.sum()
    o = _uniform(5, 4, O.RIBS, f, g)
    o.set_thermal(1.0)
    o.step(0.8, 0.0)
    rho, u, v = O.macros(o.f[0, 0])
    np.testing.assert_allclose(o.g[2, 3], thermal_eq(T, u, v), rtol=1e-14, atol=1e-16)
    o2 = _uniform(5, 4, O.RIBS, f, g)
    o2.set_thermal(1.37)
    o2.step(0.8, 0.0)
    assert o2.g[1, 1].sum() == pytest.approx(T, rel=1e-14)


def test_closed_domain_keeps_total_heat():
    tc = THERMAL_CASES["hills"]  # periodic in x and y, solid hills and top wall
    o = build_thermal_oracle(tc)
    fluid = o.obstacle == 0
    h0 = o.g[fluid].sum()
    for _ in range(50):
        o.step(tc.base.tau, tc.base.u)
    assert o.g[fluid].sum() == pytest.approx(h0, rel=1e-12)
    assert np.all(o.g[~fluid] == 0.0)  # solids are never streamed into


@pytest.mark.parametrize("u0,v0", [(0.0, 0.0), (0.07, -0.04)])
def test_buoyancy_adds_exactly_the_force(u0, v0):
    beta, gb, T, t_ref = 0.01, -1e-3, 1.2, 1.0
    f = np.array([O.equilibrium(i, 1.0, u0, v0) for i in range(9)])
    g = thermal_eq(T, u0, v0)
    o = _uniform(6, 5, O.RIBS, f, g)
    o.set_thermal(1.0, beta, gb, t_ref)
    o.step(0.9, 0.0)
    fy = -beta * (T - t_ref) * gb  # F = -rho_0 beta (T - T_ref) g, rho_0 = 1
    for y, x in ((0, 0), (3, 4)):
        fn = o.f[y, x]
        # the forcing is w_i (3 F.e_i + 9 (e_i.u)(e_i.F)): its zeroth moment is 3 F.u
        # (the -u.F part of Guo's term is absent), its first moment exactly F
        assert fn.sum() == pytest.approx(1.0 + 3.0 * fy * v0, abs=1e-15)
        assert (fn * CX).sum() == pytest.approx(u0, abs=1e-15)
        assert (fn * CY).sum() == pytest.approx(v0 + fy, abs=1e-15)
    # without beta (or without gravity) the momentum is untouched
    o2 = _uniform(6, 5, O.RIBS, f, g)
    o2.set_thermal(1.0, 0.0, gb, t_ref)
    o2.step(0.9, 0.0)
    assert (o2.f[2, 2] * CY).sum() == pytest.approx(v0, abs=1e-15)


def test_streaming_wraps_y_and_bounces_on_open_x_edges_and_solids():
    # omega_k = 0: the thermal collision is the identity, one step is pure streaming
    nx, ny = 7, 6
    rng = np.random.default_rng(5)
    f = np.array([O.equilibrium(i, 1.0, 0.0, 0.0) for i in range(9)])
    g0 = rng.uniform(0.1, 1.0, (ny, nx, 9))
    # cylinder case: y periodic, x open; one solid node
    o = _uniform(nx, ny, O.CYLINDER, f, f)
    o.set_g(g0)
    m = np.zeros((ny, nx), np.uint8)
    m[3, 3] = 1
    o.set_obstacle(m)
    o.set_thermal(0.0)
    o.step(0.8, 0.0)
    g = o.g
    opp = [0, 3, 4, 1, 2, 7, 8, 5, 6]
    for y in range(ny):
        for x in range(nx):
            if m[y, x]:
                assert np.all(g[y, x] == 0.0)
                continue
            for i in range(9):
                sx, sy = x - CX[i], (y - CY[i]) % ny
                if 0 <= sx < nx and not m[sy, sx]:
                    assert g[y, x, i] == g0[sy, sx, i]
                else:
                    assert g[y, x, i] == g0[y, x, opp[i]]
    # cavity case: the momentum bounces on every edge, g still wraps in y
    o = _uniform(nx, ny, O.CAVITY, f, f)
    o.set_g(g0)
    o.set_thermal(0.0)
    o.step(0.8, 0.0)
    assert o.g[0, 2, 2] == g0[ny - 1, 2, 2]
    assert o.g[ny - 1, 4, 4] == g0[0, 4, 4]
    assert o.g[2, 0, 1] == g0[2, 0, 3]  # x is not wrapped
