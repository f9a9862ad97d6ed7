"""GPU parity of the thermal double distribution (src/lbm.hpp:356-397, 489-528) against
the FP64 oracle, with the drivers' wall loops (src/heated_cylinder.cpp:155-164,
src/thermal_cavity.cpp:153-188) run as the library's per-node conditions.

Tolerances (absolute; g_i ~ w_i T with T up to ~9 in these cases, so g bounds scale
with max |g|; forces relative to |F| + 1e-3):
  F64 storage : 1e-12 f, 1e-12 max(1, |g|) g, 1e-10 forces   (same algorithm in FP64)
  F32 storage : 1e-5 f, 2e-5 max(1, |g|) g, 1e-4 forces      (FP32 arithmetic on f - w, g - w)
"""
import numpy as np
import pytest

import lbm2d as G
from cases import (CASES, THERMAL_CASES, build_gpu, build_gpu_thermal, build_oracle, gpu_desc,
                   gpu_thermal_rules, thermal_step)

pytestmark = pytest.mark.gpu

TOL = {"f64": (1e-12, 1e-12, 1e-10), "f32": (1e-5, 2e-5, 1e-4)}


def _tsum(g):
    """sum of g over slots in slot order (compute_temperature, src/thermal.hpp:55-60)."""
    t = g[..., 0].copy()
    for i in range(1, 9):
        t = t + g[..., i]
    return t


def _run(name, storage, graph_steps=None):
    tc = THERMAL_CASES[name]
    print(f"[thermal {name} {storage}]", flush=True)
    ens, o = build_gpu_thermal(tc, storage)
    if graph_steps:
        ens.set_launch(graph_steps=graph_steps)
    fx_o, fy_o = [], []
    for _ in range(tc.steps):
        thermal_step(o, tc)
        fx, fy = o.total_force()
        fx_o.append(fx)
        fy_o.append(fy)
    ens.step(tc.steps)
    gx, gy = ens.forces(0)
    return tc, o, ens, np.array(fx_o), np.array(fy_o), gx[:, 0], gy[:, 0]


@pytest.mark.parametrize("storage", ["f64", "f32"])
@pytest.mark.parametrize("name", sorted(THERMAL_CASES))
def test_thermal_matches_oracle(name, storage):
    tc, o, ens, fx_o, fy_o, gx, gy = _run(name, storage)
    f, g = ens.download(0), ens.download_thermal(0)
    tf, tg, tforce = TOL[storage]
    gscale = max(1.0, float(np.abs(o.g).max()))
    ef, eg = np.abs(f - o.f).max(), np.abs(g - o.g).max()
    print(f"[thermal {name} {storage}] max|df| = {ef:.3e}  max|dg| = {eg:.3e} (max|g| {gscale:.2f})")
    assert ef <= tf
    assert eg <= tg * gscale
    if tc.base.case != G.CAVITY:
        assert len(gx) == tc.steps
        scale = np.abs(fx_o).max() + 1e-3
        assert np.abs(gx - fx_o).max() <= tforce * scale
        assert np.abs(gy - fy_o).max() <= tforce * scale
    # the device temperature is the slot-order sum of what download_thermal returns
    np.testing.assert_array_equal(ens.temperature(0), _tsum(g))
    np.testing.assert_array_equal(ens.temperature(0, 3, 9), _tsum(g)[3:9])
    if storage == "f64":
        assert abs(ens.mass(0) - o.mass()) <= 1e-10 * abs(o.mass())


def test_thermal_graph_equals_single_launches():
    tc = THERMAL_CASES["heated"]
    a, _ = build_gpu_thermal(tc, "f32")
    b, _ = build_gpu_thermal(tc, "f32")
    a.set_launch(graph_steps=10)
    a.step(30)  # three graph replays
    for _ in range(30):
        b.step_lattice(0, 1)
    np.testing.assert_array_equal(a.download(0), b.download(0))
    np.testing.assert_array_equal(a.download_thermal(0), b.download_thermal(0))
    np.testing.assert_array_equal(a.forces(0)[0], b.forces(0)[0])


def test_thermal_download_before_any_step_returns_upload():
    tc = THERMAL_CASES["convection"]
    ens, o = build_gpu_thermal(tc, "f64")
    np.testing.assert_array_equal(ens.download(0), o.f)
    np.testing.assert_array_equal(ens.download_thermal(0), o.g)
    np.testing.assert_array_equal(ens.temperature(0), _tsum(o.g))


def test_thermal_parameter_changes_mid_run():
    tc = THERMAL_CASES["convection"]
    ens, o = build_gpu_thermal(tc, "f64")
    ens.step(40)
    for _ in range(40):
        thermal_step(o, tc)
    # new tau, new thermal rate and buoyancy, new wall temperatures
    ens.set_params(0, 0.75, 0.0)
    ens.set_thermal(0, 1.3, 0.0008, tc.g_buoyancy, 1.0)
    o.set_thermal(1.3, 0.0008, tc.g_buoyancy, 1.0)
    code, _ = gpu_thermal_rules(tc, o.obstacle)
    ens.set_thermal_rules(0, code, [(1.4, G.THERMAL_EQ_FLOW), (0.6, G.THERMAL_EQ_FLOW)])
    # the new conditions act on the state after the last step too
    o.thermal_cavity_walls(1.4, 0.6)
    np.testing.assert_allclose(ens.download_thermal(0), o.g, rtol=0, atol=1e-12)
    for _ in range(40):
        o.step(0.75, 0.0)
        o.thermal_cavity_walls(1.4, 0.6)
    ens.step(40)
    np.testing.assert_allclose(ens.download(0), o.f, rtol=0, atol=1e-12)
    np.testing.assert_allclose(ens.download_thermal(0), o.g, rtol=0, atol=1e-12)


def test_upload_resets_thermal_state_to_one():
    tc = THERMAL_CASES["heated"]
    ens, o = build_gpu_thermal(tc, "f64")
    ens.step(5)
    ens.upload(0, o.f.copy())
    np.testing.assert_array_equal(ens.download_thermal(0), np.ones_like(o.g))
    # and the step from there equals the oracle's with g = 1
    o.set_g(np.ones_like(o.g))
    thermal_step(o, tc)
    ens.step(1)
    np.testing.assert_allclose(ens.download_thermal(0), o.g, rtol=0, atol=1e-12)
    np.testing.assert_allclose(ens.download(0), o.f, rtol=0, atol=1e-12)


def test_thermal_errors():
    c = CASES["cylinder"]
    o = build_oracle(c)
    with G.Ensemble([gpu_desc(c)], storage="f16s") as e:
        with pytest.raises(G.LbmError):
            e.set_thermal(0, 1.0)
    with G.Ensemble([gpu_desc(c), gpu_desc(c)], storage="f32") as e:
        with pytest.raises(G.LbmError):
            e.set_thermal(0, 1.0)
    ens, _ = build_gpu(c, "f32", o=o)
    with pytest.raises(G.LbmError):
        ens.download_thermal(0)  # thermal populations off
    ens.set_thermal(0, 1.2)
    bad = np.zeros((c.ny, c.nx), np.uint8)
    bad[3, 3] = 2  # names rule 2 of 1
    with pytest.raises(G.LbmError):
        ens.set_thermal_rules(0, bad, [(1.5, G.THERMAL_EQ_REST)])
    with pytest.raises(G.LbmError):
        ens.autotune()
    ens.step(2)
    with pytest.raises(G.LbmError):
        ens.upload_thermal(0, np.ones((c.ny, c.nx, 9)))  # only right after f
    ens.close()


def test_thermal_ignores_two_steps_per_launch():
    # two steps per launch is a momentum-only kernel: a thermal handle keeps one step per
    # launch whatever the launch options say
    tc = THERMAL_CASES["heated"]
    a, _ = build_gpu_thermal(tc, "f32")
    b, _ = build_gpu_thermal(tc, "f32")
    a.set_launch(64, 4, 1, 4, 10)
    a.step(25)
    b.step(25)
    np.testing.assert_array_equal(a.download(0), b.download(0))
    np.testing.assert_array_equal(a.download_thermal(0), b.download_thermal(0))
    assert a.steps_done(0) == 25
