"""GPU parity of the configs' boundary wording: equilibrium inlet and free-slip y edges.

The HIP step with LBM_INLET_EQUILIBRIUM and LBM_EDGE_FREE_SLIP against the FP64 oracle's
same option (pinned by the method of images, the inlet's closed form and symmetry in
tests/test_oracle_spec_bcs.py), on a cylinder with a block on the free-slip edge row,
with LES, and with Bouzidi links and a rotating wall.  Tolerances as the other parity
tests: FP64 1e-12 on populations and 1e-10 relative on forces, FP32 2e-6 / 1e-4.
"""
import numpy as np
import pytest

import lbm2d as G
from cases import SPEC_CASES, apply_gpu_geometry, build_gpu, build_oracle, gpu_desc

pytestmark = pytest.mark.gpu

TOL = {"f64": (1e-12, 1e-10), "f32": (2e-6, 1e-4)}


@pytest.mark.parametrize("storage", ["f64", "f32"])
@pytest.mark.parametrize("name", sorted(SPEC_CASES))
def test_spec_boundaries_match_oracle(name, storage):
    c = SPEC_CASES[name]
    ens, o = build_gpu(c, storage)
    fx_o = []
    for _ in range(c.steps):
        o.step(c.tau, c.u)
        fx_o.append(o.total_force())
    ens.step(c.steps)
    f = ens.download(0)
    err = float(np.abs(f - o.f).max())
    fx, fy = ens.forces(0)
    fo = np.array(fx_o)
    scale = np.abs(fo[:, 0]).max() + 1e-3
    ferr = max(np.abs(fx[:, 0] - fo[:, 0]).max(), np.abs(fy[:, 0] - fo[:, 1]).max()) / scale
    print(f"\n[{name} {storage}] max|df| = {err:.2e}, force {ferr:.2e}")
    assert err <= TOL[storage][0] and ferr <= TOL[storage][1]
    # the inlet column is feq(1, u_in, 0)
    feq = G.equilibrium(np.array(1.0), np.array(c.u), np.array(0.0))
    fluid0 = o.obstacle[:, 0] == 0
    assert np.abs(f[fluid0, 0] - feq).max() <= (1e-15 if storage == "f64" else 1e-7)


def test_spec_uniform_stream_is_a_fixed_point():
    d = G.desc_for_case(G.CYLINDER, 200, 64, tau=0.6, u_in=0.1)
    d.inlet, d.y_edge = G.INLET_EQUILIBRIUM, G.EDGE_FREE_SLIP
    e = G.Ensemble([d], storage="f64")
    e.set_obstacle(0, np.zeros((64, 200), np.uint8))
    e.init_equilibrium(0, 1.0, 0.1, 0.0)
    e.step(100)
    feq = G.equilibrium(np.array(1.0), np.array(0.1), np.array(0.0))
    assert np.abs(e.download(0) - feq).max() <= 1e-14


def test_spec_ensemble_and_graphs():
    cs = [SPEC_CASES["spec_cylinder"], SPEC_CASES["spec_bouzidi"]]
    ens = G.Ensemble([gpu_desc(c) for c in cs], storage="f32")
    for l, c in enumerate(cs):
        o = build_oracle(c)
        apply_gpu_geometry(ens, l, c, o.obstacle.copy())
        ens.upload(l, o.f.copy())
    ens.set_launch(graph_steps=10)
    ens.step(30)
    for l, c in enumerate(cs):
        alone, _ = build_gpu(c, "f32")
        for _ in range(30):
            alone.step_lattice(0, 1)
        np.testing.assert_array_equal(ens.download(l), alone.download(0))
        np.testing.assert_array_equal(ens.forces(l)[0], alone.forces(0)[0])


def test_spec_refusals():
    c = SPEC_CASES["spec_cylinder"]
    ens, _ = build_gpu(c, "f32")
    with pytest.raises(G.LbmError, match="free-slip"):
        ens.set_launch(64, 8, 1, 4)  # two steps per launch
    with pytest.raises(G.LbmError, match="free-slip"):
        ens.set_thermal(0, 1.5)
    shape, _ = ens.autotune()
    assert not shape[3] & 4
    d = G.desc_for_case(G.CAVITY, 32, 32)
    d.y_edge = G.EDGE_FREE_SLIP  # the cavity's x edges are walls
    with pytest.raises(G.LbmError, match="open x edge"):
        G.Ensemble([d])
