"""lbm_step_noalloc: the allocation-free step against lbm_step through the C ABI.

The entry point runs the same kernels as lbm_step with its force partials in the
library's fixed device scratch instead of handle-allocated blocks, so state and
per-step forces must equal lbm_step's bit for bit.  Two calls in a row on two handles
(the second lattice smaller than the first) share the one scratch array.
"""
import numpy as np
import pytest

import lbm2d as G
from cases import CASES, build_gpu

pytestmark = pytest.mark.gpu


def _porous(nx, ny, density, seed, storage):
    rng = np.random.default_rng(seed)
    mask = (rng.random((ny, nx)) < density).astype(np.uint8)
    d = G.desc_for_case(G.CYLINDER, nx, ny, tau=0.7, u_in=0.05)
    e = G.Ensemble([d], storage=storage)
    e.set_obstacle(0, mask)
    e.init_equilibrium(0, 1.0, 0.05, 0.0)
    return e


def test_noalloc_matches_step_on_two_calls():
    # call 1: 1024 x 768 porous lattice, ~250k boundary-link nodes: the partials of four
    # launches fill the 16 MiB scratch, so 9 steps run as batches of 4, 4 and 1
    a = _porous(1024, 768, 0.05, 3, "f32")
    b = _porous(1024, 768, 0.05, 3, "f32")
    n_links = a.info().link_cells
    per_launch = (16 * 2 + 2 * n_links) * 8
    assert 2 <= (16 << 20) // per_launch <= 4, per_launch
    a.step(9)
    b.step(0)
    b.step_noalloc(9)
    np.testing.assert_array_equal(b.download(0), a.download(0))
    np.testing.assert_array_equal(b.forces(0)[0], a.forces(0)[0])
    np.testing.assert_array_equal(b.forces(0)[1], a.forces(0)[1])
    assert b.steps_done(0) == 9
    a.close()
    b.close()
    # call 2: a smaller lattice, FP64 with Bouzidi links and a rotating wall (the Ladd
    # term, written solids): equal to lbm_step and to the oracle
    c = CASES["bouzidi"]
    e1, o = build_gpu(c, "f64")
    e2, _ = build_gpu(c, "f64")
    e1.step(37)
    e2.step(0)
    e2.step_noalloc(37)
    o.step(c.tau, c.u, 37)
    np.testing.assert_array_equal(e2.download(0), e1.download(0))
    fx1, fy1 = e1.forces(0)
    fx2, fy2 = e2.forces(0)
    np.testing.assert_array_equal(fx2, fx1)
    np.testing.assert_array_equal(fy2, fy1)
    np.testing.assert_allclose(e2.download(0), o.f, rtol=0, atol=1e-12)
    fxo, _ = o.total_force()
    assert abs(fx2[-1, 0] - fxo) <= 1e-10 * (abs(fxo) + 1e-3)


def test_noalloc_two_steps_per_launch_and_ensembles():
    c = CASES["cylinder"]
    e1, _ = build_gpu(c, "f32")
    e2, _ = build_gpu(c, "f32")
    for e in (e1, e2):
        e.set_launch(64, 8, 1, 4)  # two steps per launch
    e1.step(9)
    e2.step(0)
    e2.step_noalloc(9)
    np.testing.assert_array_equal(e2.download(0), e1.download(0))
    np.testing.assert_array_equal(e2.forces(0)[0], e1.forces(0)[0])
    # an ensemble of two uneven lattices: one launch per step over both
    from cases import apply_gpu_geometry, build_oracle, gpu_desc
    cs = [CASES["cylinder"], CASES["plate"]]
    hs = []
    for _ in range(2):
        os_ = [build_oracle(k) for k in cs]
        h = G.Ensemble([gpu_desc(k) for k in cs], storage="f32")
        for l, (k, o) in enumerate(zip(cs, os_)):
            apply_gpu_geometry(h, l, k, o.obstacle.copy())
            h.upload(l, o.f.copy())
        hs.append(h)
    hs[0].step(13)
    hs[1].step(0)
    hs[1].step_noalloc(13)
    for l in range(2):
        np.testing.assert_array_equal(hs[1].download(l), hs[0].download(l))
        np.testing.assert_array_equal(hs[1].forces(l)[0], hs[0].forces(l)[0])


def test_noalloc_refusals():
    c = CASES["cylinder"]
    e, _ = build_gpu(c, "f32")
    # state not yet committed after an upload / a parameter change
    with pytest.raises(G.LbmError, match="lbm_step"):
        e.step_noalloc(1)
    e.step(0)
    e.step_noalloc(2)
    e.set_params(0, 0.58, 0.09)
    with pytest.raises(G.LbmError, match="lbm_step"):
        e.step_noalloc(1)
    e.step(0)
    e.step_noalloc(1)
    # the force history holds 4096 steps
    with pytest.raises(G.LbmError, match="history"):
        e.step_noalloc(4096)
    e.forces(0)
    e.step_noalloc(10)
    assert e.steps_done(0) == 13
    # one launch's partials beyond the scratch: ~1.5 M boundary-link nodes
    big = _porous(2048, 1024, 0.3, 5, "f16s")
    big.step(0)
    assert big.info().link_cells * 16 > (16 << 20)
    with pytest.raises(G.LbmError, match="scratch"):
        big.step_noalloc(1)
