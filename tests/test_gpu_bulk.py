"""Two steps per launch with bulk tiles: the 64 x rows tiles whose node ring lies inside
the lattice and holds no general-path node go to the bulk kernel, the whole tiles whose
ring holds besides plain nodes only inlet, outlet and periodic-edge nodes to its edge form,
the others to the general two-step kernel.  The split must not change a bit: populations
and force totals equal one step per launch, in every storage form, with and without
stochastic rounding, for every boundary kind next to the bulk and edge tiles."""
from __future__ import annotations

import dataclasses

import numpy as np
import pytest

import lbm2d as G
import oracle as O
from cases import (CASES, _hills, _mask_setup, _step_geometry, apply_gpu_geometry, build_gpu, build_oracle, gpu_desc,
                   split_counts)

pytestmark = pytest.mark.gpu

HILLS_BIG = _hills(384, 144)
STEP_BIG = _step_geometry(384, 144)

# the small cases scaled up until most of their tiles are bulk tiles
BULK_CASES = {
    "cylinder": dataclasses.replace(CASES["cylinder"], name="cylinder_big", nx=448, ny=160, pert=(100, 300),
                                    setup=lambda s: s.place_cylinder(112, 81, 12),
                                    gpu_geom=(("circle", 112, 81, 12, G.BB_MEI),)),
    "les": dataclasses.replace(CASES["les"], name="les_big", nx=448, ny=160, pert=(100, 300),
                               setup=lambda s: s.place_cylinder(112, 81, 12),
                               gpu_geom=(("circle", 112, 81, 12, G.BB_MEI),)),
    "bgk": dataclasses.replace(CASES["bgk"], name="bgk_big", nx=448, ny=160, pert=(100, 300),
                               setup=lambda s: s.place_cylinder(112, 81, 12),
                               gpu_geom=(("circle", 112, 81, 12, G.BB_MEI),)),
    "bouzidi": dataclasses.replace(CASES["bouzidi"], name="bouzidi_big", nx=448, ny=160,
                                   setup=lambda s: (s.place_cylinder(112, 80, 12), s.set_moving_wall(0.01, 112, 80),
                                                    s.set_mei(False)),
                                   gpu_geom=(("circle", 112, 80, 12, G.BB_BOUZIDI), ("moving", 0.01, 112, 80))),
    "hills": dataclasses.replace(CASES["hills"], name="hills_big", nx=384, ny=144, setup=_mask_setup(HILLS_BIG)),
    "cavity": dataclasses.replace(CASES["cavity"], name="cavity_big", nx=256, ny=256),
    # last tile column and row band cut off by the lattice (partial edge tiles)
    "partial": dataclasses.replace(CASES["les"], name="les_300x150", nx=300, ny=150, pert=(100, 250),
                                   setup=lambda s: s.place_cylinder(100, 75, 12),
                                   gpu_geom=(("circle", 100, 75, 12, G.BB_MEI),)),
    # a disc holding whole 64 x 8 / 64 x 16 / 64 x 32 tiles of solid nodes: empty tiles
    "disc": dataclasses.replace(CASES["les"], name="disc_big", nx=448, ny=160, pert=(100, 300),
                                setup=lambda s: s.place_cylinder(170, 80, 52),
                                gpu_geom=(("circle", 170, 80, 52, G.BB_MEI),)),
    # parabolic inlet from row ny/3 (the rows below it are general-path nodes), periodic y
    "step": dataclasses.replace(CASES["step"], name="step_big", nx=384, ny=144, setup=_mask_setup(STEP_BIG)),
}


def _written_solids(c):
    """Lattices with Bouzidi links write every solid node (no tile of them is empty)."""
    return any(g[0] == "circle" and g[4] == G.BB_BOUZIDI for g in c.gpu_geom)


FORMS = [("f32", False), ("f16s", False), ("f16s", True), ("bf16s", False), ("bf16s", True)]


@pytest.mark.parametrize("storage,sr", FORMS, ids=[f"{s}{'_sr' if r else ''}" for s, r in FORMS])
@pytest.mark.parametrize("name", sorted(BULK_CASES))
def test_bulk_tiles_bitwise_equal_one_step_per_launch(name, storage, sr):
    c = BULK_CASES[name]
    a, o = build_gpu(c, storage)
    b, _ = build_gpu(c, storage, o=o)
    if sr:
        a.set_stochastic_rounding(True, 77)
        b.set_stochastic_rounding(True, 77)
    a.set_launch(64, 1, 1, 0)
    desc = a.info()
    assert (desc.bulk_tiles, desc.edge_tiles, desc.other_tiles, desc.empty_tiles) == (0, 0, 0, 0)
    d = gpu_desc(c)
    written = _written_solids(c)
    seen = seen_edge = 0
    for rows, opt in ((16, 5), (32, 7), (8, 6), (32, 4), (16, 7)):
        b.set_launch(64, rows, 1, opt, graph_steps=10)
        info = b.info()
        exp = split_counts(o.obstacle, d, rows, written_solids=written)
        assert (info.bulk_tiles, info.edge_tiles, info.other_tiles, info.empty_tiles) == exp, f"rows {rows}"
        seen += info.bulk_tiles
        seen_edge += info.edge_tiles
        for n in (1, 5, 12):
            a.step(n)
            b.step(n)
            np.testing.assert_array_equal(b.download(0), a.download(0), err_msg=f"rows {rows} options {opt} n {n}")
    assert seen > 0
    # edge tiles wherever an edge is periodic or carries an inlet or outlet column
    assert (seen_edge > 0) == (d.x_edge == G.EDGE_PERIODIC or d.y_edge == G.EDGE_PERIODIC or d.outlet != 0), seen_edge
    fa, fya = a.forces(0)
    fb, fyb = b.forces(0)
    np.testing.assert_array_equal(fb, fa)
    np.testing.assert_array_equal(fyb, fya)
    assert not b.nan_flag(0)
    a.close()
    b.close()


def test_bulk_tiles_fp32_against_oracle():
    """The split launch in FP32 against the FP64 oracle on the LES case."""
    c = BULK_CASES["les"]
    ens, o = build_gpu(c, "f32")
    ens.set_launch(64, 32, 1, 5, graph_steps=10)
    ens.step(30)
    assert ens.info().bulk_tiles > 0
    o.step(c.tau, c.u, 30)
    assert np.abs(ens.download(0) - o.f).max() <= 2e-6
    fx, fy = ens.forces(0)
    ofx, ofy = o.total_force()
    assert abs(fx[-1, 0] - ofx) <= 1e-4 * (abs(ofx) + 1e-3)
    ens.close()


def test_bulk_plan_follows_geometry_and_refusals():
    """A new obstacle rebuilds the tile classes; 64 x 32 tiles need the split (FP32 or a
    16-bit form)."""
    c = BULK_CASES["cylinder"]
    ens, o = build_gpu(c, "f32")
    ens.set_launch(64, 16, 1, 5)
    ens.step(2)
    d = gpu_desc(c)
    first = ens.info().bulk_tiles
    m = o.obstacle.copy()
    m[40:120, 300:310] = 1
    m[0:30, 440:448] = 1  # on the outlet column: its tiles leave the edge class
    ens.set_obstacle(0, m)
    ens.upload(0, o.f.copy())
    ens.step(2)
    info = ens.info()
    exp = split_counts(m, d, 16)
    assert (info.bulk_tiles, info.edge_tiles, info.other_tiles, info.empty_tiles) == exp
    assert info.bulk_tiles < first
    ens.close()
    f64, _ = build_gpu(c, "f64")
    with pytest.raises(G.LbmError):
        f64.set_launch(64, 32, 1, 5)
    f64.close()
    two = G.Ensemble([G.desc_for_case(G.CYLINDER, 256, 128)] * 2, storage="f64")
    with pytest.raises(G.LbmError):
        two.set_launch(64, 32, 1, 5)
    two.close()


# an ensemble of the scaled-up cases plus a channel whose last column block holds only the
# outlet column (nx % 64 == 1: its two-step tiles are right-aligned)
NARROW_BIG = dataclasses.replace(BULK_CASES["les"], name="les_big_321", nx=321, ny=130, pert=(100, 200),
                                 setup=lambda s: s.place_cylinder(80, 65, 10),
                                 gpu_geom=(("circle", 80, 65, 10, G.BB_MEI),))


def _ens_cases():
    return [BULK_CASES[n] for n in sorted(BULK_CASES)] + [NARROW_BIG]


def _ensemble(cs, orcs, storage):
    ens = G.Ensemble([gpu_desc(c) for c in cs], storage=storage)
    for l, (c, o) in enumerate(zip(cs, orcs)):
        apply_gpu_geometry(ens, l, c, o.obstacle.copy())
        ens.upload(l, o.f.copy())
    return ens


@pytest.mark.parametrize("storage,sr", [("f32", False), ("f16s", True), ("bf16s", False)],
                         ids=["f32", "f16s_sr", "bf16s"])
def test_ensemble_bulk_tiles_equal_one_step_per_launch(storage, sr):
    """An ensemble's two-step launches split into bulk and other tiles too: every lattice
    holds exactly what the same ensemble holds stepped one step per launch (populations,
    forces), and without stochastic rounding what the lattice holds stepped alone (the
    random words of stochastic rounding follow a node's place in its handle)."""
    cs = _ens_cases()
    orcs = [build_oracle(c) for c in cs]
    ens = _ensemble(cs, orcs, storage)
    ref = _ensemble(cs, orcs, storage)
    ref.set_launch(64, 1, 1, 0)
    alone = []
    if not sr:
        for c, o in zip(cs, orcs):
            a, _ = build_gpu(c, storage, o=o)
            a.set_launch(64, 1, 1, 0)
            alone.append(a)
    else:
        ens.set_stochastic_rounding(True, 5)
        ref.set_stochastic_rounding(True, 5)
    seen = seen_edge = 0
    for rows, opt in ((16, 5), (8, 6), (32, 7)):
        ens.set_launch(64, rows, 1, opt, graph_steps=10)
        seen += ens.info().bulk_tiles
        seen_edge += ens.info().edge_tiles
        for n in (1, 5, 12):
            ens.step(n)
            ref.step(n)
            for a in alone:
                a.step(n)
            for l in range(len(cs)):
                got = ens.download(l)
                np.testing.assert_array_equal(got, ref.download(l), err_msg=f"lattice {l} rows {rows} n {n}")
                if alone:
                    np.testing.assert_array_equal(got, alone[l].download(0), err_msg=f"alone {l} rows {rows} n {n}")
    assert seen > 0 and seen_edge > 0
    for l in range(len(cs)):
        fe, fye = ens.forces(l)
        fr, fyr = ref.forces(l)
        np.testing.assert_array_equal(fe, fr)
        np.testing.assert_array_equal(fye, fyr)
        if alone:
            fa, fya = alone[l].forces(0)
            np.testing.assert_array_equal(fe, fa)
        assert not ens.nan_flag(l)
    for a in alone:
        a.close()
    ens.close()
    ref.close()


@pytest.mark.parametrize("storage,sr", [("f32", False), ("f16s", True)], ids=["f32", "f16s_sr"])
def test_state_after_a_batch_ending_on_two_steps(storage, sr):
    """A batch of an even number of steps ends on a two-step launch, whose first step's
    populations stay in LDS; whatever reads the state or changes what a step computes
    recomputes them first.  Macros, the mass, a change of tau, a one-lattice step and a
    download after even batches give what one step per launch gives."""
    c = BULK_CASES["les"]
    a, o = build_gpu(c, storage)
    b, _ = build_gpu(c, storage, o=o)
    if sr:
        a.set_stochastic_rounding(True, 9)
        b.set_stochastic_rounding(True, 9)
    a.set_launch(64, 1, 1, 0)
    b.set_launch(64, 16, 1, 5, graph_steps=10)
    for n in (2, 10, 4):
        a.step(n)
        b.step(n)
    for got, want in zip(b.macros(0), a.macros(0)):
        np.testing.assert_array_equal(got, want)
    assert b.mass(0) == a.mass(0)
    a.set_params(0, 0.53, 0.1)
    b.set_params(0, 0.53, 0.1)
    a.step(6)
    b.step(6)
    a.step_lattice(0, 2)
    b.step_lattice(0, 2)
    a.step(8)
    b.step(8)
    np.testing.assert_array_equal(b.download(0), a.download(0))
    fa, fya = a.forces(0)
    fb, fyb = b.forces(0)
    np.testing.assert_array_equal(fb, fa)
    np.testing.assert_array_equal(fyb, fya)
    assert not b.nan_flag(0)
    a.close()
    b.close()


def test_ensemble_state_after_even_batches_with_mixed_buffer_parity():
    """An ensemble whose lattices hold their state in different buffers (one lattice stepped
    alone an odd number of times) and an even batch of two-step launches: the recomputed
    previous populations of every lattice, and the state after more steps, equal one step
    per launch bit for bit."""
    cs = _ens_cases()
    orcs = [build_oracle(c) for c in cs]
    ens = _ensemble(cs, orcs, "f32")
    ref = _ensemble(cs, orcs, "f32")
    ref.set_launch(64, 1, 1, 0)
    ens.set_launch(64, 16, 1, 5, graph_steps=10)
    for e in (ens, ref):
        e.step_lattice(2, 3)
        e.step(4)
    for l in range(len(cs)):
        np.testing.assert_array_equal(ens.download(l), ref.download(l), err_msg=f"lattice {l}")
    for e in (ens, ref):
        e.step(10)
        e.step_lattice(1, 1)
        e.step(2)
    for l in range(len(cs)):
        np.testing.assert_array_equal(ens.download(l), ref.download(l), err_msg=f"lattice {l} later")
        fe, fye = ens.forces(l)
        fr, fyr = ref.forces(l)
        np.testing.assert_array_equal(fe, fr)
        assert not ens.nan_flag(l)
    ens.close()
    ref.close()


@pytest.mark.parametrize("storage,sr", [("f32", False), ("f16s", True), ("bf16s", False)],
                         ids=["f32", "f16s_sr", "bf16s"])
def test_equilibrium_inlet_edge_tiles(storage, sr):
    """The equilibrium inlet (every slot of an inlet node replaced by feq(1, u_in, 0)) with
    periodic y edges: its column's tiles are edge tiles, whose closure runs in stored units;
    populations and forces equal one step per launch bit for bit."""
    c = BULK_CASES["les"]
    o = build_oracle(c)
    d = gpu_desc(c)
    d.inlet = G.INLET_EQUILIBRIUM
    hs = []
    for launch in ((64, 1, 1, 0), (64, 16, 1, 5)):
        e = G.Ensemble([d], storage=storage)
        apply_gpu_geometry(e, 0, c, o.obstacle.copy())
        e.upload(0, o.f.copy())
        if sr:
            e.set_stochastic_rounding(True, 3)
        e.set_launch(*launch, graph_steps=10)
        hs.append(e)
    a, b = hs
    assert b.info().edge_tiles > 0
    for n in (3, 10, 8):
        a.step(n)
        b.step(n)
        np.testing.assert_array_equal(b.download(0), a.download(0), err_msg=f"n {n}")
    fa, _ = a.forces(0)
    fb, _ = b.forces(0)
    np.testing.assert_array_equal(fb, fa)
    assert not b.nan_flag(0)
    a.close()
    b.close()
