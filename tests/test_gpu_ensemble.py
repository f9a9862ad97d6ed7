"""GPU: ensemble launches against the one-lattice kernel, bit for bit.

An ensemble launch covers every lattice of a handle: its leading list blocks take the
general-path nodes of every lattice (boundary-link records, edge nodes, written solids)
and its bulk tiles skip them; with two steps per launch each lattice is cut into 64 x rows
tiles of its own (rows of a lattice with an outlet and nx % 64 == 1 right-aligned).  The
arithmetic per node is the one-lattice kernel's, so every lattice of the ensemble must
hold exactly the populations and force totals it holds when stepped alone, whatever the
launch form, batch split or graph use.
"""
import numpy as np
import pytest

import lbm2d as G
from cases import CASES, NARROW_CASES, SPEC_CASES, apply_gpu_geometry, build_gpu, build_oracle, gpu_desc

pytestmark = pytest.mark.gpu

NAMES = ["cylinder", "cavity", "step", "plate", "rotating", "bouzidi", "les", "bgk", "hills", "near_wall"]
# (block_x, rows, nodes per thread, options, graph_steps or 0): one-step shapes with and
# without XCD order and streaming stores, then two steps per launch on 64 x 4 / 8 / 16
LAUNCHES = [(256, 1, 1, 0, 10), (64, 1, 1, 3, 10), (128, 1, 2, 1, 0), (256, 2, 1, 2, 10),
            (64, 4, 1, 4, 10), (64, 8, 1, 5, 10), (64, 8, 1, 6, 0), (64, 4, 1, 7, 10), (64, 16, 1, 5, 10)]


def _lattices():
    # with nx % 64 == 1 and an outlet (right-aligned two-step tiles)
    return [CASES[n] for n in NAMES] + [NARROW_CASES[n] for n in sorted(NARROW_CASES)]


def _ensemble(cs, storage):
    orcs = [build_oracle(c) for c in cs]
    ens = G.Ensemble([gpu_desc(c) for c in cs], storage=storage)
    for l, (c, o) in enumerate(zip(cs, orcs)):
        apply_gpu_geometry(ens, l, c, o.obstacle.copy())
        ens.upload(l, o.f.copy())
    return ens, orcs


def _set(ens, launch):
    bx, rows, nodes, opt, graph = launch
    if graph:
        ens.set_launch(bx, rows, nodes, opt, graph_steps=graph)
    else:
        ens.set_launch(bx, rows, nodes, opt)


@pytest.mark.parametrize("storage", ["f64", "f32", "f16s"])
def test_ensemble_launch_forms_equal_each_lattice_alone(storage):
    cs = _lattices()
    alone = []
    for c in cs:
        a, _ = build_gpu(c, storage)
        a.set_launch(64, 1, 1, 0)
        alone.append(a)
    counts = (1, 5, 12)
    states = []
    for n in counts:
        for a in alone:
            a.step(n)
        states.append([a.download(0) for a in alone])
    forces = [a.forces(0) for a in alone]
    for launch in LAUNCHES:
        ens, _ = _ensemble(cs, storage)
        _set(ens, launch)
        for k, n in enumerate(counts):
            ens.step(n)
            for l, c in enumerate(cs):
                np.testing.assert_array_equal(ens.download(l), states[k][l],
                                              err_msg=f"{c.name}, launch {launch}, after {sum(counts[:k + 1])} steps")
        for l, c in enumerate(cs):
            fx, fy = ens.forces(l)
            np.testing.assert_array_equal(fx, forces[l][0], err_msg=f"{c.name}, launch {launch}")
            np.testing.assert_array_equal(fy, forces[l][1], err_msg=f"{c.name}, launch {launch}")
            assert ens.steps_done(l) == sum(counts)
            assert not ens.nan_flag(l)
        ens.close()


def test_right_aligned_lattices_match_oracle():
    """The lattices whose two-step rows are right-aligned, in FP64 inside an ensemble
    stepped two steps per launch, against the FP64 oracle (populations and forces)."""
    cs = _lattices()
    ens, orcs = _ensemble(cs, "f64")
    ens.set_launch(64, 8, 1, 5, graph_steps=10)
    n = 25
    ens.step(n)
    for l, (c, o) in enumerate(zip(cs, orcs)):
        if c.name not in ("cylinder_129", "les_65", "cylinder", "bouzidi"):
            continue
        fx_o = []
        for _ in range(n):
            o.step(c.tau, c.u)
            fx_o.append(o.total_force()[0])
        np.testing.assert_allclose(ens.download(l), o.f, rtol=0, atol=1e-12, err_msg=c.name)
        fx = ens.forces(l)[0][:, 0]
        assert np.abs(fx - np.array(fx_o)).max() <= 1e-10 * (np.abs(fx_o).max() + 1e-3), c.name


@pytest.mark.parametrize("storage", ["f16s", "bf16s"])
def test_ensemble_stochastic_rounding_repeats_in_every_launch_form(storage):
    """Stochastic rounding keys on (seed, the lattice's step, the node): the ensemble's
    one-step and two-step launches, graphs and single launches give the same bits."""
    cs = _lattices()
    ref = None
    for launch in ((256, 1, 1, 0, 10), (64, 8, 1, 5, 6), (64, 4, 1, 6, 0), (64, 1, 4, 3, 10), (64, 16, 1, 7, 6)):
        ens, _ = _ensemble(cs, storage)
        ens.set_stochastic_rounding(True, 77)
        _set(ens, launch)
        for n in (3, 8):
            ens.step(n)
        got = [ens.download(l) for l in range(len(cs))]
        if ref is None:
            ref = got
        else:
            for l, c in enumerate(cs):
                np.testing.assert_array_equal(got[l], ref[l], err_msg=f"{c.name}, launch {launch}")
        ens.close()


def test_two_steps_refused_with_free_slip_edges_and_autotune_keeps_state():
    cs = _lattices()[:3] + [SPEC_CASES["spec_cylinder"]]
    ens, _ = _ensemble(cs, "f32")
    with pytest.raises(G.LbmError):
        ens.set_launch(64, 8, 1, 5)
    ens.close()
    cs = _lattices()
    ens, _ = _ensemble(cs, "f32")
    ens.step(3)
    before = [ens.download(l) for l in range(len(cs))]
    (bx, rows, nodes, options), ms = ens.autotune()
    assert ms > 0 and options in range(8)
    for l in range(len(cs)):
        np.testing.assert_array_equal(ens.download(l), before[l])
    ens.step(7)  # the picked form steps on
    for l in range(len(cs)):
        assert np.all(np.isfinite(ens.download(l)))
    ens.close()
