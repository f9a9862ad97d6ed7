"""Every launch writes every slot it owns, and reads nothing of what it writes.

Before each launch the buffer that launch writes is filled with NaN bytes
(lbm_debug_poison); the run must then equal, bit for bit, an identical run without the
poison -- populations, per-step forces and mass.  A slot a launch should write but does
not, or a value a launch reads from its own destination (a read-after-write race across
workgroups), would bring a NaN or a stale value into the next launch's gather.  Covered:
every small case in FP64 and FP32, the 16-bit forms, two steps per launch, the thermal
kernel, an ensemble of uneven lattices, and every launch shape of the one-step kernel on
one case.
"""
import numpy as np
import pytest

import lbm2d as G
from cases import CASES, THERMAL_CASES, apply_gpu_geometry, build_gpu, build_gpu_thermal, build_oracle, gpu_desc

pytestmark = pytest.mark.gpu


def _twin_runs(make, launches, steps_per_call=1, lattices=(0,)):
    clean, dirty = make(), make()
    for e in (clean, dirty):
        e.step(0)
    for _ in range(launches):
        clean.step(steps_per_call)
        for l in lattices:
            dirty.debug_poison(l)
        dirty.step(steps_per_call)
    for l in lattices:
        a, b = clean.download(l), dirty.download(l)
        assert np.isfinite(b).all()
        np.testing.assert_array_equal(b, a)
        np.testing.assert_array_equal(dirty.forces(l)[0], clean.forces(l)[0])
        assert dirty.mass(l) == clean.mass(l)
        assert not dirty.nan_flag(l)
    return clean, dirty


@pytest.mark.parametrize("storage", ["f64", "f32"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_poisoned_destination_changes_nothing(name, storage):
    c = CASES[name]
    _twin_runs(lambda: build_gpu(c, storage)[0], 6)


@pytest.mark.parametrize("storage", ["f16s", "bf16s"])
@pytest.mark.parametrize("name", ["cylinder", "bouzidi", "les"])
def test_poisoned_destination_16bit(name, storage):
    c = CASES[name]
    _twin_runs(lambda: build_gpu(c, storage)[0], 5)


@pytest.mark.parametrize("rows", [4, 16])
@pytest.mark.parametrize("name", ["cylinder", "bouzidi", "step"])
def test_poisoned_destination_two_steps_per_launch(name, rows):
    """Calls of 3 steps: a two-step launch, then a one-step launch; the poison precedes
    the two-step launch of each call."""
    c = CASES[name]

    def make():
        e, _ = build_gpu(c, "f32")
        e.set_launch(64, rows, 1, 4)
        return e
    _twin_runs(make, 4, steps_per_call=3)


@pytest.mark.parametrize("name", ["heated", "convection"])
def test_poisoned_destination_thermal(name):
    tc = THERMAL_CASES[name]
    _twin_runs(lambda: build_gpu_thermal(tc, "f32")[0], 5)


def test_poisoned_destination_ensemble():
    names = ["cylinder", "cavity", "step", "plate", "bouzidi"]
    cs = [CASES[n] for n in names]

    def make():
        ens = G.Ensemble([gpu_desc(c) for c in cs], storage="f32")
        for l, c in enumerate(cs):
            o = build_oracle(c)
            apply_gpu_geometry(ens, l, c, o.obstacle.copy())
            ens.upload(l, o.f.copy())
        return ens
    _twin_runs(make, 5, lattices=tuple(range(len(cs))))


def test_poisoned_destination_every_launch_shape():
    c = CASES["bouzidi"]
    e, _ = build_gpu(c, "f32")
    shapes = [(bx, rows, v) for bx in (64, 128, 256, 512) for rows in (1, 2, 4, 8) for v in (1, 2, 4)]
    tried = 0
    for bx, rows, v in shapes:
        def make():
            k, _ = build_gpu(c, "f32")
            k.set_launch(bx, rows, v, 3)
            return k
        try:
            make()
        except G.LbmError:
            continue  # not a compiled shape
        _twin_runs(make, 3)
        tried += 1
    assert tried >= 8
