"""The FP64 oracle against the reference itself (oracle/_ref, compiled from the
reference's own headers by oracle/Makefile with the same -O2 -ffp-contract=off).

Bit-for-bit equality of populations, per-node forces and wall distance over whole
trajectories of every case.  Skipped where oracle/_ref was not built (the
reference sources are not present); tests/test_golden.py then carries the check.
"""
import numpy as np
import pytest

import oracle as O
from cases import CASES, build_oracle

pytestmark = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")


@pytest.mark.parametrize("name", sorted(CASES))
def test_trajectory_bitwise(name):
    c = CASES[name]
    o = build_oracle(c)
    r = build_oracle(c, cls=O.RefLattice)
    assert np.array_equal(o.f, r.f)
    for s in range(c.steps):
        o.step(c.tau, c.u)
        r.step(c.tau, c.u)
        if not np.array_equal(o.f, r.f):
            d = np.abs(o.f - r.f)
            pytest.fail(f"{name}: step {s}: populations differ, max {d.max():.3e} at "
                        f"{np.unravel_index(d.argmax(), d.shape)}")
    assert np.array_equal(o.fx_body, r.fx_body)
    assert np.array_equal(o.fy_body, r.fy_body)
    if c.les:
        assert np.array_equal(o.wall_dist, r.wall_dist)


def test_faithful_wall_distance_every_step_is_the_same():
    c = CASES["les"]
    o = O.Oracle(c.nx, c.ny, c.case, les=True, cs=c.cs, faithful=True)
    c.setup(o)
    o.init_equilibrium(*c.init)
    r = build_oracle(c, cls=O.RefLattice)
    o.perturb_seed42(*c.pert)
    for _ in range(5):
        o.step(c.tau, c.u)
        r.step(c.tau, c.u)
    assert np.array_equal(o.f, r.f)


def test_perturbation_draws_match_libstdcxx():
    ref = O.ref_uniform_draws(1000, -1e-4, 1e-4)
    g = O.MT19937(42)
    mine = np.array([g.uniform(-1e-4, 1e-4) for _ in range(1000)])
    assert np.array_equal(mine, ref)


def test_mrt_and_les_bitwise_on_random_nodes():
    rng = np.random.default_rng(11)
    for _ in range(200):
        rho = rng.uniform(0.9, 1.1)
        u, v = rng.uniform(-0.15, 0.15, 2)
        f = np.array([O.equilibrium(i, rho, u, v) for i in range(9)]) * (1 + 0.02 * rng.standard_normal(9))
        tau = rng.uniform(0.501, 2.0)
        cs2 = rng.choice([0.0, 0.01, 0.0144])
        yp = rng.choice([0.0, rng.uniform(0.1, 100)])
        assert np.array_equal(O.mrt_collide(f, tau, cs_sq=cs2, y_plus=yp),
                              O.ref_mrt_collide(f, tau, cs_sq=cs2, y_plus=yp))


def test_q_bitwise_circle_and_polygon():
    rng = np.random.default_rng(12)
    poly = np.array([[10.2, 10.1], [20.7, 12.4], [18.3, 19.9], [9.1, 17.2]])
    for _ in range(300):
        xf, yf = rng.uniform(5, 25, 2).round()
        i = int(rng.integers(1, 9))
        assert O.q_circle(15.0, 15.0, 4.5, xf, yf, i) == O.ref_q(xf, yf, i, circle=(15.0, 15.0, 4.5))
        assert O.q_polygon(poly, xf, yf, i) == O.ref_q(xf, yf, i, polygon=poly)
