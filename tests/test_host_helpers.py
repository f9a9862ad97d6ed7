"""Host-side helpers of the package that rebuild the drivers' initial state: the
mt19937 / uniform_real_distribution pair and the seed-42 perturbation."""
import numpy as np
import pytest

import lbm2d as G
import oracle as O


def test_mt19937_known_answer():
    # the C++ standard fixes the 10000th output of a default-seeded mt19937
    g = G.Mt19937()
    for _ in range(9999):
        g.next()
    assert g.next() == 4123659995


def test_mt19937_matches_oracle_generator():
    a, b = G.Mt19937(42), O.MT19937(42)
    assert [a.next() for _ in range(1500)] == [b.next() for _ in range(1500)]
    assert [a.uniform(-1e-4, 1e-4) for _ in range(700)] == [b.uniform(-1e-4, 1e-4) for _ in range(700)]


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")
def test_uniform_draws_match_reference_library():
    g = G.Mt19937(42)
    ours = np.array([g.uniform(-1e-4, 1e-4) for _ in range(2000)])
    np.testing.assert_array_equal(ours, O.ref_uniform_draws(2000, -1e-4, 1e-4))


def test_perturbation_matches_oracle_bitwise():
    nx, ny, cx, cy, r = 120, 40, 30, 21, 4
    o = O.Oracle(nx, ny, O.CYLINDER)
    o.place_cylinder(cx, cy, r)
    o.init_equilibrium(1.0, 0.1, 0.0)
    o.perturb_seed42(cx + 5, cx + 60)
    mask = G.circle_mask(nx, ny, cx, cy, r)
    np.testing.assert_array_equal(mask, o.obstacle)
    f = G.equilibrium(np.ones((ny, nx)), np.full((ny, nx), 0.1), np.zeros((ny, nx)))
    G.perturb_seed42(f, mask, cx + 5, cx + 60)
    np.testing.assert_array_equal(f, o.f)


def test_seed42_draws_are_the_perturbation_values():
    """seed42_draws holds exactly the values the driver adds to v, in its (x, y) order:
    feq(rho, u, v + dv) with them reproduces the oracle's perturbation bit for bit, and a
    window past the edge or with solids draws nothing for those nodes."""
    nx, ny, cx, cy, r = 120, 40, 30, 21, 4
    o = O.Oracle(nx, ny, O.CYLINDER)
    o.place_cylinder(cx, cy, r)
    o.init_equilibrium(1.0, 0.1, 0.0)
    o.perturb_seed42(cx - 8, nx)
    mask = G.circle_mask(nx, ny, cx, cy, r)
    dv = G.seed42_draws(mask, cx - 8, nx + 30)
    assert dv.shape == (nx - cx + 8, ny)
    solid = mask[:, cx - 8:].T.astype(bool)
    assert np.all(dv[solid] == 0.0)
    assert np.count_nonzero(dv) == np.count_nonzero(~solid)
    f = G.equilibrium(np.ones((ny, nx)), np.full((ny, nx), 0.1), np.zeros((ny, nx)))
    for x in range(cx - 8, nx):
        for y in range(ny):
            if mask[y, x]:
                continue
            c = f[y, x]
            rr = 0.0
            mx = 0.0
            my = 0.0
            for i in range(9):
                rr += c[i]
                mx += c[i] * G.CXV[i]
                my += c[i] * G.CYV[i]
            f[y, x] = G.equilibrium(np.array(rr), np.array(mx / rr), np.array(my / rr + dv[x - cx + 8, y]))
    np.testing.assert_array_equal(f, o.f)


def test_welch_matches_scipy():
    signal = pytest.importorskip("scipy.signal")
    rng = np.random.default_rng(3)
    x = np.sin(2 * np.pi * 0.013 * np.arange(5000)) + 0.3 * rng.standard_normal(5000)
    for nperseg in (256, 1000, 1249):
        f0, p0 = signal.welch(x, fs=1.0, nperseg=nperseg, window="hann", noverlap=nperseg // 2)
        f1, p1 = G.welch_psd(x, nperseg, nperseg // 2)
        np.testing.assert_allclose(f1, f0, rtol=0, atol=1e-15)
        np.testing.assert_allclose(p1, p0, rtol=1e-10, atol=1e-14)


def test_strouhal_of_a_pure_tone():
    # Cl = sin(2 pi f t) with f = St U / L: both estimators return St
    L, U, st = 60.0, 0.1, 0.165
    t = np.arange(60000)
    cl = 0.3 * np.sin(2 * np.pi * st * U / L * t + 0.4)
    assert abs(G.strouhal_crossings(cl, L, U) - st) < 1e-6
    # Welch resolves St to one bin: L / (U * nperseg)
    n = len(cl) - len(cl) // 5
    assert abs(G.strouhal_welch(cl, L, U) - st) <= L / (U * min(n // 4, 16384))


def test_coefficients_and_late_stats():
    fx = np.full(10, 0.3)
    fy = np.array([0, 1, -1, 2, -2, 0.5, -0.25, 0.1, 0.2, -0.1]) * 0.01
    cd, cl = G.coefficients(fx, fy, 60.0, 0.1)
    np.testing.assert_allclose(cd, 0.3 * 2 / (60 * 0.01))
    m, a = G.late_stats(cd, cl)
    assert m == pytest.approx(1.0)
    # latter half: cl[5:] = [0.5,-0.25,0.1,0.2,-0.1] * 2/0.6 * 0.01 -> (max - min) / 2
    assert a == pytest.approx((0.5 + 0.25) / 2 * 0.01 * 2 / 0.6)


def test_strouhal_crossings_ignore_noise_at_the_mean():
    # a tone with step-to-step noise: plain mean crossings would multiply near zero
    L, U, st = 60.0, 0.1, 0.2
    rng = np.random.default_rng(0)
    t = np.arange(80000)
    cl = 0.3 * np.sin(2 * np.pi * st * U / L * t) + 2e-3 * rng.standard_normal(len(t))
    assert abs(G.strouhal_crossings(cl, L, U) - st) / st < 2e-3


@pytest.mark.parametrize("seed", range(4))
def test_strouhal_crossings_ignore_noise_near_the_thresholds(seed):
    # white noise at 7% of the amplitude: a zero crossing that noise makes while the
    # signal still sits near -h must not count (a plain "below -h since the last one"
    # rule counts it and the next real crossing too)
    L, U, st = 60.0, 0.1, 0.2
    rng = np.random.default_rng(seed)
    t = np.arange(80000)
    cl = 0.3 * np.sin(2 * np.pi * st * U / L * t) + 0.02 * rng.standard_normal(len(t))
    assert abs(G.strouhal_crossings(cl, L, U) - st) / st < 2e-3
