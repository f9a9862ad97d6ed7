"""The C++ force analysis (lbm-2d_amd/include/analysis.hpp) against the Python one in
lbm2d.py (whose Welch estimator is checked against scipy in test_host_helpers.py) and
its FFT against numpy's, on synthetic signals and on the FP64 reference run's force
history (tests/golden/cylinder800x300_re100_forces.npz)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import lbm2d as G

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
_dp = C.POINTER(C.c_double)


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    so = str(tmp_path_factory.mktemp("analysis") / "libanalysis_harness.so")
    out = subprocess.run(["g++", "-std=c++17", "-O2", "-shared", "-fPIC", "-I", os.path.join(ROOT, "lbm-2d_amd", "include"),
                          os.path.join(HERE, "native", "analysis_harness.cpp"), "-o", so],
                         capture_output=True, text=True)
    assert out.returncode == 0, out.stderr
    L = C.CDLL(so)
    L.ah_dft.argtypes = [_dp, _dp, C.c_int, _dp, _dp]
    L.ah_welch.argtypes = [_dp, C.c_int, C.c_int, C.c_int, _dp, _dp]
    L.ah_coefficients.argtypes = [_dp, _dp, C.c_int, C.c_double, C.c_double, _dp, _dp]
    L.ah_late_stats.argtypes = [_dp, _dp, C.c_int, _dp]
    L.ah_st_welch.restype = C.c_double
    L.ah_st_welch.argtypes = [_dp, C.c_int, C.c_double, C.c_double]
    L.ah_st_crossings.restype = C.c_double
    L.ah_st_crossings.argtypes = [_dp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]
    return L


def _p(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(_dp)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 12, 100, 1000, 1024, 4097])
def test_dft_matches_numpy(lib, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    re, im = np.ascontiguousarray(x.real), np.ascontiguousarray(x.imag)
    ore, oim = np.zeros(n), np.zeros(n)
    lib.ah_dft(_p(re), _p(im), n, _p(ore), _p(oim))
    ref = np.fft.fft(x)
    assert np.abs(ore + 1j * oim - ref).max() <= 1e-11 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("nperseg", [256, 300, 1001])
def test_welch_matches_python(lib, nperseg):
    rng = np.random.default_rng(nperseg)
    x = np.sin(0.05 * np.arange(9000)) + 0.3 * rng.standard_normal(9000)
    nf = nperseg // 2 + 1
    f, p = np.zeros(nf), np.zeros(nf)
    lib.ah_welch(_p(x), len(x), nperseg, nperseg // 2, _p(f), _p(p))
    fr, pr = G.welch_psd(x, nperseg, nperseg // 2)
    np.testing.assert_array_equal(f, fr)
    np.testing.assert_allclose(p, pr, rtol=1e-9, atol=1e-12 * pr.max())


def test_reference_run_statistics_match_python(lib):
    z = np.load(os.path.join(HERE, "golden", "cylinder800x300_re100_forces.npz"))
    fx, fy = z["fx"], z["fy"]
    n = len(fx)
    cd, cl = np.zeros(n), np.zeros(n)
    lib.ah_coefficients(_p(fx), _p(fy), n, 60.0, 0.1, _p(cd), _p(cl))
    pcd, pcl = G.coefficients(fx, fy, 60.0, 0.1)
    np.testing.assert_array_equal(cd, pcd)
    np.testing.assert_array_equal(cl, pcl)
    out = np.zeros(2)
    lib.ah_late_stats(_p(cd), _p(cl), n, _p(out))
    assert tuple(out) == G.late_stats(pcd, pcl)
    st_w = lib.ah_st_welch(_p(cl), n, 60.0, 0.1)
    assert st_w == pytest.approx(G.strouhal_welch(pcl, 60.0, 0.1), rel=1e-12)
    st_x = lib.ah_st_crossings(_p(cl), n, 60.0, 0.1, 0.5, 0.5)
    assert st_x == pytest.approx(G.strouhal_crossings(pcl, 60.0, 0.1), rel=1e-12)
    print(f"\nreference run: Cd {out[0]:.6f}  Cl amp {out[1]:.6f}  St (Welch) {st_w:.5f}  St (crossings) {st_x:.5f}")


def test_tone_and_short_histories(lib):
    L, U, st = 60.0, 0.1, 0.2
    t = np.arange(80000)
    rng = np.random.default_rng(7)
    cl = 0.3 * np.sin(2 * np.pi * st * U / L * t) + 0.02 * rng.standard_normal(len(t))
    assert lib.ah_st_crossings(_p(cl), len(cl), L, U, 0.5, 0.5) == pytest.approx(G.strouhal_crossings(cl, L, U),
                                                                               rel=1e-12)
    assert abs(lib.ah_st_welch(_p(cl), len(cl), L, U) - st) <= L / (U * 16000)
    short = np.zeros(50)
    assert np.isnan(lib.ah_st_welch(_p(short), 50, L, U))
    assert np.isnan(lib.ah_st_crossings(_p(short), 50, L, U, 0.5, 0.5))
