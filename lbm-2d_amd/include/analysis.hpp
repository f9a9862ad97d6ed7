// analysis.hpp -- force-history analysis for case drivers without Python: drag and lift
// coefficients, their late-time statistics, and the Strouhal number from the lift
// history, as the reference computes them in its cylinder driver (src/main.cpp:112-115,
// 185-195) and post-processing script (scripts/postprocess.py:386-441, Welch's method).
// Header-only, host code, no dependencies.
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstddef>
#include <optional>
#include <vector>

namespace lbm_analysis {

// Cd = 2 Fx / (L u^2), Cl = 2 Fy / (L u^2), per step.
inline void coefficients(const std::vector<double>& fx, const std::vector<double>& fy, double length, double u,
                         std::vector<double>& cd, std::vector<double>& cl)
{
    const double k = 2.0 / (length * u * u);
    cd.resize(fx.size());
    cl.resize(fy.size());
    for (size_t i = 0; i < fx.size(); ++i) cd[i] = fx[i] * k;
    for (size_t i = 0; i < fy.size(); ++i) cl[i] = fy[i] * k;
}

// Mean Cd over the latter half of the history (summed in step order) and the Cl
// amplitude (max - min) / 2 there, with the extremes started at 0.
inline void late_stats(const std::vector<double>& cd, const std::vector<double>& cl, double& cd_mean,
                       double& cl_amp)
{
    const size_t n = cd.size(), late = static_cast<size_t>(static_cast<double>(n) * 0.5);
    double s = 0.0, hi = 0.0, lo = 0.0;
    for (size_t i = late; i < n; ++i) s += cd[i];
    for (size_t i = late; i < cl.size(); ++i) {
        hi = std::max(hi, cl[i]);
        lo = std::min(lo, cl[i]);
    }
    cd_mean = n > late ? s / static_cast<double>(n - late) : 0.0;
    cl_amp = (hi - lo) / 2.0;
}

namespace detail {

using cplx = std::complex<double>;

// In-place iterative radix-2 FFT (length a power of two); inverse without the 1/n.
inline void fft_pow2(std::vector<cplx>& a, bool inverse)
{
    const size_t n = a.size();
    for (size_t i = 1, j = 0; i < n; ++i) {
        size_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(a[i], a[j]);
    }
    const double pi = std::acos(-1.0);
    for (size_t len = 2; len <= n; len <<= 1) {
        const double ang = 2.0 * pi / static_cast<double>(len) * (inverse ? 1.0 : -1.0);
        for (size_t i = 0; i < n; i += len)
            for (size_t k = 0; k < len / 2; ++k) {
                const cplx w = std::polar(1.0, ang * static_cast<double>(k));
                const cplx u = a[i + k], v = a[i + k + len / 2] * w;
                a[i + k] = u + v;
                a[i + k + len / 2] = u - v;
            }
    }
}

// Discrete Fourier transform of any length: radix-2 directly, otherwise Bluestein's
// chirp-z form as a convolution of power-of-two length.
inline std::vector<cplx> dft(const std::vector<cplx>& x)
{
    const size_t n = x.size();
    if (n == 0) return {};
    if ((n & (n - 1)) == 0) {
        std::vector<cplx> a = x;
        fft_pow2(a, false);
        return a;
    }
    const double pi = std::acos(-1.0);
    size_t m = 1;
    while (m < 2 * n - 1) m <<= 1;
    std::vector<cplx> chirp(n), a(m, 0.0), b(m, 0.0);
    for (size_t k = 0; k < n; ++k) {
        // exp(-i pi k^2 / n), with k^2 reduced mod 2n to keep the angle small
        const size_t k2 = (k * k) % (2 * n);
        chirp[k] = std::polar(1.0, -pi * static_cast<double>(k2) / static_cast<double>(n));
    }
    for (size_t k = 0; k < n; ++k) a[k] = x[k] * chirp[k];
    b[0] = std::conj(chirp[0]);
    for (size_t k = 1; k < n; ++k) b[k] = b[m - k] = std::conj(chirp[k]);
    fft_pow2(a, false);
    fft_pow2(b, false);
    for (size_t k = 0; k < m; ++k) a[k] *= b[k];
    fft_pow2(a, true);
    std::vector<cplx> out(n);
    for (size_t k = 0; k < n; ++k) out[k] = a[k] * chirp[k] / static_cast<double>(m);
    return out;
}

}  // namespace detail

// One-sided Welch power spectral density at unit sampling rate: segments of nperseg
// values overlapping by noverlap, each with its mean removed and a periodic Hann window,
// averaged |FFT|^2 scaled to a density.  freqs[k] = k * (1 / nperseg).
inline void welch_psd(const std::vector<double>& x, size_t nperseg, size_t noverlap, std::vector<double>& freqs,
                      std::vector<double>& psd)
{
    const size_t step = nperseg - noverlap;
    const size_t n_seg = x.size() >= nperseg ? (x.size() - noverlap) / step : 0;
    const size_t nf = nperseg / 2 + 1;
    const double pi = std::acos(-1.0);
    std::vector<double> win(nperseg);
    double wsum = 0.0;
    for (size_t i = 0; i < nperseg; ++i) {
        win[i] = 0.5 - 0.5 * std::cos(2.0 * pi * static_cast<double>(i) / static_cast<double>(nperseg));
        wsum += win[i] * win[i];
    }
    psd.assign(nf, 0.0);
    freqs.resize(nf);
    const double df = 1.0 / static_cast<double>(nperseg);
    for (size_t k = 0; k < nf; ++k) freqs[k] = static_cast<double>(k) * df;
    std::vector<detail::cplx> seg(nperseg);
    for (size_t s = 0; s < n_seg; ++s) {
        const double* p = x.data() + s * step;
        double mean = 0.0;
        for (size_t i = 0; i < nperseg; ++i) mean += p[i];
        mean /= static_cast<double>(nperseg);
        for (size_t i = 0; i < nperseg; ++i) seg[i] = (p[i] - mean) * win[i];
        const std::vector<detail::cplx> f = detail::dft(seg);
        for (size_t k = 0; k < nf; ++k) psd[k] += std::norm(f[k]);
    }
    for (size_t k = 0; k < nf; ++k) {
        psd[k] /= static_cast<double>(n_seg ? n_seg : 1) * wsum;
        const bool edge = k == 0 || (nperseg % 2 == 0 && k == nf - 1);
        if (!edge) psd[k] *= 2.0;
    }
}

// Strouhal number from the lift history as the post-processing script computes it:
// the first fifth dropped, Welch with nperseg = min(n/4, 16384) and 50% overlap, the
// spectral peak in [0.01, 0.5] u / length, St = f length / u.  Empty when the history is
// too short or the band holds no frequency.
inline std::optional<double> strouhal_welch(const std::vector<double>& cl, double length, double u)
{
    std::vector<double> t(cl.begin() + static_cast<std::ptrdiff_t>(cl.size() / 5), cl.end());
    if (t.size() < 64) return std::nullopt;
    const size_t nperseg = std::min<size_t>(t.size() / 4, 16384);
    if (nperseg < 4) return std::nullopt;
    double mean = 0.0;
    for (double v : t) mean += v;
    mean /= static_cast<double>(t.size());
    for (double& v : t) v -= mean;
    std::vector<double> f, psd;
    welch_psd(t, nperseg, nperseg / 2, f, psd);
    const double lo = 0.01 * u / length, hi = 0.5 * u / length;
    std::optional<size_t> best;
    for (size_t k = 0; k < f.size(); ++k)
        if (f[k] >= lo && f[k] <= hi && (!best || psd[k] > psd[*best])) best = k;
    if (!best) return std::nullopt;
    return f[*best] * length / u;
}

// Strouhal number from the mean period between upward crossings of Cl's mean over the
// last `tail` of the history, counted by a Schmitt trigger: a crossing counts when the
// signal, having been below -h, rises above +h (h = hysteresis * rms), and is placed by
// linear interpolation at the last upward zero crossing before that rise.  Resolves St
// far finer than one Welch bin when the history holds few periods.
inline std::optional<double> strouhal_crossings(const std::vector<double>& cl, double length, double u,
                                                double tail = 0.5, double hysteresis = 0.5)
{
    const size_t start = static_cast<size_t>(static_cast<double>(cl.size()) * (1.0 - tail));
    std::vector<double> c(cl.begin() + static_cast<std::ptrdiff_t>(start), cl.end());
    if (c.size() < 2) return std::nullopt;
    double mean = 0.0;
    for (double v : c) mean += v;
    mean /= static_cast<double>(c.size());
    double ss = 0.0;
    for (double& v : c) {
        v -= mean;
        ss += v * v;
    }
    const double h = hysteresis * std::sqrt(ss / static_cast<double>(c.size()));
    std::vector<double> times;
    bool armed = false;
    std::ptrdiff_t last_up = -1;
    for (size_t j = 0; j < c.size(); ++j) {
        if (j > 0 && c[j - 1] < 0.0 && c[j] >= 0.0) last_up = static_cast<std::ptrdiff_t>(j) - 1;
        if (c[j] < -h) {
            armed = true;
        } else if (armed && c[j] > h) {
            const size_t i = static_cast<size_t>(last_up);
            times.push_back(static_cast<double>(i) + (-c[i]) / (c[i + 1] - c[i]));
            armed = false;
        }
    }
    if (times.size() < 2) return std::nullopt;
    const double period = (times.back() - times.front()) / static_cast<double>(times.size() - 1);
    return length / period / u;
}

}  // namespace lbm_analysis
