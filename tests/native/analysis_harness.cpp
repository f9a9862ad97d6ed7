// C entry points over lbm-2d_amd/include/analysis.hpp for tests/test_analysis_cpp.py.
#include <cmath>
#include <vector>

#include "analysis.hpp"

extern "C" {

void ah_dft(const double* re, const double* im, int n, double* ore, double* oim)
{
    std::vector<std::complex<double>> x(n);
    for (int i = 0; i < n; ++i) x[i] = {re[i], im[i]};
    const auto y = lbm_analysis::detail::dft(x);
    for (int i = 0; i < n; ++i) {
        ore[i] = y[i].real();
        oim[i] = y[i].imag();
    }
}

void ah_welch(const double* x, int n, int nperseg, int noverlap, double* freqs, double* psd)
{
    std::vector<double> f, p;
    lbm_analysis::welch_psd(std::vector<double>(x, x + n), nperseg, noverlap, f, p);
    for (size_t k = 0; k < f.size(); ++k) {
        freqs[k] = f[k];
        psd[k] = p[k];
    }
}

void ah_coefficients(const double* fx, const double* fy, int n, double length, double u, double* cd, double* cl)
{
    std::vector<double> a, b;
    lbm_analysis::coefficients(std::vector<double>(fx, fx + n), std::vector<double>(fy, fy + n), length, u, a, b);
    for (int i = 0; i < n; ++i) {
        cd[i] = a[i];
        cl[i] = b[i];
    }
}

void ah_late_stats(const double* cd, const double* cl, int n, double* out)
{
    lbm_analysis::late_stats(std::vector<double>(cd, cd + n), std::vector<double>(cl, cl + n), out[0], out[1]);
}

double ah_st_welch(const double* cl, int n, double length, double u)
{
    const auto s = lbm_analysis::strouhal_welch(std::vector<double>(cl, cl + n), length, u);
    return s ? *s : NAN;
}

double ah_st_crossings(const double* cl, int n, double length, double u, double tail, double hysteresis)
{
    const auto s = lbm_analysis::strouhal_crossings(std::vector<double>(cl, cl + n), length, u, tail, hysteresis);
    return s ? *s : NAN;
}

}  // extern "C"
