// cylinder_wake -- a case driver written against the lbm-2d solver interface
// (lbm.hpp): cylinder in a channel with the reference driver's setup
// (src/main.cpp:51-115: D = NY/5 circle, equilibrium start at u = 0.1, seed-42
// transverse perturbation, Cd/Cl from the summed node forces every step).
//
//   cylinder_wake NX NY CX CY R TAU STEPS STORAGE MODE
//     STORAGE: f64 | f32 | f16s | bf16s
//     MODE:    step (execute_time_step per step) | run (lbm_gpu::run_steps)
//
// Prints one JSON line per step ({"step","fx","fy"}) and a final line with the
// population checksums and the force analysis of analysis.hpp (mean Cd and Cl amplitude
// over the latter half, Strouhal number from Cl's crossings; -1 where undefined).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>

#include "analysis.hpp"
#include "lbm.hpp"

int main(int argc, char** argv)
{
    if (argc != 10) {
        std::fprintf(stderr, "usage: %s NX NY CX CY R TAU STEPS STORAGE MODE\n", argv[0]);
        return 2;
    }
    NX = std::atoi(argv[1]);
    NY = std::atoi(argv[2]);
    const int ccx = std::atoi(argv[3]), ccy = std::atoi(argv[4]), r = std::atoi(argv[5]);
    const double tau = std::atof(argv[6]);
    const int steps = std::atoi(argv[7]);
    const std::string storage = argv[8], mode = argv[9];
    lbm_gpu::config().storage = storage == "f64" ? LBM_STORE_F64
                              : storage == "f32" ? LBM_STORE_F32
                              : storage == "f16s" ? LBM_STORE_F16S
                                                  : LBM_STORE_BF16S;
    g_case = CaseType::CYLINDER;
    const double u_in = 0.1;

    LBMCapabilities sys;
    place_cylinder(sys, ccx, ccy, r);
    for (int n = 0; n < NX * NY; ++n)
        for (int i = 0; i < 9; ++i) sys.f[static_cast<size_t>(n) * 9 + i] = compute_equilibrium(i, 1.0, u_in, 0.0);
    std::mt19937 rng(42);
    std::uniform_real_distribution<double> pert(-1e-4, 1e-4);
    for (int x = ccx + 5; x < std::min(NX, ccx + 60); ++x)
        for (int y = 0; y < NY; ++y) {
            const int n = node_index(x, y);
            if (sys.obstacle[n]) continue;
            double* fn = &sys.f[static_cast<size_t>(n) * 9];
            const double dv = pert(rng);
            double rho, u, v;
            compute_macros(fn, rho, u, v);
            for (int i = 0; i < 9; ++i) fn[i] = compute_equilibrium(i, rho, u, v + dv);
        }

    std::vector<double> fx, fy;
    try {
        if (mode == "step") {
            for (int s = 0; s < steps; ++s) {
                execute_time_step(sys, tau, u_in);
                double sx = 0.0, sy = 0.0;
                for (int n = 0; n < NX * NY; ++n) {
                    sx += sys.fx_body[n];
                    sy += sys.fy_body[n];
                }
                fx.push_back(sx);
                fy.push_back(sy);
                std::printf("{\"step\":%d,\"fx\":%.17g,\"fy\":%.17g}\n", s, sx, sy);
            }
        } else {
            lbm_gpu::run_steps(sys, steps, tau, u_in, &fx, &fy);
            for (size_t s = 0; s < fx.size(); ++s) std::printf("{\"step\":%zu,\"fx\":%.17g,\"fy\":%.17g}\n", s, fx[s], fy[s]);
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    double sum = 0.0, sq = 0.0;
    for (double v : sys.f) {
        sum += v;
        sq += v * v;
    }
    std::vector<double> cd, cl;
    lbm_analysis::coefficients(fx, fy, 2.0 * r, u_in, cd, cl);
    double cd_mean = 0.0, cl_amp = 0.0;
    lbm_analysis::late_stats(cd, cl, cd_mean, cl_amp);
    const auto st = lbm_analysis::strouhal_crossings(cl, 2.0 * r, u_in);
    std::printf("{\"sum\":%.17g,\"sumsq\":%.17g,\"cd_mean\":%.17g,\"cl_amp\":%.17g,\"st\":%.17g}\n", sum, sq,
                cd_mean, cl_amp, st ? *st : -1.0);
    return 0;
}
