// natural_convection -- a thermal case driver written against the lbm-2d solver
// interface (lbm.hpp, thermal.hpp): the differentially heated square box of the
// reference's thermal-cavity driver (src/thermal_cavity.cpp: walls as solid nodes,
// hot x = 0 and cold x = N-1, adiabatic top and bottom, tau = 0.8, Pr = 0.71,
// beta = 0.001 with g from the Rayleigh number, no inflow).
//
//   natural_convection N RA STEPS STORAGE MODE
//     STORAGE: f64 | f32
//     MODE:    step (execute_time_step, then the wall loop on the host, as the driver)
//              run  (the wall loop as device conditions, lbm_gpu::run_steps)
//
// Prints the mean Nusselt number at the hot wall and the checksums of f and g.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "lbm.hpp"

int main(int argc, char** argv)
{
    if (argc != 6) {
        std::fprintf(stderr, "usage: %s N RA STEPS STORAGE MODE\n", argv[0]);
        return 2;
    }
    NX = NY = std::atoi(argv[1]);
    const double ra = std::atof(argv[2]);
    const int steps = std::atoi(argv[3]);
    const std::string storage = argv[4], mode = argv[5];
    lbm_gpu::config().storage = storage == "f64" ? LBM_STORE_F64 : LBM_STORE_F32;
    g_case = CaseType::CYLINDER;  // the thermal driver keeps the default case

    const double th = 1.5, tc = 0.5, tau = 0.8, L = NY;
    const double nu = (tau - 0.5) / 3.0, alpha = nu / 0.71;
    const double beta = 0.001, g_accel = ra * nu * alpha / ((th - tc) * L * L * L) / beta;

    LBMCapabilities sys;
    sys.use_thermal = true;
    sys.T_ref = 1.0;
    sys.alpha = alpha;
    sys.omega_k = 1.0 / (0.5 + 3.0 * alpha);
    sys.beta = beta;
    sys.g_buoyancy = -g_accel;
    for (int y = 0; y < NY; ++y)
        for (int x = 0; x < NX; ++x)
            if (x == 0 || x == NX - 1 || y == 0 || y == NY - 1) sys.obstacle[node_index(x, y)] = 1;
    for (int n = 0; n < NX * NY; ++n)
        for (int i = 0; i < 9; ++i) {
            sys.f[static_cast<size_t>(n) * 9 + i] = compute_equilibrium(i, 1.0, 0.0, 0.0);
            sys.g_thermal[static_cast<size_t>(n) * 9 + i] = thermal_equilibrium(i, 1.0, 0.0, 0.0);
        }
    // wall populations before the loop: isothermal sides, mirrored top and bottom
    for (int y = 0; y < NY; ++y) {
        apply_thermal_isothermal(&sys.g_thermal[static_cast<size_t>(node_index(0, y)) * 9], th);
        apply_thermal_isothermal(&sys.g_thermal[static_cast<size_t>(node_index(NX - 1, y)) * 9], tc);
    }
    for (int x = 1; x < NX - 1; ++x)
        for (int y : {0, NY - 1}) {
            double* g = &sys.g_thermal[static_cast<size_t>(node_index(x, y)) * 9];
            for (int i = 0; i < 9; ++i) g[i] = g[bounce_back[i]];
        }

    auto wall_loop = [&]() {
        for (int y = 1; y < NY - 1; ++y)
            for (auto [x, t] : {std::pair<int, double>{1, th}, {NX - 2, tc}}) {
                const size_t n = static_cast<size_t>(node_index(x, y));
                double rho, u, v;
                compute_macros(&sys.f[n * 9], rho, u, v);
                for (int i = 0; i < 9; ++i) sys.g_thermal[n * 9 + i] = thermal_equilibrium(i, t, u, v);
            }
        for (int x = 1; x < NX - 1; ++x)
            for (int y : {1, NY - 2}) {
                double* g = &sys.g_thermal[static_cast<size_t>(node_index(x, y)) * 9];
                for (int i = 0; i < 9; ++i) g[i] = g[bounce_back[i]];
            }
    };

    try {
        if (mode == "step") {
            for (int s = 0; s < steps; ++s) {
                execute_time_step(sys, tau, 0.0);
                wall_loop();
            }
        } else {
            std::vector<uint8_t> code(static_cast<size_t>(NX) * NY, 0);
            for (int y = 1; y < NY - 1; ++y) {
                code[node_index(1, y)] = 1;
                code[node_index(NX - 2, y)] = 2;
            }
            for (int x = 1; x < NX - 1; ++x) {
                code[node_index(x, 1)] |= LBM_THERMAL_MIRROR;
                code[node_index(x, NY - 2)] |= LBM_THERMAL_MIRROR;
            }
            lbm_gpu::set_thermal_conditions(code, {{th, LBM_THERMAL_EQ_FLOW, 0}, {tc, LBM_THERMAL_EQ_FLOW, 0}});
            lbm_gpu::run_steps(sys, steps, tau, 0.0);
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    // Nusselt number at the hot wall from the first two fluid columns (the driver's)
    double nu_avg = 0.0;
    for (int y = 1; y < NY - 1; ++y) {
        double t1, t2;
        compute_temperature(&sys.g_thermal[static_cast<size_t>(node_index(1, y)) * 9], t1);
        compute_temperature(&sys.g_thermal[static_cast<size_t>(node_index(2, y)) * 9], t2);
        nu_avg += -(t2 - t1) * L / (th - tc);
    }
    nu_avg /= NY - 2;
    double fs = 0.0, fq = 0.0, gs = 0.0, gq = 0.0;
    for (size_t k = 0; k < sys.f.size(); ++k) {
        fs += sys.f[k];
        fq += sys.f[k] * sys.f[k];
        gs += sys.g_thermal[k];
        gq += sys.g_thermal[k] * sys.g_thermal[k];
    }
    std::printf("{\"nu\":%.17g,\"fsum\":%.17g,\"fsumsq\":%.17g,\"gsum\":%.17g,\"gsumsq\":%.17g}\n", nu_avg, fs, fq,
                gs, gq);
    return 0;
}
