// writers_check.cpp -- TEST INFRASTRUCTURE ONLY.  Writes every output format of the
// solver interface from one fixed state, for tests/test_writers.py: oracle/Makefile
// builds it against the reference's headers into _ref/writers_ref, the test builds it
// against lbm-2d_amd/include, and the two sets of files are compared byte for byte.
// Nothing of the reference is copied: it only calls the interface's functions
// (place_cylinder, compute_equilibrium, thermal_equilibrium, node_index, save_*).
#include <cmath>
#include <filesystem>
#include <string>

#include "lbm.hpp"

int main(int argc, char** argv)
{
    if (argc != 2) return 2;
    const std::string out = argv[1];
    std::filesystem::create_directories(out);
    NX = 230;  // > 100 so the frames are sampled every 2 nodes
    NY = 41;
    LBMCapabilities sys;
    place_cylinder(sys, 60, 20, 9);
    for (int n = 0; n < NX * NY; ++n) {
        const double rho = 1.0 + 0.01 * std::sin(0.37 * n), u = 0.05 * std::cos(0.11 * n), v = 0.02 * std::sin(0.07 * n);
        const double T = 1.0 + 0.3 * std::cos(0.05 * n);
        for (int i = 0; i < 9; ++i) {
            sys.f[static_cast<size_t>(n) * 9 + i] = compute_equilibrium(i, rho, u, v);
            sys.g_thermal[static_cast<size_t>(n) * 9 + i] = thermal_equilibrium(i, T, u, v);
        }
    }
    // one node with a non-finite density exercises the writers' NaN guards
    sys.f[static_cast<size_t>(node_index(150, 30)) * 9 + 2] = std::nan("");
    save_json_frame(sys, 7, out);
    save_json_frame_thermal(sys, 8, out, 1.5);
    for (int s = 0; s < 3; ++s) save_forces_jsonl(out, s, 1.25 + 0.1 * s, -0.5 + 0.05 * s);
    save_meta_json(out, 100.0, 0.68, 0.1, 60.0, "cylinder", NX, NY);
    save_vtk_frame(sys, 3, out);
    return 0;
}
