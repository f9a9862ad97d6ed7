// ref_harness.cpp -- C-ABI around the reference solver, for checking the oracle.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile against the reference's own
// headers where they lie (-I <reference>/src); the result goes to oracle/_ref/.
// Nothing of the reference is copied: this file only calls its API
// (LBMCapabilities, place_*, execute_time_step: src/lbm.hpp, src/lbm_types.hpp)
// and reproduces the case drivers' initialisation with the same library calls
// (std::mt19937 + std::uniform_real_distribution, src/main.cpp:60-83).
#include "lbm.hpp"

#include <chrono>
#include <memory>
#include <random>

namespace {
std::unique_ptr<LBMCapabilities> g_sys;
}

extern "C" {

// Sets the process globals (src/lbm_types.hpp:15-32) and builds a fresh state.
void ref_setup(int nx, int ny, int case_type, int collision, int use_les, double cs)
{
    NX = nx;
    NY = ny;
    g_case = static_cast<CaseType>(case_type);
    g_collision = static_cast<CollisionType>(collision);
    g_use_les = use_les != 0;
    g_cs = cs;
    g_sys = std::make_unique<LBMCapabilities>();
}

void ref_place_walls() { place_walls(*g_sys); }
void ref_place_cylinder(int cx_cyl, int cy_cyl, int radius) { place_cylinder(*g_sys, cx_cyl, cy_cyl, radius); }

void ref_place_polygon(const double* xy, int n)
{
    std::vector<std::pair<double, double>> poly;
    for (int k = 0; k < n; ++k) poly.push_back({xy[2 * k], xy[2 * k + 1]});
    place_polygon(*g_sys, poly);
}

void ref_set_mei(int on) { g_sys->bb_geom.use_mei_bb = on != 0; }

void ref_set_moving_wall(int on, double omega, double rcx, double rcy)
{
    g_sys->bb_geom.has_moving_wall = on != 0;
    g_sys->bb_geom.omega = omega;
    g_sys->bb_geom.rot_cx = rcx;
    g_sys->bb_geom.rot_cy = rcy;
}

void ref_set_body_force(double fx) { g_sys->body_force_x = fx; }

void ref_init_equilibrium(double rho, double u, double v)
{
    for (int n = 0; n < NX * NY; ++n)
        for (int i = 0; i < 9; ++i) g_sys->f[n * 9 + i] = compute_equilibrium(i, rho, u, v);
}

// The cylinder drivers' seed-42 perturbation loop (src/main.cpp:68-83).
void ref_perturb_seed42(int x_begin, int x_end)
{
    std::mt19937 rng(42);
    std::uniform_real_distribution<double> dist(-1e-4, 1e-4);
    for (int x = x_begin; x < x_end; ++x)
        for (int y = 0; y < NY; ++y) {
            int n = node_index(x, y);
            if (g_sys->obstacle[n]) continue;
            double* fn = &g_sys->f[n * 9];
            double dv = dist(rng);
            double rho, u, v;
            compute_macros(fn, rho, u, v);
            for (int i = 0; i < 9; ++i) fn[i] = compute_equilibrium(i, rho, u, v + dv);
        }
}

void ref_uniform_draws(int n, double a, double b, double* out)
{
    std::mt19937 rng(42);
    std::uniform_real_distribution<double> dist(a, b);
    for (int k = 0; k < n; ++k) out[k] = dist(rng);
}

void ref_step(int n, double tau, double u_in)
{
    for (int s = 0; s < n; ++s) execute_time_step(*g_sys, tau, u_in);
}

// Wall-clock seconds for n calls of execute_time_step (no I/O).
double ref_time_steps(int n, double tau, double u_in)
{
    auto t0 = std::chrono::steady_clock::now();
    for (int s = 0; s < n; ++s) execute_time_step(*g_sys, tau, u_in);
    auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

double* ref_f() { return g_sys->f.data(); }
unsigned char* ref_obstacle() { return g_sys->obstacle.data(); }
double* ref_fx_body() { return g_sys->fx_body.data(); }
double* ref_fy_body() { return g_sys->fy_body.data(); }
double* ref_wall_dist() { return g_sys->wall_dist.data(); }

// Thermal double distribution: the LBMCapabilities fields (src/lbm_types.hpp:193-201).
void ref_set_thermal(int on, double omega_k, double beta, double g_buoyancy, double t_ref)
{
    g_sys->use_thermal = on != 0;
    g_sys->omega_k = omega_k;
    g_sys->beta = beta;
    g_sys->g_buoyancy = g_buoyancy;
    g_sys->T_ref = t_ref;
}

double* ref_g() { return g_sys->g_thermal.data(); }

// The heated-cylinder driver's wall loop (src/heated_cylinder.cpp:155-164), through
// apply_thermal_isothermal.
void ref_thermal_isothermal_solids(double t_wall)
{
    for (int n = 0; n < NX * NY; ++n)
        if (g_sys->obstacle[n]) apply_thermal_isothermal(&g_sys->g_thermal[n * 9], t_wall);
}

// The thermal-cavity driver's loop after each step (src/thermal_cavity.cpp:153-188),
// through compute_macros, thermal_equilibrium and bounce_back.
void ref_thermal_cavity_walls(double t_hot, double t_cold)
{
    for (int y = 1; y < NY - 1; ++y) {
        const int col[2] = {1, NX - 2};
        const double t[2] = {t_hot, t_cold};
        for (int k = 0; k < 2; ++k) {
            const int n = node_index(col[k], y);
            double rho, u, v;
            compute_macros(&g_sys->f[n * 9], rho, u, v);
            for (int i = 0; i < 9; ++i) g_sys->g_thermal[n * 9 + i] = thermal_equilibrium(i, t[k], u, v);
        }
    }
    for (int x = 1; x < NX - 1; ++x) {
        const int row[2] = {1, NY - 2};
        for (int k = 0; k < 2; ++k) {
            double* g = &g_sys->g_thermal[node_index(x, row[k]) * 9];
            for (int i = 0; i < 9; ++i) g[i] = g[bounce_back[i]];
        }
    }
}

void ref_mrt_collide(double* f9, double rho, double u, double v, double tau, double cs_sq, double y_plus)
{
    MRTParams p = MRTParams::from_tau(tau);
    mrt_collide(f9, rho, u, v, p, cs_sq, tau, y_plus);
}

double ref_q(int is_polygon, double cx_, double cy_, double r, const double* xy, int n, double xf, double yf, int i)
{
    BounceBackGeometry g;
    g.cx = cx_;
    g.cy = cy_;
    g.radius = r;
    if (is_polygon) {
        g.is_polygon = true;
        for (int k = 0; k < n; ++k) g.poly_vertices.push_back({xy[2 * k], xy[2 * k + 1]});
    }
    return g.compute_q(xf, yf, i);
}

}  // extern "C"
