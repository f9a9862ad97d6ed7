// thermal.hpp -- the thermal double-distribution helpers of the lbm-2d solver interface
// (src/thermal.hpp of the reference), for case drivers that set up and post-process the
// temperature populations on the host.  The step itself (collision, Boussinesq forcing,
// streaming of g) runs in the GPU library (lbm_set_thermal, include/lbm_gpu.h).
#pragma once

#include <array>
#include <cmath>

#include "lbm_types.hpp"

// D2Q9 weights of the thermal populations (the momentum weights).
constexpr std::array<double, 9> thermal_weights = {4.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0, 1.0 / 9.0,
                                                   1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};

// g_i^eq = w_i T (1 + 3 e.u + 4.5 (e.u)^2 - 1.5 u.u), reference temperature 1.
inline double thermal_equilibrium(int i, double T, double u, double v)
{
    const double eu = cx[i] * u + cy[i] * v;
    const double uu = u * u + v * v;
    return thermal_weights[i] * T * (1.0 + 3.0 * eu + 4.5 * eu * eu - 1.5 * uu);
}

// Single-relaxation-time collision of one node's g toward the equilibrium at (T, u, v).
inline void thermal_collide(double* g_node, double T, double u, double v, double omega_k)
{
    for (int i = 0; i < 9; ++i) {
        const double geq = thermal_equilibrium(i, T, u, v);
        g_node[i] -= omega_k * (g_node[i] - geq);
    }
}

// Boussinesq buoyancy F = -rho_0 beta (T - T_ref) g.
inline void boussinesq_force(double T, double T_ref, double rho_0, double beta, double g_x, double g_y, double& fx,
                             double& fy)
{
    const double dT = T - T_ref;
    fx = -rho_0 * beta * dT * g_x;
    fy = -rho_0 * beta * dT * g_y;
}

// Temperature of a node: the sum of its thermal populations in slot order.
inline void compute_temperature(const double* g_node, double& T)
{
    T = 0.0;
    for (int i = 0; i < 9; ++i) T += g_node[i];
}

enum class ThermalBC { ISOTHERMAL, ADIABATIC, HEAT_FLUX };

// Dirichlet wall: the node's g set to the equilibrium at the wall temperature and velocity.
inline void apply_thermal_isothermal(double* g_node, double T_wall, double u_wall = 0.0, double v_wall = 0.0)
{
    for (int i = 0; i < 9; ++i) g_node[i] = thermal_equilibrium(i, T_wall, u_wall, v_wall);
}

// Adiabatic wall: g[i] taken from g_opp[opposite(i)].
inline void apply_thermal_adiabatic(double* g_node, const double* g_opp)
{
    for (int i = 0; i < 9; ++i) g_node[i] = g_opp[bounce_back[i]];
}

// Nu = -2 (dT/dr)|_wall D / (T_wall - T_inf); zero when the walls and the stream are at
// one temperature.
inline double compute_nusselt(double dTdr_wall, double D, double T_wall, double T_inf)
{
    if (std::abs(T_wall - T_inf) < 1e-12) return 0.0;
    return -2.0 * dTdr_wall * D / (T_wall - T_inf);
}
