// lbm_types.hpp -- lattice constants, process globals and state of the lbm-2d
// solver interface (src/lbm_types.hpp of the reference), backed by the MI355X
// library of include/lbm_gpu.h.  Names, types and meanings follow the reference so
// that its case drivers compile unchanged; the time step itself runs on the GPU
// (lbm.hpp, execute_time_step).
#pragma once

#include <array>
#include <cmath>
#include <cstdint>
#include <deque>
#include <string>
#include <utility>
#include <vector>

// Grid dimensions, set by the case driver before LBMCapabilities is built
// (src/lbm_types.hpp:15-16).
inline int NX = 800;
inline int NY = 300;
constexpr int NUM_DIRECTIONS = 9;

// Case and collision selection (src/lbm_types.hpp:20-28); the enumerator order is the
// C ABI's case numbering (lbm_desc_for_case).
enum class CaseType { CYLINDER, CAVITY, STEP, RIBS, URBAN_CANYON, DOWNWASH,
                      SQUARE_CYLINDER, FLAT_PLATE, SPORTS_BALL,
                      PERIODIC_HILLS, CYLINDER_NEAR_WALL, SIDE_BY_SIDE,
                      ROTATING_CYLINDER, ORIFICE_PLATE };
enum class CollisionType { MRT, BGK };
inline CollisionType g_collision = CollisionType::MRT;
inline CaseType g_case = CaseType::CYLINDER;

// Smagorinsky LES switch and constant (src/lbm_types.hpp:31-32).
inline bool g_use_les = false;
inline double g_cs = 0.12;

// D2Q9 velocity set, weights and opposite directions (src/lbm_types.hpp:41-55):
//   6 2 5
//   3 0 1
//   7 4 8
constexpr std::array<int, 9> cx = {0, 1, 0, -1, 0, 1, -1, -1, 1};
constexpr std::array<int, 9> cy = {0, 0, 1, 0, -1, 1, 1, -1, -1};
constexpr std::array<double, 9> weights = {4.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0, 1.0 / 9.0,
                                           1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
constexpr std::array<int, 9> bounce_back = {0, 3, 4, 1, 2, 7, 8, 5, 6};

// MRT rates from tau (src/lbm_types.hpp:60-77): s_shear = clamp(1/tau, 0.5, 1.99),
// s_bulk = 1.2, s_normal = 1.0.
struct MRTParams {
    double s_shear;
    double s_bulk;
    double s_normal;
    static MRTParams from_tau(double tau)
    {
        const double s = 1.0 / tau;
        return {s < 0.5 ? 0.5 : (s > 1.99 ? 1.99 : s), 1.2, 1.0};
    }
};

inline int node_index(int x, int y) { return y * NX + x; }

// The interpolated bounce-back geometry: one circle or one polygon
// (src/lbm_types.hpp:90-182).  q is evaluated once per link by the library.
struct BounceBackGeometry {
    double cx = 0.0, cy = 0.0, radius = 0.0;
    std::vector<std::pair<double, double>> poly_vertices;
    bool is_polygon = false;
    bool has_moving_wall = false;
    double omega = 0.0;
    double rot_cx = 0.0, rot_cy = 0.0;
    bool use_mei_bb = true;

    bool is_valid() const { return radius > 0.0 || is_polygon; }

    // Tangential wall velocity omega x r at a point (src/lbm_types.hpp:106-114).
    void compute_wall_velocity(double px, double py, double& u_wall_x, double& u_wall_y) const
    {
        if (!has_moving_wall) {
            u_wall_x = u_wall_y = 0.0;
            return;
        }
        u_wall_x = -omega * (py - rot_cy);
        u_wall_y = omega * (px - rot_cx);
    }
};

// Simulation state (src/lbm_types.hpp:187-224), host copies in the reference layout.
// After execute_time_step, f holds the post-stream, post-boundary populations and
// fx_body/fy_body the per-node momentum exchange, as in the reference.  f_next is
// kept for source compatibility; the library does not use it.
struct LBMCapabilities {
    std::vector<double> f;
    std::vector<double> f_next;
    std::vector<uint8_t> obstacle;
    std::vector<double> wall_dist;

    std::vector<double> g_thermal;
    std::vector<double> g_thermal_next;
    bool use_thermal = false;  // thermal double distribution g (stepped on the GPU with f)
    double T_ref = 1.0;
    double alpha = 0.1666666667;
    double omega_k = 1.0;
    double beta = 0.0;
    double g_buoyancy = 0.0;

    std::vector<double> fx_body;
    std::vector<double> fy_body;

    BounceBackGeometry bb_geom;
    double body_force_x = 0.0;

    LBMCapabilities()
    {
        const size_t n = static_cast<size_t>(NX) * NY;
        f.assign(n * NUM_DIRECTIONS, 0.0);
        f_next.assign(n * NUM_DIRECTIONS, 0.0);
        obstacle.assign(n, 0);
        wall_dist.assign(n, 0.0);
        g_thermal.assign(n * NUM_DIRECTIONS, 1.0);  // T = 9 until a driver sets it
        g_thermal_next.assign(n * NUM_DIRECTIONS, 1.0);
        reset_forces();
    }

    void reset_forces()
    {
        fx_body.assign(static_cast<size_t>(NX) * NY, 0.0);
        fy_body.assign(static_cast<size_t>(NX) * NY, 0.0);
    }
};

// f_i^eq = w_i rho (1 + 3 e.u + 4.5 (e.u)^2 - 1.5 u.u)  (src/lbm_types.hpp:269-273).
inline double compute_equilibrium(int i, double rho, double u, double v)
{
    const double eu = cx[i] * u + cy[i] * v;
    const double uu = u * u + v * v;
    return weights[i] * rho * (1.0 + 3.0 * eu + 4.5 * eu * eu - 1.5 * uu);
}

// Density and velocity moments of one node (src/lbm_types.hpp:278-288).
inline void compute_macros(const double* f_node, double& rho, double& u, double& v)
{
    double r = 0.0, mx = 0.0, my = 0.0;
    for (int i = 0; i < 9; ++i) {
        r += f_node[i];
        mx += f_node[i] * cx[i];
        my += f_node[i] * cy[i];
    }
    rho = r;
    u = mx / r;
    v = my / r;
}

// Distance of every node to the nearest solid node along 4-neighbour paths
// (src/lbm_types.hpp:231-264), by breadth-first search; 1e9 where none is reachable.
inline void compute_wall_distance(LBMCapabilities& sys)
{
    const int n = NX * NY;
    std::vector<double>& d = sys.wall_dist;
    d.assign(n, 1e9);
    std::deque<int> q;
    for (int c = 0; c < n; ++c)
        if (sys.obstacle[c]) {
            d[c] = 0.0;
            q.push_back(c);
        }
    while (!q.empty()) {
        const int c = q.front();
        q.pop_front();
        const int x = c % NX, y = c / NX;
        const int nb[4][2] = {{x - 1, y}, {x + 1, y}, {x, y - 1}, {x, y + 1}};
        for (const auto& p : nb) {
            if (p[0] < 0 || p[0] >= NX || p[1] < 0 || p[1] >= NY) continue;
            const int m = node_index(p[0], p[1]);
            if (d[m] > d[c] + 1.0) {
                d[m] = d[c] + 1.0;
                q.push_back(m);
            }
        }
    }
}
