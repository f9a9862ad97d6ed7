// lbm.hpp -- the lbm-2d solver interface (src/lbm.hpp of the reference) on the MI355X
// library of include/lbm_gpu.h.
//
// execute_time_step(sys, tau, u_inflow) keeps its contract: on return sys.f holds the
// post-stream, post-boundary populations and, for every case but the cavity,
// sys.fx_body / sys.fy_body the per-node momentum exchange.  Each call moves sys.f to
// the GPU, takes one step there and brings the populations back, so drivers that read
// or write sys.f between steps keep working unchanged.  lbm_gpu::run_steps() is the
// device-resident form for many steps.  With sys.use_thermal the thermal double
// distribution sys.g_thermal moves with sys.f and steps on the GPU with it (omega_k,
// beta, g_buoyancy, T_ref from sys); the drivers' own loops on g between steps keep
// working, or lbm_gpu::set_thermal_conditions() runs them on the device in run_steps.
//
// Link against liblbm2d.so (lbm-2d_amd/Makefile).  lbm_gpu::config().storage selects
// the population storage of the GPU lattice (FP64 by default here, for results equal
// to the reference's FP64 step to rounding).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/lbm_gpu.h"
#include "geometry.hpp"
#include "lbm_types.hpp"
#include "thermal.hpp"

namespace lbm_gpu {

struct Config {
    int storage = LBM_STORE_F64;  // enum lbm_storage
    int scale_exp = 8;            // 16-bit forms
    int graph_steps = 100;        // run_steps: hipGraph length
    double shift_u = 0.0;         // 16-bit forms: store f - feq(1, shift_u, 0), e.g. the inflow
    bool stochastic_rounding = false;  // 16-bit forms: zero-mean rounding of stored values
    uint32_t rounding_seed = 0;
};

// The storage form of unchanged drivers comes from the environment: LBM2D_STORAGE =
// f64 (default), f32, f16s or bf16s; LBM2D_STOCHASTIC_ROUNDING=1 for the 16-bit forms.
inline Config& config()
{
    static Config c = [] {
        Config k;
        if (const char* s = std::getenv("LBM2D_STORAGE")) {
            const std::string v(s);
            if (v == "f64") k.storage = LBM_STORE_F64;
            else if (v == "f32") k.storage = LBM_STORE_F32;
            else if (v == "f16s") k.storage = LBM_STORE_F16S;
            else if (v == "bf16s") k.storage = LBM_STORE_BF16S;
            else throw std::runtime_error("LBM2D_STORAGE must be f64, f32, f16s or bf16s, not " + v);
        }
        if (const char* s = std::getenv("LBM2D_STOCHASTIC_ROUNDING")) k.stochastic_rounding = std::string(s) == "1";
        return k;
    }();
    return c;
}

inline void check(int rc, const char* what)
{
    if (rc != LBM_OK) throw std::runtime_error(std::string(what) + ": " + lbm_last_error());
}

// One GPU lattice mirroring the process-global lattice of the reference.
struct Context {
    lbm_handle h = nullptr;
    // what the handle was built for
    int nx = 0, ny = 0, case_type = -1, collision = -1, les = -1, storage = -1;
    double cs = 0.0, body_force = 0.0, shift_u = 0.0;
    bool sr = false;
    uint32_t sr_seed = 0;
    // geometry last given to the handle
    std::vector<uint8_t> obstacle;
    BounceBackGeometry geom;
    // fluid nodes with solid neighbours, for the per-node forces: (node, direction)
    std::vector<std::pair<int, int>> links;
    bool have_geometry = false;
    // thermal populations on the handle; device-side conditions for run_steps
    bool thermal = false;
    std::vector<uint8_t> thermal_code;
    std::vector<lbm_thermal_rule> thermal_rules;

    ~Context()
    {
        if (h) lbm_destroy(h);
    }
};

inline Context& context()
{
    static Context c;
    return c;
}

inline bool same_geometry(const BounceBackGeometry& a, const BounceBackGeometry& b)
{
    return a.cx == b.cx && a.cy == b.cy && a.radius == b.radius && a.is_polygon == b.is_polygon &&
           a.poly_vertices == b.poly_vertices && a.has_moving_wall == b.has_moving_wall && a.omega == b.omega &&
           a.rot_cx == b.rot_cx && a.rot_cy == b.rot_cy && a.use_mei_bb == b.use_mei_bb;
}

// Links counted by the momentum exchange, with the reference's neighbour rules
// (src/lbm.hpp:545-573): periodic y everywhere, periodic x for the ribs/hills cases,
// near-wall filter on the cylinder.
inline void build_force_links(Context& C, const LBMCapabilities& sys)
{
    C.links.clear();
    if (g_case == CaseType::CAVITY) return;
    const bool px = g_case == CaseType::RIBS || g_case == CaseType::PERIODIC_HILLS;
    for (int y = 0; y < NY; ++y)
        for (int x = 0; x < NX; ++x) {
            const int n = node_index(x, y);
            if (sys.obstacle[n]) continue;
            for (int i = 0; i < 9; ++i) {
                int tx = x + cx[i], ty = y + cy[i];
                ty = (ty + NY) % NY;
                if (px) tx = (tx + NX) % NX;
                else if (tx < 0 || tx >= NX) continue;
                if (!sys.obstacle[node_index(tx, ty)]) continue;
                if (g_case == CaseType::CYLINDER_NEAR_WALL) {
                    const double dx = tx - sys.bb_geom.cx, dy = ty - sys.bb_geom.cy;
                    if (std::sqrt(dx * dx + dy * dy) > sys.bb_geom.radius) continue;
                }
                C.links.emplace_back(n, i);
            }
        }
}

// Makes the handle match the process globals and sys; uploads sys.f.
inline void prepare(LBMCapabilities& sys, double tau, double u_inflow)
{
    Context& C = context();
    const Config& cfg = config();
    const int case_type = static_cast<int>(g_case);
    const int collision = g_collision == CollisionType::MRT ? LBM_MRT : LBM_BGK;
    const int les = g_use_les ? 1 : 0;
    if (!C.h || C.nx != NX || C.ny != NY || C.case_type != case_type || C.collision != collision || C.les != les ||
        C.cs != g_cs || C.body_force != sys.body_force_x || C.storage != cfg.storage || C.shift_u != cfg.shift_u ||
        C.sr != cfg.stochastic_rounding || C.sr_seed != cfg.rounding_seed || (C.thermal && !sys.use_thermal)) {
        if (C.h) {
            lbm_destroy(C.h);
            C.h = nullptr;
        }
        lbm_desc d;
        check(lbm_desc_for_case(case_type, NX, NY, &d), "lbm_desc_for_case");
        d.collision = collision;
        d.les = les;
        d.cs = g_cs;
        d.body_force_x = sys.body_force_x;
        d.tau = tau;
        d.u_in = u_inflow;
        check(lbm_create(&d, 1, cfg.storage, cfg.scale_exp, nullptr, &C.h), "lbm_create");
        if (cfg.shift_u != 0.0) check(lbm_set_storage_shift(C.h, 0, cfg.shift_u, 0.0), "lbm_set_storage_shift");
        if (cfg.stochastic_rounding)
            check(lbm_set_stochastic_rounding(C.h, 1, cfg.rounding_seed), "lbm_set_stochastic_rounding");
        C.shift_u = cfg.shift_u;
        C.sr = cfg.stochastic_rounding;
        C.sr_seed = cfg.rounding_seed;
        C.nx = NX;
        C.ny = NY;
        C.case_type = case_type;
        C.collision = collision;
        C.les = les;
        C.cs = g_cs;
        C.body_force = sys.body_force_x;
        C.storage = cfg.storage;
        C.have_geometry = false;
        C.thermal = false;
    }
    if (!C.have_geometry || C.obstacle != sys.obstacle || !same_geometry(C.geom, sys.bb_geom)) {
        check(lbm_set_obstacle(C.h, 0, sys.obstacle.data()), "lbm_set_obstacle");
        const BounceBackGeometry& g = sys.bb_geom;
        const bool interp = g.is_valid() && g_case != CaseType::RIBS && g_case != CaseType::PERIODIC_HILLS;
        if (interp) {
            const int scheme = g.use_mei_bb ? LBM_BB_MEI : LBM_BB_BOUZIDI;
            if (g.is_polygon) {
                std::vector<double> xy;
                for (const auto& p : g.poly_vertices) {
                    xy.push_back(p.first);
                    xy.push_back(p.second);
                }
                check(lbm_set_bb_polygon(C.h, 0, xy.data(), static_cast<int32_t>(g.poly_vertices.size()), scheme),
                      "lbm_set_bb_polygon");
            } else {
                check(lbm_set_bb_circle(C.h, 0, g.cx, g.cy, g.radius, scheme), "lbm_set_bb_circle");
            }
            if (g.has_moving_wall) check(lbm_set_moving_wall(C.h, 0, g.omega, g.rot_cx, g.rot_cy), "lbm_set_moving_wall");
        }
        if (g_case == CaseType::CYLINDER_NEAR_WALL)
            check(lbm_set_force_radius(C.h, 0, g.cx, g.cy, g.radius), "lbm_set_force_radius");
        C.obstacle = sys.obstacle;
        C.geom = g;
        build_force_links(C, sys);
        if (g_use_les) {
            sys.wall_dist.resize(sys.obstacle.size());
            check(lbm_wall_distance(C.h, 0, sys.wall_dist.data()), "lbm_wall_distance");
        }
        C.have_geometry = true;
    }
    check(lbm_set_params(C.h, 0, tau, u_inflow), "lbm_set_params");
    check(lbm_upload(C.h, 0, sys.f.data()), "lbm_upload");
    if (sys.use_thermal) {
        check(lbm_set_thermal(C.h, 0, sys.omega_k, sys.beta, sys.g_buoyancy, sys.T_ref), "lbm_set_thermal");
        C.thermal = true;
        check(lbm_upload_thermal(C.h, 0, sys.g_thermal.data()), "lbm_upload_thermal");
        check(lbm_set_thermal_rules(C.h, 0, nullptr, nullptr, 0), "lbm_set_thermal_rules");
    }
}

// Device-side conditions on g after every step for run_steps (lbm_set_thermal_rules):
// code[y*NX + x], rules as in include/lbm_gpu.h.  execute_time_step leaves them to the
// driver's own host loop.
inline void set_thermal_conditions(std::vector<uint8_t> code, std::vector<lbm_thermal_rule> rules)
{
    context().thermal_code = std::move(code);
    context().thermal_rules = std::move(rules);
}

// Per-node momentum exchange from the post-boundary state on the host, over the
// precomputed links only (src/lbm.hpp:542-581).
inline void node_forces(LBMCapabilities& sys)
{
    sys.reset_forces();
    for (const auto& [n, i] : context().links) {
        const double fb = sys.f[static_cast<size_t>(n) * 9 + bounce_back[i]];
        sys.fx_body[n] += cx[i] * 2.0 * fb;
        sys.fy_body[n] += cy[i] * 2.0 * fb;
    }
}

// Device-resident stepping: n steps with sys.f uploaded once and brought back at
// the end; per-step total forces (FP64, one value per step) go to fx/fy if given.
inline void run_steps(LBMCapabilities& sys, int n, double tau, double u_inflow, std::vector<double>* fx = nullptr,
                      std::vector<double>* fy = nullptr)
{
    prepare(sys, tau, u_inflow);
    Context& C = context();
    if (sys.use_thermal && !C.thermal_code.empty())
        check(lbm_set_thermal_rules(C.h, 0, C.thermal_code.data(), C.thermal_rules.data(),
                                    static_cast<int32_t>(C.thermal_rules.size())),
              "lbm_set_thermal_rules");
    check(lbm_set_launch(C.h, 0, 0, 0, -1, config().graph_steps), "lbm_set_launch");
    check(lbm_step(C.h, n), "lbm_step");
    check(lbm_download(C.h, 0, sys.f.data()), "lbm_download");
    if (sys.use_thermal) check(lbm_download_thermal(C.h, 0, sys.g_thermal.data()), "lbm_download_thermal");
    std::vector<double> bx(static_cast<size_t>(n)), by(static_cast<size_t>(n));
    int32_t got = 0;
    check(lbm_forces(C.h, 0, bx.data(), by.data(), n, &got), "lbm_forces");
    if (fx) fx->assign(bx.begin(), bx.begin() + got);
    if (fy) fy->assign(by.begin(), by.begin() + got);
    if (g_case != CaseType::CAVITY) node_forces(sys);
}

}  // namespace lbm_gpu

// One complete time step (src/lbm.hpp:293-582) on the GPU.
inline void execute_time_step(LBMCapabilities& sys, double tau, double u_inflow)
{
    lbm_gpu::prepare(sys, tau, u_inflow);
    lbm_gpu::Context& C = lbm_gpu::context();
    lbm_gpu::check(lbm_step(C.h, 1), "lbm_step");
    lbm_gpu::check(lbm_download(C.h, 0, sys.f.data()), "lbm_download");
    if (sys.use_thermal) lbm_gpu::check(lbm_download_thermal(C.h, 0, sys.g_thermal.data()), "lbm_download_thermal");
    if (g_case != CaseType::CAVITY) lbm_gpu::node_forces(sys);
}

// ------------------------------------------------------------------ geometry placement

// Cavity walls: x = 0, x = NX-1 and y = 0 solid (src/lbm.hpp:79-87).
inline void place_walls(LBMCapabilities& sys)
{
    for (int x = 0; x < NX; ++x) sys.obstacle[node_index(x, 0)] = 1;
    for (int y = 0; y < NY; ++y) {
        sys.obstacle[node_index(0, y)] = 1;
        sys.obstacle[node_index(NX - 1, y)] = 1;
    }
}

// Solid where sqrt(dx^2 + dy^2) < radius; the circle becomes the interpolated
// bounce-back geometry (src/lbm.hpp:931-944).
inline void place_cylinder(LBMCapabilities& sys, int cx_cyl, int cy_cyl, int radius)
{
    sys.bb_geom.cx = cx_cyl;
    sys.bb_geom.cy = cy_cyl;
    sys.bb_geom.radius = radius;
    for (int y = 0; y < NY; ++y)
        for (int x = 0; x < NX; ++x) {
            const double dx = x - cx_cyl, dy = y - cy_cyl;
            if (std::sqrt(dx * dx + dy * dy) < radius) sys.obstacle[node_index(x, y)] = 1;
        }
}

// Solid inside the polygon; the polygon becomes the interpolated bounce-back
// geometry (src/lbm.hpp:951-964).
inline void place_polygon(LBMCapabilities& sys, const std::vector<std::pair<double, double>>& poly)
{
    sys.bb_geom.poly_vertices = poly;
    sys.bb_geom.is_polygon = true;
    for (int y = 0; y < NY; ++y)
        for (int x = 0; x < NX; ++x)
            if (point_in_polygon(x, y, poly)) sys.obstacle[node_index(x, y)] = 1;
}

// Mean of the last window against the window before (src/lbm.hpp:93-104).
inline bool check_convergence(const std::vector<double>& values, int window = 1000, double threshold = 1e-4)
{
    const int n = static_cast<int>(values.size());
    if (n < 2 * window) return false;
    double recent = 0.0, earlier = 0.0;
    for (int i = n - window; i < n; ++i) recent += values[i];
    for (int i = n - 2 * window; i < n - window; ++i) earlier += values[i];
    return std::abs(recent / window - earlier / window) < threshold;
}

// ------------------------------------------------------------------ output writers
// File formats of the reference's writers (src/lbm.hpp:588-926), consumed by its
// post-processing scripts and web viewer.

namespace lbm_gpu {
// The frame writers' common body; with a thermal channel when t_wall is non-null
// (solid samples take *t_wall, fluid ones sum(g), NaN read as 1).
inline void write_frame(const LBMCapabilities& sys, int step, const std::string& output_dir, const double* t_wall)
{
    const int ds = std::max(1, NX / 100);
    const int mx = (NX + ds - 1) / ds, my = (NY + ds - 1) / ds;
    const int m = mx * my;
    std::vector<double> speed(m, 0.0), uu(m, 0.0), vv(m, 0.0), rr(m, 0.0), om(m, 0.0), tt(m, 0.0);
    std::vector<int> solid(m, 0);
    for (int j = 0; j < my; ++j)
        for (int i = 0; i < mx; ++i) {
            const int k = j * mx + i, n = node_index(i * ds, j * ds);
            if (sys.obstacle[n]) {
                solid[k] = 1;
                if (t_wall) tt[k] = *t_wall;
                continue;
            }
            if (t_wall) {
                double T;
                compute_temperature(&sys.g_thermal[static_cast<size_t>(n) * 9], T);
                tt[k] = std::isnan(T) ? 1.0 : T;
            }
            double rho, u, v;
            compute_macros(&sys.f[static_cast<size_t>(n) * 9], rho, u, v);
            if (rho < 1e-12 || std::isnan(rho)) {
                rho = 1.0;
                u = v = 0.0;
            }
            if (std::isnan(u)) u = 0.0;
            if (std::isnan(v)) v = 0.0;
            double s = std::sqrt(u * u + v * v);
            if (std::isnan(s)) s = 0.0;
            speed[k] = s;
            uu[k] = u;
            vv[k] = v;
            rr[k] = rho;
        }
    for (int j = 0; j < my; ++j)
        for (int i = 0; i < mx; ++i) {
            const int k = j * mx + i;
            if (solid[k]) continue;
            const int il = std::max(0, i - 1), ir = std::min(mx - 1, i + 1);
            const int jd = std::max(0, j - 1), ju = std::min(my - 1, j + 1);
            const double dvdx = (vv[j * mx + ir] - vv[j * mx + il]) / (2.0 * (ir - il) * ds);
            const double dudy = (uu[ju * mx + i] - uu[jd * mx + i]) / (2.0 * (ju - jd) * ds);
            om[k] = dvdx - dudy;
        }
    const std::string dir = output_dir + "/frames";
    std::filesystem::create_directories(dir);
    std::ofstream out(dir + "/frame_" + std::to_string(step) + ".json");
    out.precision(6);
    out << std::fixed;
    auto field = [&](const char* name, auto&& value, bool first = false) {
        out << (first ? "" : "],") << "\"" << name << "\":[";
        for (int k = 0; k < m; ++k) out << (k ? "," : "") << value(k);
    };
    out << "{\"nx\":" << mx << ",\"ny\":" << my << ",";
    field("velocity", [&](int k) { return speed[k]; }, true);
    field("u", [&](int k) { return uu[k]; });
    field("v", [&](int k) { return vv[k]; });
    field("rho", [&](int k) { return rr[k]; });
    field("p", [&](int k) { return rr[k] / 3.0; });
    field("omega", [&](int k) { return om[k]; });
    if (t_wall) field("temperature", [&](int k) { return tt[k]; });
    field("obstacle", [&](int k) { return solid[k]; });
    out << "]}";
}
}  // namespace lbm_gpu

// output_dir/frames/frame_<step>.json: fields sampled every max(1, NX/100) nodes:
// |u|, u, v, rho, p = rho/3, vorticity by central differences, obstacle flag
// (src/lbm.hpp:588-700).
inline void save_json_frame(const LBMCapabilities& sys, int step, const std::string& output_dir)
{
    lbm_gpu::write_frame(sys, step, output_dir, nullptr);
}

// The same frame with a "temperature" channel before "obstacle": sum(g) on fluid
// samples, T_wall on solid ones (src/lbm.hpp:704-816).
inline void save_json_frame_thermal(const LBMCapabilities& sys, int step, const std::string& output_dir,
                                    double T_wall)
{
    lbm_gpu::write_frame(sys, step, output_dir, &T_wall);
}

// output_dir/forces.jsonl, one {"step","cd","cl"} line per call; truncated when the
// directory changes (src/lbm.hpp:822-834).
inline void save_forces_jsonl(const std::string& output_dir, int step, double cd, double cl)
{
    static std::string open_dir;
    static std::ofstream out;
    if (open_dir != output_dir) {
        if (out.is_open()) out.close();
        out.open(output_dir + "/forces.jsonl", std::ios::trunc);
        out.precision(6);
        out << std::fixed;
        open_dir = output_dir;
    }
    out << "{\"step\":" << step << ",\"cd\":" << cd << ",\"cl\":" << cl << "}\n";
}

// output_dir/meta.json with the run parameters (src/lbm.hpp:839-855).
inline void save_meta_json(const std::string& output_dir, double re, double tau, double u_inflow,
                           double length_scale, const std::string& shape_type, int nx, int ny)
{
    std::ofstream out(output_dir + "/meta.json");
    out.precision(6);
    out << std::fixed << "{\n"
        << "  \"nx\": " << nx << ",\n"
        << "  \"ny\": " << ny << ",\n"
        << "  \"re\": " << re << ",\n"
        << "  \"tau\": " << tau << ",\n"
        << "  \"u_inflow\": " << u_inflow << ",\n"
        << "  \"length_scale\": " << length_scale << ",\n"
        << "  \"shape_type\": \"" << shape_type << "\"\n"
        << "}\n";
}

// Legacy ASCII VTK structured points: |u|, velocity, density, x-force per node
// (src/lbm.hpp:860-926).
inline void save_vtk_frame(const LBMCapabilities& sys, int frame, const std::string& output_dir = "output")
{
    std::ofstream out(output_dir + "/frame_" + std::to_string(frame) + ".vtk");
    out << "# vtk DataFile Version 3.0\nLBM Fluid Grid\nASCII\nDATASET STRUCTURED_POINTS\n"
        << "DIMENSIONS " << NX << " " << NY << " 1\n"
        << "ORIGIN 0 0 0\nSPACING 1 1 1\n"
        << "POINT_DATA " << NX * NY << "\n";
    auto node = [&](int n, double& rho, double& u, double& v) {
        compute_macros(&sys.f[static_cast<size_t>(n) * 9], rho, u, v);
    };
    out << "SCALARS VelocityMagnitude double 1\nLOOKUP_TABLE default\n";
    for (int n = 0; n < NX * NY; ++n) {
        if (sys.obstacle[n]) {
            out << "0.0\n";
            continue;
        }
        double rho, u, v;
        node(n, rho, u, v);
        out << std::sqrt(u * u + v * v) << "\n";
    }
    out << "VECTORS Velocity double\n";
    for (int n = 0; n < NX * NY; ++n) {
        if (sys.obstacle[n]) {
            out << "0 0 0\n";
            continue;
        }
        double rho, u, v;
        node(n, rho, u, v);
        out << u << " " << v << " 0\n";
    }
    out << "SCALARS Density double 1\nLOOKUP_TABLE default\n";
    for (int n = 0; n < NX * NY; ++n) {
        if (sys.obstacle[n]) {
            out << "0.0\n";
            continue;
        }
        double rho, u, v;
        node(n, rho, u, v);
        out << rho << "\n";
    }
    out << "SCALARS DragForce double 1\nLOOKUP_TABLE default\n";
    for (int n = 0; n < NX * NY; ++n) {
        if (sys.fx_body[n] != 0.0 || sys.fy_body[n] != 0.0) out << sys.fx_body[n] << "\n";
        else out << "0.0\n";
    }
}
