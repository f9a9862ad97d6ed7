/*
 * lbm_oracle.h -- FP64 CPU restatement of the lbm-2d per-timestep lattice update.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the GPU path is compared
 * against (tests/, __graft_entry__.smoke(), and the cpu_baseline leg of bench.py).
 * It is never linked into, loaded by, or used as a fallback of the product library.
 *
 * Layout follows the reference: AoS FP64 populations f[(y*nx + x)*9 + i], one
 * obstacle byte per node, per-node momentum-exchange force arrays.
 *
 * Every function cites the reference file:line it restates.  The pinning status
 * of each function (what known answer or independent check holds it) is listed
 * at the top of lbm_oracle.c.
 */
#ifndef LBM_ORACLE_H
#define LBM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* CaseType, same enumerator order as src/lbm_types.hpp:20-23. */
enum orc_case {
    ORC_CYLINDER = 0, ORC_CAVITY, ORC_STEP, ORC_RIBS, ORC_URBAN_CANYON, ORC_DOWNWASH,
    ORC_SQUARE_CYLINDER, ORC_FLAT_PLATE, ORC_SPORTS_BALL, ORC_PERIODIC_HILLS,
    ORC_CYLINDER_NEAR_WALL, ORC_SIDE_BY_SIDE, ORC_ROTATING_CYLINDER, ORC_ORIFICE_PLATE
};

/* CollisionType, src/lbm_types.hpp:26. */
enum orc_collision { ORC_MRT = 0, ORC_BGK = 1 };

typedef struct orc_lattice orc_lattice;

/* ---- lattice constants and per-node closed forms (src/lbm_types.hpp:41-55, 266-288) ---- */
int    orc_cx(int i);
int    orc_cy(int i);
double orc_weight(int i);
int    orc_opposite(int i);
double orc_equilibrium(int i, double rho, double u, double v);
void   orc_macros(const double* f9, double* rho, double* u, double* v);
/* MRTParams::from_tau: out[0]=s_shear, out[1]=s_bulk, out[2]=s_normal. */
void   orc_mrt_rates(double tau, double out[3]);
void   orc_mrt_collide(double* f9, double rho, double u, double v, const double rates[3],
                       double cs_sq, double tau_base, double y_plus);

/* ---- q geometry (src/lbm_types.hpp:118-170) ---- */
double orc_q_circle(double ccx, double ccy, double radius, double xf, double yf, int i);
double orc_q_polygon(const double* xy, int n_vertices, double xf, double yf, int i);
int    orc_point_in_polygon(double px, double py, const double* xy, int n_vertices);

/* ---- lattice object ---- */
orc_lattice* orc_create(int nx, int ny);
void   orc_destroy(orc_lattice* L);
int    orc_nx(const orc_lattice* L);
int    orc_ny(const orc_lattice* L);
double*  orc_f(orc_lattice* L);          /* current populations, AoS n*9 */
uint8_t* orc_obstacle(orc_lattice* L);   /* n bytes */
double*  orc_wall_dist(orc_lattice* L);  /* n doubles */
double*  orc_fx_body(orc_lattice* L);    /* n doubles */
double*  orc_fy_body(orc_lattice* L);    /* n doubles */

void orc_set_case(orc_lattice* L, int case_type);
void orc_set_collision(orc_lattice* L, int collision);
void orc_set_les(orc_lattice* L, int use_les, double cs);
void orc_set_body_force(orc_lattice* L, double fx);
void orc_set_mei(orc_lattice* L, int use_mei_bb);
void orc_set_moving_wall(orc_lattice* L, int on, double omega, double rot_cx, double rot_cy);
/* faithful=1: per-step wall-distance relaxation sweep and serial memset/forces, as in
 * the reference; faithful=0: wall distance by BFS once per geometry (same integers). */
void orc_set_faithful(orc_lattice* L, int faithful);
/* The configs' boundary wording (cylinder-type cases): equilibrium inlet feq(1, u_in, 0)
 * and free-slip (specular) y edges; no counterpart in the reference. */
void orc_set_spec(orc_lattice* L, int inlet_eq, int free_slip);

/* Obstacle placement (src/lbm.hpp:79-87, 931-964). */
void orc_place_walls(orc_lattice* L);
void orc_place_cylinder(orc_lattice* L, int cx_cyl, int cy_cyl, int radius);
void orc_place_polygon(orc_lattice* L, const double* xy, int n_vertices);
/* Mark the obstacle geometry changed (recomputes the hoisted wall distance). */
void orc_obstacle_changed(orc_lattice* L);

/* Initial conditions used by the case drivers (src/main.cpp:60-83, src/cavity.cpp:78-84). */
void orc_init_equilibrium(orc_lattice* L, double rho, double u, double v);
void orc_perturb_seed42(orc_lattice* L, int x_begin, int x_end);

/* One call of execute_time_step (src/lbm.hpp:293-582). */
void orc_step(orc_lattice* L, double tau, double u_inflow);
void orc_steps(orc_lattice* L, int n, double tau, double u_inflow);

/* Total force over all nodes in node order, as the drivers sum it (src/main.cpp:106-110). */
void   orc_total_force(const orc_lattice* L, double* fx, double* fy);
/* FP64 mass over non-solid nodes (src/lbm_test.cpp:173-180). */
double orc_mass(const orc_lattice* L);

/* ---- thermal double distribution (src/lbm.hpp:356-397, 489-528; src/thermal.hpp) ---- */
double* orc_g(orc_lattice* L);           /* thermal populations, AoS n*9 (start at 1.0) */
void   orc_set_thermal(orc_lattice* L, int on, double omega_k, double beta, double g_buoyancy, double t_ref);
double orc_thermal_equilibrium(int i, double T, double u, double v);
/* The thermal drivers' loops run after each step (src/heated_cylinder.cpp:155-164,
 * src/thermal_cavity.cpp:153-188). */
void   orc_thermal_isothermal_solids(orc_lattice* L, double t_wall);
void   orc_thermal_cavity_walls(orc_lattice* L, double t_hot, double t_cold);

/* Wall distance: reference relaxation sweeps (src/lbm_types.hpp:231-264) and BFS. */
void orc_wall_distance_sweep(const uint8_t* obstacle, int nx, int ny, double* dist);
void orc_wall_distance_bfs(const uint8_t* obstacle, int nx, int ny, double* dist);

/* std::mt19937 (32-bit Mersenne Twister, seed 42 path) and libstdc++'s
 * uniform_real_distribution<double>(a,b), as used at src/main.cpp:69-76. */
typedef struct { uint32_t mt[624]; int idx; } orc_mt19937;
void     orc_mt_seed(orc_mt19937* g, uint32_t seed);
uint32_t orc_mt_next(orc_mt19937* g);
double   orc_uniform_real(orc_mt19937* g, double a, double b);

#ifdef __cplusplus
}
#endif
#endif
