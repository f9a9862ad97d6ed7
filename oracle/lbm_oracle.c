/*
 * lbm_oracle.c -- FP64 CPU restatement of the lbm-2d lattice update (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker.  Never part of the product path.
 *
 * Reference citations are file:line in ajeet-krish/lbm-2d.  Floating-point
 * operations are written in the reference's evaluation order so that, compiled
 * with -ffp-contract=off, results are bit-identical to the reference built the
 * same way (tests/test_oracle_vs_ref.py checks this where oracle/_ref is built).
 *
 * Pinning of each part (tests/test_oracle_*.py):
 *   equilibrium / macros / constants   reference unit tests src/lbm_test.cpp:12-111
 *                                      restated, plus moment closed forms
 *   mrt_rates                          clamp table (closed form)
 *   mrt_collide                        conservation of rho,j; equilibrium fixed point;
 *                                      all-rates-equal == BGK; M*Minv == I brute force;
 *                                      bitwise vs oracle/_ref
 *   LES (Smagorinsky + Van Driest)     closed form of tau_eff for a hand-built state;
 *                                      bitwise vs oracle/_ref
 *   q_circle / q_polygon               exact ray/circle and ray/segment intersections
 *                                      evaluated independently (numpy, long double)
 *   wall distance                      sweep == textbook BFS == brute-force L1 geodesic
 *   streaming / bounce-back / BCs      resting-fluid invariance, periodic mass
 *                                      conservation, Zou/He algebra, outflow copy,
 *                                      brute-force push on tiny grids; bitwise vs _ref
 *   Mei / Bouzidi / Ladd               q=1/2 and q=1 special cases; bitwise vs _ref
 *   momentum exchange                  closed form on a uniform state; bitwise vs _ref
 *   mt19937 + uniform_real             10000th output of mt19937 seeded 5489 equals
 *                                      4123659995 (C++ standard [rand.predef]);
 *                                      golden fixture of the seed-42 perturbation
 *   thermal DDF (collision, Boussinesq thermal equilibrium moments (sum = T, first
 *   forcing, adiabatic streaming,      moment = T u); collision keeps T; closed domain
 *   drivers' thermal wall loops)       keeps total heat; forcing adds exactly F per node
 *                                      to the momentum; bitwise vs oracle/_ref
 */
#include "lbm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* D2Q9 velocities, weights, opposites: src/lbm_types.hpp:41-55.
 *   6 2 5
 *   3 0 1
 *   7 4 8                                                                  */
static const int CX[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
static const int CY[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
static const double W[9] = {4.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0, 1.0 / 9.0,
                            1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
static const int OPP[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
/* same x velocity, opposite y velocity: specular reflection off a y edge */
static const int MIRY[9] = {0, 1, 4, 3, 2, 8, 7, 6, 5};

struct orc_lattice {
    int nx, ny;
    int case_type;
    int collision;
    int use_les;
    double cs;
    double body_force_x;
    int faithful;
    int wall_dist_valid;
    /* the configs' wording where the reference differs (cylinder-type cases only):
     * equilibrium inlet feq(1, u_in, 0) instead of Zou/He, free-slip y edges instead of
     * periodic y.  No counterpart in the reference; pinned by invariants. */
    int spec_inlet_eq, spec_free_slip;
    /* single interpolated-BB geometry (BounceBackGeometry, src/lbm_types.hpp:90-182) */
    double g_cx, g_cy, g_radius;
    double* poly;
    int n_poly;
    int is_polygon;
    int has_moving_wall;
    double omega, rot_cx, rot_cy;
    int use_mei_bb;
    /* state (LBMCapabilities, src/lbm_types.hpp:187-224) */
    double* f;
    double* f_next;
    uint8_t* obstacle;
    double* wall_dist;
    double* fx_body;
    double* fy_body;
    /* thermal double distribution (src/lbm_types.hpp:193-201) */
    int use_thermal;
    double* g;
    double* g_next;
    double t_ref, omega_k, beta, g_buoyancy;
};

int orc_cx(int i) { return CX[i]; }
int orc_cy(int i) { return CY[i]; }
double orc_weight(int i) { return W[i]; }
int orc_opposite(int i) { return OPP[i]; }

/* compute_equilibrium, src/lbm_types.hpp:269-273. */
double orc_equilibrium(int i, double rho, double u, double v)
{
    double eu = (double)CX[i] * u + (double)CY[i] * v;
    double uu = u * u + v * v;
    return W[i] * rho * (1.0 + 3.0 * eu + 4.5 * eu * eu - 1.5 * uu);
}

/* compute_macros, src/lbm_types.hpp:278-288. */
void orc_macros(const double* f9, double* rho, double* u, double* v)
{
    double r = 0.0, mx = 0.0, my = 0.0;
    for (int i = 0; i < 9; ++i) {
        r += f9[i];
        mx += f9[i] * (double)CX[i];
        my += f9[i] * (double)CY[i];
    }
    *rho = r;
    *u = mx / r;
    *v = my / r;
}

static double clamp_rate(double s) { return s < 0.5 ? 0.5 : (s > 1.99 ? 1.99 : s); }

/* MRTParams::from_tau, src/lbm_types.hpp:72-76. */
void orc_mrt_rates(double tau, double out[3])
{
    out[0] = clamp_rate(1.0 / tau);
    out[1] = 1.2;
    out[2] = 1.0;
}

/* Inverse of the d'Humieres D2Q9 moment matrix, moments ordered
 * (rho, e, eps, jx, qx, jy, qy, pxx, pxy).  Row sums are taken left to right over
 * the nonzero entries, which is the reference's order at src/lbm.hpp:193-217. */
static const double MINV[9][9] = {
    {1.0 / 9, -1.0 / 9, 1.0 / 9, 0, 0, 0, 0, 0, 0},
    {1.0 / 9, -1.0 / 36, -1.0 / 18, 1.0 / 6, -1.0 / 6, 0, 0, 1.0 / 4, 0},
    {1.0 / 9, -1.0 / 36, -1.0 / 18, 0, 0, 1.0 / 6, -1.0 / 6, -1.0 / 4, 0},
    {1.0 / 9, -1.0 / 36, -1.0 / 18, -1.0 / 6, 1.0 / 6, 0, 0, 1.0 / 4, 0},
    {1.0 / 9, -1.0 / 36, -1.0 / 18, 0, 0, -1.0 / 6, 1.0 / 6, -1.0 / 4, 0},
    {1.0 / 9, 1.0 / 18, 1.0 / 36, 1.0 / 6, 1.0 / 12, 1.0 / 6, 1.0 / 12, 0, 1.0 / 4},
    {1.0 / 9, 1.0 / 18, 1.0 / 36, -1.0 / 6, -1.0 / 12, 1.0 / 6, 1.0 / 12, 0, -1.0 / 4},
    {1.0 / 9, 1.0 / 18, 1.0 / 36, -1.0 / 6, -1.0 / 12, -1.0 / 6, -1.0 / 12, 0, 1.0 / 4},
    {1.0 / 9, 1.0 / 18, 1.0 / 36, 1.0 / 6, 1.0 / 12, -1.0 / 6, -1.0 / 12, 0, -1.0 / 4},
};

/* tau_eff root of the Smagorinsky closure, src/lbm.hpp:161-165 and 177-181. */
static double les_rate(double cs_sq, double pxx_neq, double pxy_neq, double rho, double tau)
{
    double Q = sqrt(2.0 * (pxx_neq * pxx_neq + pxy_neq * pxy_neq));
    double A = 9.0 * cs_sq * Q / (2.0 * rho);
    double tau_eff = (tau + sqrt(tau * tau + 4.0 * A)) / 2.0;
    return clamp_rate(1.0 / tau_eff);
}

/* mrt_collide, src/lbm.hpp:128-218 (moments 136-141, equilibria 148-153,
 * Smagorinsky 155-166, Van Driest 171-182, relaxation 185-190, inverse 193-217). */
void orc_mrt_collide(double* f, double rho, double u, double v, const double rates[3],
                     double cs_sq, double tau_base, double y_plus)
{
    double e = -4 * f[0] - f[1] - f[2] - f[3] - f[4] + 2 * (f[5] + f[6] + f[7] + f[8]);
    double eps = 4 * f[0] - 2 * (f[1] + f[2] + f[3] + f[4]) + (f[5] + f[6] + f[7] + f[8]);
    double qx = -2 * f[1] + 2 * f[3] + f[5] - f[6] - f[7] + f[8];
    double qy = -2 * f[2] + 2 * f[4] + f[5] + f[6] - f[7] - f[8];
    double pxx = f[1] - f[2] + f[3] - f[4];
    double pxy = f[5] - f[6] + f[7] - f[8];
    double jx = rho * u, jy = rho * v;
    double usq = u * u + v * v;

    double e_eq = -2 * rho + 3 * rho * usq;
    double eps_eq = rho - 3 * rho * usq;
    double qx_eq = -jx, qy_eq = -jy;
    double pxx_eq = rho * (u * u - v * v);
    double pxy_eq = rho * u * v;

    double s_shear = rates[0];
    if (cs_sq > 0.0 && tau_base > 0.0)
        s_shear = les_rate(cs_sq, pxx - pxx_eq, pxy - pxy_eq, rho, tau_base);
    if (cs_sq > 0.0 && tau_base > 0.0 && y_plus > 0.0) {
        double damp = 1.0 - exp(-y_plus / 26.0);
        double cs_d = sqrt(cs_sq) * damp;
        s_shear = les_rate(cs_d * cs_d, pxx - pxx_eq, pxy - pxy_eq, rho, tau_base);
    }

    e -= rates[1] * (e - e_eq);
    eps -= rates[1] * (eps - eps_eq);
    qx -= rates[2] * (qx - qx_eq);
    qy -= rates[2] * (qy - qy_eq);
    pxx -= s_shear * (pxx - pxx_eq);
    pxy -= s_shear * (pxy - pxy_eq);

    const double m[9] = {rho, e, eps, jx, qx, jy, qy, pxx, pxy};
    for (int i = 0; i < 9; ++i) {
        double acc = MINV[i][0] * m[0];
        for (int k = 1; k < 9; ++k)
            if (MINV[i][k] != 0.0) acc = acc + MINV[i][k] * m[k];
        f[i] = acc;
    }
}

/* BounceBackGeometry::q_cylinder, src/lbm_types.hpp:118-138: first root t in (0,1]
 * of |p + t e_i - c| = r along the link, 1.0 when the ray misses. */
double orc_q_circle(double ccx, double ccy, double radius, double xf, double yf, int i)
{
    double dx = xf - ccx, dy = yf - ccy;
    double ex = (double)CX[i], ey = (double)CY[i];
    double a = ex * ex + ey * ey;
    double b = 2.0 * (dx * ex + dy * ey);
    double c = dx * dx + dy * dy - radius * radius;
    double disc = b * b - 4.0 * a * c;
    if (disc < 0.0) return 1.0;
    double sd = sqrt(disc);
    double t1 = (-b - sd) / (2.0 * a);
    double t2 = (-b + sd) / (2.0 * a);
    if (t1 > 0.0 && t1 <= 1.0) return t1;
    if (t2 > 0.0 && t2 <= 1.0) return t2;
    return 1.0;
}

/* BounceBackGeometry::q_polygon, src/lbm_types.hpp:143-170: smallest t in (0,1]
 * over the polygon edges hit by the link segment; near-parallel edges skipped. */
double orc_q_polygon(const double* xy, int n, double xf, double yf, int i)
{
    double ex = (double)CX[i], ey = (double)CY[i];
    if (n < 3) return 1.0;
    double best = 1.0;
    for (int v = 0; v < n; ++v) {
        int w = (v + 1) % n;
        double x1 = xy[2 * v], y1 = xy[2 * v + 1];
        double sx = xy[2 * w] - x1, sy = xy[2 * w + 1] - y1;
        double den = ex * sy - ey * sx;
        if (fabs(den) < 1e-15) continue;
        double t = ((x1 - xf) * sy - (y1 - yf) * sx) / den;
        double s = ((x1 - xf) * ey - (y1 - yf) * ex) / den;
        if (t > 0.0 && t <= best && s >= 0.0 && s <= 1.0) best = t;
    }
    return best;
}

/* point_in_polygon, src/geometry.hpp:120-134 (even-odd crossing rule). */
int orc_point_in_polygon(double px, double py, const double* xy, int n)
{
    int inside = 0;
    for (int i = 0, j = n - 1; i < n; j = i++) {
        double xi = xy[2 * i], yi = xy[2 * i + 1];
        double xj = xy[2 * j], yj = xy[2 * j + 1];
        if (((yi > py) != (yj > py)) && (px < (xj - xi) * (py - yi) / (yj - yi) + xi))
            inside = !inside;
    }
    return inside;
}

/* ---------------------------------------------------------------- lattice object */

orc_lattice* orc_create(int nx, int ny)
{
    orc_lattice* L = (orc_lattice*)calloc(1, sizeof(orc_lattice));
    size_t n = (size_t)nx * (size_t)ny;
    L->nx = nx;
    L->ny = ny;
    L->case_type = ORC_CYLINDER;
    L->collision = ORC_MRT;
    L->cs = 0.12; /* g_cs default, src/lbm_types.hpp:32 */
    L->use_mei_bb = 1; /* src/lbm_types.hpp:181 */
    L->faithful = 1;
    L->f = (double*)calloc(n * 9, sizeof(double));
    L->f_next = (double*)calloc(n * 9, sizeof(double));
    L->obstacle = (uint8_t*)calloc(n, 1);
    L->wall_dist = (double*)calloc(n, sizeof(double));
    L->fx_body = (double*)calloc(n, sizeof(double));
    L->fy_body = (double*)calloc(n, sizeof(double));
    /* g starts at 1.0 in every slot (LBMCapabilities ctor, src/lbm_types.hpp:215-216) */
    L->g = (double*)malloc(n * 9 * sizeof(double));
    L->g_next = (double*)malloc(n * 9 * sizeof(double));
    for (size_t k = 0; k < n * 9; ++k) L->g[k] = L->g_next[k] = 1.0;
    L->t_ref = 1.0;
    L->omega_k = 1.0;
    return L;
}

void orc_destroy(orc_lattice* L)
{
    if (!L) return;
    free(L->f);
    free(L->f_next);
    free(L->obstacle);
    free(L->wall_dist);
    free(L->fx_body);
    free(L->fy_body);
    free(L->g);
    free(L->g_next);
    free(L->poly);
    free(L);
}

int orc_nx(const orc_lattice* L) { return L->nx; }
double* orc_g(orc_lattice* L) { return L->g; }

void orc_set_thermal(orc_lattice* L, int on, double omega_k, double beta, double g_buoyancy, double t_ref)
{
    L->use_thermal = on;
    L->omega_k = omega_k;
    L->beta = beta;
    L->g_buoyancy = g_buoyancy;
    L->t_ref = t_ref;
}

/* thermal_equilibrium, src/thermal.hpp:24-30 (T_0 = 1). */
double orc_thermal_equilibrium(int i, double T, double u, double v)
{
    double edotc = (double)CX[i] * u + (double)CY[i] * v;
    double u2 = u * u + v * v;
    return W[i] * T * (1.0 + 3.0 * edotc + 4.5 * edotc * edotc - 1.5 * u2);
}

/* The heated-cylinder driver's wall loop after every step (src/heated_cylinder.cpp:
 * 155-164, apply_thermal_isothermal src/thermal.hpp:70-75): every solid node takes the
 * thermal equilibrium at t_wall and rest. */
void orc_thermal_isothermal_solids(orc_lattice* L, double t_wall)
{
    size_t n = (size_t)L->nx * L->ny;
    for (size_t c = 0; c < n; ++c) {
        if (!L->obstacle[c]) continue;
        for (int i = 0; i < 9; ++i) L->g[c * 9 + i] = orc_thermal_equilibrium(i, t_wall, 0.0, 0.0);
    }
}

/* The thermal-cavity driver's loop after every step (src/thermal_cavity.cpp:153-188):
 * fluid columns x = 1 and x = nx-2 (rows 1..ny-2) take the thermal equilibrium at t_hot
 * / t_cold with the node's own velocity; then rows y = 1 and y = ny-2 (columns
 * 1..nx-2) copy g[opp(i)] into g[i] in place, in slot order. */
void orc_thermal_cavity_walls(orc_lattice* L, double t_hot, double t_cold)
{
    const int nx = L->nx, ny = L->ny;
    for (int y = 1; y < ny - 1; ++y) {
        const int cols[2] = {1, nx - 2};
        const double temps[2] = {t_hot, t_cold};
        for (int k = 0; k < 2; ++k) {
            size_t c = (size_t)y * nx + cols[k];
            double rho, u, v;
            orc_macros(&L->f[c * 9], &rho, &u, &v);
            for (int i = 0; i < 9; ++i) L->g[c * 9 + i] = orc_thermal_equilibrium(i, temps[k], u, v);
        }
    }
    for (int x = 1; x < nx - 1; ++x) {
        const int rows[2] = {1, ny - 2};
        for (int k = 0; k < 2; ++k) {
            double* gc = &L->g[((size_t)rows[k] * nx + x) * 9];
            for (int i = 0; i < 9; ++i) gc[i] = gc[OPP[i]];
        }
    }
}
int orc_ny(const orc_lattice* L) { return L->ny; }
double* orc_f(orc_lattice* L) { return L->f; }
uint8_t* orc_obstacle(orc_lattice* L) { return L->obstacle; }
double* orc_wall_dist(orc_lattice* L) { return L->wall_dist; }
double* orc_fx_body(orc_lattice* L) { return L->fx_body; }
double* orc_fy_body(orc_lattice* L) { return L->fy_body; }
void orc_set_case(orc_lattice* L, int c) { L->case_type = c; }
void orc_set_collision(orc_lattice* L, int c) { L->collision = c; }
void orc_set_les(orc_lattice* L, int on, double cs) { L->use_les = on; L->cs = cs; }
void orc_set_body_force(orc_lattice* L, double fx) { L->body_force_x = fx; }
void orc_set_mei(orc_lattice* L, int on) { L->use_mei_bb = on; }
void orc_set_faithful(orc_lattice* L, int on) { L->faithful = on; }
void orc_set_spec(orc_lattice* L, int inlet_eq, int free_slip)
{
    L->spec_inlet_eq = inlet_eq;
    L->spec_free_slip = free_slip;
}
void orc_obstacle_changed(orc_lattice* L) { L->wall_dist_valid = 0; }

void orc_set_moving_wall(orc_lattice* L, int on, double omega, double rcx, double rcy)
{
    L->has_moving_wall = on;
    L->omega = omega;
    L->rot_cx = rcx;
    L->rot_cy = rcy;
}

/* place_walls, src/lbm.hpp:79-87. */
void orc_place_walls(orc_lattice* L)
{
    for (int x = 0; x < L->nx; ++x) L->obstacle[x] = 1;
    for (int y = 0; y < L->ny; ++y) {
        L->obstacle[(size_t)y * L->nx] = 1;
        L->obstacle[(size_t)y * L->nx + L->nx - 1] = 1;
    }
    L->wall_dist_valid = 0;
}

/* place_cylinder, src/lbm.hpp:931-944: strict sqrt(dx^2+dy^2) < r. */
void orc_place_cylinder(orc_lattice* L, int cxc, int cyc, int r)
{
    L->g_cx = cxc;
    L->g_cy = cyc;
    L->g_radius = r;
    for (int y = 0; y < L->ny; ++y)
        for (int x = 0; x < L->nx; ++x) {
            double dx = (double)(x - cxc), dy = (double)(y - cyc);
            if (sqrt(dx * dx + dy * dy) < r) L->obstacle[(size_t)y * L->nx + x] = 1;
        }
    L->wall_dist_valid = 0;
}

/* place_polygon, src/lbm.hpp:951-964. */
void orc_place_polygon(orc_lattice* L, const double* xy, int n)
{
    free(L->poly);
    L->poly = (double*)malloc(sizeof(double) * 2 * (size_t)(n > 0 ? n : 1));
    memcpy(L->poly, xy, sizeof(double) * 2 * (size_t)n);
    L->n_poly = n;
    L->is_polygon = 1;
    for (int y = 0; y < L->ny; ++y)
        for (int x = 0; x < L->nx; ++x)
            if (orc_point_in_polygon((double)x, (double)y, xy, n))
                L->obstacle[(size_t)y * L->nx + x] = 1;
    L->wall_dist_valid = 0;
}

/* Driver initialisation: f = feq(rho,u,v) on every node, src/main.cpp:60-66. */
void orc_init_equilibrium(orc_lattice* L, double rho, double u, double v)
{
    size_t n = (size_t)L->nx * L->ny;
    for (size_t c = 0; c < n; ++c)
        for (int i = 0; i < 9; ++i) L->f[c * 9 + i] = orc_equilibrium(i, rho, u, v);
}

/* Seed-42 transverse-velocity perturbation, src/main.cpp:68-83: x outer, y inner,
 * solids skipped, one uniform(-1e-4,1e-4) draw per fluid node. */
void orc_perturb_seed42(orc_lattice* L, int x_begin, int x_end)
{
    orc_mt19937 g;
    orc_mt_seed(&g, 42u);
    for (int x = x_begin; x < x_end; ++x)
        for (int y = 0; y < L->ny; ++y) {
            size_t c = (size_t)y * L->nx + x;
            if (L->obstacle[c]) continue;
            double dv = orc_uniform_real(&g, -1e-4, 1e-4);
            double rho, u, v;
            orc_macros(&L->f[c * 9], &rho, &u, &v);
            for (int i = 0; i < 9; ++i) L->f[c * 9 + i] = orc_equilibrium(i, rho, u, v + dv);
        }
}

/* ---------------------------------------------------------------- wall distance */

/* compute_wall_distance, src/lbm_types.hpp:231-264: Gauss-Seidel relaxation of
 * d = min(d, d_nb + 1) over 4-neighbours (non-periodic) until nothing changes. */
void orc_wall_distance_sweep(const uint8_t* obs, int nx, int ny, double* d)
{
    size_t n = (size_t)nx * ny;
    for (size_t c = 0; c < n; ++c) d[c] = obs[c] ? 0.0 : 1e9;
    int changed = 1;
    while (changed) {
        changed = 0;
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                size_t c = (size_t)y * nx + x;
                if (obs[c]) continue;
                double best = d[c];
                const int nbx[4] = {x - 1, x + 1, x, x};
                const int nby[4] = {y, y, y - 1, y + 1};
                for (int k = 0; k < 4; ++k) {
                    if (nbx[k] < 0 || nbx[k] >= nx || nby[k] < 0 || nby[k] >= ny) continue;
                    double cand = d[(size_t)nby[k] * nx + nbx[k]] + 1.0;
                    if (cand < best) {
                        best = cand;
                        changed = 1;
                    }
                }
                d[c] = best;
            }
    }
}

/* Multi-source breadth-first search from all solid nodes; yields the same fixpoint
 * as the relaxation (4-neighbour graph distance, 1e9 where no solid is reachable). */
void orc_wall_distance_bfs(const uint8_t* obs, int nx, int ny, double* d)
{
    size_t n = (size_t)nx * ny;
    size_t* queue = (size_t*)malloc(sizeof(size_t) * (n ? n : 1));
    size_t head = 0, tail = 0;
    for (size_t c = 0; c < n; ++c) {
        if (obs[c]) {
            d[c] = 0.0;
            queue[tail++] = c;
        } else {
            d[c] = 1e9;
        }
    }
    while (head < tail) {
        size_t c = queue[head++];
        int x = (int)(c % nx), y = (int)(c / nx);
        const int nbx[4] = {x - 1, x + 1, x, x};
        const int nby[4] = {y, y, y - 1, y + 1};
        for (int k = 0; k < 4; ++k) {
            if (nbx[k] < 0 || nbx[k] >= nx || nby[k] < 0 || nby[k] >= ny) continue;
            size_t m = (size_t)nby[k] * nx + nbx[k];
            if (d[m] > d[c] + 1.0) {
                d[m] = d[c] + 1.0;
                queue[tail++] = m;
            }
        }
    }
    free(queue);
}

/* ---------------------------------------------------------------- boundary terms */

/* apply_bouzidi_bb, src/lbm.hpp:226-288: value bounced into f_next[node][opp(i)] for
 * the link from fluid (x,y) towards the solid node (x,y)+e_i. */
static double bounce_value(const orc_lattice* L, int x, int y, int i, const double* fn, size_t node)
{
    double q = L->is_polygon ? orc_q_polygon(L->poly, L->n_poly, (double)x, (double)y, i)
                             : orc_q_circle(L->g_cx, L->g_cy, L->g_radius, (double)x, (double)y, i);
    double fbb;
    if (L->use_mei_bb) {
        /* Mei / Filippova-Hanel form, src/lbm.hpp:234-256 */
        double rho, u, v;
        orc_macros(fn, &rho, &u, &v);
        double uwx = 0.0, uwy = 0.0;
        if (L->has_moving_wall) {
            /* compute_wall_velocity at the solid node, src/lbm_types.hpp:106-114 */
            double px = (double)(x + CX[i]), py = (double)(y + CY[i]);
            uwx = -L->omega * (py - L->rot_cy);
            uwy = L->omega * (px - L->rot_cx);
        }
        double feq_w = orc_equilibrium(i, rho, uwx, uwy);
        double eu = (double)CX[i] * uwx + (double)CY[i] * uwy;
        fbb = q * feq_w + (1.0 - q) * fn[i] + (2.0 * q - 1.0) * W[i] * rho * eu * 3.0;
    } else {
        /* Bouzidi (2001), src/lbm.hpp:258-273 */
        if (q < 0.5) {
            int sx = x - CX[i], sy = y - CY[i];
            if (sx >= 0 && sx < L->nx && sy >= 0 && sy < L->ny)
                fbb = 2.0 * q * fn[i] + (1.0 - 2.0 * q) * L->f[((size_t)sy * L->nx + sx) * 9 + i];
            else
                fbb = fn[i];
        } else {
            double h = 1.0 / (2.0 * q);
            fbb = h * fn[i] + (1.0 - h) * L->f[node * 9 + OPP[i]];
        }
        if (L->has_moving_wall) {
            /* Ladd (1994) term, src/lbm.hpp:276-285 */
            double px = (double)(x + CX[i]), py = (double)(y + CY[i]);
            double uwx = -L->omega * (py - L->rot_cy);
            double uwy = L->omega * (px - L->rot_cx);
            double rho, u, v;
            orc_macros(fn, &rho, &u, &v);
            double eu = (double)CX[i] * uwx + (double)CY[i] * uwy;
            fbb -= 2.0 * W[i] * rho * eu * 3.0;
        }
    }
    return fbb;
}

/* Zou/He velocity inlet without transverse correction, src/lbm.hpp:21-34 and the
 * parabolic variant src/lbm.hpp:40-58 share this closure. */
static void zou_he_west(double* fn, double u_in)
{
    double rho = (fn[0] + fn[2] + fn[4] + 2.0 * (fn[3] + fn[6] + fn[7])) / (1.0 - u_in);
    fn[1] = fn[3] + (2.0 / 3.0) * rho * u_in;
    fn[5] = fn[7] + (1.0 / 6.0) * rho * u_in;
    fn[8] = fn[6] + (1.0 / 6.0) * rho * u_in;
}

static int is_periodic_x_case(int c) { return c == ORC_RIBS || c == ORC_PERIODIC_HILLS; }
static int has_geometry(const orc_lattice* L) { return L->g_radius > 0.0 || L->is_polygon; }

/* ---------------------------------------------------------------- time step */

/* execute_time_step, src/lbm.hpp:293-582. */
void orc_step(orc_lattice* L, double tau, double u_in)
{
    const int nx = L->nx, ny = L->ny;
    const size_t n = (size_t)nx * ny;
    double* f = L->f;
    double* fn = L->f_next;
    const uint8_t* obs = L->obstacle;

    double Fscale = (L->body_force_x != 0.0) ? (1.0 - 1.0 / (2.0 * tau)) * 3.0 * L->body_force_x : 0.0;

    /* collision, src/lbm.hpp:300-354 */
    if (L->collision == ORC_MRT) {
        double rates[3];
        orc_mrt_rates(tau, rates);
        double cs_sq = L->use_les ? L->cs * L->cs : 0.0;
        if (cs_sq > 0.0) {
            if (L->faithful) {
                orc_wall_distance_sweep(obs, nx, ny, L->wall_dist); /* every step, :304 */
            } else if (!L->wall_dist_valid) {
                orc_wall_distance_bfs(obs, nx, ny, L->wall_dist);
                L->wall_dist_valid = 1;
            }
        }
        double nu = (tau - 0.5) / 3.0;
#pragma omp parallel for collapse(2) schedule(static)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                size_t c = (size_t)y * nx + x;
                if (obs[c]) continue;
                double* fc = &f[c * 9];
                double rho, u, v;
                orc_macros(fc, &rho, &u, &v);
                double y_plus = 0.0;
                if (cs_sq > 0.0 && nu > 0.0) {
                    double speed = sqrt(u * u + v * v);
                    double u_tau = sqrt(speed * speed + 1e-12);
                    y_plus = L->wall_dist[c] * u_tau / nu;
                }
                orc_mrt_collide(fc, rho, u, v, rates, cs_sq, tau, y_plus);
                if (Fscale != 0.0)
                    for (int i = 0; i < 9; ++i) fc[i] += W[i] * CX[i] * Fscale;
            }
    } else {
#pragma omp parallel for collapse(2) schedule(static)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                size_t c = (size_t)y * nx + x;
                if (obs[c]) continue;
                double* fc = &f[c * 9];
                double rho, u, v;
                orc_macros(fc, &rho, &u, &v);
                double inv_tau = 1.0 / tau;
                for (int i = 0; i < 9; ++i) {
                    double feq = orc_equilibrium(i, rho, u, v);
                    fc[i] -= inv_tau * (fc[i] - feq);
                    if (Fscale != 0.0) fc[i] += W[i] * CX[i] * Fscale;
                }
            }
    }

    /* thermal collision and Boussinesq forcing, src/lbm.hpp:356-397: macros of the
     * post-collision f, T of the pre-collision g (compute_temperature, src/thermal.hpp:
     * 55-60), thermal_collide (src/thermal.hpp:35-41), boussinesq_force (:45-52) with
     * g = (0, g_buoyancy), added to f as w_i (3 F.e_i + 9 (e_i.u)(e_i.F)) */
    if (L->use_thermal) {
        const double rho_0 = 1.0;
#pragma omp parallel for collapse(2) schedule(static)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                size_t c = (size_t)y * nx + x;
                if (obs[c]) continue;
                double* fc = &f[c * 9];
                double* gc = &L->g[c * 9];
                double rho, u, v;
                orc_macros(fc, &rho, &u, &v);
                double T = 0.0;
                for (int i = 0; i < 9; ++i) T += gc[i];
                for (int i = 0; i < 9; ++i) {
                    double g_eq = orc_thermal_equilibrium(i, T, u, v);
                    gc[i] -= L->omega_k * (gc[i] - g_eq);
                }
                if (L->beta > 0.0 && L->g_buoyancy != 0.0) {
                    double dT = T - L->t_ref;
                    double fxb = -rho_0 * L->beta * dT * 0.0;
                    double fyb = -rho_0 * L->beta * dT * L->g_buoyancy;
                    for (int i = 0; i < 9; ++i) {
                        double edotu = (double)CX[i] * u + (double)CY[i] * v;
                        double Fdot = (double)CX[i] * fxb + (double)CY[i] * fyb;
                        fc[i] += W[i] * (3.0 * Fdot + 9.0 * edotu * ((double)CX[i] * fxb + (double)CY[i] * fyb));
                    }
                }
            }
    }

    /* zero f_next (and g_next), src/lbm.hpp:400-403 */
    memset(fn, 0, sizeof(double) * n * 9);
    if (L->use_thermal) memset(L->g_next, 0, sizeof(double) * n * 9);

    /* streaming + bounce-back, src/lbm.hpp:406-484 */
    const int interp = has_geometry(L);
    const int cse = L->case_type;
#pragma omp parallel for collapse(2) schedule(static)
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            size_t c = (size_t)y * nx + x;
            if (obs[c]) continue;
            const double* fc = &f[c * 9];
            for (int i = 0; i < 9; ++i) {
                int tx = x + CX[i], ty = y + CY[i];
                if (cse == ORC_CAVITY) {
                    /* src/lbm.hpp:408-433 */
                    if (tx < 0 || tx >= nx || ty < 0 || ty >= ny) {
                        fn[c * 9 + OPP[i]] = fc[i];
                        continue;
                    }
                } else if (is_periodic_x_case(cse)) {
                    /* src/lbm.hpp:434-456 */
                    if (tx < 0) tx += nx;
                    if (tx >= nx) tx -= nx;
                    if (ty < 0) ty += ny;
                    if (ty >= ny) ty -= ny;
                    size_t t = (size_t)ty * nx + tx;
                    if (obs[t]) fn[c * 9 + OPP[i]] = fc[i];
                    else fn[t * 9 + i] = fc[i];
                    continue;
                } else {
                    if (L->spec_free_slip && (ty < 0 || ty >= ny)) {
                        /* free slip: the population returns into the node beside its
                         * emitter along x, y velocity reversed (the emitter's mirror
                         * image); dropped beyond the open x edge; bounced back into the
                         * emitter when that node is solid (it met the solid's image) */
                        if (tx < 0 || tx >= nx) continue;
                        size_t t = (size_t)y * nx + tx;
                        if (obs[t]) fn[c * 9 + OPP[i]] = fc[i];
                        else fn[t * 9 + MIRY[i]] = fc[i];
                        continue;
                    }
                    /* src/lbm.hpp:457-484: y periodic, x-outgoing dropped */
                    if (ty < 0) ty += ny;
                    if (ty >= ny) ty -= ny;
                    if (tx < 0 || tx >= nx) continue;
                }
                size_t t = (size_t)ty * nx + tx;
                if (obs[t]) {
                    if (interp) fn[c * 9 + OPP[i]] = bounce_value(L, x, y, i, fc, c);
                    else fn[c * 9 + OPP[i]] = fc[i];
                } else {
                    fn[t * 9 + i] = fc[i];
                }
            }
        }

    /* swap, src/lbm.hpp:487 */
    L->f = fn;
    L->f_next = f;
    f = L->f;

    /* thermal streaming, src/lbm.hpp:489-528: y periodic (x too for ribs / hills); a
     * population leaving the lattice or hitting a solid node bounces back into the
     * opposite slot of its own node (adiabatic) */
    if (L->use_thermal) {
        const double* g = L->g;
        double* gn = L->g_next;
#pragma omp parallel for collapse(2) schedule(static)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                size_t c = (size_t)y * nx + x;
                if (obs[c]) continue;
                for (int i = 0; i < 9; ++i) {
                    int tx = x + CX[i], ty = y + CY[i];
                    if (is_periodic_x_case(cse)) {
                        if (tx < 0) tx += nx;
                        if (tx >= nx) tx -= nx;
                    }
                    if (ty < 0) ty += ny;
                    if (ty >= ny) ty -= ny;
                    if (tx < 0 || tx >= nx || obs[(size_t)ty * nx + tx]) gn[c * 9 + OPP[i]] = g[c * 9 + i];
                    else gn[((size_t)ty * nx + tx) * 9 + i] = g[c * 9 + i];
                }
            }
        L->g = gn;
        L->g_next = (double*)g;
    }

    /* open boundaries, src/lbm.hpp:530-539 */
    if (cse == ORC_CAVITY) {
        /* enforce_lid, src/lbm.hpp:110-120 */
        for (int x = 0; x < nx; ++x) {
            double* fc = &f[((size_t)(ny - 1) * nx + x) * 9];
            double rho, u, v;
            orc_macros(fc, &rho, &u, &v);
            for (int i = 0; i < 9; ++i) fc[i] = orc_equilibrium(i, rho, u_in, 0.0);
        }
    } else if (!is_periodic_x_case(cse)) {
        if (cse == ORC_STEP) {
            /* enforce_step_inflow(NY/3, NY-1-NY/3), src/lbm.hpp:40-58 and :534 */
            int h_step = ny / 3, h_inlet = ny - 1 - ny / 3;
            for (int y = h_step; y < ny; ++y) {
                size_t c = (size_t)y * nx;
                if (obs[c]) continue;
                double yy = (double)(y - h_step);
                double h = (double)h_inlet;
                double u_loc = u_in * 4.0 * yy * (h - yy) / (h * h);
                zou_he_west(&f[c * 9], u_loc);
            }
        } else if (L->spec_inlet_eq) {
            /* equilibrium inlet: every fluid node of x = 0 becomes feq(1, u_in, 0) */
            for (int y = 0; y < ny; ++y) {
                size_t c = (size_t)y * nx;
                if (obs[c]) continue;
                for (int i = 0; i < 9; ++i) f[c * 9 + i] = orc_equilibrium(i, 1.0, u_in, 0.0);
            }
        } else {
            for (int y = 0; y < ny; ++y) zou_he_west(&f[(size_t)y * nx * 9], u_in);
        }
        /* enforce_outflow, src/lbm.hpp:63-73 */
        for (int y = 0; y < ny; ++y) {
            double* dst = &f[((size_t)y * nx + nx - 1) * 9];
            const double* src = &f[((size_t)y * nx + nx - 2) * 9];
            for (int i = 0; i < 9; ++i) dst[i] = src[i];
        }
    }

    /* momentum exchange, src/lbm.hpp:542-581 (serial, node order) */
    if (cse != ORC_CAVITY) {
        memset(L->fx_body, 0, sizeof(double) * n);
        memset(L->fy_body, 0, sizeof(double) * n);
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                size_t c = (size_t)y * nx + x;
                if (obs[c]) continue;
                for (int i = 0; i < 9; ++i) {
                    int tx = x + CX[i], ty = y + CY[i];
                    if (L->spec_free_slip && (ty < 0 || ty >= ny)) continue;
                    if (ty < 0) ty += ny;
                    if (ty >= ny) ty -= ny;
                    if (is_periodic_x_case(cse)) {
                        if (tx < 0) tx += nx;
                        if (tx >= nx) tx -= nx;
                    } else if (tx < 0 || tx >= nx) {
                        continue;
                    }
                    size_t t = (size_t)ty * nx + tx;
                    if (!obs[t]) continue;
                    if (cse == ORC_CYLINDER_NEAR_WALL) {
                        double dx = (double)tx - L->g_cx, dy = (double)ty - L->g_cy;
                        if (sqrt(dx * dx + dy * dy) > L->g_radius) continue;
                    }
                    double fb = f[c * 9 + OPP[i]];
                    L->fx_body[c] += CX[i] * 2.0 * fb;
                    L->fy_body[c] += CY[i] * 2.0 * fb;
                }
            }
    }
}

void orc_steps(orc_lattice* L, int n, double tau, double u_in)
{
    for (int s = 0; s < n; ++s) orc_step(L, tau, u_in);
}

void orc_total_force(const orc_lattice* L, double* fx, double* fy)
{
    size_t n = (size_t)L->nx * L->ny;
    double sx = 0.0, sy = 0.0;
    for (size_t c = 0; c < n; ++c) {
        sx += L->fx_body[c];
        sy += L->fy_body[c];
    }
    *fx = sx;
    *fy = sy;
}

double orc_mass(const orc_lattice* L)
{
    size_t n = (size_t)L->nx * L->ny;
    double m = 0.0;
    for (size_t c = 0; c < n; ++c) {
        if (L->obstacle[c]) continue;
        for (int i = 0; i < 9; ++i) m += L->f[c * 9 + i];
    }
    return m;
}

/* ---------------------------------------------------------------- random numbers */

/* MT19937 (Matsumoto & Nishimura 1998) with the C++ standard's seeding
 * ([rand.eng.mers]: x_i = 1812433253 * (x_{i-1} ^ (x_{i-1} >> 30)) + i). */
void orc_mt_seed(orc_mt19937* g, uint32_t seed)
{
    g->mt[0] = seed;
    for (int i = 1; i < 624; ++i)
        g->mt[i] = 1812433253u * (g->mt[i - 1] ^ (g->mt[i - 1] >> 30)) + (uint32_t)i;
    g->idx = 624;
}

uint32_t orc_mt_next(orc_mt19937* g)
{
    if (g->idx >= 624) {
        for (int k = 0; k < 624; ++k) {
            uint32_t y = (g->mt[k] & 0x80000000u) | (g->mt[(k + 1) % 624] & 0x7fffffffu);
            uint32_t nv = g->mt[(k + 397) % 624] ^ (y >> 1);
            if (y & 1u) nv ^= 0x9908b0dfu;
            g->mt[k] = nv;
        }
        g->idx = 0;
    }
    uint32_t y = g->mt[g->idx++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

/* generate_canonical<double,53> over a 32-bit engine takes two draws,
 * (g1 + g2 * 2^32) / 2^64, clamped below 1; the distribution then maps
 * x -> x * (b - a) + a  (libstdc++ as shipped with g++ 11). */
double orc_uniform_real(orc_mt19937* g, double a, double b)
{
    double lo = (double)orc_mt_next(g);
    double hi = (double)orc_mt_next(g);
    double sum = lo + hi * 4294967296.0;
    double r = sum / 18446744073709551616.0;
    if (r >= 1.0) r = nextafter(1.0, 0.0);
    return r * (b - a) + a;
}
