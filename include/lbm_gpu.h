/*
 * lbm_gpu.h -- C ABI of the MI355X (gfx950) D2Q9 lattice-update library.
 *
 * The reference has no FFI for this path: its boundary is the header-only C++
 * call  execute_time_step(LBMCapabilities&, double tau, double u_inflow)
 * (src/lbm.hpp:293) operating on the process-global lattice of
 * src/lbm_types.hpp:15-32 and 187-224.  The entry points below replace that call
 * and the state it touches, with the reference's implicit inputs made explicit
 * per lattice so that many lattices of different sizes can live in one handle.
 *
 * Conventions
 *   - Every function returns LBM_OK (0) or a nonzero status; lbm_last_error()
 *     then holds a message (per host thread).  No exception crosses the ABI.
 *   - Host buffers belong to the caller; device memory belongs to the handle.
 *   - Populations cross the boundary in the reference's layout: AoS FP64,
 *     f[(y*nx + x)*9 + i], directions as src/lbm_types.hpp:41-42.
 *   - A handle is used from one host thread at a time; its kernels run in
 *     order on one HIP stream (the caller's, or the null stream).
 *   - The state visible after lbm_step is the reference's sys.f after
 *     execute_time_step returns: post-stream, post-boundary, pre-collision.
 */
#ifndef LBM_GPU_H
#define LBM_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBM_OK 0
#define LBM_ERR_INVALID 1 /* bad argument or call order */
#define LBM_ERR_DEVICE 2  /* HIP runtime error */
#define LBM_ERR_NOMEM 3   /* device or host allocation failed */

typedef struct lbm_ensemble* lbm_handle;

/* Population storage in HBM.  All forms store the deviation h_i = (f_i - w_i) * 2^k;
 * F64 and F32 use k = 0, the 16-bit forms use the handle's scale exponent. */
enum lbm_storage { LBM_STORE_F64 = 0, LBM_STORE_F32 = 1, LBM_STORE_F16S = 2, LBM_STORE_BF16S = 3 };

/* CollisionType, src/lbm_types.hpp:26. */
enum lbm_collision { LBM_MRT = 0, LBM_BGK = 1 };

/* Treatment of each lattice edge (src/lbm.hpp:408-484): populations leaving through
 * an OPEN edge are dropped (zero) and an inlet/outlet may overwrite them; a WALL
 * edge bounces them back into the same node; PERIODIC wraps.  FREE_SLIP (y edges of a
 * lattice with an open x edge; no counterpart in the reference, which wraps y): a
 * population leaving through the edge comes back specularly, its y velocity reversed,
 * into the node beside its emitter (the method of images: dropped beyond the open x
 * edge, bounced back into its emitter where that node is solid; the momentum exchange
 * counts boundary links inside the lattice only); such lattices step one step per launch
 * and carry no thermal populations. */
enum lbm_edge { LBM_EDGE_PERIODIC = 0, LBM_EDGE_OPEN = 1, LBM_EDGE_WALL = 2, LBM_EDGE_FREE_SLIP = 3 };

/* Velocity inlet at x = 0: Zou/He on every row (src/lbm.hpp:21-34), the parabolic
 * profile on rows y >= inlet_y0 of fluid nodes (src/lbm.hpp:40-58), or EQUILIBRIUM: every
 * fluid node of x = 0 becomes feq(1, u_in, 0) (no counterpart in the reference). */
enum lbm_inlet { LBM_INLET_NONE = 0, LBM_INLET_ZOU_HE = 1, LBM_INLET_PARABOLIC = 2, LBM_INLET_EQUILIBRIUM = 3 };

/* Boundary-link schemes (src/lbm.hpp:226-288). */
enum lbm_bb { LBM_BB_SIMPLE = 0, LBM_BB_MEI = 1, LBM_BB_BOUZIDI = 2 };

/* One lattice.  Replaces the globals NX, NY, g_case, g_collision, g_use_les, g_cs
 * (src/lbm_types.hpp:15-32) and LBMCapabilities::body_force_x (:207). */
typedef struct lbm_desc {
    int32_t nx, ny;
    int32_t collision;        /* enum lbm_collision */
    int32_t les;              /* 1: Smagorinsky + Van Driest (src/lbm.hpp:155-182) */
    double cs;                /* Smagorinsky constant (g_cs) */
    int32_t x_edge, y_edge;   /* enum lbm_edge */
    int32_t inlet;            /* enum lbm_inlet */
    int32_t inlet_y0;         /* parabolic inlet: first row */
    int32_t inlet_h;          /* parabolic inlet: profile height */
    int32_t outlet;           /* 1: zero-gradient copy x=nx-2 -> nx-1 (src/lbm.hpp:63-73) */
    int32_t lid;              /* 1: row ny-1 := feq(rho, u_in, 0) (src/lbm.hpp:110-120) */
    int32_t forces;           /* 1: momentum exchange every step (src/lbm.hpp:542-581) */
    double body_force_x;      /* LBMCapabilities::body_force_x */
    double tau;               /* relaxation time */
    double u_in;              /* inlet / lid velocity (u_inflow) */
} lbm_desc;

/* One boundary link: fluid node `cell` (= y*nx + x) whose neighbour cell + e_dir is
 * solid.  Replaces the per-step q evaluation of BounceBackGeometry
 * (src/lbm_types.hpp:116-176) by a precomputed entry. */
typedef struct lbm_link {
    int32_t cell;
    int32_t dir;              /* 1..8 */
    int32_t scheme;           /* enum lbm_bb */
    int32_t slot;             /* force slot of this lattice, -1 = not counted */
    double q;                 /* wall distance fraction along the link, (0,1] */
    double uwx, uwy;          /* wall velocity at the solid node */
} lbm_link;

typedef struct lbm_info {
    int64_t cells;            /* sum of nx*ny over the lattices */
    int64_t fluid_cells;
    int64_t link_cells;       /* fluid nodes with at least one boundary link */
    int64_t les_cells;        /* cells of lattices with LES on */
    int32_t bytes_per_pop;    /* 8, 4 or 2 */
    int32_t n_lattices;
    int64_t device_bytes;     /* device memory held by the handle */
    int32_t block_x, rows;    /* launch shape chosen for the step kernel */
    int32_t nodes_per_thread;
    int32_t launch_options;
    int32_t bulk_tiles;       /* two steps per launch split by tile class: tiles taken   */
    int32_t other_tiles;      /* by the bulk kernel and by the general one, the edge     */
    int32_t edge_tiles;       /* tiles (inlet, outlet, periodic edges) taken by its edge */
    int32_t empty_tiles;      /* form, and tiles of unread solid nodes left out (0: not  */
} lbm_info;                   /* split)                                                  */

/* Message of the last failing call on this host thread. */
const char* lbm_last_error(void);

/* Fills *out for one of the reference's CaseType values (src/lbm_types.hpp:20-23),
 * reproducing the branch structure of execute_time_step: stream edges
 * (src/lbm.hpp:408-484), boundary conditions (:530-539) and forces (:543). */
int lbm_desc_for_case(int case_type, int32_t nx, int32_t ny, lbm_desc* out);

/* Creates a handle holding n lattices with one storage form.  stream may be NULL
 * (null stream) or a hipStream_t owned by the caller.  scale_exp is the k of the
 * 16-bit forms (ignored for F64/F32). */
int lbm_create(const lbm_desc* descs, int32_t n, int32_t storage, int32_t scale_exp,
               void* stream, lbm_handle* out);
int lbm_destroy(lbm_handle h);
/* Layout of the population buffers lbm_create would allocate for these lattices,
 * computed without a device (memory planning): lattice l's nine planes of ny rows of
 * pitch = nx rounded up to 64 values start at f_off[l] (n entries, may be null); each
 * buffer holds *pool_elems values followed by a guard of *guard_elems (more than one
 * row of the widest lattice: the step kernels form a node's pull addresses before they
 * know whether it lies on an edge).  Bytes per buffer: (pool + guard) * bytes per
 * population; a handle holds two, four with the thermal populations.  No counterpart in
 * the reference, whose state is two std::vector<double> of nx*ny*9. */
int lbm_population_layout(const lbm_desc* descs, int32_t n, int64_t* f_off, int64_t* pool_elems,
                          int64_t* guard_elems);

/* Geometry.  Set it before lbm_upload: a change afterwards invalidates the stored
 * state until the next upload.  lbm_set_obstacle takes nx*ny bytes (nonzero = solid), the
 * LBMCapabilities::obstacle array; it resets every link to simple bounce-back
 * counted in slot 0, the reference's behaviour without bb_geom. */
int lbm_set_obstacle(lbm_handle h, int32_t lattice, const uint8_t* mask);
/* The drivers' obstacle placement on the device: place_cylinder (src/lbm.hpp:931-944,
 * solid where sqrt(dx^2+dy^2) < r) for each of the n_circles (cx, cy, r) triples, then
 * place_polygon (src/lbm.hpp:951-964, even-odd rule of src/geometry.hpp:120-134) for each
 * polygon: poly_n[k] vertices, consecutive (x, y) pairs in poly_xy.  Cells only become
 * solid; each obstacle costs its bounding box, not the lattice.  The mask changes as with
 * lbm_set_obstacle (bounce-back geometry, moving wall, force radius and explicit links
 * are cleared: set them afterwards).  The reference's place_* also set bb_geom: here that
 * is lbm_set_bb_circle / lbm_set_bb_polygon. */
int lbm_place_obstacles(lbm_handle h, int32_t lattice, const double* circles, int32_t n_circles,
                        const double* poly_xy, const int32_t* poly_n, int32_t n_polys);
/* The lattice's obstacle mask, nx*ny bytes (1 = solid): LBMCapabilities::obstacle
 * (src/lbm_types.hpp:190), which the drivers read for their output frames. */
int lbm_get_obstacle(lbm_handle h, int32_t lattice, uint8_t* mask);
/* The reference's single interpolated-BB geometry (BounceBackGeometry with a circle
 * or polygon, src/lbm_types.hpp:90-176): q of every link from that one shape, 1.0
 * where the link misses it; scheme LBM_BB_MEI (the reference default,
 * src/lbm_types.hpp:181) or LBM_BB_BOUZIDI. */
int lbm_set_bb_circle(lbm_handle h, int32_t lattice, double cx, double cy, double r, int32_t scheme);
int lbm_set_bb_polygon(lbm_handle h, int32_t lattice, const double* xy, int32_t n_vertices,
                       int32_t scheme);
/* Rotating wall (BounceBackGeometry::compute_wall_velocity, src/lbm_types.hpp:106-114). */
int lbm_set_moving_wall(lbm_handle h, int32_t lattice, double omega, double rot_cx, double rot_cy);
/* Counts only links whose solid node lies within r of (cx,cy) (src/lbm.hpp:569-573). */
int lbm_set_force_radius(lbm_handle h, int32_t lattice, double cx, double cy, double r);
/* Explicit per-link entries (per-obstacle q, scheme, wall velocity, slot).  Every
 * link must exist in the obstacle mask. */
int lbm_set_links(lbm_handle h, int32_t lattice, const lbm_link* links, int32_t n);
/* Number of force slots of a lattice (default 1). */
int lbm_set_force_slots(lbm_handle h, int32_t lattice, int32_t n_slots);

/* Relaxation time and inlet/lid velocity for the following steps (the collision of
 * the state already stored is redone with the new tau). */
int lbm_set_params(lbm_handle h, int32_t lattice, double tau, double u_in);

/* State transfer in the reference layout (AoS FP64, nx*ny*9 doubles). */
int lbm_upload(lbm_handle h, int32_t lattice, const double* f_aos);
int lbm_download(lbm_handle h, int32_t lattice, double* f_aos);
/* Initial state feq(rho, u, v) on every node, built on the device: the drivers'
 * initialisation loop (src/main.cpp:60-66, src/cavity.cpp:78-84) without a host array,
 * for lattices too large to stage in AoS FP64.  Resets the step count like an upload. */
int lbm_init_equilibrium(lbm_handle h, int32_t lattice, double rho, double u, double v);
/* 16-bit storage forms only: store f_i - feq_i(1, u, v) instead of f_i - w_i (the
 * default, u = v = 0).  With (u, v) the free-stream velocity the stored deviations of
 * the undisturbed flow are near zero, so binary16 / bfloat16 rounding stays far below
 * the flow's own variations there.  No counterpart in the reference, whose state is
 * FP64.  Invalidates the stored state: upload or initialise afterwards. */
int lbm_set_storage_shift(lbm_handle h, int32_t lattice, double u, double v);
/* 16-bit storage forms only: round each stored population stochastically (add uniform
 * random bits below the last kept bit, then truncate) instead of to nearest even.  The
 * bits come from a hash of the seed, the lattice's step and the node, so runs repeat
 * exactly; rounding errors have zero mean, which removes the drift that identical
 * nodes rounding identically cause.  No counterpart in the reference (FP64 state). */
int lbm_set_stochastic_rounding(lbm_handle h, int32_t on, uint32_t seed);
/* The cylinder drivers' shedding trigger (src/main.cpp:67-83) on the device: every fluid
 * node of columns [x0, x1) becomes feq(rho, u, v + dv[(x - x0)*ny + y]) of its own moments
 * (solid nodes' entries are ignored; the drivers draw dv from std::mt19937(42) with
 * uniform_real_distribution(-1e-4, 1e-4), x outer, y inner, fluid nodes only).  Only
 * right after lbm_upload / lbm_init_equilibrium, before stepping: with the device-built
 * initial state, the driver's setup without a 9*N FP64 host array. */
int lbm_perturb_velocity(lbm_handle h, int32_t lattice, int32_t x0, int32_t x1, const double* dv);
/* Density and velocity of rows [y0, y1) of the current state, row-major, nx*(y1-y0)
 * FP64 each (any output may be null): compute_macros (src/lbm_types.hpp:278-288) on
 * what lbm_download would return, computed on the device -- the fields the frame
 * writers take (src/lbm.hpp:588-700) at a third of the transfer. */
int lbm_macros(lbm_handle h, int32_t lattice, int32_t y0, int32_t y1, double* rho, double* u, double* v);

/* Advances every lattice by n steps: one launch per step (or per two steps) covering
 * all lattices, replayed from a hipGraph of graph_steps steps when n allows.  n = 0
 * commits pending changes and prepares those graphs without stepping. */
int lbm_step(lbm_handle h, int32_t n);
/* Debugging aid for race and coverage checks (no counterpart in the reference): fills
 * lattice l's planes of the population buffer the next step launch writes (and of the
 * thermal populations') with all-one bytes, a NaN in every storage form.  A launch that
 * writes every slot it owns leaves the results unchanged; a slot it should write but
 * does not is then read back as NaN by the following launch.  The state lbm_download
 * would return before the next step is lost. */
int lbm_debug_poison(lbm_handle h, int32_t lattice);
/* Advances one lattice by n steps with one launch per step and no graph. */
int lbm_step_lattice(lbm_handle h, int32_t lattice, int32_t n);
/* Allocation-free stepping.  No counterpart in the reference: for callers that cannot
 * allocate memory while they step (code under hipGraph stream capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Advances every lattice by n steps as lbm_step
 * does -- same kernels, same results bit for bit -- with one launch per step and no
 * hipGraph, the launches' force partials in a fixed scratch array of
 * LBM_NOALLOC_SCRATCH_BYTES in device memory that each GPU holds for the life of the
 * process.  It calls no hipMalloc or hipFree, instantiates no graph and allocates no host
 * memory.  LBM_ERR_INVALID when the handle has changes not yet on the device (call
 * lbm_step(h, 0) once after changing the geometry, parameters or launch shape, or after
 * an upload), when one launch's partials exceed the scratch (16 B per boundary-link node
 * plus 256 B per force slot: about one million link nodes), or when the per-step force
 * history (4096 steps per lattice) would overwrite steps not yet read with lbm_forces.
 * Every call on a GPU shares that GPU's scratch and leaves it holding its last partials:
 * calls that overlap on one GPU (from two host threads, or on two streams) race. */
#define LBM_NOALLOC_SCRATCH_BYTES (16u << 20)
int lbm_step_noalloc(lbm_handle h, int32_t n);
/* Launch options of the step kernel (bit field). */
#define LBM_LAUNCH_XCD_ORDER 1     /* consecutive tiles stay on one XCD (one L2) */
#define LBM_LAUNCH_STREAM_STORES 2 /* nontemporal stores of the new populations */
/* Two steps per launch: tiles of 64 x rows nodes (shape (64, rows, 1), rows 4, 8 or 16)
 * with a one-node ring, the intermediate state in LDS.  Every lattice of the handle
 * (an ensemble too; not thermal handles, not free-slip y edges, and on a one-lattice
 * handle not an outlet with nx % 64 == 1); a batch of an odd number of steps ends with a
 * one-step launch, and after an even one the calls that read the state recompute the
 * populations of the step before the last first (one one-step launch, a temporary copy of
 * the population buffer).  The results are those of one step per launch, bit for bit. */
#define LBM_LAUNCH_TWO_STEPS 4

/* Launch shape of the step kernel (threads per block, rows per block, consecutive
 * nodes per thread), option bits and hipGraph length; 0 (-1 for options) keeps the
 * current value.  Shapes are a compiled set; a group of nodes moves at most 16
 * bytes per population plane. */
int lbm_set_launch(lbm_handle h, int32_t block_x, int32_t rows, int32_t nodes_per_thread,
                   int32_t options, int32_t graph_steps);
/* Times every compiled launch shape with every option set on the current state and
 * keeps the fastest.  Each candidate steps two device copies of the population buffers
 * in turn (as stepping does), so the call needs twice the population memory free for
 * its duration.  The lattice state and force history are unchanged. */
int lbm_autotune(lbm_handle h, int32_t* best_block_x, int32_t* best_rows, int32_t* best_nodes,
                 int32_t* best_options, float* best_ms);

/* Total force per step, per slot, since the previous call (FP64), in step order.
 * fx/fy receive up to max_steps*n_slots values (step-major); *n_steps the count. */
int lbm_forces(lbm_handle h, int32_t lattice, double* fx, double* fy, int32_t max_steps,
               int32_t* n_steps);
/* FP64 mass: sum of all populations over non-solid nodes of the current state. */
int lbm_mass(lbm_handle h, int32_t lattice, double* mass);
/* Nonzero when a non-finite density appeared since the last call. */
int lbm_nan_flag(lbm_handle h, int32_t lattice, int32_t* flag);
/* Steps taken by a lattice since its last upload. */
int lbm_steps_done(lbm_handle h, int32_t lattice, int64_t* steps);
int lbm_get_info(lbm_handle h, lbm_info* out);
/* Waits for all work of the handle. */
int lbm_sync(lbm_handle h);

/* ---- Thermal double distribution (src/lbm.hpp:356-397, 489-528; src/thermal.hpp:16-60).
 * Nine more populations g per node carry the temperature T = sum g_i; they replace
 * LBMCapabilities::use_thermal, g_thermal, T_ref, omega_k, beta and g_buoyancy
 * (src/lbm_types.hpp:193-201).  In each step, after the momentum collision, g relaxes at
 * omega_k to the thermal equilibrium w_i T (1 + 3 e.u + 4.5 (e.u)^2 - 1.5 u^2) with
 * the velocity of the post-collision f, and, when beta > 0 and g_buoyancy != 0, f gains
 * w_i (3 F.e_i + 9 (e_i.u)(e_i.F)) with F = (0, -beta (T - t_ref) g_buoyancy) and T
 * taken before the thermal collision; g then streams periodic in y (and in x on a
 * periodic x edge) and bounces back into its own node's opposite slot off the other x
 * edges and off solid nodes, which receive no g.  One-lattice handles, F64 or F32. */
/* Turns the thermal populations on (g = 1.0 in every slot, as the reference's
 * constructor) or changes their parameters.  Turning them on needs a lattice that has
 * not stepped since its last upload. */
int lbm_set_thermal(lbm_handle h, int32_t lattice, double omega_k, double beta, double g_buoyancy,
                    double t_ref);
/* Per-node conditions on the thermal state after every step: what the thermal drivers
 * do on the host between steps (src/heated_cylinder.cpp:155-164,
 * src/thermal_cavity.cpp:153-188).  code[y*nx + x]: low 4 bits k > 0 set g to the
 * thermal equilibrium at rules[k-1].t with the node's own velocity (LBM_THERMAL_EQ_FLOW)
 * or at rest (LBM_THERMAL_EQ_REST); bit LBM_THERMAL_MIRROR then sets g[i] = g[opp(i)]
 * in slot order, in place.  code null: none.  The conditions do not touch an uploaded
 * state; they act from the first step on. */
#define LBM_THERMAL_MIRROR 16
enum lbm_thermal_kind { LBM_THERMAL_EQ_FLOW = 0, LBM_THERMAL_EQ_REST = 1 };
typedef struct lbm_thermal_rule {
    double t;
    int32_t kind;
    int32_t pad;
} lbm_thermal_rule;
int lbm_set_thermal_rules(lbm_handle h, int32_t lattice, const uint8_t* code, const lbm_thermal_rule* rules,
                          int32_t n_rules);
/* Thermal state in the reference layout (AoS FP64 like f).  lbm_upload and
 * lbm_init_equilibrium reset g to 1.0; lbm_upload_thermal follows them, before any step. */
int lbm_upload_thermal(lbm_handle h, int32_t lattice, const double* g_aos);
int lbm_download_thermal(lbm_handle h, int32_t lattice, double* g_aos);
/* Temperature sum_i g_i (compute_temperature, src/thermal.hpp:55-60) of rows [y0, y1),
 * nx*(y1-y0) FP64, computed on the device. */
int lbm_temperature(lbm_handle h, int32_t lattice, int32_t y0, int32_t y1, double* t);

/* Wall distance used by the LES (src/lbm_types.hpp:231-264), computed once per
 * geometry: out[nx*ny] in lattice units, 1e9 where no solid is reachable. */
int lbm_wall_distance(lbm_handle h, int32_t lattice, double* out);

#ifdef __cplusplus
}
#endif
#endif /* LBM_GPU_H */
