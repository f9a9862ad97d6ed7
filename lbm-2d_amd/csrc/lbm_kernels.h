// lbm_kernels.h -- launch interface between the host runtime and the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "lbm_internal.h"

namespace lbm {

// Arguments of the fused step kernel (passed by value; graph capture keeps a copy).
// Step-kernel launch options (bit field).
enum : int32_t {
    LAUNCH_XCD_ORDER = 1,       // consecutive tiles on one XCD (blocks go round-robin over 8)
    LAUNCH_STREAM_STORES = 2,   // nontemporal stores of the new populations
    LAUNCH_TWO_STEPS = 4,       // one-lattice launches: two steps per launch (tile 64 x rows)
};

struct StepArgs {
    // single-lattice launches: everything the kernel needs before its first population
    // load is here, by value
    LatticeDev L0;
    int32_t lattice0;
    int32_t tiles_x, tiles_y;   // tiles per row band, row bands
    int32_t options;            // LAUNCH_* bits
    const Tile* bulk_tiles;     // two-step bulk launches: one tile per block
    // general-path nodes, taken by the first special_blocks blocks: the lattice's
    // boundary-link records, its edge nodes, then its written solids (list of
    // in-lattice indices y*pitch+x)
    int32_t n_links;
    int32_t n_edges;
    const int32_t* special;
    int32_t n_special;
    int32_t special_blocks;
    const void* src;            // post-collision populations of the previous step
    void* dst;                  // new post-collision populations (ensemble mode: non-null
                                // overrides the per-lattice buffer, for the autotuner; with a
                                // non-null src both replace the per-lattice buffers -- the
                                // other tiles of a one-lattice handle's split launch)
    // ensemble launches
    const LatticeDev* lat;
    const Tile* tiles;          // one entry per block
    int32_t n_tiles;
    int32_t step_arg;           // step index inside the batch (stochastic-rounding keys)
    int32_t flips;              // buffer swaps before this launch in the batch (parity)
    const int32_t* par;         // per-lattice index of the buffer holding the state
    void* buf0;
    void* buf1;
    // both
    const uint8_t* flags;
    const int32_t* link_slot;
    const LinkRec* links;
    const uint16_t* wdist;
    double* acc;                // this launch's fallback accumulators [kAccCopies][total_slots][2]
    double* acc2;               // two-step launches: the second step's
    double* nf;                 // this launch's force terms per link record [records][2]
    double* nf2;                // two-step launches: the second step's
    const int64_t* steps;       // per-lattice step count at the start of the batch
    int32_t sr;                 // 16-bit forms: stochastic rounding of stored values
    uint32_t sr_seed;
    int32_t total_slots;
    int32_t* nan_flags;         // per lattice
    float scale, inv_scale;     // 2^k and 2^-k of the 16-bit forms
};

// Force totals of a batch of n launches: accumulator blocks -> history ring.
struct FlushArgs {
    double* acc;                // [n][kAccCopies][total_slots][2], cleared on the way
    const double* nf;           // [n][n_recs][2] force terms per link record
    int32_t n_recs;
    const int32_t* slot_start;  // records of slot s: slot_recs[slot_start[s] .. slot_start[s+1])
    const int32_t* slot_recs;
    double* hist;               // [total_slots][R][2]
    const int64_t* steps;       // per-lattice step count at the start of the batch
    const int32_t* slot_lattice;  // lattice of each force slot
    int32_t slot_begin, slot_count;
    int32_t n;
    int32_t total_slots;
    int32_t R;
};

// Materialise the reference-visible state of rows [y0,y1) of one lattice as AoS FP64
// and accumulate sum(f - w) over fluid nodes into *mass_dev.
struct StateArgs {
    const LatticeDev* lat;
    int32_t lattice;
    int32_t y0, y1;
    int32_t from_state;         // 1: src holds the state itself (no step taken yet)
    const void* src;
    const uint8_t* flags;
    const int32_t* link_slot;
    const LinkRec* links;
    double* out;                // AoS FP64, (y1-y0)*nx*9; null: not written
    double* macros;             // planes rho, u, v, each (y1-y0)*nx FP64; null: not written
    double* mass_dev;           // null: no mass sum
    void* coll_dst;             // non-null: write the state's collision here (re-collide)
    const uint16_t* wdist;
    float scale, inv_scale;
};

// Load rows [y0,y1) of AoS FP64 state: keep it in `state_dst`, its collision in `coll_dst`.
struct PrologueArgs {
    const LatticeDev* lat;
    int32_t lattice;
    int32_t y0, y1;
    const double* in;           // null: uniform equilibrium feq(eq_rho, eq_u, eq_v)
    double eq_rho, eq_u, eq_v;
    void* state_dst;
    void* coll_dst;
    const uint8_t* flags;
    const uint16_t* wdist;
    float scale, inv_scale;
};

// Thermal double distribution of one lattice (src/lbm.hpp:356-397, 489-528): what the
// thermal kernel does over rows [y0, y1) of StepArgs.L0.
enum : int32_t {
    THERMAL_STEP = 0,     // one step: f and g from src / gsrc (post-collision) to dst / gdst, forces
    THERMAL_COLLIDE = 1,  // the stored post-collision f and g of the current state (no forces)
    THERMAL_OUT_G = 2,    // the state's g as AoS FP64 rows into out
    THERMAL_OUT_T = 3,    // the state's temperature sum(g) into out
};
constexpr int kThermalRules = 15;
constexpr uint8_t THERMAL_MIRROR = 16;  // per-node code bit: g[i] = g[opp(i)] in slot order

struct ThermalArgs {
    const void* gsrc;           // thermal planes: post-collision of the previous step, or the state
    void* gdst;
    const uint8_t* code;        // per-node conditions after each step (pitch layout), or null
    double rule_t[kThermalRules];
    uint32_t rule_rest;         // bit k: rule k is the equilibrium at rest (else at the node's velocity)
    int32_t mode;               // THERMAL_*
    int32_t from_state;         // the sources hold the state itself (no step taken yet)
    int32_t y0, y1;
    double omega_k, beta, g_buoyancy, t_ref;
    double* out;
};

// Rows [y0, y1) of AoS FP64 thermal state (null: 1.0 in every slot) into the state planes.
struct ThermalLoadArgs {
    const LatticeDev* lat;
    int32_t lattice;
    int32_t y0, y1;
    const double* in;
    void* state_dst;
};

// Wall distance (uint16, 0xffff = none within 65534) of one lattice's nodes from its
// flag bytes; both pointers at the lattice's first cell.
hipError_t launch_wall_distance(const uint8_t* flags, uint16_t* wdist, int nx, int ny, int pitch, hipStream_t s);
// Transverse velocity perturbation of the stored state (columns [x0, x1) of one lattice):
// fluid node (x, y) becomes feq(rho, u, v + dv[(x - x0) * ny + y]) of its own moments.
struct PerturbArgs {
    const LatticeDev* lat;
    int32_t lattice;
    int32_t x0, x1;
    const double* dv;
    void* state;
    const uint8_t* flags;
    float scale, inv_scale;
};

hipError_t launch_perturb(int storage, const PerturbArgs& a, int ny, hipStream_t s);

// One obstacle rasterised into an nx-wide mask over its bounding box [x0,x1) x [y0,y1):
// a circle (kind 0: cx, cy, r) or a polygon (kind 1: n vertices xy[2n]).
struct PlaceArgs {
    uint8_t* mask;
    int32_t nx;
    int32_t x0, y0, x1, y1;
    int32_t kind;
    double cx, cy, r;
    const double* xy;
    int32_t n;
};

hipError_t launch_place(const PlaceArgs& a, hipStream_t s);

// Cell flags and boundary-link records of one lattice from its obstacle mask.  Pass 1
// (launch_geom_cells): flags (pitch layout, zeroed beforehand), each cell's link mask
// bits 1..8 shifted down by one, link cells per row (zeroed) and the fluid count (zeroed).
// Pass 2 (launch_geom_links, after the host's scan of the row counts into row_off): the
// link cells in row-major order -> rec_cell (y*pitch+x), rec_mask (bits 1..8 shifted down
// by one, as in lmask) and lslot (pitch layout, -1 beforehand).
struct GeomArgs {
    const uint8_t* obs;
    uint8_t* flags;
    uint8_t* lmask;
    int32_t* row_links;
    unsigned long long* fluid;
    const int32_t* row_off;
    int32_t* lslot;
    int32_t* rec_cell;
    uint8_t* rec_mask;
    int32_t nx, ny, pitch, x_edge, y_edge;
};

hipError_t launch_geom_cells(const GeomArgs& g, hipStream_t s);
// Inlet and outlet column marks (CELL_COLUMN) of rows 1..ny-2 of an open-x lattice:
// the inlet node of rows >= inlet_from, the outlet node when outlet is set, each only
// without SOLID or LINKS (at the outlet, on the copied node nx-2 as well).
hipError_t launch_geom_columns(uint8_t* flags, int nx, int ny, int pitch, int inlet_from, int outlet,
                               hipStream_t s);
// Lattices with Bouzidi links: every solid node is written (CELL_WRITE); row_count[y]
// (zeroed beforehand) counts the solid nodes off the edges, and the second pass lists
// them, y * pitch + x in row-major order from row_off[y].
hipError_t launch_geom_write(uint8_t* flags, int nx, int ny, int pitch, int32_t* row_count, hipStream_t s);
hipError_t launch_geom_solids(const uint8_t* flags, int nx, int ny, int pitch, const int32_t* row_off,
                              int32_t* out, hipStream_t s);
hipError_t launch_geom_links(const GeomArgs& g, hipStream_t s);
hipError_t launch_thermal(int storage, const StepArgs& a, const ThermalArgs& t, hipStream_t s);
hipError_t launch_thermal_load(int storage, const ThermalLoadArgs& a, int nx, hipStream_t s);

// single: the one-lattice kernel (StepArgs.L0, src, dst); else the ensemble kernel.
hipError_t launch_step(int storage, int block_x, int rows, int vec, bool single, int n_blocks, const StepArgs& a,
                       hipStream_t s);
hipError_t launch_flush(const FlushArgs& f, hipStream_t s);
// Two steps per launch: tiles of 64 x tile_rows (4, 8 or 16), 256 threads; single: one
// lattice (StepArgs.L0, src, dst, tiles_x), else the ensemble's two-step tile table.
hipError_t launch_step2(int storage, int tile_rows, bool single, int n_blocks, const StepArgs& a, hipStream_t s);
// Two steps per launch on the bulk tiles of one lattice (single: StepArgs.L0, src, dst) or of
// an ensemble's lattices (StepArgs.lat, parity): StepArgs.bulk_tiles, 64 x tile_rows tiles
// (8, 16 or 32 rows) whose node ring lies inside their lattice and holds no general-path
// node; FP32 and the 16-bit forms.  edge: StepArgs.bulk_tiles holds edge tiles instead
// (class 2 of launch_tile_class), whose lattices' planes take 32-bit byte offsets.
bool bulk_supported(int storage, int tile_rows);
hipError_t launch_step2_bulk(int storage, int tile_rows, bool single, bool edge, int n_blocks, const StepArgs& a,
                             hipStream_t s);
// One workgroup that does nothing (orders the kernels of a split launch's streams).
hipError_t launch_hold(hipStream_t s);
// Classes of the n_tiles tiles of 64 x ty (lattice descriptors on the device): out[t] = 1
// for a bulk tile, 2 for an edge tile (a whole tile whose ring holds no general-path node
// but inlet, outlet and periodic-edge nodes), 3 for a tile of solid nodes nothing reads, 0
// for the others.
hipError_t launch_tile_class(const Tile* tiles, int n_tiles, const LatticeDev* lat, const uint8_t* flags, int ty,
                             uint8_t* out, hipStream_t s);
// Blocks of block_x threads that take the general-path list nodes: n_links link
// records (kListLanes per wave), then n_edges edge nodes (every lane).
int list_blocks(int block_x, int n_links, int n_edges);
hipError_t launch_state(int storage, const StateArgs& a, int nx, hipStream_t s);
hipError_t launch_prologue(int storage, const PrologueArgs& a, int nx, hipStream_t s);
// p[0 .. n) = 0 on stream s.
hipError_t launch_fill_zero(double* p, int64_t n, hipStream_t s);
// steps += inc, buffer parity ^= flips & 1 for lattices [begin, begin+count)
hipError_t launch_add_steps(int64_t* steps, int32_t* par, int32_t begin, int32_t count, int64_t inc,
                            int32_t flips, hipStream_t s);

bool step_shape_supported(int storage, int block_x, int rows, int vec);
// Fills up to max compiled shapes {block_x, rows, nodes per thread} for a storage
// form; returns their count.
int step_shapes(int storage, int (*out)[3], int max);

}  // namespace lbm
