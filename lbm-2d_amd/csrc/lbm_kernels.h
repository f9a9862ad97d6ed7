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
};

struct StepArgs {
    const LatticeDev* lat;
    const Tile* tiles;          // ensemble mode: one entry per block
    int32_t n_tiles;
    int32_t single;             // single-lattice mode: lattice index (tiles == nullptr)
    int32_t tiles_x;            // single-lattice mode: tiles per row band
    int32_t tiles_y;
    void* buf0;
    void* buf1;
    void* dst_override;         // non-null: write here instead of the other buffer
    const uint8_t* flags;
    const int32_t* link_slot;
    const LinkRec* links;
    const uint16_t* wdist;
    const int64_t* steps;       // per-lattice step count at the start of the batch
    const int32_t* par;         // per-lattice index of the buffer holding the state
    double* acc;                // [2][kAccCopies][total_slots][2] force accumulators
    double* hist;               // [total_slots][R][2] force history ring
    int32_t R;
    int32_t total_slots;
    int32_t* nan_flags;         // per lattice
    int32_t step_arg;           // launch index inside the batch
    int32_t abs_mode;           // 1: k = k_abs and src = buf[src_par] (one lattice)
    int64_t k_abs;
    int32_t src_par;
    int32_t flush_begin, flush_count;
    int32_t options;            // LAUNCH_* bits
    float scale, inv_scale;     // 2^k and 2^-k of the 16-bit forms
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

hipError_t launch_step(int storage, int block_x, int rows, int vec, int n_blocks, const StepArgs& a,
                       hipStream_t s);
hipError_t launch_state(int storage, const StateArgs& a, int nx, hipStream_t s);
hipError_t launch_prologue(int storage, const PrologueArgs& a, int nx, hipStream_t s);
hipError_t launch_add_steps(int64_t* steps, int32_t* par, int32_t begin, int32_t count, int64_t inc,
                            hipStream_t s);

bool step_shape_supported(int storage, int block_x, int rows, int vec);
// Fills up to max compiled shapes {block_x, rows, nodes per thread} for a storage
// form; returns their count.
int step_shapes(int storage, int (*out)[3], int max);

}  // namespace lbm
