// lbm_internal.h -- structures shared by the host runtime and the gfx950 kernels.
#pragma once

#include <cstddef>
#include <cstdint>

namespace lbm {

// Force accumulators are kept in this many copies (chosen by workgroup) so that the
// FP64 atomics of different workgroups rarely meet on one address; the copies are
// summed in fixed order when a step's totals are moved to the history.
constexpr int kAccCopies = 16;

// Per-cell flag byte (pitch layout, one per node).
enum : uint8_t {
    CELL_SOLID = 1u << 0,    // obstacle node
    CELL_SPECIAL = 1u << 1,  // needs the general path (edge, boundary link, solid)
    CELL_LINKS = 1u << 2,    // fluid node with >= 1 boundary link (link record exists)
    CELL_WRITE = 1u << 3,    // solid node whose stored value must be kept current
    CELL_COLUMN = 1u << 4,   // inlet / outlet node off the corners whose general path reduces
                             // to plain pulls (+ the inlet closure, or shifted one column)
};

// Population storage forms (match enum lbm_storage).
enum Storage : int { ST_F64 = 0, ST_F32 = 1, ST_F16S = 2, ST_BF16S = 3 };

// Device descriptor of one lattice (read by every block of that lattice).
struct alignas(16) LatticeDev {
    int32_t nx, ny, pitch, lattice;
    int64_t f_off;        // element offset of population plane 0 in the pool
    int64_t plane;        // elements per population plane (ny * pitch)
    int64_t cell_off;     // offset of this lattice in per-cell arrays
    int32_t link_off;     // first link record of this lattice
    int32_t slot_base;    // first force slot of this lattice
    int32_t n_slots;
    int32_t collision, les, x_edge, y_edge, inlet, inlet_y0, inlet_h, outlet, lid, forces;
    int32_t mei_static;  // every boundary link is Mei with a resting wall
    double s_shear;                    // MRT shear rate; s_bulk = 1.2, s_normal = 1 are fixed
                                       // (MRTParams::from_tau, src/lbm_types.hpp:72-76)
    double tau, inv_tau, u_in, cs, cs2, nu, inv_nu, fscale;
    double half_tau, qtau2;            // tau / 2, tau^2 / 4 (the LES relaxation time)
    // 16-bit storage shift: stored h_i = 2^k (f_i - w_i - shift[i]); shift_s = 2^k shift.
    // Decoding adds shift, encoding adds nshift = -shift.  Without a shift (shifted == 0)
    // both hold -0.0f, for which x + (-0) = x for every x: a kernel may then skip the
    // additions and still give the bits of one that performs them.
    float shift[9], shift_s[9];
    float nshift[9], nshift_s[9];
    int32_t shifted;
    float pad1;
};

constexpr uint32_t LINK_MOVING = 1u << 31;

// Boundary-link record for one fluid node with links (indexed by link_slot[cell]).
struct alignas(16) LinkRec {
    uint32_t mask;       // bit i: the neighbour along e_i is solid (link i)
    uint32_t scheme;     // 2 bits per direction: enum lbm_bb; LINK_MOVING: some wall velocity != 0
    int16_t slot[9];     // force slot per link, -1 = excluded
    int16_t pad;         // (slot and pad are read as 5 dwords)
    int32_t cell;        // the node, y * pitch + x in its lattice
    double q[9];
    double uwx[9];
    double uwy[9];
};

static_assert(offsetof(LinkRec, slot) == 8 && offsetof(LinkRec, cell) == 28, "LinkRec layout");

// Lanes of a wave that take boundary-link list nodes (their loads scatter, and a wave's
// memory requests issue one address at a time).
constexpr int kListLanes = 16;

// One block of work of an ensemble launch: lattice, first row, first column.  Two-step
// tiles may start at a negative column (right-aligned rows: the nodes left of x = 0 are
// not the tile's).
struct Tile {
    int32_t lattice, y0, x0, pad;
};


}  // namespace lbm
