// lbm_host.hip -- host runtime behind the C ABI of include/lbm_gpu.h.
//
// Owns the device pools of an ensemble of lattices:
//   populations  two buffers (A-B), SoA planes of pitch-aligned rows per lattice
//   flags        one byte per node (solid / special / links / write)
//   link_slot    int32 per node: index of the node's boundary-link record
//   links        one LinkRec per fluid node with links (q, scheme, wall velocity, slot)
//   wdist        uint16 wall distance per node, computed on the device from the flags
//   forces       per launch of a batch: FP64 force terms per link record and fallback
//                accumulators per slot; a per-slot history ring filled by the flush
//   thermal      (one-lattice handles) two buffers of g planes with the populations'
//                layout and parity, one condition byte per node
// and builds them from the reference's obstacle/geometry description
// (src/lbm_types.hpp:90-182, src/lbm.hpp:79-87, 931-964): the per-node pass on the
// device, the per-link values here.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/lbm_gpu.h"
#include "lbm_internal.h"
#include "lbm_kernels.h"

namespace lbm {
namespace {

thread_local std::string g_error;

int fail(int code, const std::string& msg)
{
    g_error = msg;
    return code;
}

#define LBM_HIP(call)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail(LBM_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr int CX[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
constexpr int CY[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};

constexpr int kHistoryRows = 4096;
// Device time per candidate launch shape in lbm_autotune (ms).
constexpr float kTuneMs = 1.5f;
constexpr int kStripTiles = 16;  // bulk-tile strip width in 64-node tiles
constexpr int kMinBatch = 32;  // accumulator blocks when graphs are shorter or off
constexpr int64_t kStagingBytes = 256ll << 20;

// q along link i from fluid node (xf,yf) to a circle: first root t in (0,1] of the
// ray/circle intersection, 1.0 when the ray misses (src/lbm_types.hpp:118-138).
double q_circle(double ccx, double ccy, double r, double xf, double yf, int i)
{
    const double dx = xf - ccx, dy = yf - ccy;
    const double ex = CX[i], ey = CY[i];
    const double a = ex * ex + ey * ey;
    const double b = 2.0 * (dx * ex + dy * ey);
    const double c = dx * dx + dy * dy - r * r;
    const double disc = b * b - 4.0 * a * c;
    if (disc < 0.0) return 1.0;
    const double sd = std::sqrt(disc);
    const double t1 = (-b - sd) / (2.0 * a), t2 = (-b + sd) / (2.0 * a);
    if (t1 > 0.0 && t1 <= 1.0) return t1;
    if (t2 > 0.0 && t2 <= 1.0) return t2;
    return 1.0;
}

// q to a closed polygon: smallest t in (0,1] over the edges crossed by the link
// segment, near-parallel edges skipped (src/lbm_types.hpp:143-170).
double q_polygon(const std::vector<double>& xy, double xf, double yf, int i)
{
    const int n = static_cast<int>(xy.size() / 2);
    if (n < 3) return 1.0;
    const double ex = CX[i], ey = CY[i];
    double best = 1.0;
    for (int v = 0; v < n; ++v) {
        const int w = (v + 1) % n;
        const double x1 = xy[2 * v], y1 = xy[2 * v + 1];
        const double sx = xy[2 * w] - x1, sy = xy[2 * w + 1] - y1;
        const double den = ex * sy - ey * sx;
        if (std::fabs(den) < 1e-15) continue;
        const double t = ((x1 - xf) * sy - (y1 - yf) * sx) / den;
        const double s = ((x1 - xf) * ey - (y1 - yf) * ex) / den;
        if (t > 0.0 && t <= best && s >= 0.0 && s <= 1.0) best = t;
    }
    return best;
}

struct HostLattice {
    lbm_desc d{};
    int32_t pitch = 0;
    int64_t f_off = 0, cell_off = 0;
    int32_t link_off = 0, slot_base = 0, n_slots = 1;
    // geometry description
    std::vector<uint8_t> obstacle;
    int bb_kind = 0;  // 0 none, 1 circle, 2 polygon
    double bcx = 0, bcy = 0, br = 0;
    std::vector<double> poly;
    int bb_scheme = 1;
    bool moving = false;
    double omega = 0, rcx = 0, rcy = 0;
    bool frad = false;
    double fcx = 0, fcy = 0, fr = 0;
    std::vector<lbm_link> explicit_links;
    double shift_u = 0.0, shift_v = 0.0;  // 16-bit storage shift: feq(1, shift_u, shift_v) - w
    bool geom_dirty = true;
    // built geometry
    std::vector<LinkRec> recs;
    int32_t* d_lslot_new = nullptr;  // flags and link slots of the last build (device, pitch
    uint8_t* d_flags_new = nullptr;  // layout), moved into the pools by commit
    std::vector<int32_t> special;    // y*pitch+x of the general-path nodes
    int64_t special_off = 0;
    int64_t fluid = 0;
    bool uses_bouzidi = false;
    // run state
    bool uploaded = false;
    int64_t steps = 0;
    int32_t par = 0;
    int64_t read_pos = 0;
    std::vector<double> pend_fx, pend_fy;
};

}  // namespace

struct Ensemble {
    int storage = ST_F32;
    int scale_exp = 0;
    float scale = 1.f, inv_scale = 1.f;
    int bpp = 4;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::vector<HostLattice> lat;
    int64_t pool_elems = 0, cell_elems = 0;
    int64_t guard_elems = 0;  // population buffers end with this many more elements: the
                              // step kernel's pull addresses reach past the pool by < one row
    void* buf[2] = {nullptr, nullptr};
    uint8_t* d_flags = nullptr;
    int32_t* d_lslot = nullptr;
    uint16_t* d_wdist = nullptr;
    LinkRec* d_links = nullptr;
    size_t links_cap = 0;
    LatticeDev* d_lat = nullptr;
    Tile* d_tiles = nullptr;       // ensembles: the one-step tiles
    size_t tiles_cap = 0;
    int n_tiles = 0;
    Tile* d_tiles2 = nullptr;      // ensembles with two steps per launch: 64 x rows tiles
    size_t tiles2_cap = 0;
    int n_tiles2 = 0;
    // two-step launches, per tile height (8, 16, 32): the bulk tiles (step2_bulk_kernel),
    // the edge tiles (its edge form) and the others (the ensemble form of step2_kernel on
    // 64 x rest_rows tiles, rest_rows = min(rows, 16)), classified from the flags
    struct BulkPlan {
        int rows = 0;  // 0: not built
        int rest_rows = 0;
        Tile* d = nullptr;     // bulk tiles
        Tile* de = nullptr;    // edge tiles
        Tile* dr = nullptr;    // the others, as an ensemble tile table
        int n_bulk = 0, n_edge = 0, n_rest = 0;
        int n_empty = 0;       // tiles of solid nodes nothing reads: left out
    };
    BulkPlan plan[3];
    // two more streams for the edge and the other tiles of a split launch, so that they run
    // beside the bulk tiles (fork and join by events; graph capture records the branches)
    hipStream_t side = nullptr, side2 = nullptr;
    hipEvent_t fork_ev = nullptr, join_ev = nullptr, join2_ev = nullptr;
    // the last batch ended with a two-step launch: the buffer that is not current holds
    // C_{k-2} instead of C_{k-1} (complete_prev)
    bool prev_pending = false;
    int64_t* d_steps = nullptr;
    int32_t* d_par = nullptr;
    double* d_acc = nullptr;       // [acc_blocks][kAccCopies][total_slots][2]
    int acc_blocks = 0;            // launches a batch may hold
    int32_t* d_slot_lat = nullptr; // lattice of each force slot
    int32_t* d_special = nullptr;  // every lattice's special-node list
    size_t special_cap = 0;
    double* d_nf = nullptr;        // [acc_blocks][n_recs][2] force terms per link record
    double* d_dummy_nf = nullptr;
    int32_t* d_slot_start = nullptr;  // records per force slot (CSR), in record order
    int32_t* d_slot_recs = nullptr;
    int n_recs = 0;
    std::vector<LatticeDev> host_lat;
    double* d_hist = nullptr;
    double* d_dummy_acc = nullptr;
    int32_t* d_nan = nullptr;
    int32_t* d_dummy_nan = nullptr;
    double* d_mass = nullptr;
    double* d_stage = nullptr;
    int64_t stage_bytes = 0;
    int total_slots = 0;
    int R = kHistoryRows;
    int bx = 256, rows = 1, vec = 1, options = 0, graph_steps = 100;
    // the one-step launch that ends a batch of two-step launches: the autotuner's fastest
    // one-step shape, else (64, 1, 1) with the two-step launch's order and store options
    int tail[4] = {0, 0, 0, 0};
    int sr = 0;                    // stochastic rounding (16-bit forms)
    uint32_t sr_seed = 0;
    bool geom_dirty = true, desc_dirty = true, tiles_dirty = true, counters_stale = false;
    hipGraphExec_t graph[2] = {nullptr, nullptr};  // by the starting buffer parity (one lattice)
    bool graph_dirty = true;
    int64_t device_bytes = 0;
    // thermal double distribution (one-lattice handles, F64 / F32 storage): g planes laid
    // out as the f planes, the same buffer parity; per-node conditions and their rules
    bool thermal = false;
    void* gbuf[2] = {nullptr, nullptr};
    uint8_t* d_tcode = nullptr;
    bool has_rules = false;
    double omega_k = 1.0, beta = 0.0, g_buoyancy = 0.0, t_ref = 1.0;
    double rule_t[kThermalRules] = {};
    uint32_t rule_rest = 0;
};

namespace {

int lattice_ok(Ensemble* E, int l)
{
    if (!E) return fail(LBM_ERR_INVALID, "null handle");
    if (l < 0 || l >= static_cast<int>(E->lat.size())) return fail(LBM_ERR_INVALID, "lattice index out of range");
    return LBM_OK;
}

template <typename T> int dev_alloc(Ensemble* E, T** p, size_t count)
{
    if (*p) {
        (void)hipFree(*p);
        *p = nullptr;
    }
    if (count == 0) count = 1;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(LBM_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    E->device_bytes += static_cast<int64_t>(count * sizeof(T));
    return LBM_OK;
}

// Layout of the population buffers for lattices descs[0..n): each lattice's 9 planes of
// ny rows of pitch (nx rounded up to 64) values from f_off[l], blocks aligned to 256
// values; then a guard.  The step kernels form a node's nine pull addresses before they
// know whether the node is on an edge, and the general path gathers sources one row
// beyond the lattice on open edges: the furthest element any kernel forms for lattice l
// is f_off + 9 ny pitch + nx - 2 (the south-east source of the last row's last node,
// lbm_kernels.hip cell_update / preload), which lands in the next lattice's block or,
// for the last lattice, in the guard of one row of the widest lattice plus 256 values.
struct PoolLayout {
    std::vector<int64_t> f_off;
    int64_t pool = 0;   // values of every lattice's planes
    int64_t guard = 0;  // values after them
    int64_t elems() const { return pool + guard; }
};

PoolLayout pool_layout(const lbm_desc* descs, int n)
{
    PoolLayout P;
    int64_t guard = 256;
    for (int l = 0; l < n; ++l) {
        const int64_t pitch = (descs[l].nx + 63) / 64 * 64;
        const int64_t plane = static_cast<int64_t>(descs[l].ny) * pitch;
        P.f_off.push_back(P.pool);
        P.pool += (9 * plane + 255) / 256 * 256;
        guard = std::max<int64_t>(guard, pitch + 256);
    }
    P.guard = guard;
    return P;
}

// Every population buffer of a handle -- the two pools, the thermal populations' two,
// the autotuner's copies -- is allocated here, with the guard.
int alloc_population(Ensemble* E, void** p, const char* what)
{
    *p = nullptr;
    const size_t bytes = static_cast<size_t>(E->pool_elems + E->guard_elems) * E->bpp;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(LBM_ERR_NOMEM, std::string(what) + ": " + hipGetErrorString(e));
    }
    E->device_bytes += static_cast<int64_t>(bytes);
    return LBM_OK;
}

// Host-to-device copy ordered on the handle's stream, complete on return.
int h2d(Ensemble* E, void* dst, const void* src, size_t bytes)
{
    if (bytes == 0) return LBM_OK;
    LBM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, E->stream));
    LBM_HIP(hipStreamSynchronize(E->stream));
    return LBM_OK;
}

int d2h(Ensemble* E, void* dst, const void* src, size_t bytes)
{
    if (bytes == 0) return LBM_OK;
    LBM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, E->stream));
    LBM_HIP(hipStreamSynchronize(E->stream));
    return LBM_OK;
}

bool neighbour(const lbm_desc& d, int x, int y, int i, int& tx, int& ty)
{
    tx = x + CX[i];
    ty = y + CY[i];
    if (tx < 0 || tx >= d.nx) {
        if (d.x_edge != LBM_EDGE_PERIODIC) return false;
        tx = tx < 0 ? tx + d.nx : tx - d.nx;
    }
    if (ty < 0 || ty >= d.ny) {
        if (d.y_edge != LBM_EDGE_PERIODIC) return false;
        ty = ty < 0 ? ty + d.ny : ty - d.ny;
    }
    return true;
}

// Frees a lattice's built-but-not-committed device geometry.
void drop_new_geometry(HostLattice& H)
{
    if (H.d_lslot_new) (void)hipFree(H.d_lslot_new);
    if (H.d_flags_new) (void)hipFree(H.d_flags_new);
    H.d_lslot_new = nullptr;
    H.d_flags_new = nullptr;
}

// Device scratch freed on every return path of the geometry build.
struct Scratch {
    void* p = nullptr;
    ~Scratch()
    {
        if (p) (void)hipFree(p);
    }
    template <typename T> int alloc(T*& out, size_t count)
    {
        if (hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) {
            p = nullptr;
            return fail(LBM_ERR_NOMEM, "geometry build scratch");
        }
        out = static_cast<T*>(p);
        return LBM_OK;
    }
};

// Flags, link records and wall distance of one lattice from its description.  The
// per-node pass (flags, link masks, the records' order) runs on the device; the
// per-link values (q, schemes, slots, wall velocities) are set here on the records.
int build_geometry(Ensemble* E, HostLattice& H)
{
    const lbm_desc& d = H.d;
    const int nx = d.nx, ny = d.ny;
    const size_t ncell = static_cast<size_t>(nx) * ny, npitch = static_cast<size_t>(ny) * H.pitch;
    if (H.obstacle.size() != ncell) H.obstacle.assign(ncell, 0);
    Scratch s_obs, s_lmask, s_rows, s_off, s_fluid, s_cell, s_mask;
    GeomArgs g{};
    int rc;
    uint8_t* obs = nullptr;
    int32_t* rows = nullptr;
    if ((rc = s_obs.alloc(obs, ncell)) || (rc = s_lmask.alloc(g.lmask, ncell)) ||
        (rc = s_rows.alloc(rows, ny)) || (rc = s_fluid.alloc(g.fluid, 1)))
        return rc;
    // flags and link slots stay on the device: commit copies them into the pools
    drop_new_geometry(H);
    if (hipMalloc(&H.d_lslot_new, std::max<size_t>(npitch, 1) * 4) != hipSuccess ||
        hipMalloc(&H.d_flags_new, std::max<size_t>(npitch, 1)) != hipSuccess) {
        drop_new_geometry(H);
        return fail(LBM_ERR_NOMEM, "geometry build scratch");
    }
    g.lslot = H.d_lslot_new;
    g.flags = H.d_flags_new;
    g.obs = obs;
    g.row_links = rows;
    g.nx = nx;
    g.ny = ny;
    g.pitch = H.pitch;
    g.x_edge = d.x_edge;
    g.y_edge = d.y_edge;
    if ((rc = h2d(E, obs, H.obstacle.data(), ncell)) != LBM_OK) return rc;
    LBM_HIP(hipMemsetAsync(g.flags, 0, npitch, E->stream));
    LBM_HIP(hipMemsetAsync(g.lslot, 0xff, npitch * 4, E->stream));
    LBM_HIP(hipMemsetAsync(rows, 0, static_cast<size_t>(ny) * 4, E->stream));
    LBM_HIP(hipMemsetAsync(g.fluid, 0, 8, E->stream));
    LBM_HIP(launch_geom_cells(g, E->stream));
    std::vector<int32_t> row_off(static_cast<size_t>(ny) + 1, 0);
    unsigned long long fluid = 0;
    if ((rc = d2h(E, row_off.data() + 1, rows, static_cast<size_t>(ny) * 4)) != LBM_OK) return rc;
    if ((rc = d2h(E, &fluid, g.fluid, 8)) != LBM_OK) return rc;
    for (int y = 0; y < ny; ++y) row_off[y + 1] += row_off[y];
    const int32_t n_links = row_off[ny];
    std::vector<int32_t> rec_cell(n_links);
    std::vector<uint8_t> rec_mask(n_links);
    if (n_links > 0) {
        int32_t* off = nullptr;
        if ((rc = s_off.alloc(off, ny)) || (rc = s_cell.alloc(g.rec_cell, n_links)) ||
            (rc = s_mask.alloc(g.rec_mask, n_links)))
            return rc;
        g.row_off = off;
        if ((rc = h2d(E, off, row_off.data(), static_cast<size_t>(ny) * 4)) != LBM_OK) return rc;
        LBM_HIP(launch_geom_links(g, E->stream));
        if ((rc = d2h(E, rec_cell.data(), g.rec_cell, static_cast<size_t>(n_links) * 4)) != LBM_OK) return rc;
        if ((rc = d2h(E, rec_mask.data(), g.rec_mask, static_cast<size_t>(n_links))) != LBM_OK) return rc;
    }
    H.fluid = static_cast<int64_t>(fluid);
    // records in row-major order of their nodes: simple bounce-back counted in slot 0
    // until the geometry below says otherwise
    H.recs.assign(static_cast<size_t>(n_links), LinkRec{});
    for (int32_t r = 0; r < n_links; ++r) {
        LinkRec& k = H.recs[r];
        k.mask = static_cast<uint32_t>(rec_mask[r]) << 1;
        k.scheme = 0;
        for (int i = 0; i < 9; ++i) {
            k.slot[i] = 0;
            k.q[i] = 0.5;
            k.uwx[i] = k.uwy[i] = 0.0;
        }
        k.cell = rec_cell[r];
    }
    auto each_link = [&](auto&& fn) {
        for (LinkRec& r : H.recs) {
            const int y = r.cell / H.pitch, x = r.cell - y * H.pitch;
            for (int i = 1; i < 9; ++i)
                if (r.mask & (1u << i)) fn(x, y, i, r);
        }
    };
    if (H.bb_kind != 0) {
        each_link([&](int x, int y, int i, LinkRec& r) {
            r.q[i] = H.bb_kind == 1 ? q_circle(H.bcx, H.bcy, H.br, x, y, i) : q_polygon(H.poly, x, y, i);
            r.scheme = (r.scheme & ~(3u << (2 * i))) | (static_cast<uint32_t>(H.bb_scheme) << (2 * i));
        });
    }
    if (H.moving) {
        // compute_wall_velocity at the (unwrapped) solid node (src/lbm_types.hpp:106-114, lbm.hpp:244)
        each_link([&](int x, int y, int i, LinkRec& r) {
            const double px = x + CX[i], py = y + CY[i];
            r.uwx[i] = -H.omega * (py - H.rcy);
            r.uwy[i] = H.omega * (px - H.rcx);
        });
    }
    if (H.frad) {
        // keep links whose (wrapped) solid node is within r of the centre (src/lbm.hpp:569-573)
        each_link([&](int x, int y, int i, LinkRec& r) {
            int tx, ty;
            neighbour(d, x, y, i, tx, ty);
            const double dx = tx - H.fcx, dy = ty - H.fcy;
            if (std::sqrt(dx * dx + dy * dy) > H.fr) r.slot[i] = -1;
        });
    }
    for (const lbm_link& k : H.explicit_links) {
        if (k.cell < 0 || k.cell >= static_cast<int64_t>(nx) * ny || k.dir < 1 || k.dir > 8)
            return fail(LBM_ERR_INVALID, "link cell/dir out of range");
        // the node's record: records are in row-major order of their nodes
        const int32_t node = static_cast<int32_t>(static_cast<int64_t>(k.cell / nx) * H.pitch + k.cell % nx);
        const auto it = std::lower_bound(H.recs.begin(), H.recs.end(), node,
                                         [](const LinkRec& r, int32_t v) { return r.cell < v; });
        const int32_t ri = (it != H.recs.end() && it->cell == node) ? static_cast<int32_t>(it - H.recs.begin()) : -1;
        if (ri < 0 || !(H.recs[ri].mask & (1u << k.dir)))
            return fail(LBM_ERR_INVALID, "link " + std::to_string(k.cell) + "/" + std::to_string(k.dir) +
                                             " is not a fluid-to-solid link of the obstacle mask");
        if (k.scheme < 0 || k.scheme > 2) return fail(LBM_ERR_INVALID, "bad link scheme");
        if (k.slot >= H.n_slots) return fail(LBM_ERR_INVALID, "link slot beyond the lattice's force slots");
        LinkRec& r = H.recs[ri];
        r.q[k.dir] = k.q;
        r.uwx[k.dir] = k.uwx;
        r.uwy[k.dir] = k.uwy;
        r.slot[k.dir] = static_cast<int16_t>(k.slot);
        r.scheme = (r.scheme & ~(3u << (2 * k.dir))) | (static_cast<uint32_t>(k.scheme) << (2 * k.dir));
    }
    for (LinkRec& r : H.recs)
        for (int i = 1; i < 9; ++i)
            if ((r.mask & (1u << i)) && (r.uwx[i] != 0.0 || r.uwy[i] != 0.0)) r.scheme |= LINK_MOVING;
    H.uses_bouzidi = false;
    for (const LinkRec& r : H.recs)
        for (int i = 1; i < 9; ++i)
            if ((r.mask & (1u << i)) && ((r.scheme >> (2 * i)) & 3) == LBM_BB_BOUZIDI) H.uses_bouzidi = true;
    H.special.clear();
    // inlet and outlet nodes the step kernel's bulk tiles take: off the corners, fluid,
    // without links; at the outlet the copied node (nx-2) as well.  Their state is the
    // plain pulls plus the inlet closure (Zou/He overwrites exactly the slots whose
    // sources lie off the lattice, the equilibrium inlet every slot), or the plain pulls
    // of the copied node.
    if (d.x_edge == LBM_EDGE_OPEN && nx >= 4) {
        const int inlet_from = (d.inlet == LBM_INLET_ZOU_HE || d.inlet == LBM_INLET_EQUILIBRIUM) ? 1
                               : d.inlet == LBM_INLET_PARABOLIC ? std::max(1, d.inlet_y0)
                                                                : ny;
        LBM_HIP(launch_geom_columns(H.d_flags_new, nx, ny, H.pitch, inlet_from, d.outlet, E->stream));
    }
    if (H.uses_bouzidi) {
        // Bouzidi reads the stored value of upstream nodes, which may be solid: every
        // solid node is written, and the ones off the edges are listed for the
        // one-lattice kernel's list blocks (its link and edge blocks cover the other
        // general-path nodes), in row-major order
        LBM_HIP(hipMemsetAsync(rows, 0, static_cast<size_t>(ny) * 4, E->stream));
        LBM_HIP(launch_geom_write(H.d_flags_new, nx, ny, H.pitch, rows, E->stream));
        std::vector<int32_t> solid_off(static_cast<size_t>(ny) + 1, 0);
        if ((rc = d2h(E, solid_off.data() + 1, rows, static_cast<size_t>(ny) * 4)) != LBM_OK) return rc;
        for (int y = 0; y < ny; ++y) solid_off[y + 1] += solid_off[y];
        const int32_t n_solid = solid_off[ny];
        if (n_solid > 0) {
            Scratch s_soff, s_list;
            int32_t* soff = nullptr;
            int32_t* list = nullptr;
            if ((rc = s_soff.alloc(soff, ny)) || (rc = s_list.alloc(list, n_solid))) return rc;
            if ((rc = h2d(E, soff, solid_off.data(), static_cast<size_t>(ny) * 4)) != LBM_OK) return rc;
            LBM_HIP(launch_geom_solids(H.d_flags_new, nx, ny, H.pitch, soff, list, E->stream));
            H.special.resize(static_cast<size_t>(n_solid));
            if ((rc = d2h(E, H.special.data(), list, static_cast<size_t>(n_solid) * 4)) != LBM_OK) return rc;
        }
    }
    LBM_HIP(hipStreamSynchronize(E->stream));
    H.geom_dirty = false;
    return LBM_OK;
}

LatticeDev make_dev(const HostLattice& H, int index, float scale)
{
    LatticeDev L{};
    const lbm_desc& d = H.d;
    L.nx = d.nx;
    L.ny = d.ny;
    L.pitch = H.pitch;
    L.lattice = index;
    L.f_off = H.f_off;
    L.plane = static_cast<int64_t>(d.ny) * H.pitch;
    L.cell_off = H.cell_off;
    L.link_off = H.link_off;
    L.slot_base = H.slot_base;
    L.n_slots = H.n_slots;
    L.collision = d.collision;
    L.les = d.les;
    L.x_edge = d.x_edge;
    L.y_edge = d.y_edge;
    L.inlet = d.inlet;
    L.inlet_y0 = d.inlet_y0;
    L.inlet_h = d.inlet_h;
    L.outlet = d.outlet;
    L.lid = d.lid;
    L.forces = d.forces;
    L.mei_static = 1;
    for (const LinkRec& r : H.recs) {
        if (r.scheme & LINK_MOVING) L.mei_static = 0;
        for (int i = 1; i < 9; ++i)
            if ((r.mask >> i) & 1u && ((r.scheme >> (2 * i)) & 3u) != static_cast<uint32_t>(LBM_BB_MEI)) L.mei_static = 0;
    }
    // MRTParams::from_tau (src/lbm_types.hpp:72-76)
    double s = 1.0 / d.tau;
    L.s_shear = s < 0.5 ? 0.5 : (s > 1.99 ? 1.99 : s);
    L.tau = d.tau;
    L.inv_tau = 1.0 / d.tau;
    L.half_tau = 0.5 * d.tau;
    L.qtau2 = 0.25 * d.tau * d.tau;
    L.u_in = d.u_in;
    L.cs = d.les ? d.cs : 0.0;
    L.cs2 = d.les ? d.cs * d.cs : 0.0;
    L.nu = (d.tau - 0.5) / 3.0;
    L.inv_nu = L.nu > 0.0 ? 1.0 / L.nu : 0.0;
    // Fscale (src/lbm.hpp:296-297)
    L.fscale = d.body_force_x != 0.0 ? (1.0 - 1.0 / (2.0 * d.tau)) * 3.0 * d.body_force_x : 0.0;
    // storage shift feq_i(1, u_s, v_s) - w_i (compute_equilibrium, src/lbm_types.hpp:269-273)
    L.shifted = 0;
    for (int i = 0; i < 9; ++i) {
        const double w = i == 0 ? 4.0 / 9.0 : (i < 5 ? 1.0 / 9.0 : 1.0 / 36.0);
        const double eu = CX[i] * H.shift_u + CY[i] * H.shift_v;
        const double uu = H.shift_u * H.shift_u + H.shift_v * H.shift_v;
        L.shift[i] = static_cast<float>(w * (3.0 * eu + 4.5 * eu * eu - 1.5 * uu));
        if (L.shift[i] != 0.0f) L.shifted = 1;
    }
    for (int i = 0; i < 9; ++i) {
        // no shift: -0.0f both ways (adding -0 leaves every value, -0 included, as it is)
        if (!L.shifted) L.shift[i] = -0.0f;
        L.shift_s[i] = L.shift[i] * scale;  // exact: scale is a power of two
        L.nshift[i] = L.shifted ? -L.shift[i] : -0.0f;
        L.nshift_s[i] = L.shifted ? -L.shift_s[i] : -0.0f;
    }
    return L;
}

// Rebuilds the geometry pools, slot layout and descriptors that changed.
// Nodes on the lattice's edges: rows 0 and ny-1, columns 0 and nx-1 between them.
int edge_nodes(int nx, int ny)
{
    return (ny == 1 ? nx : 2 * nx) + (nx == 1 ? 1 : 2) * std::max(ny - 2, 0);
}

// The shape and options of one-step launches {block_x, rows, nodes per thread, options}:
// the handle's, or with two steps per launch that of the launches ending a batch (the
// autotuner's fastest one-step shape, else (64, 1, 1) with the order and store options).
void one_step_shape(const Ensemble* E, int s[4])
{
    if (!(E->options & LAUNCH_TWO_STEPS)) {
        s[0] = E->bx;
        s[1] = E->rows;
        s[2] = E->vec;
        s[3] = E->options;
        return;
    }
    const bool tuned = E->tail[0] != 0;
    s[0] = tuned ? E->tail[0] : 64;
    s[1] = tuned ? E->tail[1] : 1;
    s[2] = tuned ? E->tail[2] : 1;
    s[3] = (tuned ? E->tail[3] : E->options) & (LAUNCH_XCD_ORDER | LAUNCH_STREAM_STORES);
}

// First column of a lattice's two-step tiles in an ensemble: 0, or -63 when the last of
// the 64-wide tiles would hold only the outlet column (its outflow copy gathers the
// column before it, which must lie inside the tile).
int two_step_x_start(const lbm_desc& d) { return d.outlet && d.nx % 64 == 1 ? -63 : 0; }

int plan_slot(int rows) { return rows == 8 ? 0 : (rows == 16 ? 1 : (rows == 32 ? 2 : -1)); }

// Whether two-step launches of tile height rows split into bulk and other tiles: a
// momentum handle (one lattice or an ensemble), FP32 or a 16-bit form, rows 8, 16 or 32,
// and every lattice's rows short enough for the bulk kernel's 32-bit offsets within a
// tile's ring and its 24-bit row products.
bool bulk_eligible(const Ensemble* E, int rows)
{
    if (E->thermal || !bulk_supported(E->storage, rows)) return false;
    for (const auto& H : E->lat) {
        const int64_t row_bytes = static_cast<int64_t>(H.pitch) * E->bpp;
        if ((rows + 3) * row_bytes >= (int64_t(1) << 31) || H.pitch >= (1 << 24)) return false;
    }
    return true;
}

// Whether a lattice's edge tiles can take the bulk kernel's edge form: its pull addresses
// are 32-bit byte offsets within a plane and 24-bit row products.
bool edge_tiles_fit(const Ensemble* E, const HostLattice& H)
{
    return static_cast<int64_t>(H.d.ny + 1) * H.pitch * E->bpp < (int64_t(1) << 32) && H.d.ny < (1 << 24);
}

void drop_plans(Ensemble* E)
{
    for (auto& p : E->plan) {
        if (p.d) (void)hipFree(p.d);
        if (p.de) (void)hipFree(p.de);
        if (p.dr) (void)hipFree(p.dr);
        p = Ensemble::BulkPlan{};
    }
}

// Classifies the 64 x rows two-step tiles of every lattice (rows of an ensemble lattice
// with an outlet and nx % 64 == 1 right-aligned, as its two-step tile table) from the
// device flags, and uploads the bulk tiles, the edge tiles, and the other tiles cut to
// rest_rows rows as a tile table for the ensemble form of the two-step kernel (which takes
// inlet, outlet and periodic-edge nodes on its plain path and queues the remaining
// general-path nodes of a tile: the other tiles are those holding such nodes, or partial
// tiles).
int build_plan(Ensemble* E, int rows)
{
    if (!E->side) {
        if (hipStreamCreateWithFlags(&E->side, hipStreamNonBlocking) != hipSuccess) {
            E->side = nullptr;
            return fail(LBM_ERR_DEVICE, "hipStreamCreateWithFlags: side stream");
        }
        if (hipStreamCreateWithFlags(&E->side2, hipStreamNonBlocking) != hipSuccess) {
            E->side2 = nullptr;
            return fail(LBM_ERR_DEVICE, "hipStreamCreateWithFlags: side stream");
        }
        if (hipEventCreateWithFlags(&E->fork_ev, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&E->join_ev, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&E->join2_ev, hipEventDisableTiming) != hipSuccess)
            return fail(LBM_ERR_DEVICE, "hipEventCreateWithFlags: fork / join events");
    }
    Ensemble::BulkPlan& P = E->plan[plan_slot(rows)];
    if (P.d) (void)hipFree(P.d);
    if (P.de) (void)hipFree(P.de);
    if (P.dr) (void)hipFree(P.dr);
    P = Ensemble::BulkPlan{};
    std::vector<Tile> tiles;
    for (size_t l = 0; l < E->lat.size(); ++l) {
        const lbm_desc& d = E->lat[l].d;
        const int xs = E->lat.size() > 1 ? two_step_x_start(d) : 0;
        for (int y0 = 0; y0 < d.ny; y0 += rows)
            for (int x0 = xs; x0 < d.nx; x0 += 64) tiles.push_back(Tile{static_cast<int32_t>(l), y0, x0, 0});
    }
    const size_t n = tiles.size();
    Tile* d_tiles = nullptr;
    uint8_t* d_cls = nullptr;
    if (hipMalloc(reinterpret_cast<void**>(&d_tiles), n * sizeof(Tile)) != hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&d_cls), n) != hipSuccess) {
        if (d_tiles) (void)hipFree(d_tiles);
        return fail(LBM_ERR_NOMEM, "hipMalloc: two-step tile classes");
    }
    std::vector<uint8_t> cls(n);
    int rc = h2d(E, d_tiles, tiles.data(), n * sizeof(Tile));
    if (rc == LBM_OK) {
        hipError_t e = launch_tile_class(d_tiles, static_cast<int>(n), E->d_lat, E->d_flags, rows, d_cls, E->stream);
        rc = e == hipSuccess ? d2h(E, cls.data(), d_cls, n)
                             : fail(LBM_ERR_DEVICE, std::string("tile classes: ") + hipGetErrorString(e));
    }
    (void)hipFree(d_tiles);
    (void)hipFree(d_cls);
    if (rc != LBM_OK) return rc;
    const int rest_rows = std::min(rows, 16);
    std::vector<Tile> bulk, edge, rest;
    for (size_t k = 0; k < n; ++k) {
        const Tile& t = tiles[k];
        if (cls[k] == 1) {
            bulk.push_back(t);
        } else if (cls[k] == 2 && edge_tiles_fit(E, E->lat[t.lattice])) {
            edge.push_back(t);
        } else if (cls[k] == 3) {
            ++P.n_empty;
        } else {
            const int ny = E->lat[t.lattice].d.ny;
            for (int y = t.y0; y < std::min(t.y0 + rows, ny); y += rest_rows) rest.push_back(Tile{t.lattice, y, t.x0, 0});
        }
    }
    // Bulk tiles in vertical strips of kStripTiles tile columns, row band after row band
    // inside a strip: a tile's ring rows are read again by the tiles above and below it,
    // which then follow a strip's width later (in L2) rather than a lattice row of tiles
    // later.  Blocks go to the XCDs in contiguous ranges of this order (LAUNCH_XCD_ORDER).
    std::stable_sort(bulk.begin(), bulk.end(), [](const Tile& p, const Tile& q) {
        if (p.lattice != q.lattice) return p.lattice < q.lattice;
        const int sp = p.x0 / (64 * kStripTiles), sq = q.x0 / (64 * kStripTiles);
        if (sp != sq) return sp < sq;
        if (p.y0 != q.y0) return p.y0 < q.y0;
        return p.x0 < q.x0;
    });
    if ((rc = dev_alloc(E, &P.d, bulk.size())) != LBM_OK) return rc;
    if ((rc = h2d(E, P.d, bulk.data(), bulk.size() * sizeof(Tile))) != LBM_OK) return rc;
    if ((rc = dev_alloc(E, &P.de, edge.size())) != LBM_OK) return rc;
    if ((rc = h2d(E, P.de, edge.data(), edge.size() * sizeof(Tile))) != LBM_OK) return rc;
    if ((rc = dev_alloc(E, &P.dr, rest.size())) != LBM_OK) return rc;
    if ((rc = h2d(E, P.dr, rest.data(), rest.size() * sizeof(Tile))) != LBM_OK) return rc;
    P.rows = rows;
    P.rest_rows = rest_rows;
    P.n_bulk = static_cast<int>(bulk.size());
    P.n_edge = static_cast<int>(edge.size());
    P.n_rest = static_cast<int>(rest.size());
    return LBM_OK;
}

// The plan of the handle's two-step tile height, or null (not split or not built).
const Ensemble::BulkPlan* current_plan(const Ensemble* E)
{
    if (!(E->options & LAUNCH_TWO_STEPS) || !bulk_eligible(E, E->rows)) return nullptr;
    const Ensemble::BulkPlan& P = E->plan[plan_slot(E->rows)];
    return P.rows == E->rows ? &P : nullptr;
}

// One two-step launch of a batch: with a plan, the bulk tiles, the edge tiles and the
// others beside one another; else the grid of tiles (one lattice) or the ensemble's tile
// table.
hipError_t launch_pair(Ensemble* E, const StepArgs& a2, bool single, int blocks2)
{
    const Ensemble::BulkPlan* P = current_plan(E);
    if (!P) return launch_step2(E->storage, E->rows, single, blocks2, a2, E->stream);
    StepArgs ab = a2;
    ab.bulk_tiles = P->d;
    StepArgs ae = a2;
    ae.bulk_tiles = P->de;
    StepArgs ar = a2;
    ar.tiles = P->dr;
    ar.n_tiles = P->n_rest;
    // the kernels with tiles, the bulk tiles' first: that one on the handle's stream, the
    // others (few blocks: the edge tiles, the other tiles' long general-path chains) on the
    // side streams, beside it; all read the source buffer and write disjoint nodes of the
    // destination, and the next launch on the stream waits for all.  The dispatcher keeps
    // serving the kernel that reached it first: if the bulk kernel's thousands of blocks
    // come first, the other kernels' blocks wait for its last ones and the launch runs as
    // long as their chains after it (on the 6400 x 2400 lattice 205 against 60 us).  A
    // workgroup with no work goes ahead of the bulk kernel on the handle's stream (graph
    // replays start their parallel branches in no fixed order), so that the side streams'
    // kernels are dispatched while it runs.
    auto launch = [&](int kind, hipStream_t s) {
        if (kind == 0) return launch_step2_bulk(E->storage, P->rows, single, false, P->n_bulk, ab, s);
        if (kind == 1) return launch_step2_bulk(E->storage, P->rows, single, true, P->n_edge, ae, s);
        return launch_step2(E->storage, P->rest_rows, false, P->n_rest, ar, s);
    };
    int kinds[3], n = 0;
    if (P->n_bulk) kinds[n++] = 0;
    if (P->n_edge) kinds[n++] = 1;
    if (P->n_rest) kinds[n++] = 2;
    if (n == 0) return hipSuccess;
    hipError_t e;
    if (n > 1) {
        if ((e = hipEventRecord(E->fork_ev, E->stream)) != hipSuccess) return e;
        hipStream_t sides[2] = {E->side, E->side2};
        hipEvent_t joins[2] = {E->join_ev, E->join2_ev};
        for (int k = 1; k < n; ++k) {
            if ((e = hipStreamWaitEvent(sides[k - 1], E->fork_ev, 0)) != hipSuccess) return e;
            if ((e = launch(kinds[k], sides[k - 1])) != hipSuccess) return e;
            if ((e = hipEventRecord(joins[k - 1], sides[k - 1])) != hipSuccess) return e;
        }
        if ((e = launch_hold(E->stream)) != hipSuccess) return e;
        if ((e = launch(kinds[0], E->stream)) != hipSuccess) return e;
        for (int k = 1; k < n; ++k)
            if ((e = hipStreamWaitEvent(E->stream, joins[k - 1], 0)) != hipSuccess) return e;
        return hipSuccess;
    }
    return launch(kinds[0], E->stream);
}

int commit(Ensemble* E)
{
    int rc;
    if (E->geom_dirty) {
        bool any = false;
        std::vector<char> rebuilt(E->lat.size(), 0);
        for (size_t l = 0; l < E->lat.size(); ++l)
            if (E->lat[l].geom_dirty) {
                if ((rc = build_geometry(E, E->lat[l])) != LBM_OK) return rc;
                any = true;
                rebuilt[l] = 1;
            }
        if (any) {
            size_t nrec = 0;
            int slots = 0;
            for (auto& H : E->lat) {
                H.link_off = static_cast<int32_t>(nrec);
                nrec += H.recs.size();
                H.slot_base = slots;
                slots += H.n_slots;
            }
            if (nrec > E->links_cap) {
                if ((rc = dev_alloc(E, &E->d_links, nrec)) != LBM_OK) return rc;
                E->links_cap = nrec;
                E->graph_dirty = true;
            }
            std::vector<LinkRec> all;
            all.reserve(nrec);
            for (auto& H : E->lat) {
                all.insert(all.end(), H.recs.begin(), H.recs.end());
            }
            if ((rc = h2d(E, E->d_links, all.data(), all.size() * sizeof(LinkRec))) != LBM_OK) return rc;
            // each record's force terms go to its primary slot: that of its first counted link
            std::vector<std::vector<int32_t>> per_slot(slots);
            for (auto& H : E->lat) {
                if (!H.d.forces) continue;
                for (size_t r = 0; r < H.recs.size(); ++r)
                    for (int i = 1; i < 9; ++i)
                        if ((H.recs[r].mask & (1u << i)) && H.recs[r].slot[i] >= 0) {
                            per_slot[H.slot_base + H.recs[r].slot[i]].push_back(H.link_off + static_cast<int32_t>(r));
                            break;
                        }
            }
            std::vector<int32_t> start(1, 0), recs_of;
            for (auto& v : per_slot) {
                recs_of.insert(recs_of.end(), v.begin(), v.end());
                start.push_back(static_cast<int32_t>(recs_of.size()));
            }
            if ((rc = dev_alloc(E, &E->d_slot_start, start.size())) != LBM_OK) return rc;
            if ((rc = dev_alloc(E, &E->d_slot_recs, recs_of.size())) != LBM_OK) return rc;
            if ((rc = h2d(E, E->d_slot_start, start.data(), start.size() * 4)) != LBM_OK) return rc;
            if ((rc = h2d(E, E->d_slot_recs, recs_of.data(), recs_of.size() * 4)) != LBM_OK) return rc;
            if (static_cast<int>(nrec) != E->n_recs || !E->d_dummy_nf) {
                E->n_recs = static_cast<int>(nrec);
                if ((rc = dev_alloc(E, &E->d_dummy_nf, nrec * 2)) != LBM_OK) return rc;
                E->acc_blocks = 0;  // the per-record force blocks are reallocated below
            }
            std::vector<int32_t> special;
            for (auto& H : E->lat) {
                H.special_off = static_cast<int64_t>(special.size());
                special.insert(special.end(), H.special.begin(), H.special.end());
            }
            if (special.size() > E->special_cap || !E->d_special) {
                if ((rc = dev_alloc(E, &E->d_special, special.size())) != LBM_OK) return rc;
                E->special_cap = special.size();
            }
            if ((rc = h2d(E, E->d_special, special.data(), special.size() * 4)) != LBM_OK) return rc;
            E->graph_dirty = true;
            drop_plans(E);  // the tile classes follow the flags
            // flags and link slots of the lattices just built, device to device (the others
            // are in place)
            for (auto& H : E->lat)
                if (H.d_lslot_new && H.d_flags_new) {
                    const size_t np = static_cast<size_t>(H.d.ny) * H.pitch;
                    LBM_HIP(hipMemcpyAsync(E->d_flags + H.cell_off, H.d_flags_new, np, hipMemcpyDeviceToDevice,
                                           E->stream));
                    LBM_HIP(hipMemcpyAsync(E->d_lslot + H.cell_off, H.d_lslot_new, np * 4, hipMemcpyDeviceToDevice,
                                           E->stream));
                    LBM_HIP(hipStreamSynchronize(E->stream));
                    drop_new_geometry(H);
                }
            // wall distance of each lattice just built from its flags, on the device (read
            // by the LES lattices' collisions and by lbm_wall_distance)
            for (size_t l = 0; l < E->lat.size(); ++l) {
                if (!rebuilt[l]) continue;
                const HostLattice& H = E->lat[l];
                LBM_HIP(launch_wall_distance(E->d_flags + H.cell_off, E->d_wdist + H.cell_off, H.d.nx, H.d.ny,
                                             H.pitch, E->stream));
            }
            if (slots != E->total_slots) {
                E->total_slots = slots;
                const size_t dummy_n = static_cast<size_t>(slots) * kAccCopies * 2;
                if ((rc = dev_alloc(E, &E->d_dummy_acc, dummy_n)) != LBM_OK) return rc;
                if ((rc = dev_alloc(E, &E->d_hist, static_cast<size_t>(slots) * E->R * 2)) != LBM_OK) return rc;
                if ((rc = dev_alloc(E, &E->d_slot_lat, slots)) != LBM_OK) return rc;
                LBM_HIP(hipMemsetAsync(E->d_hist, 0, static_cast<size_t>(slots) * E->R * 16, E->stream));
                E->acc_blocks = 0;  // reallocated below
                E->graph_dirty = true;
            }
            std::vector<int32_t> slot_lat;
            for (size_t l = 0; l < E->lat.size(); ++l)
                for (int k = 0; k < E->lat[l].n_slots; ++k) slot_lat.push_back(static_cast<int32_t>(l));
            if ((rc = h2d(E, E->d_slot_lat, slot_lat.data(), slot_lat.size() * 4)) != LBM_OK) return rc;
        }
        E->geom_dirty = false;
        E->desc_dirty = true;
    }
    // accumulator blocks: one per launch of the longest batch (a graph), all zero
    // between batches (flush_kernel clears what it reads)
    const int want_blocks = std::max(E->graph_steps, kMinBatch);
    if (E->acc_blocks != want_blocks) {
        const size_t acc_n = static_cast<size_t>(want_blocks) * kAccCopies * E->total_slots * 2;
        if ((rc = dev_alloc(E, &E->d_acc, acc_n)) != LBM_OK) return rc;
        LBM_HIP(hipMemsetAsync(E->d_acc, 0, acc_n * 8, E->stream));
        if ((rc = dev_alloc(E, &E->d_nf, static_cast<size_t>(want_blocks) * E->n_recs * 2)) != LBM_OK) return rc;
        E->acc_blocks = want_blocks;
        E->graph_dirty = true;
    }
    if (E->desc_dirty) {
        E->host_lat.clear();
        for (size_t l = 0; l < E->lat.size(); ++l)
            E->host_lat.push_back(make_dev(E->lat[l], static_cast<int>(l), E->scale));
        if ((rc = h2d(E, E->d_lat, E->host_lat.data(), E->host_lat.size() * sizeof(LatticeDev))) != LBM_OK)
            return rc;
        E->desc_dirty = false;
        E->graph_dirty = true;  // one-lattice graphs hold the descriptor in their arguments
    }
    if (E->tiles_dirty && E->lat.size() > 1) {
        // one-step launches: the tiles of every lattice (general-path nodes in place: list
        // blocks here raised the kernel's registers from 93 to 136 and slowed the
        // 256-lattice ensemble from 0.58 to 0.74 ms per step)
        int s1[4];
        one_step_shape(E, s1);
        const int rows = s1[1], width = s1[0] * s1[2];
        std::vector<Tile> tiles;
        for (size_t l = 0; l < E->lat.size(); ++l) {
            const lbm_desc& d = E->lat[l].d;
            for (int y0 = 0; y0 < d.ny; y0 += rows)
                for (int x0 = 0; x0 < d.nx; x0 += width) tiles.push_back(Tile{static_cast<int32_t>(l), y0, x0, 0});
        }
        if (tiles.size() > E->tiles_cap) {
            if ((rc = dev_alloc(E, &E->d_tiles, tiles.size())) != LBM_OK) return rc;
            E->tiles_cap = tiles.size();
        }
        if ((rc = h2d(E, E->d_tiles, tiles.data(), tiles.size() * sizeof(Tile))) != LBM_OK) return rc;
        E->n_tiles = static_cast<int>(tiles.size());
        // two-step launches: 64 x rows tiles per lattice; rows of a lattice with an outlet
        // and nx % 64 == 1 are right-aligned (the first tile starts at x = -63), so that the
        // outflow copy's source column lies inside the last tile
        std::vector<Tile> t2;
        if (E->options & LAUNCH_TWO_STEPS) {
            for (size_t l = 0; l < E->lat.size(); ++l) {
                const lbm_desc& d = E->lat[l].d;
                const int xs = two_step_x_start(d);
                for (int y0 = 0; y0 < d.ny; y0 += E->rows)
                    for (int x0 = xs; x0 < d.nx; x0 += 64) t2.push_back(Tile{static_cast<int32_t>(l), y0, x0, 0});
            }
            if (t2.size() > E->tiles2_cap) {
                if ((rc = dev_alloc(E, &E->d_tiles2, t2.size())) != LBM_OK) return rc;
                E->tiles2_cap = t2.size();
            }
            if ((rc = h2d(E, E->d_tiles2, t2.data(), t2.size() * sizeof(Tile))) != LBM_OK) return rc;
        }
        E->n_tiles2 = static_cast<int>(t2.size());
        E->tiles_dirty = false;
        E->graph_dirty = true;
    }
    if ((E->options & LAUNCH_TWO_STEPS) && bulk_eligible(E, E->rows) && E->plan[plan_slot(E->rows)].rows != E->rows) {
        if ((rc = build_plan(E, E->rows)) != LBM_OK) return rc;
        E->graph_dirty = true;
    }
    if (E->counters_stale) {
        std::vector<int64_t> st;
        std::vector<int32_t> pr;
        for (auto& H : E->lat) {
            st.push_back(H.steps);
            pr.push_back(H.par);
        }
        if ((rc = h2d(E, E->d_steps, st.data(), st.size() * 8)) != LBM_OK) return rc;
        if ((rc = h2d(E, E->d_par, pr.data(), pr.size() * 4)) != LBM_OK) return rc;
        E->counters_stale = false;
    }
    return LBM_OK;
}

StepArgs base_args(Ensemble* E)
{
    StepArgs a{};
    a.lat = E->d_lat;
    a.buf0 = E->buf[0];
    a.buf1 = E->buf[1];
    a.flags = E->d_flags;
    a.link_slot = E->d_lslot;
    a.links = E->d_links;
    a.wdist = E->d_wdist;
    a.par = E->d_par;
    a.acc = E->d_acc;
    a.total_slots = E->total_slots;
    a.nan_flags = E->d_nan;
    a.scale = E->scale;
    a.inv_scale = E->inv_scale;
    a.options = E->options;
    a.steps = E->d_steps;
    a.sr = E->sr;
    a.sr_seed = E->sr_seed;
    return a;
}

// Grid for one launch over one lattice (single >= 0, or an ensemble of one) or over
// every lattice; returns whether the one-lattice kernel applies.
bool grid_for(Ensemble* E, int single, StepArgs& a, int& blocks, int bx = 0, int rows = 0, int vec = 0)
{
    if (bx == 0) {
        bx = E->bx;
        rows = E->rows;
        vec = E->vec;
    }
    if (single >= 0 || E->lat.size() == 1) {
        const int l = single >= 0 ? single : 0;
        const lbm_desc& d = E->lat[l].d;
        const HostLattice& H = E->lat[l];
        a.L0 = E->host_lat[l];
        a.lattice0 = l;
        a.tiles = nullptr;
        a.tiles_x = (d.nx + bx * vec - 1) / (bx * vec);
        a.tiles_y = (d.ny + rows - 1) / rows;
        a.n_links = static_cast<int32_t>(H.recs.size());
        a.n_edges = edge_nodes(d.nx, d.ny);
        a.special = E->d_special + H.special_off;
        a.n_special = static_cast<int32_t>(H.special.size());
        a.special_blocks = list_blocks(bx, a.n_links, a.n_edges) + (a.n_special + bx - 1) / bx;
        blocks = a.special_blocks + a.tiles_x * a.tiles_y;
        return true;
    }
    a.tiles = E->d_tiles;
    a.n_tiles = E->n_tiles;
    blocks = E->n_tiles;
    return false;
}

// Copies the flushed history rows [read_pos, steps) of lattice l to the host.
int drain(Ensemble* E, int l)
{
    HostLattice& H = E->lat[l];
    if (H.steps <= H.read_pos) return LBM_OK;
    LBM_HIP(hipStreamSynchronize(E->stream));
    std::vector<double> ring(static_cast<size_t>(H.n_slots) * E->R * 2);
    int rc = d2h(E, ring.data(), E->d_hist + static_cast<int64_t>(H.slot_base) * E->R * 2, ring.size() * 8);
    if (rc != LBM_OK) return rc;
    for (int64_t k = H.read_pos; k < H.steps; ++k)
        for (int s = 0; s < H.n_slots; ++s) {
            const int64_t o = (static_cast<int64_t>(s) * E->R + k % E->R) * 2;
            H.pend_fx.push_back(ring[o]);
            H.pend_fy.push_back(ring[o + 1]);
        }
    H.read_pos = H.steps;
    return LBM_OK;
}

// A batch of n steps writes history rows steps .. steps+n-1 (mod R): drain first when
// it would overwrite rows not yet read.
int ensure_history_room(Ensemble* E, int l, int64_t n)
{
    HostLattice& H = E->lat[l];
    if (H.steps + n - H.read_pos > E->R) return drain(E, l);
    return LBM_OK;
}

bool one_lattice(Ensemble* E, int l) { return l >= 0 || E->lat.size() == 1; }

// Two-step launches in a batch of n steps (LAUNCH_TWO_STEPS: a one-lattice handle, or
// every lattice of an ensemble; one lattice of an ensemble alone steps one step per
// launch); an odd batch ends with a one-step launch.  A batch that ends with a two-step
// launch leaves C_{k-1} unstored (its intermediate state stays in LDS): prev_pending, and
// complete_prev recomputes it before anything reads it.
int two_step_pairs(Ensemble* E, int n, int l)
{
    if (E->thermal || !(E->options & LAUNCH_TWO_STEPS) || n < 2) return 0;
    if (l >= 0 && E->lat.size() > 1) return 0;
    return n / 2;
}

// Whether a batch of n steps ends with a two-step launch.
bool ends_on_pair(Ensemble* E, int n, int l)
{
    const int p = two_step_pairs(E, n, l);
    return p > 0 && 2 * p == n;
}

// Whether two steps per launch apply: tile 64 x (4, 8 or 16) with 256 threads (block_x 64,
// one node per thread), or 64 x 32 where the tiles split into bulk and other tiles; on a
// one-lattice handle the outflow copy's source column (nx-2 and its neighbours) inside the
// last tile's ring (ensembles right-align such rows).
int two_step_ok(Ensemble* E, int bx, int rows, int vec)
{
    if (bx != 64 || vec != 1 || (rows != 4 && rows != 8 && rows != 16 && rows != 32))
        return fail(LBM_ERR_INVALID, "two steps per launch take shapes (64, 4, 1), (64, 8, 1), (64, 16, 1) and "
                                     "(64, 32, 1)");
    if (rows == 32 && !bulk_eligible(E, 32))
        return fail(LBM_ERR_INVALID, "two steps per launch on 64 x 32 tiles: FP32 or a 16-bit form");
    for (const auto& H : E->lat) {
        const lbm_desc& d = H.d;
        if (E->lat.size() == 1 && d.outlet && d.nx % 64 == 1)
            return fail(LBM_ERR_INVALID, "two steps per launch: nx % 64 == 1 with an outlet");
        // the second step gathers from LDS; a free-slip edge's reflection reads the buffer
        if (d.y_edge == LBM_EDGE_FREE_SLIP) return fail(LBM_ERR_INVALID, "two steps per launch: free-slip y edges");
    }
    return LBM_OK;
}

// Buffer swaps of a batch: one per launch.
int batch_flips(Ensemble* E, int n, int l) { return n - two_step_pairs(E, n, l); }

// Thermal kernel arguments from the handle's parameters and node conditions.
ThermalArgs thermal_args(Ensemble* E, int mode, int from_state, int y0, int y1, double* out)
{
    ThermalArgs t{};
    t.code = E->has_rules ? E->d_tcode : nullptr;
    for (int k = 0; k < kThermalRules; ++k) t.rule_t[k] = E->rule_t[k];
    t.rule_rest = E->rule_rest;
    t.mode = mode;
    t.from_state = from_state;
    t.y0 = y0;
    t.y1 = y1;
    t.omega_k = E->omega_k;
    t.beta = E->beta;
    t.g_buoyancy = E->g_buoyancy;
    t.t_ref = E->t_ref;
    t.out = out;
    return t;
}

// n steps over lattice l (l < 0: every lattice), then the force flush and the counter
// update.  The force partials of launch i go to block i of acc_base (fallback
// accumulators, zero on entry and cleared again by the flush) and of nf_base (per-record
// terms): the handle's blocks (n <= acc_blocks) unless the caller brings others.
// One-lattice launches take their buffers from the host's parity; graph capture records
// them for that starting parity.
int launch_batch(Ensemble* E, int n, int l, double* acc_base = nullptr, double* nf_base = nullptr)
{
    StepArgs a = base_args(E);
    int blocks;
    const bool single = grid_for(E, l, a, blocks);
    const int li = l >= 0 ? l : 0;
    const int32_t par = E->lat[li].par;
    const int pairs = two_step_pairs(E, n, l);
    if (!acc_base) acc_base = E->d_acc;
    if (!nf_base) nf_base = E->d_nf;
    auto block_of = [&](int i) { return acc_base + static_cast<int64_t>(i) * kAccCopies * E->total_slots * 2; };
    auto terms_of = [&](int i) { return nf_base + static_cast<int64_t>(i) * E->n_recs * 2; };
    int swaps = 0;
    if (E->thermal) {
        // one launch of the thermal kernel per step (one lattice; f and g share the parity)
        for (int i = 0; i < n; ++i, ++swaps) {
            const int sp = par ^ (i & 1);
            a.src = E->buf[sp];
            a.dst = E->buf[1 - sp];
            a.acc = block_of(i);
            a.nf = terms_of(i);
            a.step_arg = i;
            ThermalArgs t = thermal_args(E, THERMAL_STEP, 0, 0, E->lat[0].d.ny, nullptr);
            t.gsrc = E->gbuf[sp];
            t.gdst = E->gbuf[1 - sp];
            LBM_HIP(launch_thermal(E->storage, a, t, E->stream));
        }
    } else {
        // two-step launches on 64 x rows tiles (one lattice: a grid of tiles; an ensemble:
        // its two-step tile table), then one-step launches
        if (pairs > 0) {
            StepArgs a2 = a;
            int blocks2;
            if (single) {
                const lbm_desc& d = E->lat[li].d;
                a2.tiles_x = (d.nx + 63) / 64;
                a2.tiles_y = (d.ny + E->rows - 1) / E->rows;
                blocks2 = a2.tiles_x * a2.tiles_y;
            } else {
                a2.tiles = E->d_tiles2;
                a2.n_tiles = E->n_tiles2;
                blocks2 = E->n_tiles2;
            }
            for (int i = 0; i < pairs; ++i, ++swaps) {
                if (single) {
                    const int sp = par ^ (swaps & 1);
                    a2.src = E->buf[sp];
                    a2.dst = E->buf[1 - sp];
                }
                a2.acc = block_of(2 * i);
                a2.acc2 = block_of(2 * i + 1);
                a2.nf = terms_of(2 * i);
                a2.nf2 = terms_of(2 * i + 1);
                a2.step_arg = 2 * i;
                a2.flips = swaps;
                LBM_HIP(launch_pair(E, a2, single, blocks2));
            }
        }
        int s1[4];
        one_step_shape(E, s1);
        StepArgs a1 = base_args(E);
        int blocks1;
        grid_for(E, l, a1, blocks1, s1[0], s1[1], s1[2]);
        a1.options = s1[3];
        for (int i = 2 * pairs; i < n; ++i, ++swaps) {
            if (single) {
                const int sp = par ^ (swaps & 1);
                a1.src = E->buf[sp];
                a1.dst = E->buf[1 - sp];
            }
            a1.acc = block_of(i);
            a1.nf = terms_of(i);
            a1.step_arg = i;
            a1.flips = swaps;
            LBM_HIP(launch_step(E->storage, s1[0], s1[1], s1[2], single, blocks1, a1, E->stream));
        }
    }
    FlushArgs f{};
    f.acc = acc_base;
    f.nf = nf_base;
    f.n_recs = E->n_recs;
    f.slot_start = E->d_slot_start;
    f.slot_recs = E->d_slot_recs;
    f.hist = E->d_hist;
    f.steps = E->d_steps;
    f.slot_lattice = E->d_slot_lat;
    f.slot_begin = l >= 0 ? E->lat[l].slot_base : 0;
    f.slot_count = l >= 0 ? E->lat[l].n_slots : E->total_slots;
    f.n = n;
    f.total_slots = E->total_slots;
    f.R = E->R;
    LBM_HIP(launch_flush(f, E->stream));
    const int begin = l >= 0 ? l : 0, count = l >= 0 ? 1 : static_cast<int>(E->lat.size());
    LBM_HIP(launch_add_steps(E->d_steps, E->d_par, begin, count, n, swaps, E->stream));
    return LBM_OK;
}

// Completes the state after a batch that ended with a two-step launch: C_{k-1}, which the
// two-step kernel kept in LDS, is recomputed from C_{k-2} (still in the buffer that is not
// current) by one one-step launch, from a copy of that buffer into it, with the random words
// of its step and dummy force and NaN accumulators -- the bits of the two-step launch's first
// step (every launch form rounds alike).  Downloads, macros, masses, uploads, parameter,
// geometry and rounding changes, single-lattice stepping and the autotuner call it first.
int complete_prev(Ensemble* E)
{
    if (!E->prev_pending) return LBM_OK;
    void* T = nullptr;
    const int64_t held = E->device_bytes;
    int rc = alloc_population(E, &T, "a copy of the previous population buffer");
    E->device_bytes = held;  // freed before returning
    if (rc != LBM_OK) return rc;
    int s1[4];
    one_step_shape(E, s1);
    StepArgs a = base_args(E);
    int blocks;
    const bool single = grid_for(E, -1, a, blocks, s1[0], s1[1], s1[2]);
    a.options = s1[3];
    a.acc = E->d_dummy_acc;
    a.nf = E->d_dummy_nf;
    a.nan_flags = E->d_dummy_nan;
    a.step_arg = -2;  // the device step counts are k + 1: step k - 1's keys
    hipError_t e = hipSuccess;
    if (single) {
        const int par = E->lat[0].par;
        e = hipMemcpyAsync(T, E->buf[1 - par], static_cast<size_t>(E->pool_elems) * E->bpp, hipMemcpyDeviceToDevice,
                           E->stream);
        a.src = T;
        a.dst = E->buf[1 - par];
        if (e == hipSuccess) e = launch_step(E->storage, s1[0], s1[1], s1[2], true, blocks, a, E->stream);
    } else {
        // per buffer parity p: the lattices of parity p read their C_{k-2} from T and write
        // C_{k-1} into buf[1 - p]; the others read buf[1 - p] and write into T (unused)
        for (int p = 0; p < 2 && e == hipSuccess; ++p) {
            bool any = false;
            for (size_t l = 0; l < E->lat.size() && e == hipSuccess; ++l) {
                const HostLattice& H = E->lat[l];
                if (H.par != p) continue;
                any = true;
                const size_t off = static_cast<size_t>(H.f_off) * E->bpp;
                const size_t bytes = static_cast<size_t>(9) * E->host_lat[l].plane * E->bpp;
                e = hipMemcpyAsync(static_cast<char*>(T) + off, static_cast<const char*>(E->buf[1 - p]) + off, bytes,
                                   hipMemcpyDeviceToDevice, E->stream);
            }
            if (!any || e != hipSuccess) continue;
            a.buf0 = p == 0 ? T : E->buf[0];
            a.buf1 = p == 1 ? T : E->buf[1];
            a.dst = nullptr;
            a.flips = 0;
            e = launch_step(E->storage, s1[0], s1[1], s1[2], false, blocks, a, E->stream);
        }
    }
    const hipError_t se = hipStreamSynchronize(E->stream);
    (void)hipFree(T);
    if (e != hipSuccess || se != hipSuccess)
        return fail(LBM_ERR_DEVICE, std::string("completing the state after two-step launches: ") +
                                        hipGetErrorString(e != hipSuccess ? e : se));
    E->prev_pending = false;
    return LBM_OK;
}

void drop_graphs(Ensemble* E)
{
    for (auto& g : E->graph)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
    E->graph_dirty = false;
}

// Captures a batch of graph_steps launches over every lattice into graph[p] (p: the
// starting parity of a one-lattice handle; 0 for ensembles, whose kernel reads it).
int build_graph(Ensemble* E, int p)
{
    hipGraph_t g;
    LBM_HIP(hipStreamBeginCapture(E->stream, hipStreamCaptureModeThreadLocal));
    int rc = launch_batch(E, E->graph_steps, -1);
    hipError_t e = hipStreamEndCapture(E->stream, &g);
    if (rc != LBM_OK) return rc;
    if (e != hipSuccess) return fail(LBM_ERR_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&E->graph[p], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        E->graph[p] = nullptr;
        return fail(LBM_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    return LBM_OK;
}

int require_uploaded(Ensemble* E)
{
    for (size_t l = 0; l < E->lat.size(); ++l)
        if (!E->lat[l].uploaded)
            return fail(LBM_ERR_INVALID, "lattice " + std::to_string(l) + " has no state (call lbm_upload)");
    return LBM_OK;
}

// Runs the state kernel over lattice l in row chunks; out may be null (mass only).
// Runs the state kernel over lattice l in row chunks: host_out receives the state
// (AoS FP64) when non-null, mass_out its FP64 mass; recollide rewrites the stored
// post-collision buffer from the state with the current descriptor.
int materialize(Ensemble* E, int l, double* host_out, double* mass_out, bool recollide = false)
{
    HostLattice& H = E->lat[l];
    const lbm_desc& d = H.d;
    int rc;
    if ((rc = commit(E)) != LBM_OK) return rc;
    const double zero = 0.0;
    if ((rc = h2d(E, E->d_mass, &zero, 8)) != LBM_OK) return rc;
    const int64_t row_bytes = static_cast<int64_t>(d.nx) * 9 * 8;
    const int chunk = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(d.ny, kStagingBytes / row_bytes)));
    const bool from_state = H.steps == 0;
    // state = buf[1-par] after upload; C_{k-1} = buf[1-par] after k >= 1 steps
    const void* src = E->buf[1 - H.par];
    for (int y0 = 0; y0 < d.ny; y0 += chunk) {
        const int y1 = std::min(d.ny, y0 + chunk);
        StateArgs s{};
        s.lat = E->d_lat;
        s.lattice = l;
        s.y0 = y0;
        s.y1 = y1;
        s.from_state = from_state ? 1 : 0;
        s.src = src;
        s.flags = E->d_flags;
        s.link_slot = E->d_lslot;
        s.links = E->d_links;
        s.out = host_out ? E->d_stage : nullptr;
        s.mass_dev = mass_out ? E->d_mass : nullptr;
        s.coll_dst = recollide ? E->buf[H.par] : nullptr;
        s.wdist = E->d_wdist;
        s.scale = E->scale;
        s.inv_scale = E->inv_scale;
        LBM_HIP(launch_state(E->storage, s, d.nx, E->stream));
        if (host_out) {
            rc = d2h(E, host_out + static_cast<int64_t>(y0) * d.nx * 9, E->d_stage,
                     static_cast<size_t>(y1 - y0) * row_bytes);
            if (rc != LBM_OK) return rc;
        }
    }
    if (mass_out) {
        double m;
        if ((rc = d2h(E, &m, E->d_mass, 8)) != LBM_OK) return rc;
        *mass_out = static_cast<double>(H.fluid) + m;
    }
    return LBM_OK;
}

// Density and velocity of rows [y0, y1) of the current state, each as nx*(y1-y0) FP64.
int fields(Ensemble* E, int l, int y0, int y1, double* rho, double* u, double* v)
{
    HostLattice& H = E->lat[l];
    const lbm_desc& d = H.d;
    int rc;
    if ((rc = commit(E)) != LBM_OK) return rc;
    const int64_t row_bytes = static_cast<int64_t>(d.nx) * 3 * 8;
    const int chunk = static_cast<int>(std::min<int64_t>(y1 - y0, E->stage_bytes / row_bytes));
    if (chunk < 1) return fail(LBM_ERR_INVALID, "staging buffer smaller than one row of fields");
    const void* src = E->buf[1 - H.par];
    for (int c0 = y0; c0 < y1; c0 += chunk) {
        const int c1 = std::min(y1, c0 + chunk);
        StateArgs s{};
        s.lat = E->d_lat;
        s.lattice = l;
        s.y0 = c0;
        s.y1 = c1;
        s.from_state = H.steps == 0 ? 1 : 0;
        s.src = src;
        s.flags = E->d_flags;
        s.link_slot = E->d_lslot;
        s.links = E->d_links;
        s.macros = E->d_stage;
        s.wdist = E->d_wdist;
        s.scale = E->scale;
        s.inv_scale = E->inv_scale;
        LBM_HIP(launch_state(E->storage, s, d.nx, E->stream));
        const int64_t n = static_cast<int64_t>(c1 - c0) * d.nx;
        const int64_t off = static_cast<int64_t>(c0 - y0) * d.nx;
        double* dst[3] = {rho, u, v};
        for (int k = 0; k < 3; ++k)
            if (dst[k] && (rc = d2h(E, dst[k] + off, E->d_stage + k * n, n * 8)) != LBM_OK) return rc;
    }
    return LBM_OK;
}

// Thermal state rows of lattice 0 from AoS FP64 g (null: 1.0 in every slot) into the
// state planes gbuf[1 - par].
int thermal_load(Ensemble* E, const double* g)
{
    HostLattice& H = E->lat[0];
    const lbm_desc& d = H.d;
    const int64_t row_bytes = static_cast<int64_t>(d.nx) * 72;
    const int chunk = g ? static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(d.ny, E->stage_bytes / row_bytes)))
                        : d.ny;
    for (int y0 = 0; y0 < d.ny; y0 += chunk) {
        const int y1 = std::min(d.ny, y0 + chunk);
        int rc;
        if (g && (rc = h2d(E, E->d_stage, g + static_cast<int64_t>(y0) * d.nx * 9, (y1 - y0) * row_bytes)) != LBM_OK)
            return rc;
        ThermalLoadArgs t{};
        t.lat = E->d_lat;
        t.lattice = 0;
        t.y0 = y0;
        t.y1 = y1;
        t.in = g ? E->d_stage : nullptr;
        t.state_dst = E->gbuf[1 - H.par];
        LBM_HIP(launch_thermal_load(E->storage, t, d.nx, E->stream));
    }
    LBM_HIP(hipStreamSynchronize(E->stream));
    return LBM_OK;
}

// Rewrites the stored post-collision f and g of lattice 0 from its current state (after
// an upload, or a change of the parameters or node conditions): the pull form keeps
// the next step's collisions, done with the old values.
int thermal_recollide(Ensemble* E)
{
    int rc;
    if ((rc = commit(E)) != LBM_OK) return rc;
    HostLattice& H = E->lat[0];
    StepArgs a = base_args(E);
    int blocks;
    grid_for(E, 0, a, blocks);
    a.src = E->buf[1 - H.par];
    a.dst = E->buf[H.par];
    a.nf = nullptr;
    ThermalArgs t = thermal_args(E, THERMAL_COLLIDE, H.steps == 0 ? 1 : 0, 0, H.d.ny, nullptr);
    t.gsrc = E->gbuf[1 - H.par];
    t.gdst = E->gbuf[H.par];
    LBM_HIP(launch_thermal(E->storage, a, t, E->stream));
    LBM_HIP(hipStreamSynchronize(E->stream));
    return LBM_OK;
}

// Rows [y0, y1) of lattice 0's thermal state: AoS FP64 g (mode THERMAL_OUT_G) or T
// (THERMAL_OUT_T) into out, in chunks through the staging buffer.
int thermal_out(Ensemble* E, int mode, int y0, int y1, double* out)
{
    int rc;
    if ((rc = commit(E)) != LBM_OK) return rc;
    HostLattice& H = E->lat[0];
    const int per_node = mode == THERMAL_OUT_G ? 9 : 1;
    const int64_t row_bytes = static_cast<int64_t>(H.d.nx) * per_node * 8;
    const int chunk = static_cast<int>(std::min<int64_t>(y1 - y0, E->stage_bytes / row_bytes));
    if (chunk < 1) return fail(LBM_ERR_INVALID, "staging buffer smaller than one row");
    for (int c0 = y0; c0 < y1; c0 += chunk) {
        const int c1 = std::min(y1, c0 + chunk);
        StepArgs a = base_args(E);
        int blocks;
        grid_for(E, 0, a, blocks);
        a.src = E->buf[1 - H.par];
        a.nf = nullptr;
        ThermalArgs t = thermal_args(E, mode, H.steps == 0 ? 1 : 0, c0, c1, E->d_stage);
        t.gsrc = E->gbuf[1 - H.par];
        LBM_HIP(launch_thermal(E->storage, a, t, E->stream));
        rc = d2h(E, out + static_cast<int64_t>(c0 - y0) * H.d.nx * per_node, E->d_stage,
                 static_cast<size_t>(c1 - c0) * row_bytes);
        if (rc != LBM_OK) return rc;
    }
    return LBM_OK;
}

int thermal_ok(Ensemble* E, int l)
{
    int rc = lattice_ok(E, l);
    if (rc) return rc;
    if (!E->thermal) return fail(LBM_ERR_INVALID, "the thermal populations are off (lbm_set_thermal)");
    return LBM_OK;
}

}  // namespace
}  // namespace lbm

// Device scratch of lbm_step_noalloc: the force partials of one call's launches
// (fallback accumulators of every launch, then the per-record terms of every launch).
// Each GPU holds its own copy for the life of the process.  In host code the name
// stands for no device address: host code uses only what scratch_on_device returns.
static __device__ __align__(256) double scratch[LBM_NOALLOC_SCRATCH_BYTES / sizeof(double)];

namespace lbm {
namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_dev[kMaxDevices] = {};

// Address of the current device's scratch, looked up once per device (with that
// device current) by hipGetSymbolAddress; each device's copy has its own address.
int scratch_on_device(double** out)
{
    int dev = -1;
    LBM_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(LBM_ERR_DEVICE, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        LBM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(LBM_ERR_DEVICE, "hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_dev[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_dev[dev];
    return LBM_OK;
}

}  // namespace
}  // namespace lbm

using lbm::Ensemble;
using lbm::fail;

struct lbm_ensemble : Ensemble {};

extern "C" {

const char* lbm_last_error(void) { return lbm::g_error.c_str(); }

int lbm_desc_for_case(int case_type, int32_t nx, int32_t ny, lbm_desc* out)
{
    if (!out) return fail(LBM_ERR_INVALID, "null desc");
    if (case_type < 0 || case_type > 13) return fail(LBM_ERR_INVALID, "unknown case type");
    lbm_desc d{};
    d.nx = nx;
    d.ny = ny;
    d.collision = LBM_MRT;
    d.les = 0;
    d.cs = 0.12;  // g_cs default (src/lbm_types.hpp:32)
    d.tau = 1.0;
    d.u_in = 0.1;
    enum { CYLINDER = 0, CAVITY = 1, STEP = 2, RIBS = 3, PERIODIC_HILLS = 9 };
    if (case_type == CAVITY) {
        // src/lbm.hpp:408-433 stream, :531-532 lid, :543 no forces
        d.x_edge = d.y_edge = LBM_EDGE_WALL;
        d.lid = 1;
    } else if (case_type == RIBS || case_type == PERIODIC_HILLS) {
        // src/lbm.hpp:434-456 fully periodic, no open boundaries
        d.x_edge = d.y_edge = LBM_EDGE_PERIODIC;
        d.forces = 1;
    } else {
        // src/lbm.hpp:457-484 periodic y, open x; inlet + outflow (:533-539)
        d.x_edge = LBM_EDGE_OPEN;
        d.y_edge = LBM_EDGE_PERIODIC;
        d.outlet = 1;
        d.forces = 1;
        if (case_type == STEP) {
            d.inlet = LBM_INLET_PARABOLIC;
            d.inlet_y0 = ny / 3;
            d.inlet_h = ny - 1 - ny / 3;
        } else {
            d.inlet = LBM_INLET_ZOU_HE;
        }
    }
    *out = d;
    return LBM_OK;
}

int lbm_create(const lbm_desc* descs, int32_t n, int32_t storage, int32_t scale_exp, void* stream,
               lbm_handle* out)
{
    using namespace lbm;
    if (!descs || n <= 0 || !out) return fail(LBM_ERR_INVALID, "lbm_create: bad arguments");
    if (storage < ST_F64 || storage > ST_BF16S) return fail(LBM_ERR_INVALID, "unknown storage form");
    if ((storage == ST_F16S || storage == ST_BF16S) && (scale_exp < 0 || scale_exp > 14))
        return fail(LBM_ERR_INVALID, "scale exponent must be in [0,14]");
    for (int l = 0; l < n; ++l) {
        const lbm_desc& d = descs[l];
        if (d.nx < 3 || d.ny < 3) return fail(LBM_ERR_INVALID, "lattice must be at least 3x3");
        // node indices y * pitch + x (pitch: nx rounded up to 64) are int32
        if (static_cast<int64_t>((d.nx + 63) / 64 * 64) * d.ny > (1ll << 31))
            return fail(LBM_ERR_INVALID, "lattice too large: its rows padded to 64 nodes exceed 2^31 nodes");
        if (d.collision < 0 || d.collision > 1 || d.x_edge < 0 || d.x_edge > 2 || d.y_edge < 0 || d.y_edge > 3 ||
            d.inlet < 0 || d.inlet > 3)
            return fail(LBM_ERR_INVALID, "bad enum in lbm_desc");
        if (d.tau <= 0.0) return fail(LBM_ERR_INVALID, "tau must be positive");
        if (d.y_edge == LBM_EDGE_FREE_SLIP && d.x_edge != LBM_EDGE_OPEN)
            return fail(LBM_ERR_INVALID, "free-slip y edges need an open x edge");
        if (d.inlet == LBM_INLET_PARABOLIC && d.inlet_h <= 0)
            return fail(LBM_ERR_INVALID, "parabolic inlet needs inlet_h > 0");
    }
    auto* E = new lbm_ensemble();
    E->storage = storage;
    E->bpp = storage == ST_F64 ? 8 : (storage == ST_F32 ? 4 : 2);
    E->scale_exp = (storage == ST_F16S || storage == ST_BF16S) ? scale_exp : 0;
    E->scale = std::ldexp(1.0f, E->scale_exp);
    E->inv_scale = std::ldexp(1.0f, -E->scale_exp);
    if (stream) {
        E->stream = static_cast<hipStream_t>(stream);
    } else {
        if (hipStreamCreate(&E->stream) != hipSuccess) {
            delete E;
            return fail(LBM_ERR_DEVICE, "hipStreamCreate failed");
        }
        E->own_stream = true;
    }
    const lbm::PoolLayout P = lbm::pool_layout(descs, n);
    int64_t cells = 0;
    E->lat.resize(n);
    for (int l = 0; l < n; ++l) {
        HostLattice& H = E->lat[l];
        H.d = descs[l];
        H.pitch = (descs[l].nx + 63) / 64 * 64;
        H.f_off = P.f_off[l];
        const int64_t plane = static_cast<int64_t>(descs[l].ny) * H.pitch;
        H.cell_off = cells;
        cells += (plane + 255) / 256 * 256;
    }
    E->pool_elems = P.pool;
    E->cell_elems = cells;
    E->guard_elems = P.guard;
    int rc = LBM_OK;
    for (int l = 0; l < n && rc == LBM_OK; ++l) {
        const HostLattice& H = E->lat[l];
        const int64_t furthest = H.f_off + 9 * static_cast<int64_t>(H.d.ny) * H.pitch + H.d.nx - 2;
        if (furthest >= P.elems()) rc = fail(LBM_ERR_INVALID, "population layout: pulls reach past the guard");
    }
    for (int b = 0; b < 2 && rc == LBM_OK; ++b) rc = lbm::alloc_population(E, &E->buf[b], "population buffers");
    if (rc == LBM_OK) rc = dev_alloc(E, &E->d_flags, cells);
    if (rc == LBM_OK) rc = dev_alloc(E, &E->d_lslot, cells);
    // no link record until a geometry build writes a lattice's slots
    if (rc == LBM_OK && hipMemset(E->d_lslot, 0xff, static_cast<size_t>(cells) * 4) != hipSuccess)
        rc = fail(LBM_ERR_DEVICE, "link slot pool");
    if (rc == LBM_OK) rc = dev_alloc(E, &E->d_wdist, cells);
    if (rc == LBM_OK) rc = dev_alloc(E, &E->d_lat, n);
    if (rc == LBM_OK) rc = dev_alloc(E, &E->d_steps, n);
    if (rc == LBM_OK) rc = dev_alloc(E, &E->d_par, n);
    if (rc == LBM_OK) rc = dev_alloc(E, &E->d_nan, n);
    if (rc == LBM_OK) rc = dev_alloc(E, &E->d_dummy_nan, n);
    if (rc == LBM_OK) rc = dev_alloc(E, &E->d_mass, 1);
    // staging for row chunks of AoS FP64 state: at most kStagingBytes or one row
    int64_t stage = 0;
    for (auto& H : E->lat) {
        const int64_t row = static_cast<int64_t>(H.d.nx) * 72;
        const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(H.d.ny, kStagingBytes / row));
        stage = std::max<int64_t>(stage, chunk * row);
    }
    if (rc == LBM_OK) rc = dev_alloc(E, &E->d_stage, static_cast<size_t>(stage / 8 + 1));
    E->stage_bytes = rc == LBM_OK ? stage : 0;
    if (rc == LBM_OK) {
        std::vector<int32_t> z(n, 0);
        std::vector<int64_t> z64(n, 0);
        rc = lbm::h2d(E, E->d_nan, z.data(), n * 4);
        if (rc == LBM_OK) rc = lbm::h2d(E, E->d_dummy_nan, z.data(), n * 4);
        if (rc == LBM_OK) rc = lbm::h2d(E, E->d_par, z.data(), n * 4);
        if (rc == LBM_OK) rc = lbm::h2d(E, E->d_steps, z64.data(), n * 8);
    }
    if (rc != LBM_OK) {
        std::string msg = lbm::g_error;
        lbm_destroy(E);
        return fail(rc, msg);
    }
    *out = E;
    return LBM_OK;
}

int lbm_population_layout(const lbm_desc* descs, int32_t n, int64_t* f_off, int64_t* pool_elems,
                          int64_t* guard_elems)
{
    if (!descs || n <= 0) return fail(LBM_ERR_INVALID, "lbm_population_layout: bad arguments");
    for (int l = 0; l < n; ++l)
        if (descs[l].nx < 3 || descs[l].ny < 3) return fail(LBM_ERR_INVALID, "lattice must be at least 3x3");
    const lbm::PoolLayout P = lbm::pool_layout(descs, n);
    if (f_off)
        for (int l = 0; l < n; ++l) f_off[l] = P.f_off[l];
    if (pool_elems) *pool_elems = P.pool;
    if (guard_elems) *guard_elems = P.guard;
    return LBM_OK;
}

int lbm_destroy(lbm_handle h)
{
    if (!h) return LBM_OK;
    Ensemble* E = h;
    if (E->stream) (void)hipStreamSynchronize(E->stream);
    lbm::drop_graphs(E);
    void* ptrs[] = {E->buf[0], E->buf[1], E->d_flags, E->d_lslot, E->d_wdist, E->d_links, E->d_lat,
                    E->d_tiles, E->d_tiles2, E->d_steps, E->d_par, E->d_acc, E->d_hist, E->d_dummy_acc, E->d_nan,
                    E->d_dummy_nan, E->d_mass, E->d_stage, E->d_slot_lat, E->d_special, E->d_nf,
                    E->d_dummy_nf, E->d_slot_start, E->d_slot_recs, E->gbuf[0], E->gbuf[1], E->d_tcode};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    lbm::drop_plans(E);
    for (auto& H : E->lat) lbm::drop_new_geometry(H);
    for (hipStream_t s : {E->side, E->side2})
        if (s) {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
    for (hipEvent_t ev : {E->fork_ev, E->join_ev, E->join2_ev})
        if (ev) (void)hipEventDestroy(ev);
    if (E->own_stream) (void)hipStreamDestroy(E->stream);
    delete h;
    return LBM_OK;
}

int lbm_set_obstacle(lbm_handle h, int32_t l, const uint8_t* mask)
{
    if (h && h->prev_pending) {
        const int cp = lbm::complete_prev(h);
        if (cp != LBM_OK) return cp;
    }
    int rc = lbm::lattice_ok(h, l);
    if (rc) return rc;
    if (!mask) return fail(LBM_ERR_INVALID, "null mask");
    auto& H = h->lat[l];
    const size_t n = static_cast<size_t>(H.d.nx) * H.d.ny;
    H.obstacle.resize(n);
    for (size_t c = 0; c < n; ++c) H.obstacle[c] = mask[c] ? 1 : 0;
    H.bb_kind = 0;
    H.moving = false;
    H.frad = false;
    H.explicit_links.clear();
    H.geom_dirty = true;
    H.uploaded = false;  // the stored state belongs to the old geometry
    h->geom_dirty = true;
    return LBM_OK;
}

int lbm_place_obstacles(lbm_handle h, int32_t l, const double* circles, int32_t n_circles, const double* poly_xy,
                        const int32_t* poly_n, int32_t n_polys)
{
    if (h && h->prev_pending) {
        const int cp = lbm::complete_prev(h);
        if (cp != LBM_OK) return cp;
    }
    using namespace lbm;
    int rc = lattice_ok(h, l);
    if (rc) return rc;
    if (n_circles < 0 || n_polys < 0) return fail(LBM_ERR_INVALID, "negative obstacle count");
    if ((n_circles && !circles) || (n_polys && (!poly_xy || !poly_n))) return fail(LBM_ERR_INVALID, "null obstacle array");
    int64_t n_vert = 0;
    for (int k = 0; k < n_polys; ++k) {
        if (poly_n[k] < 1) return fail(LBM_ERR_INVALID, "a polygon needs at least one vertex");
        n_vert += poly_n[k];
    }
    for (int64_t k = 0; k < 2 * n_vert; ++k)
        if (!std::isfinite(poly_xy[k])) return fail(LBM_ERR_INVALID, "non-finite polygon vertex");
    for (int k = 0; k < 3 * n_circles; ++k)
        if (!std::isfinite(circles[k])) return fail(LBM_ERR_INVALID, "non-finite circle");
    Ensemble* E = h;
    HostLattice& H = E->lat[l];
    const int nx = H.d.nx, ny = H.d.ny;
    const size_t n = static_cast<size_t>(nx) * ny;
    if (H.obstacle.size() != n) H.obstacle.assign(n, 0);
    uint8_t* d_mask = nullptr;
    double* d_xy = nullptr;
    if (hipMalloc(&d_mask, n) != hipSuccess) return fail(LBM_ERR_NOMEM, "obstacle mask staging");
    if (n_vert && hipMalloc(&d_xy, static_cast<size_t>(n_vert) * 16) != hipSuccess) {
        (void)hipFree(d_mask);
        return fail(LBM_ERR_NOMEM, "polygon staging");
    }
    rc = h2d(E, d_mask, H.obstacle.data(), n);
    if (rc == LBM_OK && n_vert) rc = h2d(E, d_xy, poly_xy, static_cast<size_t>(n_vert) * 16);
    // each obstacle over its bounding box, widened by a cell against rounding and clipped
    auto clip = [](double lo, double hi, int size, int32_t& a, int32_t& b) {
        lo = std::floor(lo) - 1.0;
        hi = std::floor(hi) + 2.0;
        a = static_cast<int32_t>(std::max(0.0, std::min(lo, static_cast<double>(size))));
        b = static_cast<int32_t>(std::max(0.0, std::min(hi, static_cast<double>(size))));
    };
    hipError_t e = hipSuccess;
    for (int k = 0; k < n_circles && rc == LBM_OK && e == hipSuccess; ++k) {
        PlaceArgs a{};
        a.mask = d_mask;
        a.nx = nx;
        a.kind = 0;
        a.cx = circles[3 * k];
        a.cy = circles[3 * k + 1];
        a.r = circles[3 * k + 2];
        const double ra = std::fabs(a.r);
        clip(a.cx - ra, a.cx + ra, nx, a.x0, a.x1);
        clip(a.cy - ra, a.cy + ra, ny, a.y0, a.y1);
        e = launch_place(a, E->stream);
    }
    int64_t v0 = 0;
    for (int k = 0; k < n_polys && rc == LBM_OK && e == hipSuccess; ++k) {
        PlaceArgs a{};
        a.mask = d_mask;
        a.nx = nx;
        a.kind = 1;
        a.xy = d_xy + 2 * v0;
        a.n = poly_n[k];
        double x_lo = poly_xy[2 * v0], x_hi = x_lo, y_lo = poly_xy[2 * v0 + 1], y_hi = y_lo;
        for (int v = 1; v < poly_n[k]; ++v) {
            x_lo = std::min(x_lo, poly_xy[2 * (v0 + v)]);
            x_hi = std::max(x_hi, poly_xy[2 * (v0 + v)]);
            y_lo = std::min(y_lo, poly_xy[2 * (v0 + v) + 1]);
            y_hi = std::max(y_hi, poly_xy[2 * (v0 + v) + 1]);
        }
        clip(x_lo, x_hi, nx, a.x0, a.x1);
        clip(y_lo, y_hi, ny, a.y0, a.y1);
        e = launch_place(a, E->stream);
        v0 += poly_n[k];
    }
    if (rc == LBM_OK && e == hipSuccess) e = hipStreamSynchronize(E->stream);
    if (rc == LBM_OK && e != hipSuccess) rc = fail(LBM_ERR_DEVICE, std::string("rasterisation: ") + hipGetErrorString(e));
    if (rc == LBM_OK) rc = d2h(E, H.obstacle.data(), d_mask, n);
    (void)hipFree(d_mask);
    if (d_xy) (void)hipFree(d_xy);
    if (rc != LBM_OK) return rc;
    // the mask changed: as lbm_set_obstacle
    H.bb_kind = 0;
    H.moving = false;
    H.frad = false;
    H.explicit_links.clear();
    H.geom_dirty = true;
    H.uploaded = false;
    E->geom_dirty = true;
    return LBM_OK;
}

int lbm_get_obstacle(lbm_handle h, int32_t l, uint8_t* mask)
{
    int rc = lbm::lattice_ok(h, l);
    if (rc) return rc;
    if (!mask) return fail(LBM_ERR_INVALID, "null mask");
    const auto& H = h->lat[l];
    const size_t n = static_cast<size_t>(H.d.nx) * H.d.ny;
    for (size_t c = 0; c < n; ++c) mask[c] = c < H.obstacle.size() ? H.obstacle[c] : 0;
    return LBM_OK;
}

int lbm_set_storage_shift(lbm_handle h, int32_t l, double u, double v)
{
    if (h && h->prev_pending) {
        const int cp = lbm::complete_prev(h);
        if (cp != LBM_OK) return cp;
    }
    int rc = lbm::lattice_ok(h, l);
    if (rc) return rc;
    if (h->storage != lbm::ST_F16S && h->storage != lbm::ST_BF16S)
        return fail(LBM_ERR_INVALID, "the storage shift applies to the 16-bit storage forms");
    if (!(std::fabs(u) < 0.5 && std::fabs(v) < 0.5)) return fail(LBM_ERR_INVALID, "shift velocity out of range");
    auto& H = h->lat[l];
    H.shift_u = u;
    H.shift_v = v;
    h->desc_dirty = true;
    H.uploaded = false;  // the stored state is encoded against the old shift
    return LBM_OK;
}

int lbm_set_stochastic_rounding(lbm_handle h, int32_t on, uint32_t seed)
{
    if (h && h->prev_pending) {
        const int cp = lbm::complete_prev(h);
        if (cp != LBM_OK) return cp;
    }
    if (!h) return fail(LBM_ERR_INVALID, "null handle");
    if (h->storage != lbm::ST_F16S && h->storage != lbm::ST_BF16S)
        return fail(LBM_ERR_INVALID, "stochastic rounding applies to the 16-bit storage forms");
    h->sr = on ? 1 : 0;
    h->sr_seed = seed;
    h->graph_dirty = true;
    return LBM_OK;
}

int lbm_set_bb_circle(lbm_handle h, int32_t l, double cx, double cy, double r, int32_t scheme)
{
    if (h && h->prev_pending) {
        const int cp = lbm::complete_prev(h);
        if (cp != LBM_OK) return cp;
    }
    int rc = lbm::lattice_ok(h, l);
    if (rc) return rc;
    if (scheme < 0 || scheme > 2) return fail(LBM_ERR_INVALID, "bad scheme");
    auto& H = h->lat[l];
    H.bb_kind = 1;
    H.bcx = cx;
    H.bcy = cy;
    H.br = r;
    H.bb_scheme = scheme;
    H.geom_dirty = h->geom_dirty = true;
    H.uploaded = false;  // the stored state belongs to the old geometry
    return LBM_OK;
}

int lbm_set_bb_polygon(lbm_handle h, int32_t l, const double* xy, int32_t n, int32_t scheme)
{
    if (h && h->prev_pending) {
        const int cp = lbm::complete_prev(h);
        if (cp != LBM_OK) return cp;
    }
    int rc = lbm::lattice_ok(h, l);
    if (rc) return rc;
    if (!xy || n < 0 || scheme < 0 || scheme > 2) return fail(LBM_ERR_INVALID, "bad polygon arguments");
    auto& H = h->lat[l];
    H.bb_kind = 2;
    H.poly.assign(xy, xy + 2 * n);
    H.bb_scheme = scheme;
    H.geom_dirty = h->geom_dirty = true;
    H.uploaded = false;  // the stored state belongs to the old geometry
    return LBM_OK;
}

int lbm_set_moving_wall(lbm_handle h, int32_t l, double omega, double rcx, double rcy)
{
    if (h && h->prev_pending) {
        const int cp = lbm::complete_prev(h);
        if (cp != LBM_OK) return cp;
    }
    int rc = lbm::lattice_ok(h, l);
    if (rc) return rc;
    auto& H = h->lat[l];
    H.moving = true;
    H.omega = omega;
    H.rcx = rcx;
    H.rcy = rcy;
    H.geom_dirty = h->geom_dirty = true;
    H.uploaded = false;  // the stored state belongs to the old geometry
    return LBM_OK;
}

int lbm_set_force_radius(lbm_handle h, int32_t l, double cx, double cy, double r)
{
    if (h && h->prev_pending) {
        const int cp = lbm::complete_prev(h);
        if (cp != LBM_OK) return cp;
    }
    int rc = lbm::lattice_ok(h, l);
    if (rc) return rc;
    auto& H = h->lat[l];
    H.frad = true;
    H.fcx = cx;
    H.fcy = cy;
    H.fr = r;
    H.geom_dirty = h->geom_dirty = true;
    H.uploaded = false;  // the stored state belongs to the old geometry
    return LBM_OK;
}

int lbm_set_links(lbm_handle h, int32_t l, const lbm_link* links, int32_t n)
{
    if (h && h->prev_pending) {
        const int cp = lbm::complete_prev(h);
        if (cp != LBM_OK) return cp;
    }
    int rc = lbm::lattice_ok(h, l);
    if (rc) return rc;
    if (n < 0 || (n > 0 && !links)) return fail(LBM_ERR_INVALID, "bad links");
    auto& H = h->lat[l];
    H.explicit_links.assign(links, links + n);
    H.geom_dirty = h->geom_dirty = true;
    H.uploaded = false;  // the stored state belongs to the old geometry
    return LBM_OK;
}

int lbm_set_force_slots(lbm_handle h, int32_t l, int32_t n_slots)
{
    if (h && h->prev_pending) {
        const int cp = lbm::complete_prev(h);
        if (cp != LBM_OK) return cp;
    }
    int rc = lbm::lattice_ok(h, l);
    if (rc) return rc;
    if (n_slots < 1 || n_slots > 32767) return fail(LBM_ERR_INVALID, "n_slots out of range");
    auto& H = h->lat[l];
    if (H.steps > 0 || H.uploaded) return fail(LBM_ERR_INVALID, "set force slots before lbm_upload");
    H.n_slots = n_slots;
    H.geom_dirty = h->geom_dirty = true;
    H.uploaded = false;  // the stored state belongs to the old geometry
    return LBM_OK;
}

int lbm_set_params(lbm_handle h, int32_t l, double tau, double u_in)
{
    if (h && h->prev_pending) {
        const int cp = lbm::complete_prev(h);
        if (cp != LBM_OK) return cp;
    }
    int rc = lbm::lattice_ok(h, l);
    if (rc) return rc;
    if (!(tau > 0.0)) return fail(LBM_ERR_INVALID, "tau must be positive");
    auto& H = h->lat[l];
    if (H.d.tau != tau && H.uploaded) {
        // The stored buffer already holds the next step's collision (pull form), done
        // with the old tau: redo it from the current state.  The state itself was
        // produced with the old inlet velocity, so that one changes afterwards.
        H.d.tau = tau;
        h->desc_dirty = true;
        rc = h->thermal ? lbm::thermal_recollide(h) : lbm::materialize(h, l, nullptr, nullptr, true);
        if (rc != LBM_OK) return rc;
    }
    if (H.d.tau != tau || H.d.u_in != u_in) {
        H.d.tau = tau;
        H.d.u_in = u_in;
        h->desc_dirty = true;
    }
    return LBM_OK;
}

namespace lbm {
namespace {
// Prologue over all rows of lattice l (f null: uniform equilibrium), then resets the
// lattice's step count, force history and NaN flag.
int load_state(Ensemble* E, int l, const double* f, double rho, double u, double v)
{
    int rc;
    if ((rc = commit(E)) != LBM_OK) return rc;
    HostLattice& H = E->lat[l];
    const lbm_desc& d = H.d;
    const int64_t row_bytes = static_cast<int64_t>(d.nx) * 72;
    const int chunk = f ? static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(d.ny, kStagingBytes / row_bytes)))
                        : d.ny;
    for (int y0 = 0; y0 < d.ny; y0 += chunk) {
        const int y1 = std::min(d.ny, y0 + chunk);
        if (f && (rc = h2d(E, E->d_stage, f + static_cast<int64_t>(y0) * d.nx * 9, (y1 - y0) * row_bytes)) != LBM_OK)
            return rc;
        PrologueArgs p{};
        p.lat = E->d_lat;
        p.lattice = l;
        p.y0 = y0;
        p.y1 = y1;
        p.in = f ? E->d_stage : nullptr;
        p.eq_rho = rho;
        p.eq_u = u;
        p.eq_v = v;
        p.state_dst = E->buf[1 - H.par];
        p.coll_dst = E->buf[H.par];
        p.flags = E->d_flags;
        p.wdist = E->d_wdist;
        p.scale = E->scale;
        p.inv_scale = E->inv_scale;
        LBM_HIP(launch_prologue(E->storage, p, d.nx, E->stream));
    }
    LBM_HIP(hipStreamSynchronize(E->stream));
    H.steps = 0;
    H.read_pos = 0;
    H.pend_fx.clear();
    H.pend_fy.clear();
    H.uploaded = true;
    // the accumulators are clear between batches; reset the NaN flag
    const int32_t zero = 0;
    if ((rc = h2d(E, E->d_nan + l, &zero, 4)) != LBM_OK) return rc;
    E->counters_stale = true;
    if (E->thermal) {
        // g returns to the reference's initial 1.0 (src/lbm_types.hpp:215): an upload of
        // g follows; the post-collision f then carries the forcing of that g
        if ((rc = thermal_load(E, nullptr)) != LBM_OK) return rc;
        if ((rc = thermal_recollide(E)) != LBM_OK) return rc;
    }
    return LBM_OK;
}
}  // namespace
}  // namespace lbm

int lbm_upload(lbm_handle h, int32_t l, const double* f)
{
    if (h && h->prev_pending) {
        const int cp = lbm::complete_prev(h);
        if (cp != LBM_OK) return cp;
    }
    int rc = lbm::lattice_ok(h, l);
    if (rc) return rc;
    if (!f) return fail(LBM_ERR_INVALID, "null population array");
    return lbm::load_state(h, l, f, 0.0, 0.0, 0.0);
}

int lbm_init_equilibrium(lbm_handle h, int32_t l, double rho, double u, double v)
{
    if (h && h->prev_pending) {
        const int cp = lbm::complete_prev(h);
        if (cp != LBM_OK) return cp;
    }
    int rc = lbm::lattice_ok(h, l);
    if (rc) return rc;
    if (!(rho > 0.0)) return fail(LBM_ERR_INVALID, "density must be positive");
    return lbm::load_state(h, l, nullptr, rho, u, v);
}

int lbm_perturb_velocity(lbm_handle h, int32_t l, int32_t x0, int32_t x1, const double* dv)
{
    if (h && h->prev_pending) {
        const int cp = lbm::complete_prev(h);
        if (cp != LBM_OK) return cp;
    }
    using namespace lbm;
    int rc = lattice_ok(h, l);
    if (rc) return rc;
    Ensemble* E = h;
    HostLattice& H = E->lat[l];
    if (!dv) return fail(LBM_ERR_INVALID, "null perturbation array");
    if (!H.uploaded || H.steps > 0)
        return fail(LBM_ERR_INVALID, "perturb right after lbm_upload / lbm_init_equilibrium, before stepping");
    if (x0 < 0 || x1 > H.d.nx || x0 >= x1) return fail(LBM_ERR_INVALID, "column range out of bounds");
    if ((rc = commit(E)) != LBM_OK) return rc;
    const size_t n = static_cast<size_t>(x1 - x0) * H.d.ny;
    double* d_dv = nullptr;
    if (hipMalloc(&d_dv, n * 8) != hipSuccess) return fail(LBM_ERR_NOMEM, "perturbation staging");
    rc = h2d(E, d_dv, dv, n * 8);
    if (rc == LBM_OK) {
        PerturbArgs a{};
        a.lat = E->d_lat;
        a.lattice = l;
        a.x0 = x0;
        a.x1 = x1;
        a.dv = d_dv;
        a.state = E->buf[1 - H.par];
        a.flags = E->d_flags;
        a.scale = E->scale;
        a.inv_scale = E->inv_scale;
        hipError_t e = launch_perturb(E->storage, a, H.d.ny, E->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(E->stream);
        if (e != hipSuccess) rc = fail(LBM_ERR_DEVICE, std::string("perturbation: ") + hipGetErrorString(e));
    }
    (void)hipFree(d_dv);
    if (rc != LBM_OK) return rc;
    // the stored post-collision state follows the new state
    return E->thermal ? thermal_recollide(E) : materialize(E, l, nullptr, nullptr, true);
}

int lbm_macros(lbm_handle h, int32_t l, int32_t y0, int32_t y1, double* rho, double* u, double* v)
{
    if (h && h->prev_pending) {
        const int cp = lbm::complete_prev(h);
        if (cp != LBM_OK) return cp;
    }
    int rc = lbm::lattice_ok(h, l);
    if (rc) return rc;
    if (!h->lat[l].uploaded) return fail(LBM_ERR_INVALID, "no state uploaded");
    if (y0 < 0 || y1 > h->lat[l].d.ny || y0 >= y1) return fail(LBM_ERR_INVALID, "row range out of bounds");
    return lbm::fields(h, l, y0, y1, rho, u, v);
}

int lbm_download(lbm_handle h, int32_t l, double* f)
{
    if (h && h->prev_pending) {
        const int cp = lbm::complete_prev(h);
        if (cp != LBM_OK) return cp;
    }
    int rc = lbm::lattice_ok(h, l);
    if (rc) return rc;
    if (!f) return fail(LBM_ERR_INVALID, "null output array");
    if (!h->lat[l].uploaded) return fail(LBM_ERR_INVALID, "lattice has no state");
    return lbm::materialize(h, l, f, nullptr);
}

int lbm_mass(lbm_handle h, int32_t l, double* mass)
{
    if (h && h->prev_pending) {
        const int cp = lbm::complete_prev(h);
        if (cp != LBM_OK) return cp;
    }
    int rc = lbm::lattice_ok(h, l);
    if (rc) return rc;
    if (!mass) return fail(LBM_ERR_INVALID, "null output");
    if (!h->lat[l].uploaded) return fail(LBM_ERR_INVALID, "lattice has no state");
    return lbm::materialize(h, l, nullptr, mass);
}

int lbm_step(lbm_handle h, int32_t n)
{
    using namespace lbm;
    if (!h) return fail(LBM_ERR_INVALID, "null handle");
    if (n < 0) return fail(LBM_ERR_INVALID, "negative step count");
    Ensemble* E = h;
    int rc;
    if ((rc = require_uploaded(E)) != LBM_OK) return rc;
    if ((rc = commit(E)) != LBM_OK) return rc;
    if (E->graph_dirty) drop_graphs(E);
    if (n == 0 && E->graph_steps >= 2) {
        // prepare the graphs later calls replay: one per starting buffer parity of a
        // one-lattice handle (a batch may leave either), one for an ensemble
        const int np = E->lat.size() == 1 ? 2 : 1;
        const int32_t par = E->lat[0].par;
        for (int p = 0; p < np && rc == LBM_OK; ++p)
            if (!E->graph[p]) {
                if (np == 2) E->lat[0].par = p;  // the capture takes the buffers of parity p
                rc = build_graph(E, p);
            }
        E->lat[0].par = par;
        if (rc != LBM_OK) return rc;
    }
    while (n > 0) {
        const bool use_graph = E->graph_steps >= 2 && n >= E->graph_steps;
        const int batch = use_graph ? E->graph_steps : std::min(n, E->acc_blocks);
        for (size_t l = 0; l < E->lat.size(); ++l)
            if ((rc = ensure_history_room(E, static_cast<int>(l), batch)) != LBM_OK) return rc;
        if (use_graph) {
            const int p = E->lat.size() == 1 ? E->lat[0].par : 0;
            if (!E->graph[p] && (rc = build_graph(E, p)) != LBM_OK) return rc;
            LBM_HIP(hipGraphLaunch(E->graph[p], E->stream));
        } else {
            if ((rc = launch_batch(E, batch, -1)) != LBM_OK) return rc;
        }
        const int flips = batch_flips(E, batch, -1);
        for (auto& H : E->lat) {
            H.steps += batch;
            H.par ^= flips & 1;
        }
        E->prev_pending = ends_on_pair(E, batch, -1);
        n -= batch;
    }
    return LBM_OK;
}

int lbm_step_noalloc(lbm_handle h, int32_t n)
{
    using namespace lbm;
    if (!h) return fail(LBM_ERR_INVALID, "null handle");
    if (n < 0) return fail(LBM_ERR_INVALID, "negative step count");
    Ensemble* E = h;
    int rc;
    if ((rc = require_uploaded(E)) != LBM_OK) return rc;
    // everything commit() would write must be on the device already: it may allocate
    if (E->geom_dirty || E->desc_dirty || E->counters_stale || (E->tiles_dirty && E->lat.size() > 1) ||
        !E->d_slot_start || !E->d_hist ||
        ((E->options & LAUNCH_TWO_STEPS) && bulk_eligible(E, E->rows) && !current_plan(E)))
        return fail(LBM_ERR_INVALID, "lbm_step_noalloc: the handle has changes not yet on the device (geometry, "
                                     "parameters, launch shape or an upload): call lbm_step(h, 0) first");
    if (n == 0) return LBM_OK;
    const int64_t acc_per = static_cast<int64_t>(kAccCopies) * E->total_slots * 2;
    const int64_t nf_per = static_cast<int64_t>(E->n_recs) * 2;
    const int64_t cap = static_cast<int64_t>(LBM_NOALLOC_SCRATCH_BYTES / sizeof(double));
    if (acc_per + nf_per > cap)
        return fail(LBM_ERR_INVALID, "lbm_step_noalloc: one launch's force partials (" +
                                         std::to_string((acc_per + nf_per) * 8) + " bytes for " +
                                         std::to_string(E->n_recs) + " link records) exceed the " +
                                         std::to_string(LBM_NOALLOC_SCRATCH_BYTES) + "-byte scratch");
    // the flush writes history rows steps .. steps + n - 1 (mod R); draining them to the
    // host would allocate, so rows not yet read must not be overwritten
    for (auto& H : E->lat)
        if (H.steps + n - H.read_pos > E->R)
            return fail(LBM_ERR_INVALID, "lbm_step_noalloc: the force history holds " + std::to_string(E->R) +
                                             " steps per lattice: read it with lbm_forces first");
    const int64_t chunk = std::min<int64_t>(n, cap / (acc_per + nf_per));
    double* base = nullptr;
    if ((rc = scratch_on_device(&base)) != LBM_OK) return rc;
    double* acc_base = base;
    double* nf_base = base + chunk * acc_per;
    // the accumulators start at zero (every flush clears what it read, but an earlier
    // call may have stopped between its launches and its flush)
    LBM_HIP(launch_fill_zero(acc_base, chunk * acc_per, E->stream));
    while (n > 0) {
        const int batch = static_cast<int>(std::min<int64_t>(n, chunk));
        if ((rc = launch_batch(E, batch, -1, acc_base, nf_base)) != LBM_OK) return rc;
        const int flips = batch_flips(E, batch, -1);
        for (auto& H : E->lat) {
            H.steps += batch;
            H.par ^= flips & 1;
        }
        E->prev_pending = ends_on_pair(E, batch, -1);
        n -= batch;
    }
    return LBM_OK;
}

int lbm_debug_poison(lbm_handle h, int32_t l)
{
    if (h && h->prev_pending) {
        const int cp = lbm::complete_prev(h);
        if (cp != LBM_OK) return cp;
    }
    using namespace lbm;
    int rc = lattice_ok(h, l);
    if (rc) return rc;
    Ensemble* E = h;
    HostLattice& H = E->lat[l];
    if (!H.uploaded) return fail(LBM_ERR_INVALID, "lattice has no state");
    // the next launch over this lattice reads buf[par] and writes buf[1 - par]
    const size_t off = static_cast<size_t>(H.f_off) * E->bpp;
    const size_t bytes = static_cast<size_t>(9) * H.d.ny * H.pitch * E->bpp;
    LBM_HIP(hipMemsetAsync(static_cast<char*>(E->buf[1 - H.par]) + off, 0xff, bytes, E->stream));
    if (E->thermal) LBM_HIP(hipMemsetAsync(static_cast<char*>(E->gbuf[1 - H.par]) + off, 0xff, bytes, E->stream));
    LBM_HIP(hipStreamSynchronize(E->stream));
    return LBM_OK;
}

int lbm_step_lattice(lbm_handle h, int32_t l, int32_t n)
{
    if (h && h->prev_pending) {
        const int cp = lbm::complete_prev(h);
        if (cp != LBM_OK) return cp;
    }
    using namespace lbm;
    int rc = lattice_ok(h, l);
    if (rc) return rc;
    if (n < 0) return fail(LBM_ERR_INVALID, "negative step count");
    Ensemble* E = h;
    HostLattice& H = E->lat[l];
    if (!H.uploaded) return fail(LBM_ERR_INVALID, "lattice has no state");
    if ((rc = commit(E)) != LBM_OK) return rc;
    while (n > 0) {
        const int batch = std::min(n, E->acc_blocks);
        if ((rc = ensure_history_room(E, l, batch)) != LBM_OK) return rc;
        if ((rc = launch_batch(E, batch, l)) != LBM_OK) return rc;
        H.steps += batch;
        H.par ^= batch_flips(E, batch, l) & 1;
        E->prev_pending = ends_on_pair(E, batch, l);
        n -= batch;
    }
    return LBM_OK;
}

int lbm_set_launch(lbm_handle h, int32_t bx, int32_t rows, int32_t nodes_per_thread, int32_t options,
                   int32_t graph_steps)
{
    using namespace lbm;
    if (!h) return fail(LBM_ERR_INVALID, "null handle");
    Ensemble* E = h;
    const int nbx = bx > 0 ? bx : E->bx, nrows = rows > 0 ? rows : E->rows;
    const int nv = nodes_per_thread > 0 ? nodes_per_thread : E->vec;
    const int nopt = options >= 0 ? options : E->options;
    if (nopt & ~(LAUNCH_XCD_ORDER | LAUNCH_STREAM_STORES | LAUNCH_TWO_STEPS))
        return fail(LBM_ERR_INVALID, "unknown launch option bits");
    if (nopt & LAUNCH_TWO_STEPS) {
        int rc = two_step_ok(E, nbx, nrows, nv);
        if (rc) return rc;
    } else if (!step_shape_supported(E->storage, nbx, nrows, nv)) {
        return fail(LBM_ERR_INVALID, "launch shape not compiled for this storage form");
    }
    if (graph_steps < 0 || (graph_steps > 0 && (graph_steps % 2 != 0 || graph_steps > 2000)))
        return fail(LBM_ERR_INVALID, "graph_steps must be even and at most 2000");
    if (nbx != E->bx || nrows != E->rows || nv != E->vec || nopt != E->options) {
        for (int k = 0; k < 4; ++k) E->tail[k] = 0;
        E->bx = nbx;
        E->rows = nrows;
        E->vec = nv;
        E->options = nopt;
        E->tiles_dirty = true;
        E->graph_dirty = true;
    }
    if (graph_steps > 0 && graph_steps != E->graph_steps) {
        E->graph_steps = graph_steps;
        E->graph_dirty = true;
    }
    return LBM_OK;
}

int lbm_autotune(lbm_handle h, int32_t* best_bx, int32_t* best_rows, int32_t* best_nodes, int32_t* best_options,
                 float* best_ms)
{
    if (h && h->prev_pending) {
        const int cp = lbm::complete_prev(h);
        if (cp != LBM_OK) return cp;
    }
    using namespace lbm;
    if (!h) return fail(LBM_ERR_INVALID, "null handle");
    Ensemble* E = h;
    int rc;
    if (E->thermal) return fail(LBM_ERR_INVALID, "a thermal lattice has one launch shape: nothing to tune");
    if ((rc = require_uploaded(E)) != LBM_OK) return rc;
    if ((rc = commit(E)) != LBM_OK) return rc;
    // Candidates are timed as stepping runs them: each launch reads what the previous one
    // wrote, ping-ponging between two copies of the population buffers (a launch that
    // always reads the same source would hide what streaming stores cost the next step's
    // reads).  The state, the accumulators and the history stay untouched.
    // each copy carries the population buffers' guard: the kernels read from it
    const size_t pool_bytes = static_cast<size_t>(E->pool_elems) * E->bpp;
    void* copies[2] = {nullptr, nullptr};
    const int64_t held = E->device_bytes;
    for (int k = 0; k < 2; ++k)
        if ((rc = alloc_population(E, &copies[k], "autotune copies of the population buffers")) != LBM_OK) {
            if (copies[0]) (void)hipFree(copies[0]);
            E->device_bytes = held;
            return rc;
        }
    E->device_bytes = held;  // freed before returning
    for (int k = 0; k < 2 && rc == LBM_OK; ++k)
        if (hipMemcpyAsync(copies[k], E->buf[k], pool_bytes, hipMemcpyDeviceToDevice, E->stream) != hipSuccess)
            rc = fail(LBM_ERR_DEVICE, "autotune scratch copy");
    if (rc != LBM_OK) {
        (void)hipFree(copies[0]);
        (void)hipFree(copies[1]);
        return rc;
    }
    // launch n of a candidate: one lattice reads copies[(par + n) & 1] and writes the
    // other; an ensemble takes both from the per-lattice parity flipped n times
    const int par0 = E->lat[0].par;
    auto ping = [&](StepArgs& a, bool single, int n) {
        if (single) {
            a.src = copies[(par0 + n) & 1];
            a.dst = copies[(par0 + n + 1) & 1];
        } else {
            a.buf0 = copies[0];
            a.buf1 = copies[1];
            a.dst = nullptr;
            a.step_arg = n;
            a.flips = n;
        }
    };
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    int shapes[64][3];
    const int n_shapes = std::min(64, step_shapes(E->storage, shapes, 64));
    float best = 1e30f;
    const int saved[4] = {E->bx, E->rows, E->vec, E->options};
    int bs[4] = {E->bx, E->rows, E->vec, E->options};
    int bs1[4] = {64, 1, 1, 0};  // fastest one-step shape (the tail of two-step batches)
    float best1 = 1e30f;
    rc = LBM_OK;
    for (int si = 0; si < 4 * n_shapes; ++si) {
        E->bx = shapes[si % n_shapes][0];
        E->rows = shapes[si % n_shapes][1];
        E->vec = shapes[si % n_shapes][2];
        E->options = si / n_shapes;
        E->tiles_dirty = true;
        if ((rc = commit(E)) != LBM_OK) break;
        StepArgs a = base_args(E);
        int blocks;
        const bool single = grid_for(E, -1, a, blocks);
        // dummy accumulators keep the force history intact
        a.acc = E->d_dummy_acc;
        a.nf = E->d_dummy_nf;
        a.nan_flags = E->d_dummy_nan;
        int n = 0;
        hipError_t le = hipSuccess;
        for (int w = 0; w < 2 && le == hipSuccess; ++w, ++n) {
            ping(a, single, n);
            le = launch_step(E->storage, E->bx, E->rows, E->vec, single, blocks, a, E->stream);
        }
        // ten launches for an estimate, then enough for ~1.5 ms of device time, so that
        // shapes a few percent apart on small lattices are told apart
        float ms = 0.f;
        bool ok = le == hipSuccess;
        for (int pass = 0, reps = 10; pass < 2 && ok; ++pass) {
            (void)hipEventRecord(e0, E->stream);
            for (int r = 0; r < reps && le == hipSuccess; ++r, ++n) {
                ping(a, single, n);
                le = launch_step(E->storage, E->bx, E->rows, E->vec, single, blocks, a, E->stream);
            }
            (void)hipEventRecord(e1, E->stream);
            ok = le == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
            if (!ok) break;
            (void)hipEventElapsedTime(&ms, e0, e1);
            ms /= reps;
            if (ms * reps >= kTuneMs) break;
            reps = std::min(2000, static_cast<int>(kTuneMs / std::max(ms, 1e-4f)) + 1);
        }
        if (!ok) {
            rc = fail(LBM_ERR_DEVICE, "autotune launch failed");
            break;
        }
        if (ms < best1) {
            best1 = ms;
            bs1[0] = E->bx;
            bs1[1] = E->rows;
            bs1[2] = E->vec;
            bs1[3] = E->options;
        }
        if (ms < best) {
            best = ms;
            bs[0] = E->bx;
            bs[1] = E->rows;
            bs[2] = E->vec;
            bs[3] = E->options;
        }
    }
    // two steps per launch: per-step time of each tile height
    for (int si = 0; si < 16 && rc == LBM_OK; ++si) {
        const int rows = 4 << (si / 4), opt = LAUNCH_TWO_STEPS | (si & 3);
        if (two_step_ok(E, 64, rows, 1) != LBM_OK) continue;
        E->bx = 64;
        E->rows = rows;
        E->vec = 1;
        E->options = opt;
        E->tiles_dirty = true;
        if ((rc = commit(E)) != LBM_OK) break;
        StepArgs a = base_args(E);
        int blocks;
        const bool single = grid_for(E, -1, a, blocks);
        if (single) {
            const lbm_desc& d = E->lat[0].d;
            a.tiles_x = (d.nx + 63) / 64;
            a.tiles_y = (d.ny + rows - 1) / rows;
            blocks = a.tiles_x * a.tiles_y;
        } else {
            a.tiles = E->d_tiles2;
            a.n_tiles = E->n_tiles2;
            blocks = E->n_tiles2;
        }
        a.acc = a.acc2 = E->d_dummy_acc;
        a.nf = a.nf2 = E->d_dummy_nf;
        a.nan_flags = E->d_dummy_nan;
        int n = 0;
        hipError_t le = hipSuccess;
        for (int w = 0; w < 2 && le == hipSuccess; ++w, ++n) {
            ping(a, single, n);
            le = launch_pair(E, a, single, blocks);
        }
        float ms = 0.f;
        bool ok = le == hipSuccess;
        for (int pass = 0, reps = 5; pass < 2 && ok; ++pass) {
            (void)hipEventRecord(e0, E->stream);
            for (int r = 0; r < reps && le == hipSuccess; ++r, ++n) {
                ping(a, single, n);
                le = launch_pair(E, a, single, blocks);
            }
            (void)hipEventRecord(e1, E->stream);
            ok = le == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
            if (!ok) break;
            (void)hipEventElapsedTime(&ms, e0, e1);
            ms /= reps;
            if (ms * reps >= kTuneMs) break;
            reps = std::min(1000, static_cast<int>(kTuneMs / std::max(ms, 1e-4f)) + 1);
        }
        if (!ok) {
            rc = fail(LBM_ERR_DEVICE, "autotune launch failed");
            break;
        }
        ms /= 2;  // per step
        if (ms < best) {
            best = ms;
            bs[0] = 64;
            bs[1] = rows;
            bs[2] = 1;
            bs[3] = opt;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(copies[0]);
    (void)hipFree(copies[1]);
    const int* pick = rc == LBM_OK ? bs : saved;
    E->bx = pick[0];
    E->rows = pick[1];
    E->vec = pick[2];
    E->options = pick[3];
    for (int k = 0; k < 4; ++k) E->tail[k] = rc == LBM_OK && (bs[3] & LAUNCH_TWO_STEPS) ? bs1[k] : 0;
    E->tiles_dirty = true;
    E->graph_dirty = true;
    if (rc != LBM_OK) return rc;
    if ((rc = commit(E)) != LBM_OK) return rc;
    if (best_bx) *best_bx = bs[0];
    if (best_rows) *best_rows = bs[1];
    if (best_nodes) *best_nodes = bs[2];
    if (best_options) *best_options = bs[3];
    if (best_ms) *best_ms = best;
    return LBM_OK;
}

int lbm_forces(lbm_handle h, int32_t l, double* fx, double* fy, int32_t max_steps, int32_t* n_steps)
{
    using namespace lbm;
    int rc = lattice_ok(h, l);
    if (rc) return rc;
    Ensemble* E = h;
    HostLattice& H = E->lat[l];
    if ((rc = drain(E, l)) != LBM_OK) return rc;
    const int64_t avail = static_cast<int64_t>(H.pend_fx.size()) / H.n_slots;
    const int64_t take = std::min<int64_t>(avail, max_steps);
    if (take > 0 && (!fx || !fy)) return fail(LBM_ERR_INVALID, "null force output");
    const size_t nv = static_cast<size_t>(take) * H.n_slots;
    if (nv) {
        std::memcpy(fx, H.pend_fx.data(), nv * 8);
        std::memcpy(fy, H.pend_fy.data(), nv * 8);
        H.pend_fx.erase(H.pend_fx.begin(), H.pend_fx.begin() + nv);
        H.pend_fy.erase(H.pend_fy.begin(), H.pend_fy.begin() + nv);
    }
    if (n_steps) *n_steps = static_cast<int32_t>(take);
    return LBM_OK;
}

int lbm_nan_flag(lbm_handle h, int32_t l, int32_t* flag)
{
    int rc = lbm::lattice_ok(h, l);
    if (rc) return rc;
    if (!flag) return fail(LBM_ERR_INVALID, "null output");
    if ((rc = lbm::d2h(h, flag, h->d_nan + l, 4)) != LBM_OK) return rc;
    const int32_t zero = 0;
    return lbm::h2d(h, h->d_nan + l, &zero, 4);
}

int lbm_steps_done(lbm_handle h, int32_t l, int64_t* steps)
{
    int rc = lbm::lattice_ok(h, l);
    if (rc) return rc;
    if (!steps) return fail(LBM_ERR_INVALID, "null output");
    *steps = h->lat[l].steps;
    return LBM_OK;
}

int lbm_get_info(lbm_handle h, lbm_info* out)
{
    using namespace lbm;
    if (!h || !out) return fail(LBM_ERR_INVALID, "null argument");
    int rc;
    if ((rc = commit(h)) != LBM_OK) return rc;
    lbm_info r{};
    for (auto& H : h->lat) {
        const int64_t n = static_cast<int64_t>(H.d.nx) * H.d.ny;
        r.cells += n;
        r.fluid_cells += H.fluid;
        r.link_cells += static_cast<int64_t>(H.recs.size());
        if (H.d.les) r.les_cells += n;
    }
    r.bytes_per_pop = h->bpp;
    r.n_lattices = static_cast<int32_t>(h->lat.size());
    r.device_bytes = h->device_bytes;
    r.block_x = h->bx;
    r.rows = h->rows;
    r.nodes_per_thread = h->vec;
    r.launch_options = h->options;
    if (const Ensemble::BulkPlan* P = current_plan(h)) {
        r.bulk_tiles = P->n_bulk;
        r.other_tiles = P->n_rest;
        r.edge_tiles = P->n_edge;
        r.empty_tiles = P->n_empty;
    }
    *out = r;
    return LBM_OK;
}

int lbm_sync(lbm_handle h)
{
    if (!h) return fail(LBM_ERR_INVALID, "null handle");
    hipError_t e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return fail(LBM_ERR_DEVICE, std::string("hipStreamSynchronize: ") + hipGetErrorString(e));
    return LBM_OK;
}

int lbm_set_thermal(lbm_handle h, int32_t l, double omega_k, double beta, double g_buoyancy, double t_ref)
{
    if (h && h->prev_pending) {
        const int cp = lbm::complete_prev(h);
        if (cp != LBM_OK) return cp;
    }
    using namespace lbm;
    int rc = lattice_ok(h, l);
    if (rc) return rc;
    Ensemble* E = h;
    if (E->lat.size() != 1) return fail(LBM_ERR_INVALID, "the thermal populations need a handle of one lattice");
    if (E->lat[0].d.y_edge == LBM_EDGE_FREE_SLIP)
        return fail(LBM_ERR_INVALID, "the thermal populations stream periodic in y: no free-slip y edges");
    if (E->storage != ST_F64 && E->storage != ST_F32)
        return fail(LBM_ERR_INVALID, "the thermal populations need F64 or F32 storage");
    if (!std::isfinite(omega_k) || !std::isfinite(beta) || !std::isfinite(g_buoyancy) || !std::isfinite(t_ref))
        return fail(LBM_ERR_INVALID, "thermal parameters must be finite");
    HostLattice& H = E->lat[0];
    const bool first = !E->thermal;
    if (first) {
        if (H.uploaded && H.steps > 0)
            return fail(LBM_ERR_INVALID, "turn the thermal populations on before stepping (or upload again)");
        for (int b = 0; b < 2; ++b)
            if ((rc = alloc_population(E, &E->gbuf[b], "thermal buffers")) != LBM_OK) return rc;
        if ((rc = dev_alloc(E, &E->d_tcode, E->cell_elems)) != LBM_OK) return rc;
        E->thermal = true;
    }
    E->omega_k = omega_k;
    E->beta = beta;
    E->g_buoyancy = g_buoyancy;
    E->t_ref = t_ref;
    E->graph_dirty = true;
    if (H.uploaded) {
        if (first && (rc = thermal_load(E, nullptr)) != LBM_OK) return rc;
        return thermal_recollide(E);
    }
    return LBM_OK;
}

int lbm_set_thermal_rules(lbm_handle h, int32_t l, const uint8_t* code, const lbm_thermal_rule* rules,
                          int32_t n_rules)
{
    if (h && h->prev_pending) {
        const int cp = lbm::complete_prev(h);
        if (cp != LBM_OK) return cp;
    }
    using namespace lbm;
    int rc = thermal_ok(h, l);
    if (rc) return rc;
    Ensemble* E = h;
    if (n_rules < 0 || n_rules > kThermalRules) return fail(LBM_ERR_INVALID, "at most 15 thermal rules");
    if (n_rules > 0 && !rules) return fail(LBM_ERR_INVALID, "null rule array");
    for (int k = 0; k < n_rules; ++k) {
        if (rules[k].kind != LBM_THERMAL_EQ_FLOW && rules[k].kind != LBM_THERMAL_EQ_REST)
            return fail(LBM_ERR_INVALID, "unknown thermal rule kind");
        if (!std::isfinite(rules[k].t)) return fail(LBM_ERR_INVALID, "rule temperature must be finite");
    }
    HostLattice& H = E->lat[0];
    if (code) {
        std::vector<uint8_t> pc(E->cell_elems, 0);
        for (int y = 0; y < H.d.ny; ++y)
            for (int x = 0; x < H.d.nx; ++x) {
                const uint8_t c = code[static_cast<int64_t>(y) * H.d.nx + x];
                if ((c & ~(15u | LBM_THERMAL_MIRROR)) || (c & 15) > n_rules)
                    return fail(LBM_ERR_INVALID, "node condition code names no rule");
                pc[H.cell_off + static_cast<int64_t>(y) * H.pitch + x] = c;
            }
        if ((rc = h2d(E, E->d_tcode, pc.data(), pc.size())) != LBM_OK) return rc;
    }
    E->has_rules = code != nullptr;
    E->rule_rest = 0;
    for (int k = 0; k < kThermalRules; ++k) {
        E->rule_t[k] = k < n_rules ? rules[k].t : 0.0;
        if (k < n_rules && rules[k].kind == LBM_THERMAL_EQ_REST) E->rule_rest |= 1u << k;
    }
    E->graph_dirty = true;
    // the conditions act on the state after a step: the current one changes with them
    if (H.uploaded && H.steps > 0) return thermal_recollide(E);
    return LBM_OK;
}

int lbm_upload_thermal(lbm_handle h, int32_t l, const double* g)
{
    if (h && h->prev_pending) {
        const int cp = lbm::complete_prev(h);
        if (cp != LBM_OK) return cp;
    }
    using namespace lbm;
    int rc = thermal_ok(h, l);
    if (rc) return rc;
    if (!g) return fail(LBM_ERR_INVALID, "null population array");
    HostLattice& H = h->lat[0];
    if (!H.uploaded || H.steps > 0)
        return fail(LBM_ERR_INVALID, "upload g right after f (lbm_upload or lbm_init_equilibrium), before stepping");
    if ((rc = commit(h)) != LBM_OK) return rc;
    if ((rc = thermal_load(h, g)) != LBM_OK) return rc;
    return thermal_recollide(h);
}

int lbm_download_thermal(lbm_handle h, int32_t l, double* g)
{
    if (h && h->prev_pending) {
        const int cp = lbm::complete_prev(h);
        if (cp != LBM_OK) return cp;
    }
    using namespace lbm;
    int rc = thermal_ok(h, l);
    if (rc) return rc;
    if (!g) return fail(LBM_ERR_INVALID, "null output array");
    if (!h->lat[0].uploaded) return fail(LBM_ERR_INVALID, "lattice has no state");
    return thermal_out(h, THERMAL_OUT_G, 0, h->lat[0].d.ny, g);
}

int lbm_temperature(lbm_handle h, int32_t l, int32_t y0, int32_t y1, double* t)
{
    if (h && h->prev_pending) {
        const int cp = lbm::complete_prev(h);
        if (cp != LBM_OK) return cp;
    }
    using namespace lbm;
    int rc = thermal_ok(h, l);
    if (rc) return rc;
    if (!t) return fail(LBM_ERR_INVALID, "null output array");
    if (!h->lat[0].uploaded) return fail(LBM_ERR_INVALID, "lattice has no state");
    if (y0 < 0 || y1 > h->lat[0].d.ny || y0 >= y1) return fail(LBM_ERR_INVALID, "row range out of bounds");
    return thermal_out(h, THERMAL_OUT_T, y0, y1, t);
}

int lbm_wall_distance(lbm_handle h, int32_t l, double* out)
{
    int rc = lbm::lattice_ok(h, l);
    if (rc) return rc;
    if (!out) return fail(LBM_ERR_INVALID, "null output");
    if ((rc = lbm::commit(h)) != LBM_OK) return rc;
    auto& H = h->lat[l];
    std::vector<uint16_t> w(static_cast<size_t>(H.d.ny) * H.pitch);
    if ((rc = lbm::d2h(h, w.data(), h->d_wdist + H.cell_off, w.size() * 2)) != LBM_OK) return rc;
    for (int y = 0; y < H.d.ny; ++y)
        for (int x = 0; x < H.d.nx; ++x) {
            const uint16_t v = w[static_cast<int64_t>(y) * H.pitch + x];
            out[static_cast<int64_t>(y) * H.d.nx + x] = v == 0xffff ? 1e9 : static_cast<double>(v);
        }
    return LBM_OK;
}

}  // extern "C"
