// lbm_kernels.hip -- D2Q9 lattice update kernels for gfx950 (MI355X).
//
// One fused launch per time step, in pull form.  The population buffers hold the
// post-collision state C_k (the reference's sys.f after the collision loop of
// src/lbm.hpp:300-354).  The launch for step k+1 at node x:
//   1. gathers C_k(x - e_j) for every direction j (streaming, src/lbm.hpp:405-484),
//      or the bounce-back value of the link when x - e_j is solid (src/lbm.hpp:226-288),
//   2. applies the open boundaries to that node (src/lbm.hpp:530-539),
//   3. adds its momentum-exchange terms to the FP64 force sum (src/lbm.hpp:542-581),
//   4. collides (MRT / BGK, LES, body force) and writes C_{k+1}(x) in place of nothing:
//      reads come from one buffer, writes go to the other (A-B pattern).
// Storage keeps deviations h_i = (f_i - w_i) * 2^k (k = 0 for F64/F32) in SoA planes
// of pitch-aligned rows; arithmetic is done on deviations in FP32 (FP64 for F64).
//
// Bulk nodes (no edge, no solid neighbour) take a branch-free path of 9 shifted,
// coalesced row loads, 9 aligned stores and one flag byte.  Every other node takes
// the general path, which handles links, edges and boundary conditions exactly as
// the reference's branches do.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "lbm_kernels.h"

namespace lbm {
namespace {

constexpr int CX[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
constexpr int CY[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
constexpr int OPP[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
constexpr double WD[9] = {4.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0, 1.0 / 9.0,
                          1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};

enum { EDGE_PERIODIC = 0, EDGE_OPEN = 1, EDGE_WALL = 2 };
enum { INLET_NONE = 0, INLET_ZOU_HE = 1, INLET_PARABOLIC = 2 };
enum { BB_SIMPLE = 0, BB_MEI = 1, BB_BOUZIDI = 2 };

// ------------------------------------------------------------------ storage forms

template <int ST> struct Form;
template <> struct Form<ST_F64> {
    using store = double;
    using real = double;
    static __device__ __forceinline__ real dec(store v, float) { return v; }
    static __device__ __forceinline__ store enc(real d, float) { return d; }
};
template <> struct Form<ST_F32> {
    using store = float;
    using real = float;
    static __device__ __forceinline__ real dec(store v, float) { return v; }
    static __device__ __forceinline__ store enc(real d, float) { return d; }
};
template <> struct Form<ST_F16S> {
    using store = uint16_t;
    using real = float;
    static __device__ __forceinline__ real dec(store v, float inv_scale)
    {
        return static_cast<float>(__builtin_bit_cast(_Float16, v)) * inv_scale;
    }
    // Conversion to binary16 rounds to nearest even (v_cvt_f16_f32).
    static __device__ __forceinline__ store enc(real d, float scale)
    {
        return __builtin_bit_cast(uint16_t, static_cast<_Float16>(d * scale));
    }
};
template <> struct Form<ST_BF16S> {
    using store = uint16_t;
    using real = float;
    static __device__ __forceinline__ real dec(store v, float inv_scale)
    {
        return __uint_as_float(static_cast<uint32_t>(v) << 16) * inv_scale;
    }
    // Round to nearest even on the 16 dropped mantissa bits; NaN stays NaN.
    static __device__ __forceinline__ store enc(real d, float scale)
    {
        uint32_t u = __float_as_uint(d * scale);
        if ((u & 0x7fffffffu) > 0x7f800000u) return static_cast<uint16_t>((u >> 16) | 0x40u);
        u += 0x7fffu + ((u >> 16) & 1u);
        return static_cast<uint16_t>(u >> 16);
    }
};

template <typename R> __device__ __forceinline__ R wgt(int i) { return static_cast<R>(WD[i]); }

// ------------------------------------------------------------------ collision

template <typename R> __device__ __forceinline__ R clamp_rate(R s)
{
    return s < R(0.5) ? R(0.5) : (s > R(1.99) ? R(1.99) : s);
}

template <typename R> __device__ __forceinline__ R dsqrt(R x);
template <> __device__ __forceinline__ float dsqrt<float>(float x) { return __builtin_sqrtf(x); }
template <> __device__ __forceinline__ double dsqrt<double>(double x) { return __builtin_sqrt(x); }
template <typename R> __device__ __forceinline__ R dexp(R x);
template <> __device__ __forceinline__ float dexp<float>(float x) { return __expf(x); }
template <> __device__ __forceinline__ double dexp<double>(double x) { return exp(x); }

// MRT (src/lbm.hpp:128-218) or BGK (src/lbm.hpp:334-354) on deviations d_i = f_i - w_i,
// followed by the body force (src/lbm.hpp:327-331).  The relaxation is applied as
// f <- f - M^-1 S (m - m_eq): rho and j are conserved, so only the six
// non-conserved moment differences enter.  With w subtracted, the moments of w
// (e = -2, eps = 1, others 0) cancel against the equilibria analytically.
template <typename R>
__device__ __forceinline__ void collide(R d[9], const LatticeDev& L, R wall_dist)
{
    const R s0 = d[1] + d[2] + d[3] + d[4];
    const R s1 = d[5] + d[6] + d[7] + d[8];
    const R sum = d[0] + s0 + s1;  // rho - 1
    const R rho = R(1) + sum;
    const R jx = (d[1] - d[3]) + (d[5] - d[6]) - (d[7] - d[8]);
    const R jy = (d[2] - d[4]) + (d[5] + d[6]) - (d[7] + d[8]);
    const R inv_rho = R(1) / rho;
    const R u = jx * inv_rho, v = jy * inv_rho;
    const R usq = u * u + v * v;

    if (L.collision != 0) {
        // BGK: d_i -= (d_i - (feq_i - w_i)) / tau
        const R it = static_cast<R>(L.inv_tau);
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            const R eu = R(CX[i]) * u + R(CY[i]) * v;
            const R feq_dev = wgt<R>(i) * (sum + rho * (R(3) * eu + R(4.5) * eu * eu - R(1.5) * usq));
            d[i] -= it * (d[i] - feq_dev);
        }
    } else {
        const R e_d = R(-4) * d[0] - s0 + R(2) * s1;
        const R eps_d = R(4) * d[0] - R(2) * s0 + s1;
        const R qx_d = R(-2) * d[1] + R(2) * d[3] + d[5] - d[6] - d[7] + d[8];
        const R qy_d = R(-2) * d[2] + R(2) * d[4] + d[5] + d[6] - d[7] - d[8];
        const R pxx_d = d[1] - d[2] + d[3] - d[4];
        const R pxy_d = d[5] - d[6] + d[7] - d[8];
        const R ru = R(3) * rho * usq;
        const R ne = e_d + R(2) * sum - ru;      // e - e_eq
        const R neps = eps_d - sum + ru;         // eps - eps_eq
        const R nqx = qx_d + jx;                 // qx - qx_eq
        const R nqy = qy_d + jy;
        const R npxx = pxx_d - (jx * u - jy * v);  // pxx - rho (u^2 - v^2)
        const R npxy = pxy_d - jx * v;             // pxy - rho u v
        R ss = static_cast<R>(L.s_shear);
        if (L.les) {
            // Smagorinsky with Van Driest damping (src/lbm.hpp:155-182, y+ :314-325).
            // The damped block replaces the plain one whenever y+ > 0.
            R cs2 = static_cast<R>(L.cs2);
            const R nu = static_cast<R>(L.nu);
            if (nu > R(0)) {
                const R y_plus = wall_dist * dsqrt<R>(usq + R(1e-12)) / nu;
                if (y_plus > R(0)) {
                    const R damp = R(1) - dexp<R>(-y_plus / R(26));
                    const R csd = dsqrt<R>(cs2) * damp;
                    cs2 = csd * csd;
                }
            }
            const R Q = dsqrt<R>(R(2) * (npxx * npxx + npxy * npxy));
            const R A = R(9) * cs2 * Q / (R(2) * rho);
            const R tau = static_cast<R>(L.tau);
            const R tau_eff = (tau + dsqrt<R>(tau * tau + R(4) * A)) * R(0.5);
            ss = clamp_rate<R>(R(1) / tau_eff);
        }
        const R sb = static_cast<R>(L.s_bulk), sn = static_cast<R>(L.s_normal);
        const R De = sb * ne, Deps = sb * neps;
        const R Dqx = sn * nqx, Dqy = sn * nqy;
        const R Dpxx = ss * npxx, Dpxy = ss * npxy;
        // rows of M^-1 applied to (0, De, Deps, 0, Dqx, 0, Dqy, Dpxx, Dpxy)
        const R a = De * R(-1.0 / 36.0) - Deps * R(1.0 / 18.0);
        const R b = De * R(1.0 / 18.0) + Deps * R(1.0 / 36.0);
        const R qx6 = Dqx * R(1.0 / 6.0), qy6 = Dqy * R(1.0 / 6.0);
        const R qx12 = Dqx * R(1.0 / 12.0), qy12 = Dqy * R(1.0 / 12.0);
        const R p4 = Dpxx * R(0.25), x4 = Dpxy * R(0.25);
        d[0] -= (Deps - De) * R(1.0 / 9.0);
        d[1] -= a - qx6 + p4;
        d[2] -= a - qy6 - p4;
        d[3] -= a + qx6 + p4;
        d[4] -= a + qy6 - p4;
        d[5] -= b + qx12 + qy12 + x4;
        d[6] -= b - qx12 + qy12 - x4;
        d[7] -= b - qx12 - qy12 + x4;
        d[8] -= b + qx12 - qy12 - x4;
    }
    if (L.fscale != 0.0) {
        const R fa = static_cast<R>(WD[1] * L.fscale), fd = static_cast<R>(WD[5] * L.fscale);
        d[1] += fa;
        d[3] -= fa;
        d[5] += fd;
        d[6] -= fd;
        d[7] -= fd;
        d[8] += fd;
    }
}

// ------------------------------------------------------------------ general path

template <int ST> struct Ctx {
    using S = typename Form<ST>::store;
    const LatticeDev& L;
    const S* __restrict__ base;  // plane 0 of this lattice in the source buffer
    const uint8_t* __restrict__ flags;
    const int32_t* __restrict__ link_slot;
    const LinkRec* __restrict__ links;
    float inv_scale;
};

// Post-stream, pre-boundary deviations at node (px,py), read from the
// post-collision buffer (pull form of the streaming loops src/lbm.hpp:408-484).
template <int ST>
__device__ __forceinline__ void gather_general(const Ctx<ST>& c, int px, int py,
                                            typename Form<ST>::real d[9])
{
    using R = typename Form<ST>::real;
    const LatticeDev& L = c.L;
    const int64_t idx = static_cast<int64_t>(py) * L.pitch + px;
    const uint8_t fl = c.flags[L.cell_off + idx];
    if (fl & CELL_SOLID) {
        // solid nodes are never streamed into: f_next stays zero (src/lbm.hpp:400)
#pragma unroll
        for (int j = 0; j < 9; ++j) d[j] = -wgt<R>(j);
        return;
    }
    R own[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) own[i] = Form<ST>::dec(c.base[i * L.plane + idx], c.inv_scale);

    const LinkRec* rec = nullptr;
    uint32_t mask = 0;
    if (fl & CELL_LINKS) {
        rec = &c.links[L.link_off + c.link_slot[L.cell_off + idx]];
        mask = rec->mask;
    }
    R rho_own = R(1);
    if (mask) {
        R s = R(0);
#pragma unroll
        for (int i = 0; i < 9; ++i) s += own[i];
        rho_own += s;
    }

#pragma unroll
    for (int j = 0; j < 9; ++j) {
        const int i = OPP[j];  // push direction whose target is px - e_j
        if (mask & (1u << i)) {
            // boundary link (px,py) -> (px,py)+e_i solid: value for slot j = opp(i)
            const int scheme = (rec->scheme >> (2 * i)) & 3;
            const R q = static_cast<R>(rec->q[i]);
            const R uwx = static_cast<R>(rec->uwx[i]), uwy = static_cast<R>(rec->uwy[i]);
            const R eu = R(CX[i]) * uwx + R(CY[i]) * uwy;
            R val;
            if (scheme == BB_MEI) {
                // Mei / Filippova-Hanel (src/lbm.hpp:234-256), as a deviation from w_i
                const R uu = uwx * uwx + uwy * uwy;
                const R feq_dev =
                    wgt<R>(i) * (rho_own * (R(1) + R(3) * eu + R(4.5) * eu * eu - R(1.5) * uu) - R(1));
                val = q * feq_dev + (R(1) - q) * own[i] + (R(2) * q - R(1)) * wgt<R>(i) * rho_own * eu * R(3);
            } else if (scheme == BB_BOUZIDI) {
                // Bouzidi (src/lbm.hpp:258-273) + Ladd term (src/lbm.hpp:276-285)
                if (q < R(0.5)) {
                    const int sx = px - CX[i], sy = py - CY[i];
                    if (sx >= 0 && sx < L.nx && sy >= 0 && sy < L.ny) {
                        const R up = Form<ST>::dec(
                            c.base[i * L.plane + static_cast<int64_t>(sy) * L.pitch + sx], c.inv_scale);
                        // the reference reads f (not f - w): w_i cancels since weights sum to 1
                        val = R(2) * q * own[i] + (R(1) - R(2) * q) * up;
                    } else {
                        val = own[i];
                    }
                } else {
                    const R h = R(1) / (R(2) * q);
                    val = h * own[i] + (R(1) - h) * own[j];
                }
                val -= R(2) * wgt<R>(i) * rho_own * eu * R(3);
            } else {
                val = own[i];
            }
            d[j] = val;
            continue;
        }
        int sx = px - CX[j], sy = py - CY[j];
        bool wall = false, open = false;
        if (sx < 0 || sx >= L.nx) {
            if (L.x_edge == EDGE_PERIODIC) sx = sx < 0 ? sx + L.nx : sx - L.nx;
            else if (L.x_edge == EDGE_WALL) wall = true;
            else open = true;
        }
        if (sy < 0 || sy >= L.ny) {
            if (L.y_edge == EDGE_PERIODIC) sy = sy < 0 ? sy + L.ny : sy - L.ny;
            else if (L.y_edge == EDGE_WALL) wall = true;
            else open = true;
        }
        if (wall) {
            d[j] = own[i];  // bounced into the same node (src/lbm.hpp:418-419)
        } else if (open) {
            d[j] = -wgt<R>(j);  // dropped: f_next keeps its zero (src/lbm.hpp:470)
        } else {
            d[j] = Form<ST>::dec(c.base[j * L.plane + static_cast<int64_t>(sy) * L.pitch + sx],
                                 c.inv_scale);
        }
    }
}

// Zou/He velocity inlet on deviations (src/lbm.hpp:21-34 / 40-58); w1=w3, w5=w7=w6=w8
// so the closure keeps its form, and w0+w2+w4+2(w3+w6+w7) = 1.
template <typename R> __device__ __forceinline__ void zou_he(R d[9], R u)
{
    const R rho = (R(1) + d[0] + d[2] + d[4] + R(2) * (d[3] + d[6] + d[7])) / (R(1) - u);
    d[1] = d[3] + R(2.0 / 3.0) * rho * u;
    d[5] = d[7] + R(1.0 / 6.0) * rho * u;
    d[8] = d[6] + R(1.0 / 6.0) * rho * u;
}

// Node state after streaming and the open boundaries (src/lbm.hpp:530-539).
// Boundary order: inlet, outlet copy, lid.  Solid nodes skip the inlet and the lid,
// which leave their zero populations unchanged.
template <int ST>
__device__ __forceinline__ void state_general(const Ctx<ST>& c, int x, int y, bool solid,
                              typename Form<ST>::real d[9])
{
    using R = typename Form<ST>::real;
    const LatticeDev& L = c.L;
    const bool outlet_col = L.outlet && x == L.nx - 1;
    const int gx = outlet_col ? L.nx - 2 : x;  // enforce_outflow copies x = nx-2
    gather_general<ST>(c, gx, y, d);
    if (gx == 0 && L.inlet != INLET_NONE) {
        const bool gsolid = (c.flags[L.cell_off + static_cast<int64_t>(y) * L.pitch + gx] & CELL_SOLID) != 0;
        if (!gsolid) {
            if (L.inlet == INLET_ZOU_HE) {
                zou_he<R>(d, static_cast<R>(L.u_in));
            } else if (y >= L.inlet_y0) {
                const double yy = static_cast<double>(y - L.inlet_y0);
                const double h = static_cast<double>(L.inlet_h);
                zou_he<R>(d, static_cast<R>(L.u_in * 4.0 * yy * (h - yy) / (h * h)));
            }
        }
    }
    if (L.lid && y == L.ny - 1 && !solid) {
        // enforce_lid: feq(rho, u_lid, 0) (src/lbm.hpp:110-120)
        R sum = R(0);
#pragma unroll
        for (int i = 0; i < 9; ++i) sum += d[i];
        const R rho = R(1) + sum;
        const R u = static_cast<R>(L.u_in);
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            const R eu = R(CX[i]) * u;
            d[i] = wgt<R>(i) * (sum + rho * (R(3) * eu + R(4.5) * eu * eu - R(1.5) * u * u));
        }
    }
}

// ------------------------------------------------------------------ reductions

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Per-thread force partial keyed by slot; a second slot goes straight to memory.
struct ForcePart {
    int slot = -1;
    double fx = 0.0, fy = 0.0;
    __device__ __forceinline__ void add(int s, double x, double y, double* accp)
    {
        if (slot < 0 || slot == s) {
            slot = s;
            fx += x;
            fy += y;
        } else {
            atomicAdd(accp + 2 * s, x);
            atomicAdd(accp + 2 * s + 1, y);
        }
    }
};

// Wave-level reduction of the per-thread partials, one atomic per distinct slot.
__device__ __forceinline__ void flush_forces(ForcePart p, double* accp)
{
    int my = p.slot;
    const int lane = __lane_id();
    for (;;) {
        const uint64_t m = __ballot(my >= 0);
        if (m == 0) break;
        const int lead = __ffsll(static_cast<long long>(m)) - 1;
        const int s = __shfl(my, lead, 64);
        const double vx = wave_sum(my == s ? p.fx : 0.0);
        const double vy = wave_sum(my == s ? p.fy : 0.0);
        if (lane == lead) {
            atomicAdd(accp + 2 * s, vx);
            atomicAdd(accp + 2 * s + 1, vy);
        }
        if (my == s) my = -1;
    }
}

// ------------------------------------------------------------------ step kernel

template <int ST, int BX, int ROWS>
__global__ __launch_bounds__(BX) void step_kernel(StepArgs a)
{
    using S = typename Form<ST>::store;
    using R = typename Form<ST>::real;

    int lat_id, y0, x0;
    if (a.tiles) {
        const Tile t = a.tiles[blockIdx.x];
        lat_id = t.lattice;
        y0 = t.y0;
        x0 = t.x0;
    } else {
        lat_id = a.single;
        const int tx = blockIdx.x % a.tiles_x;
        const int ty = blockIdx.x / a.tiles_x;
        y0 = ty * ROWS;
        x0 = tx * BX;
    }
    const LatticeDev& L = a.lat[lat_id];
    const int64_t k = a.abs_mode ? a.k_abs : a.steps[lat_id] + a.step_arg;
    const int sp = a.abs_mode ? a.src_par : (a.par[lat_id] ^ (a.step_arg & 1));
    const S* src = static_cast<const S*>(sp ? a.buf1 : a.buf0);
    S* dst = static_cast<S*>(a.dst_override ? a.dst_override : (sp ? a.buf0 : a.buf1));
    double* accp = a.acc + static_cast<int64_t>(k & 1) * a.total_slots * 2;

    // Step k-1 is complete: move its force totals to the history, clear its accumulator.
    if (blockIdx.x == 0) {
        for (int l = a.flush_begin + static_cast<int>(threadIdx.x); l < a.flush_begin + a.flush_count;
             l += BX) {
            const int64_t kl = a.abs_mode ? a.k_abs : a.steps[l] + a.step_arg;
            if (kl < 1) continue;
            const LatticeDev& Ll = a.lat[l];
            double* prev = a.acc + static_cast<int64_t>((kl - 1) & 1) * a.total_slots * 2;
            const int64_t row = (kl - 1) % a.R;
            for (int s = 0; s < Ll.n_slots; ++s) {
                const int gs = Ll.slot_base + s;
                a.hist[(static_cast<int64_t>(gs) * a.R + row) * 2] = prev[2 * gs];
                a.hist[(static_cast<int64_t>(gs) * a.R + row) * 2 + 1] = prev[2 * gs + 1];
                prev[2 * gs] = 0.0;
                prev[2 * gs + 1] = 0.0;
            }
        }
    }

    const Ctx<ST> c{L, src + L.f_off, a.flags, a.link_slot, a.links, a.inv_scale};
    ForcePart part;
    bool bad = false;
    const int x = x0 + static_cast<int>(threadIdx.x);

#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
        const int y = y0 + r;
        if (x >= L.nx || y >= L.ny) continue;
        const int64_t idx = static_cast<int64_t>(y) * L.pitch + x;
        const uint8_t fl = a.flags[L.cell_off + idx];
        R d[9];
        const bool solid = (fl & CELL_SOLID) != 0;
        if (!(fl & CELL_SPECIAL)) {
            const S* b = c.base + idx;
#pragma unroll
            for (int j = 0; j < 9; ++j)
                d[j] = Form<ST>::dec(b[j * L.plane - (CY[j] * L.pitch + CX[j])], a.inv_scale);
        } else {
            if (solid && !(fl & CELL_WRITE)) continue;
            state_general<ST>(c, x, y, solid, d);
            if ((fl & CELL_LINKS) && L.forces) {
                // momentum exchange on the post-boundary state (src/lbm.hpp:542-581)
                const LinkRec& rec = a.links[L.link_off + a.link_slot[L.cell_off + idx]];
#pragma unroll
                for (int i = 1; i < 9; ++i) {
                    if (!(rec.mask & (1u << i)) || rec.slot[i] < 0) continue;
                    const double fb = WD[OPP[i]] + static_cast<double>(d[OPP[i]]);
                    part.add(L.slot_base + rec.slot[i], CX[i] * 2.0 * fb, CY[i] * 2.0 * fb, accp);
                }
            }
        }
        if (!solid) {
            R wd = R(0);
            if (L.les) {
                const uint16_t w = a.wdist[L.cell_off + idx];
                wd = w == 0xffffu ? R(1e9) : static_cast<R>(w);
            }
            collide<R>(d, L, wd);
            if (!(d[0] == d[0])) bad = true;
        }
        S* o = dst + L.f_off + idx;
#pragma unroll
        for (int j = 0; j < 9; ++j) o[j * L.plane] = Form<ST>::enc(d[j], a.scale);
    }
    if (bad) atomicOr(a.nan_flags + lat_id, 1);
    flush_forces(part, accp);
}

// ------------------------------------------------------------------ state / prologue

template <int ST>
__global__ __launch_bounds__(256) void state_kernel(StateArgs a)
{
    using S = typename Form<ST>::store;
    using R = typename Form<ST>::real;
    const LatticeDev& L = a.lat[a.lattice];
    const int x = blockIdx.x * 256 + threadIdx.x;
    const int y = a.y0 + blockIdx.y;
    double mass = 0.0;
    if (x < L.nx && y < a.y1) {
        const int64_t idx = static_cast<int64_t>(y) * L.pitch + x;
        const uint8_t fl = a.flags[L.cell_off + idx];
        const S* base = static_cast<const S*>(a.src) + L.f_off;
        R d[9];
        const bool solid = (fl & CELL_SOLID) != 0;
        if (a.from_state) {
#pragma unroll
            for (int j = 0; j < 9; ++j) d[j] = Form<ST>::dec(base[j * L.plane + idx], a.inv_scale);
        } else if (!(fl & CELL_SPECIAL)) {
#pragma unroll
            for (int j = 0; j < 9; ++j)
                d[j] = Form<ST>::dec(base[idx + j * L.plane - (CY[j] * L.pitch + CX[j])], a.inv_scale);
        } else {
            const Ctx<ST> c{L, base, a.flags, a.link_slot, a.links, a.inv_scale};
            state_general<ST>(c, x, y, solid, d);
        }
        double* o = a.out + (static_cast<int64_t>(y - a.y0) * L.nx + x) * 9;
#pragma unroll
        for (int j = 0; j < 9; ++j) {
            o[j] = WD[j] + static_cast<double>(d[j]);
            if (!solid) mass += static_cast<double>(d[j]);
        }
    }
    mass = wave_sum(mass);
    if (__lane_id() == 0 && mass != 0.0) atomicAdd(a.mass_dev, mass);
}

template <int ST>
__global__ __launch_bounds__(256) void prologue_kernel(PrologueArgs a)
{
    using S = typename Form<ST>::store;
    using R = typename Form<ST>::real;
    const LatticeDev& L = a.lat[a.lattice];
    const int x = blockIdx.x * 256 + threadIdx.x;
    const int y = a.y0 + blockIdx.y;
    if (x >= L.nx || y >= a.y1) return;
    const int64_t idx = static_cast<int64_t>(y) * L.pitch + x;
    const double* in = a.in + (static_cast<int64_t>(y - a.y0) * L.nx + x) * 9;
    R d[9];
#pragma unroll
    for (int j = 0; j < 9; ++j) d[j] = static_cast<R>(in[j] - WD[j]);
    S* st = static_cast<S*>(a.state_dst) + L.f_off + idx;
    S* co = static_cast<S*>(a.coll_dst) + L.f_off + idx;
#pragma unroll
    for (int j = 0; j < 9; ++j) st[j * L.plane] = Form<ST>::enc(d[j], a.scale);
    if (!(a.flags[L.cell_off + idx] & CELL_SOLID)) {
        R wd = R(0);
        if (L.les) {
            const uint16_t w = a.wdist[L.cell_off + idx];
            wd = w == 0xffffu ? R(1e9) : static_cast<R>(w);
        }
        collide<R>(d, L, wd);
    }
#pragma unroll
    for (int j = 0; j < 9; ++j) co[j * L.plane] = Form<ST>::enc(d[j], a.scale);
}

// Advances the per-lattice step counters after a batch of launches; an odd batch
// leaves the state in the other buffer.
__global__ void add_steps_kernel(int64_t* steps, int32_t* par, int32_t begin, int32_t count, int64_t inc)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) {
        steps[begin + i] += inc;
        par[begin + i] ^= static_cast<int32_t>(inc & 1);
    }
}

template <int ST, int BX, int ROWS>
hipError_t launch_step_t(int n_blocks, const StepArgs& a, hipStream_t s)
{
    hipLaunchKernelGGL((step_kernel<ST, BX, ROWS>), dim3(n_blocks), dim3(BX), 0, s, a);
    return hipGetLastError();
}

template <int ST>
hipError_t launch_step_st(int bx, int rows, int n_blocks, const StepArgs& a, hipStream_t s)
{
#define LBM_SHAPE(B, RW) \
    if (bx == B && rows == RW) return launch_step_t<ST, B, RW>(n_blocks, a, s);
    LBM_SHAPE(64, 1) LBM_SHAPE(64, 2) LBM_SHAPE(64, 4) LBM_SHAPE(64, 8)
    LBM_SHAPE(128, 1) LBM_SHAPE(128, 2) LBM_SHAPE(128, 4) LBM_SHAPE(128, 8)
    LBM_SHAPE(256, 1) LBM_SHAPE(256, 2) LBM_SHAPE(256, 4) LBM_SHAPE(256, 8)
    LBM_SHAPE(512, 1) LBM_SHAPE(512, 2) LBM_SHAPE(512, 4) LBM_SHAPE(512, 8)
#undef LBM_SHAPE
    return hipErrorInvalidValue;
}

}  // namespace

bool step_shape_supported(int bx, int rows)
{
    return (bx == 64 || bx == 128 || bx == 256 || bx == 512) &&
           (rows == 1 || rows == 2 || rows == 4 || rows == 8);
}

hipError_t launch_step(int storage, int bx, int rows, int n_blocks, const StepArgs& a, hipStream_t s)
{
    switch (storage) {
    case ST_F64: return launch_step_st<ST_F64>(bx, rows, n_blocks, a, s);
    case ST_F32: return launch_step_st<ST_F32>(bx, rows, n_blocks, a, s);
    case ST_F16S: return launch_step_st<ST_F16S>(bx, rows, n_blocks, a, s);
    case ST_BF16S: return launch_step_st<ST_BF16S>(bx, rows, n_blocks, a, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_state(int storage, const StateArgs& a, int nx, hipStream_t s)
{
    dim3 grid((nx + 255) / 256, a.y1 - a.y0);
    switch (storage) {
    case ST_F64: hipLaunchKernelGGL(state_kernel<ST_F64>, grid, dim3(256), 0, s, a); break;
    case ST_F32: hipLaunchKernelGGL(state_kernel<ST_F32>, grid, dim3(256), 0, s, a); break;
    case ST_F16S: hipLaunchKernelGGL(state_kernel<ST_F16S>, grid, dim3(256), 0, s, a); break;
    case ST_BF16S: hipLaunchKernelGGL(state_kernel<ST_BF16S>, grid, dim3(256), 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_prologue(int storage, const PrologueArgs& a, int nx, hipStream_t s)
{
    dim3 grid((nx + 255) / 256, a.y1 - a.y0);
    switch (storage) {
    case ST_F64: hipLaunchKernelGGL(prologue_kernel<ST_F64>, grid, dim3(256), 0, s, a); break;
    case ST_F32: hipLaunchKernelGGL(prologue_kernel<ST_F32>, grid, dim3(256), 0, s, a); break;
    case ST_F16S: hipLaunchKernelGGL(prologue_kernel<ST_F16S>, grid, dim3(256), 0, s, a); break;
    case ST_BF16S: hipLaunchKernelGGL(prologue_kernel<ST_BF16S>, grid, dim3(256), 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_add_steps(int64_t* steps, int32_t* par, int32_t begin, int32_t count, int64_t inc,
                            hipStream_t s)
{
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(add_steps_kernel, dim3((count + 255) / 256), dim3(256), 0, s, steps, par, begin, count,
                       inc);
    return hipGetLastError();
}

}  // namespace lbm
