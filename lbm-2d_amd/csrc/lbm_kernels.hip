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
//
// dec/enc convert between a stored value and the deviation f - w; dec_raw/enc_raw
// keep the stored scale (h = 2^k (f - w - s_i)), for arithmetic done in scaled units.
// The 16-bit forms store the deviation less a per-direction constant s_i (the lattice's
// storage shift, zero unless set), passed as `sh`; the FP64 and FP32 forms ignore it.

template <int ST> struct Form;
template <> struct Form<ST_F64> {
    using store = double;
    using real = double;
    static constexpr bool shifted = false;
    static __device__ __forceinline__ real dec(store v, float, float) { return v; }
    static __device__ __forceinline__ store enc(real d, float, float) { return d; }
    static __device__ __forceinline__ real dec_raw(store v) { return v; }
    static __device__ __forceinline__ store enc_raw(real d) { return d; }
};
template <> struct Form<ST_F32> {
    using store = float;
    using real = float;
    static constexpr bool shifted = false;
    static __device__ __forceinline__ real dec(store v, float, float) { return v; }
    static __device__ __forceinline__ store enc(real d, float, float) { return d; }
    static __device__ __forceinline__ real dec_raw(store v) { return v; }
    static __device__ __forceinline__ store enc_raw(real d) { return d; }
};
template <> struct Form<ST_F16S> {
    using store = uint16_t;
    using real = float;
    static constexpr bool shifted = true;
    static __device__ __forceinline__ real dec_raw(store v) { return static_cast<float>(__builtin_bit_cast(_Float16, v)); }
    // Conversion to binary16 rounds to nearest even (v_cvt_f16_f32).
    static __device__ __forceinline__ store enc_raw(real h) { return __builtin_bit_cast(uint16_t, static_cast<_Float16>(h)); }
    static __device__ __forceinline__ real dec(store v, float inv_scale, float sh) { return dec_raw(v) * inv_scale + sh; }
    static __device__ __forceinline__ store enc(real d, float scale, float sh) { return enc_raw((d - sh) * scale); }
};
template <> struct Form<ST_BF16S> {
    using store = uint16_t;
    using real = float;
    static constexpr bool shifted = true;
    static __device__ __forceinline__ real dec_raw(store v) { return __uint_as_float(static_cast<uint32_t>(v) << 16); }
    // Round to nearest even (v_cvt_pk_bf16_f32 on gfx950); NaN stays NaN.
    static __device__ __forceinline__ store enc_raw(real h) { return __builtin_bit_cast(uint16_t, static_cast<__bf16>(h)); }
    static __device__ __forceinline__ real dec(store v, float inv_scale, float sh) { return dec_raw(v) * inv_scale + sh; }
    static __device__ __forceinline__ store enc(real d, float scale, float sh) { return enc_raw((d - sh) * scale); }
};

template <typename R> __device__ __forceinline__ R wgt(int i) { return static_cast<R>(WD[i]); }

// ------------------------------------------------------------------ collision

template <typename R> __device__ __forceinline__ R clamp_rate(R s)
{
    return s < R(0.5) ? R(0.5) : (s > R(1.99) ? R(1.99) : s);
}

// FP32: the hardware square root, reciprocal and exp2 (1 ulp class); FP64: correctly rounded.
template <typename R> __device__ __forceinline__ R dsqrt(R x);
template <> __device__ __forceinline__ float dsqrt<float>(float x) { return __builtin_amdgcn_sqrtf(x); }
template <> __device__ __forceinline__ double dsqrt<double>(double x) { return __builtin_sqrt(x); }
template <typename R> __device__ __forceinline__ R recip(R x);
template <> __device__ __forceinline__ float recip<float>(float x) { return __builtin_amdgcn_rcpf(x); }
template <> __device__ __forceinline__ double recip<double>(double x) { return 1.0 / x; }
template <typename R> __device__ __forceinline__ R dexp(R x);
template <> __device__ __forceinline__ float dexp<float>(float x) { return __expf(x); }
template <> __device__ __forceinline__ double dexp<double>(double x) { return exp(x); }

// MRT (src/lbm.hpp:128-218) or BGK (src/lbm.hpp:334-354), then the body force
// (src/lbm.hpp:327-331), on d_i = sc (f_i - w_i): sc = 1 for plain deviations, 2^k
// when a 16-bit form computes in its stored units.
//
// The MRT relaxation f <- f - M^-1 S (m - m_eq) needs only the six non-conserved
// moment differences (rho, j conserved).  With w subtracted and the rates of
// MRTParams::from_tau (s_bulk = 1.2, s_normal = 1; src/lbm_types.hpp:72-76) folded in:
//   e - e_eq     = -2 d0 + S_ax + 4 S_dg - 3 J2          (S_ax = d1+d2+d3+d4,
//   eps - eps_eq = 3 (d0 - S_ax + J2)                      S_dg = d5+d6+d7+d8,
//   qx - qx_eq   = 2 jx - 3 (d1 - d3)                      J2 = rho |u|^2)
//   qy - qy_eq   = 2 jy - 3 (d2 - d4)
//   pxx - pxx_eq = (d1 + d3) - (d2 + d4) - (jx u - jy v)
//   pxy - pxy_eq = (d5 + d7) - (d6 + d8) - jx v
// and the rows of M^-1 regroup into the few sums below.
template <typename R>
__device__ __forceinline__ void collide(R d[9], const LatticeDev& L, R wall_dist, R sc)
{
    const R a13 = d[1] + d[3], s13 = d[1] - d[3];
    const R a24 = d[2] + d[4], s24 = d[2] - d[4];
    const R a57 = d[5] + d[7], s57 = d[5] - d[7];
    const R a68 = d[6] + d[8], s68 = d[6] - d[8];
    const R sax = a13 + a24, sdg = a57 + a68;
    const R sum = d[0] + sax + sdg;   // sc (rho - 1)
    const R jx = s13 + s57 - s68;     // sc rho u
    const R jy = s24 + s57 + s68;     // sc rho v
    const R rho_s = sc + sum;         // sc rho
    const R ir = recip<R>(rho_s);
    const R u = jx * ir, v = jy * ir; // velocity (unscaled)

    if (L.collision != 0) {
        // BGK: d_i -= (d_i - sc (feq_i - w_i)) / tau
        const R it = static_cast<R>(L.inv_tau);
        const R usq = u * u + v * v;
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            const R eu = R(CX[i]) * u + R(CY[i]) * v;
            const R feq_dev = wgt<R>(i) * (sum + rho_s * (R(3) * eu + R(4.5) * eu * eu - R(1.5) * usq));
            d[i] -= it * (d[i] - feq_dev);
        }
    } else {
        const R J2 = jx * u + jy * v;                                   // sc rho |u|^2
        const R ne = R(-2) * d[0] + sax + R(4) * sdg - R(3) * J2;        // e - e_eq
        const R e3 = d[0] - sax + J2;                                    // (eps - eps_eq) / 3
        const R npxx = (a13 - a24) - (jx * u - jy * v);
        const R npxy = (a57 - a68) - jx * v;
        R ss = static_cast<R>(L.s_shear);
        if (L.les) {
            // Smagorinsky with Van Driest damping (src/lbm.hpp:155-182, y+ :314-325).
            // The damped block replaces the plain one whenever y+ > 0.
            R cs2 = static_cast<R>(L.cs2);
            const R nu = static_cast<R>(L.nu);
            if (nu > R(0)) {
                const R y_plus = wall_dist * dsqrt<R>(u * u + v * v + R(1e-12)) * static_cast<R>(L.inv_nu);
                if (y_plus > R(0)) {
                    const R damp = R(1) - dexp<R>(y_plus * R(-1.0 / 26.0));
                    const R csd = static_cast<R>(L.cs) * damp;
                    cs2 = csd * csd;
                }
            }
            // A = 9 cs^2 Q / (2 rho); Q/rho is scale free
            const R Q = dsqrt<R>(R(2) * (npxx * npxx + npxy * npxy));
            const R A = R(4.5) * cs2 * Q * ir;
            const R tau = static_cast<R>(L.tau);
            const R tau_eff = (tau + dsqrt<R>(tau * tau + R(4) * A)) * R(0.5);
            ss = clamp_rate<R>(recip<R>(tau_eff));
        }
        // M^-1 S (m - m_eq), s_bulk = 1.2 and s_normal = 1
        const R a = ne * R(-1.2 / 36.0) - e3 * R(3.6 / 18.0);
        const R b = ne * R(1.2 / 18.0) + e3 * R(3.6 / 36.0);
        const R c0 = e3 * R(3.6 / 9.0) - ne * R(1.2 / 9.0);
        const R qx6 = jx * R(1.0 / 3.0) - s13 * R(0.5);  // (qx - qx_eq) / 6
        const R qy6 = jy * R(1.0 / 3.0) - s24 * R(0.5);
        const R ss4 = ss * R(0.25);
        const R p4 = ss4 * npxx, x4 = ss4 * npxy;
        const R t = a + p4, t2 = a - p4;
        const R u1 = b + x4, u2 = b - x4;
        const R qp = (qx6 + qy6) * R(0.5), qm = (qx6 - qy6) * R(0.5);
        d[0] -= c0;
        d[1] = d[1] - t + qx6;
        d[3] = d[3] - t - qx6;
        d[2] = d[2] - t2 + qy6;
        d[4] = d[4] - t2 - qy6;
        d[5] -= u1 + qp;
        d[7] -= u1 - qp;
        d[6] -= u2 - qm;
        d[8] -= u2 + qm;
    }
    if (L.fscale != 0.0) {
        const R fa = static_cast<R>(WD[1] * L.fscale) * sc, fd = static_cast<R>(WD[5] * L.fscale) * sc;
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
    for (int i = 0; i < 9; ++i) own[i] = Form<ST>::dec(c.base[i * L.plane + idx], c.inv_scale, L.shift[i]);

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
                            c.base[i * L.plane + static_cast<int64_t>(sy) * L.pitch + sx], c.inv_scale, L.shift[i]);
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
                                 c.inv_scale, L.shift[j]);
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

// Store of a value or an aligned pack of values, optionally nontemporal.
template <typename T> __device__ __forceinline__ void store_val(T* p, const T& v, bool nt)
{
    if constexpr (sizeof(T) == 16) {
        using U = uint32_t __attribute__((ext_vector_type(4)));
        const U u = __builtin_bit_cast(U, v);
        if (nt) __builtin_nontemporal_store(u, reinterpret_cast<U*>(p));
        else *reinterpret_cast<U*>(p) = u;
    } else if constexpr (sizeof(T) == 8) {
        using U = uint32_t __attribute__((ext_vector_type(2)));
        const U u = __builtin_bit_cast(U, v);
        if (nt) __builtin_nontemporal_store(u, reinterpret_cast<U*>(p));
        else *reinterpret_cast<U*>(p) = u;
    } else if constexpr (sizeof(T) == 4) {
        const uint32_t u = __builtin_bit_cast(uint32_t, v);
        if (nt) __builtin_nontemporal_store(u, reinterpret_cast<uint32_t*>(p));
        else *reinterpret_cast<uint32_t*>(p) = u;
    } else {
        static_assert(sizeof(T) == 2, "store size");
        const uint16_t u = __builtin_bit_cast(uint16_t, v);
        if (nt) __builtin_nontemporal_store(u, reinterpret_cast<uint16_t*>(p));
        else *reinterpret_cast<uint16_t*>(p) = u;
    }
}

// One node, one cell per thread: the bulk pull or the general path.
template <int ST>
__device__ __forceinline__ void cell_update(const LatticeDev& L, const StepArgs& a,
                                            const Ctx<ST>& c, typename Form<ST>::store* dst,
                                            double* accp, int x, int y, ForcePart& part, bool& bad)
{
    using S = typename Form<ST>::store;
    using R = typename Form<ST>::real;
    const int64_t idx = static_cast<int64_t>(y) * L.pitch + x;
    const uint8_t fl = a.flags[L.cell_off + idx];
    R d[9];
    const bool solid = (fl & CELL_SOLID) != 0;
    if (!(fl & CELL_SPECIAL)) {
        const S* b = c.base + idx;
#pragma unroll
        for (int j = 0; j < 9; ++j)
            d[j] = Form<ST>::dec(b[j * L.plane - (CY[j] * L.pitch + CX[j])], a.inv_scale, L.shift[j]);
    } else {
        if (solid && !(fl & CELL_WRITE)) return;
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
        collide<R>(d, L, wd, R(1));
        if (!(d[0] == d[0])) bad = true;
    }
    S* o = dst + L.f_off + idx;
#pragma unroll
    for (int j = 0; j < 9; ++j)
        store_val(o + j * L.plane, Form<ST>::enc(d[j], a.scale, L.shift[j]), a.options & LAUNCH_STREAM_STORES);
}

// Whole-wave lane shifts (DPP wave_shr:1 / wave_shl:1): lane i receives lane i-1 /
// lane i+1; lanes without a source keep their own value.
__device__ __forceinline__ uint32_t from_left_lane(uint32_t v)
{
    return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(static_cast<int>(v), static_cast<int>(v), 0x138, 0xf,
                                                              0xf, false));
}
__device__ __forceinline__ uint32_t from_right_lane(uint32_t v)
{
    return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(static_cast<int>(v), static_cast<int>(v), 0x130, 0xf,
                                                              0xf, false));
}
template <typename S> __device__ __forceinline__ S shift_in(S v, bool left)
{
    if constexpr (sizeof(S) == 8) {
        const uint64_t u = __builtin_bit_cast(uint64_t, v);
        uint32_t lo = static_cast<uint32_t>(u), hi = static_cast<uint32_t>(u >> 32);
        lo = left ? from_left_lane(lo) : from_right_lane(lo);
        hi = left ? from_left_lane(hi) : from_right_lane(hi);
        return __builtin_bit_cast(S, (static_cast<uint64_t>(hi) << 32) | lo);
    } else if constexpr (sizeof(S) == 4) {
        const uint32_t u = __builtin_bit_cast(uint32_t, v);
        return __builtin_bit_cast(S, left ? from_left_lane(u) : from_right_lane(u));
    } else {
        const uint32_t u = static_cast<uint32_t>(v);
        return static_cast<S>(left ? from_left_lane(u) : from_right_lane(u));
    }
}

template <typename S, int V> struct alignas(sizeof(S) * V) Pack {
    S v[V];
};

// V consecutive bulk nodes of one row per thread.  Every plane moves as one aligned
// vector of V values; a population arriving from x-1 (x+1) for the first (last) node
// of the group comes from the neighbouring lane through a DPP shift, and the first
// (last) lane of the wave loads it itself.  Requires an interior row.
template <int ST, int V>
__device__ __forceinline__ void group_update(const LatticeDev& L, const StepArgs& a, const Ctx<ST>& c,
                                             typename Form<ST>::store* dst, int x, int y, bool bulk,
                                             bool& bad)
{
    using S = typename Form<ST>::store;
    using R = typename Form<ST>::real;
    using P = Pack<S, V>;
    const int64_t idx = static_cast<int64_t>(y) * L.pitch + x;
    P in[9];
#pragma unroll
    for (int j = 0; j < 9; ++j)
        in[j] = *reinterpret_cast<const P*>(c.base + j * L.plane + idx - CY[j] * L.pitch);
    S edge[9];
#pragma unroll
    for (int j = 0; j < 9; ++j) {
        if (CX[j] == 1) edge[j] = shift_in<S>(in[j].v[V - 1], true);
        if (CX[j] == -1) edge[j] = shift_in<S>(in[j].v[0], false);
    }
    if (!bulk) return;
    const int lane = __lane_id();
    if (lane == 0) {
#pragma unroll
        for (int j = 0; j < 9; ++j)
            if (CX[j] == 1) edge[j] = c.base[j * L.plane + idx - CY[j] * L.pitch - 1];
    }
    if (lane == 63) {
#pragma unroll
        for (int j = 0; j < 9; ++j)
            if (CX[j] == -1) edge[j] = c.base[j * L.plane + idx - CY[j] * L.pitch + V];
    }
    Pack<uint16_t, V> wd4;
    if (L.les) wd4 = *reinterpret_cast<const Pack<uint16_t, V>*>(a.wdist + L.cell_off + idx);
    P out[9];
#pragma unroll
    for (int k = 0; k < V; ++k) {
        R d[9];
#pragma unroll
        for (int j = 0; j < 9; ++j) {
            S s;
            if (CX[j] == 0) s = in[j].v[k];
            else if (CX[j] == 1) s = k == 0 ? edge[j] : in[j].v[k - 1];
            else s = k == V - 1 ? edge[j] : in[j].v[k + 1];
            d[j] = Form<ST>::dec_raw(s);
            if constexpr (Form<ST>::shifted) d[j] += L.shift_s[j];
        }
        R wd = R(0);
        if (L.les) wd = wd4.v[k] == 0xffffu ? R(1e9) : static_cast<R>(wd4.v[k]);
        collide<R>(d, L, wd, static_cast<R>(a.scale));  // in stored units
        if (!(d[0] == d[0])) bad = true;
#pragma unroll
        for (int j = 0; j < 9; ++j) {
            if constexpr (Form<ST>::shifted) d[j] -= L.shift_s[j];
            out[j].v[k] = Form<ST>::enc_raw(d[j]);
        }
    }
#pragma unroll
    for (int j = 0; j < 9; ++j)
        store_val(reinterpret_cast<P*>(dst + L.f_off + j * L.plane + idx), out[j], a.options & LAUNCH_STREAM_STORES);
}

template <int ST, int BX, int ROWS, int V>
__global__ __launch_bounds__(BX) void step_kernel(StepArgs a)
{
    using S = typename Form<ST>::store;

    // Workgroups are handed to the 8 XCDs round-robin; with LAUNCH_XCD_ORDER XCD x
    // takes one contiguous range of tiles, so lines shared by neighbouring tiles stay
    // in one L2.  The map is a bijection on [0, gridDim.x).
    int b = static_cast<int>(blockIdx.x);
    if (a.options & LAUNCH_XCD_ORDER) {
        const int nb = static_cast<int>(gridDim.x), q = nb >> 3, r = nb & 7, xcd = b & 7;
        b = xcd * q + (xcd < r ? xcd : r) + (b >> 3);
    }
    int lat_id, y0, x0;
    if (a.tiles) {
        const Tile t = a.tiles[b];
        lat_id = t.lattice;
        y0 = t.y0;
        x0 = t.x0;
    } else {
        lat_id = a.single;
        const int tx = b % a.tiles_x;
        const int ty = b / a.tiles_x;
        y0 = ty * ROWS;
        x0 = tx * BX * V;
    }
    const LatticeDev L = a.lat[lat_id];  // by value: scalar registers, loaded before any store
    const int64_t k = a.abs_mode ? a.k_abs : a.steps[lat_id] + a.step_arg;
    const int sp = a.abs_mode ? a.src_par : (a.par[lat_id] ^ (a.step_arg & 1));
    const S* src = static_cast<const S*>(sp ? a.buf1 : a.buf0);
    S* dst = static_cast<S*>(a.dst_override ? a.dst_override : (sp ? a.buf0 : a.buf1));
    double* accp = a.acc + (static_cast<int64_t>(k & 1) * kAccCopies + blockIdx.x % kAccCopies) * a.total_slots * 2;

    // Step k-1 is complete: move its force totals to the history, clear its accumulator.
    if (blockIdx.x == 0) {
        for (int l = a.flush_begin + static_cast<int>(threadIdx.x); l < a.flush_begin + a.flush_count;
             l += BX) {
            const int64_t kl = a.abs_mode ? a.k_abs : a.steps[l] + a.step_arg;
            if (kl < 1) continue;
            const LatticeDev& Ll = a.lat[l];
            double* prev = a.acc + static_cast<int64_t>((kl - 1) & 1) * kAccCopies * a.total_slots * 2;
            const int64_t row = (kl - 1) % a.R;
            for (int s = 0; s < Ll.n_slots; ++s) {
                const int gs = Ll.slot_base + s;
                double fx = 0.0, fy = 0.0;
                for (int cp = 0; cp < kAccCopies; ++cp) {
                    double* q = prev + (static_cast<int64_t>(cp) * a.total_slots + gs) * 2;
                    fx += q[0];
                    fy += q[1];
                    q[0] = 0.0;
                    q[1] = 0.0;
                }
                a.hist[(static_cast<int64_t>(gs) * a.R + row) * 2] = fx;
                a.hist[(static_cast<int64_t>(gs) * a.R + row) * 2 + 1] = fy;
            }
        }
    }

    const Ctx<ST> c{L, src + L.f_off, a.flags, a.link_slot, a.links, a.inv_scale};
    ForcePart part;
    bool bad = false;
    const int x = x0 + V * static_cast<int>(threadIdx.x);

#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
        const int y = y0 + r;
        if (y >= L.ny) break;  // uniform across the block
        bool done = false;
        if constexpr (V > 1) {
            if (y >= 1 && y <= L.ny - 2) {  // uniform across the block
                // the group is bulk when it lies inside [1, nx-2] and no node is special
                bool bulk = false;
                if (x + V - 1 <= L.nx - 2 && x >= 1) {
                    const Pack<uint8_t, V> f =
                        *reinterpret_cast<const Pack<uint8_t, V>*>(a.flags + L.cell_off + static_cast<int64_t>(y) * L.pitch + x);
                    uint32_t any = 0;
#pragma unroll
                    for (int q = 0; q < V; ++q) any |= f.v[q];
                    bulk = !(any & CELL_SPECIAL);
                }
                // every lane inside the padded row takes part in the lane shifts
                if (x < L.pitch) group_update<ST, V>(L, a, c, dst, x, y, bulk, bad);
                done = bulk;
            }
        }
        if (!done) {
            // one copy of the general path, looped (keeps the kernel small)
#pragma unroll 1
            for (int q = 0; q < V; ++q)
                if (x + q < L.nx) cell_update<ST>(L, a, c, dst, accp, x + q, y, part, bad);
        }
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
    const LatticeDev L = a.lat[a.lattice];
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
            for (int j = 0; j < 9; ++j) d[j] = Form<ST>::dec(base[j * L.plane + idx], a.inv_scale, L.shift[j]);
        } else if (!(fl & CELL_SPECIAL)) {
#pragma unroll
            for (int j = 0; j < 9; ++j)
                d[j] = Form<ST>::dec(base[idx + j * L.plane - (CY[j] * L.pitch + CX[j])], a.inv_scale, L.shift[j]);
        } else {
            const Ctx<ST> c{L, base, a.flags, a.link_slot, a.links, a.inv_scale};
            state_general<ST>(c, x, y, solid, d);
        }
        if (a.out) {
            double* o = a.out + (static_cast<int64_t>(y - a.y0) * L.nx + x) * 9;
#pragma unroll
            for (int j = 0; j < 9; ++j) o[j] = WD[j] + static_cast<double>(d[j]);
        }
        if (a.macros) {
            // compute_macros (src/lbm_types.hpp:278-288) on the populations w + d, in order
#pragma clang fp contract(off)
            double rho = 0.0, mx = 0.0, my = 0.0;
#pragma unroll
            for (int j = 0; j < 9; ++j) {
                const double f = WD[j] + static_cast<double>(d[j]);
                rho += f;
                mx += f * CX[j];
                my += f * CY[j];
            }
            const int64_t n = static_cast<int64_t>(a.y1 - a.y0) * L.nx;
            const int64_t o = static_cast<int64_t>(y - a.y0) * L.nx + x;
            a.macros[o] = rho;
            a.macros[n + o] = mx / rho;
            a.macros[2 * n + o] = my / rho;
        }
        if (!solid) {
#pragma unroll
            for (int j = 0; j < 9; ++j) mass += static_cast<double>(d[j]);
        }
        if (a.coll_dst && (!solid || (fl & CELL_WRITE))) {
            // the stored post-collision state of this step, with the current parameters
            if (!solid) {
                R wd = R(0);
                if (L.les) {
                    const uint16_t w = a.wdist[L.cell_off + idx];
                    wd = w == 0xffffu ? R(1e9) : static_cast<R>(w);
                }
                collide<R>(d, L, wd, R(1));
            }
            S* co = static_cast<S*>(a.coll_dst) + L.f_off + idx;
#pragma unroll
            for (int j = 0; j < 9; ++j) co[j * L.plane] = Form<ST>::enc(d[j], a.scale, L.shift[j]);
        }
    }
    if (a.mass_dev) {
        mass = wave_sum(mass);
        if (__lane_id() == 0 && mass != 0.0) atomicAdd(a.mass_dev, mass);
    }
}

template <int ST>
__global__ __launch_bounds__(256) void prologue_kernel(PrologueArgs a)
{
    using S = typename Form<ST>::store;
    using R = typename Form<ST>::real;
    const LatticeDev L = a.lat[a.lattice];
    const int x = blockIdx.x * 256 + threadIdx.x;
    const int y = a.y0 + blockIdx.y;
    if (x >= L.nx || y >= a.y1) return;
    const int64_t idx = static_cast<int64_t>(y) * L.pitch + x;
    R d[9];
    if (a.in) {
        const double* in = a.in + (static_cast<int64_t>(y - a.y0) * L.nx + x) * 9;
#pragma unroll
        for (int j = 0; j < 9; ++j) d[j] = static_cast<R>(in[j] - WD[j]);
    } else {
        // uniform equilibrium feq(rho, u, v) - w, as the case drivers' initial loop,
        // unfused so that it rounds as compute_equilibrium (src/lbm_types.hpp:269-273)
#pragma clang fp contract(off)
#pragma unroll
        for (int j = 0; j < 9; ++j) {
            const double eu = CX[j] * a.eq_u + CY[j] * a.eq_v;
            const double uu = a.eq_u * a.eq_u + a.eq_v * a.eq_v;
            d[j] = static_cast<R>(WD[j] * a.eq_rho * (1.0 + 3.0 * eu + 4.5 * eu * eu - 1.5 * uu) - WD[j]);
        }
    }
    S* st = static_cast<S*>(a.state_dst) + L.f_off + idx;
    S* co = static_cast<S*>(a.coll_dst) + L.f_off + idx;
#pragma unroll
    for (int j = 0; j < 9; ++j) st[j * L.plane] = Form<ST>::enc(d[j], a.scale, L.shift[j]);
    if (!(a.flags[L.cell_off + idx] & CELL_SOLID)) {
        R wd = R(0);
        if (L.les) {
            const uint16_t w = a.wdist[L.cell_off + idx];
            wd = w == 0xffffu ? R(1e9) : static_cast<R>(w);
        }
        collide<R>(d, L, wd, R(1));
    }
#pragma unroll
    for (int j = 0; j < 9; ++j) co[j * L.plane] = Form<ST>::enc(d[j], a.scale, L.shift[j]);
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

// A group of nodes moves at most 16 bytes per plane and thread (one dwordx4).
template <int ST, int V> constexpr bool group_fits() { return sizeof(typename Form<ST>::store) * V <= 16; }

template <int ST, int BX, int ROWS, int V>
hipError_t launch_step_t(int n_blocks, const StepArgs& a, hipStream_t s)
{
    if constexpr (!group_fits<ST, V>()) {
        return hipErrorInvalidValue;
    } else {
        hipLaunchKernelGGL((step_kernel<ST, BX, ROWS, V>), dim3(n_blocks), dim3(BX), 0, s, a);
        return hipGetLastError();
    }
}

// Launch shapes (threads per block, rows per block, nodes per thread) compiled in.
#define LBM_SHAPES(X)                                                                          \
    X(64, 1, 1) X(128, 1, 1) X(256, 1, 1) X(256, 2, 1) X(64, 1, 2) X(128, 1, 2) X(256, 1, 2) \
    X(64, 1, 4) X(128, 1, 4) X(256, 1, 4) X(64, 2, 4) X(128, 2, 4) X(64, 1, 8) X(128, 1, 8)

template <int ST>
hipError_t launch_step_st(int bx, int rows, int v, int n_blocks, const StepArgs& a, hipStream_t s)
{
#define LBM_SHAPE(B, RW, VV) \
    if (bx == B && rows == RW && v == VV) return launch_step_t<ST, B, RW, VV>(n_blocks, a, s);
    LBM_SHAPES(LBM_SHAPE)
#undef LBM_SHAPE
    return hipErrorInvalidValue;
}

}  // namespace

static int store_bytes(int storage) { return storage == ST_F64 ? 8 : (storage == ST_F32 ? 4 : 2); }

bool step_shape_supported(int storage, int bx, int rows, int v)
{
    if (store_bytes(storage) * v > 16) return false;
#define LBM_SHAPE(B, RW, VV) \
    if (bx == B && rows == RW && v == VV) return true;
    LBM_SHAPES(LBM_SHAPE)
#undef LBM_SHAPE
    return false;
}

int step_shapes(int storage, int (*out)[3], int max)
{
    int n = 0;
#define LBM_SHAPE(B, RW, VV)                              \
    if (store_bytes(storage) * VV <= 16) {                \
        if (n < max) {                                    \
            out[n][0] = B;                                \
            out[n][1] = RW;                               \
            out[n][2] = VV;                               \
        }                                                 \
        ++n;                                              \
    }
    LBM_SHAPES(LBM_SHAPE)
#undef LBM_SHAPE
    return n;
}

hipError_t launch_step(int storage, int bx, int rows, int v, int n_blocks, const StepArgs& a, hipStream_t s)
{
    switch (storage) {
    case ST_F64: return launch_step_st<ST_F64>(bx, rows, v, n_blocks, a, s);
    case ST_F32: return launch_step_st<ST_F32>(bx, rows, v, n_blocks, a, s);
    case ST_F16S: return launch_step_st<ST_F16S>(bx, rows, v, n_blocks, a, s);
    case ST_BF16S: return launch_step_st<ST_BF16S>(bx, rows, v, n_blocks, a, s);
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
