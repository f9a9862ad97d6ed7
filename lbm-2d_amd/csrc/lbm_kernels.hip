// lbm_kernels.hip -- D2Q9 lattice update kernels for gfx950 (MI355X).
//
// One fused launch per time step, in pull form.  The population buffers hold the
// post-collision state C_k (the reference's sys.f after the collision loop of
// src/lbm.hpp:300-354).  The launch for step k+1 at node x:
//   1. gathers C_k(x - e_j) for every direction j (streaming, src/lbm.hpp:405-484),
//      or the bounce-back value of the link when x - e_j is solid (src/lbm.hpp:226-288),
//   2. applies the open boundaries to that node (src/lbm.hpp:530-539),
//   3. stores its momentum-exchange terms in FP64 for its link record
//      (src/lbm.hpp:542-581), summed per slot in a fixed order after the batch,
//   4. collides (MRT / BGK, LES, body force) and writes C_{k+1}(x): reads come from one
//      buffer, writes go to the other (A-B pattern).
// Storage keeps deviations h_i = (f_i - w_i - s_i) * 2^k (k = 0 and s = 0 for F64/F32)
// in SoA planes of pitch-aligned rows; arithmetic is done on deviations in FP32 (FP64
// for F64).
//
// Bulk nodes (no edge, no solid neighbour) take a branch-free path of 9 shifted,
// coalesced row loads, 9 aligned stores and one flag byte.  Every other node takes
// the general path, which handles links, edges and boundary conditions exactly as
// the reference's branches do.  step2_kernel advances one lattice two steps per launch
// through LDS.  thermal_kernel steps a lattice that carries the thermal populations g
// as well (src/lbm.hpp:356-397, 489-528): the same pull form for f, the adiabatic
// streaming of g, the thermal collision and the Boussinesq forcing of f.  Set-up
// kernels: obstacle rasterisation, the per-node geometry pass (flags, link records),
// wall distance, initial state and perturbation.
#include <hip/hip_runtime.h>

#include <cstdint>

#include <type_traits>

#include "lbm_kernels.h"

namespace lbm {
namespace {

constexpr int CX[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
constexpr int CY[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
constexpr int OPP[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
constexpr double WD[9] = {4.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0, 1.0 / 9.0,
                          1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};

enum { EDGE_PERIODIC = 0, EDGE_OPEN = 1, EDGE_WALL = 2, EDGE_FREE_SLIP = 3 };
enum { INLET_NONE = 0, INLET_ZOU_HE = 1, INLET_PARABOLIC = 2, INLET_EQUILIBRIUM = 3 };
// direction with the same x velocity and the opposite y velocity (specular reflection
// off a free-slip y edge)
constexpr int MIRY[9] = {0, 1, 4, 3, 2, 8, 7, 6, 5};
enum { BB_SIMPLE = 0, BB_MEI = 1, BB_BOUZIDI = 2 };

// ------------------------------------------------------------------ storage forms
//
// dec/enc convert between a stored value and the deviation f - w; dec_raw/enc_raw
// keep the stored scale (h = 2^k (f - w - s_i)), for arithmetic done in scaled units.
// The 16-bit forms store the deviation less a per-direction constant s_i (the lattice's
// storage shift): dec adds sh = s_i, enc adds sh = -s_i (both -0.0f without a shift, an
// exact identity); the FP64 and FP32 forms ignore them.

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
    static __device__ __forceinline__ store enc(real d, float scale, float nsh) { return enc_raw((d + nsh) * scale); }
    // Stochastic rounding (v_cvt_sr_f16_f32): in the binary16 normal range the top 13
    // bits of rnd are added below the last kept mantissa bit and the sum truncated; below
    // it the instruction keeps rounding stochastically at the subnormal spacing.
    static __device__ __forceinline__ store enc_sr_raw(real h, uint32_t rnd)
    {
        typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
        const half2_t v = __builtin_amdgcn_cvt_sr_f16_f32(half2_t{0, 0}, h, rnd, false);
        return __builtin_bit_cast(uint16_t, v.x);
    }
};
template <> struct Form<ST_BF16S> {
    using store = uint16_t;
    using real = float;
    static constexpr bool shifted = true;
    static __device__ __forceinline__ real dec_raw(store v) { return __uint_as_float(static_cast<uint32_t>(v) << 16); }
    // Round to nearest even (v_cvt_pk_bf16_f32 on gfx950); NaN stays NaN.
    static __device__ __forceinline__ store enc_raw(real h) { return __builtin_bit_cast(uint16_t, static_cast<__bf16>(h)); }
    static __device__ __forceinline__ real dec(store v, float inv_scale, float sh) { return dec_raw(v) * inv_scale + sh; }
    static __device__ __forceinline__ store enc(real d, float scale, float nsh) { return enc_raw((d + nsh) * scale); }
    // Stochastic rounding (v_cvt_sr_bf16_f32): the top 16 bits of rnd are added below the
    // last kept bit and the sum truncated; infinities stay, NaN stays NaN.
    static __device__ __forceinline__ store enc_sr_raw(real h, uint32_t rnd)
    {
        typedef short short2_t __attribute__((ext_vector_type(2)));
        const short2_t v = __builtin_amdgcn_cvt_sr_bf16_f32(short2_t{0, 0}, h, rnd, false);
        return static_cast<uint16_t>(v.x);
    }
};

// Random bits for stochastic rounding: a per-launch key (seed and step) mixed once with
// the node's global index (lowbias32); each further word is the previous one times an
// odd constant, and a population takes one 16-bit half of a word.
__device__ __forceinline__ uint32_t mix32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

struct Rounding {
    bool on;        // stochastic rounding of the 16-bit forms
    uint32_t key;   // seed and step
    __device__ __forceinline__ uint32_t node(int64_t cell) const
    {
        return mix32(key ^ (static_cast<uint32_t>(cell) * 0x9e3779b1u));
    }
};

// Random bits for the next population in their top 16 bits: the high and low halves of
// a word in turn, a new word every second call.  h carries the word and a call count
// in its caller's loop index, so the sequence is fixed per node.
__device__ __forceinline__ uint32_t next_rand(uint32_t& h, int j)
{
    // h << 16 as a byte permute: written as a shift, the compiler folds it into the
    // multiply of the next word (a second 32-bit multiply, a quarter-rate instruction)
    if (j & 1) return __builtin_amdgcn_perm(h, h, 0x01000c0cu);
    if (j) h *= 0x85ebca6bu;
    return h & 0xffff0000u;
}

// Encode deviation d for storage (nsh: minus the shift), rounding to nearest even or
// stochastically.
template <int ST>
__device__ __forceinline__ typename Form<ST>::store enc_node(typename Form<ST>::real d, float scale, float nsh,
                                                             const Rounding& rd, uint32_t& h, int j)
{
    if constexpr (Form<ST>::shifted) {
        if (rd.on) return Form<ST>::enc_sr_raw((d + nsh) * scale, next_rand(h, j));
    }
    return Form<ST>::enc(d, scale, nsh);
}

// The same for a value already in stored units (less the shift).
template <int ST>
__device__ __forceinline__ typename Form<ST>::store enc_node_raw(typename Form<ST>::real h_, const Rounding& rd,
                                                                 uint32_t& h, int j)
{
    if constexpr (Form<ST>::shifted) {
        if (rd.on) return Form<ST>::enc_sr_raw(h_, next_rand(h, j));
    }
    return Form<ST>::enc_raw(h_);
}

template <typename R> __device__ __forceinline__ R wgt(int i) { return static_cast<R>(WD[i]); }

// ------------------------------------------------------------------ collision

// max(0.5, min(s, 1.99)): one v_med3_f32 in FP32
template <typename R> __device__ __forceinline__ R clamp_rate(R s)
{
    return __builtin_fmax(R(0.5), __builtin_fmin(s, R(1.99)));
}
template <> __device__ __forceinline__ float clamp_rate<float>(float s)
{
    return __builtin_fmaxf(0.5f, __builtin_fminf(s, 1.99f));
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
template <typename R> __device__ __forceinline__ R fma_(R a, R b, R c);
template <> __device__ __forceinline__ float fma_<float>(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
template <> __device__ __forceinline__ double fma_<double>(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <typename R>
__device__ __forceinline__ void collide(R d[9], const LatticeDev& L, R wall_dist, R sc)
{
    // Fused multiply-adds are written out and nothing else is fused: the same rounding in
    // every kernel and launch shape that inlines this.
#pragma clang fp contract(off)
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
        const R usq = fma_<R>(u, u, v * v);
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            const R eu = R(CX[i]) * u + R(CY[i]) * v;
            const R shape = fma_<R>(eu, fma_<R>(R(4.5), eu, R(3)), R(-1.5) * usq);
            const R feq_dev = wgt<R>(i) * fma_<R>(rho_s, shape, sum);
            d[i] = fma_<R>(-it, d[i] - feq_dev, d[i]);
        }
    } else {
        const R J2 = fma_<R>(jx, u, jy * v);                                      // sc rho |u|^2
        const R ne = fma_<R>(R(-2), d[0], fma_<R>(R(4), sdg, fma_<R>(R(-3), J2, sax)));  // e - e_eq
        const R e3 = d[0] - sax + J2;                                             // (eps - eps_eq) / 3
        const R npxx = (a13 - a24) - fma_<R>(jx, u, -(jy * v));
        const R npxy = fma_<R>(-jx, v, a57 - a68);
        R ss = static_cast<R>(L.s_shear);
        if (L.les) {
            // Smagorinsky with Van Driest damping (src/lbm.hpp:155-182, y+ :314-325).
            // The damped block replaces the plain one whenever y+ > 0.
            R cs2 = static_cast<R>(L.cs2);
            const R nu = static_cast<R>(L.nu);
            if (nu > R(0)) {
                const R uu = fma_<R>(u, u, fma_<R>(v, v, R(1e-12)));
                const R inu = static_cast<R>(L.inv_nu);
                // Far from walls the damping 1 - exp(-y+/26) rounds to exactly 1 (once
                // y+/26 > 17.4 in FP32, 37.4 in FP64): the same bits without the square
                // root and the exponential.  The test on y+^2 is exact enough -- either
                // side of its threshold gives damp = 1 there.
                constexpr double kFar = sizeof(R) == 4 ? 18.0 * 26.0 : 38.0 * 26.0;
                const R wi = wall_dist * inu;
                if (wi * wi * uu > R(kFar * kFar)) {
                    const R csd = static_cast<R>(L.cs);
                    cs2 = csd * csd;
                } else {
                    const R y_plus = wall_dist * dsqrt<R>(uu) * inu;
                    if (y_plus > R(0)) {
                        const R damp = R(1) - dexp<R>(y_plus * R(-1.0 / 26.0));
                        const R csd = static_cast<R>(L.cs) * damp;
                        cs2 = csd * csd;
                    }
                }
            }
            // A = 9 cs^2 Q / (2 rho), Q = sqrt(2 S) (Q/rho is scale free), and
            // tau_eff = (tau + sqrt(tau^2 + 4 A)) / 2 = tau/2 + sqrt(tau^2/4 + A)
            const R S = fma_<R>(npxx, npxx, npxy * npxy);
            const R A = (cs2 * R(4.5 * 1.4142135623730950488)) * (dsqrt<R>(S) * ir);
            const R tau_eff = static_cast<R>(L.half_tau) + dsqrt<R>(static_cast<R>(L.qtau2) + A);
            ss = clamp_rate<R>(recip<R>(tau_eff));
        }
        // M^-1 S (m - m_eq), s_bulk = 1.2 and s_normal = 1
        const R a = fma_<R>(ne, R(-1.2 / 36.0), e3 * R(-3.6 / 18.0));
        const R b = fma_<R>(ne, R(1.2 / 18.0), e3 * R(3.6 / 36.0));
        const R c0 = fma_<R>(e3, R(3.6 / 9.0), ne * R(-1.2 / 9.0));
        const R qx6 = fma_<R>(jx, R(1.0 / 3.0), s13 * R(-0.5));  // (qx - qx_eq) / 6
        const R qy6 = fma_<R>(jy, R(1.0 / 3.0), s24 * R(-0.5));
        const R ss4 = ss * R(0.25);
        const R t = fma_<R>(ss4, npxx, a), t2 = fma_<R>(-ss4, npxx, a);
        const R u1 = fma_<R>(ss4, npxy, b), u2 = fma_<R>(-ss4, npxy, b);
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

// Where a launch's link records put their force terms: nf[2 r], nf[2 r + 1] for record r
// (this launch's block), accp for terms of a record's secondary slots (atomic).  A null
// nf: forces not counted.
struct ForceSink {
    double* nf;
    double* accp;
};

template <int ST> struct Ctx {
    using S = typename Form<ST>::store;
    const LatticeDev& L;
    const S* __restrict__ base;  // plane 0 of this lattice in the source buffer
    const uint8_t* __restrict__ flags;
    const int32_t* __restrict__ link_slot;
    const LinkRec* __restrict__ links;
    float inv_scale;
};

// What the general path reads at node (px,py), fetched in one round of independent
// loads: its flags and link slot, its own stored values and, per direction j, the
// stored value at the streaming source (px,py) - e_j, wrapped on periodic edges.  Off
// the lattice on other edges the address lies outside it but inside the population
// pool (whose guard covers one row past its end); those values go unused.
template <int ST> struct Pre {
    using S = typename Form<ST>::store;
    uint8_t fl;
    int32_t ls;
    S own[9];
    S pull[9];
};

template <int ST>
__device__ __forceinline__ void preload(const Ctx<ST>& c, int px, int py, Pre<ST>& p)
{
    const LatticeDev& L = c.L;
    const int64_t idx = static_cast<int64_t>(py) * L.pitch + px;
    p.fl = c.flags[L.cell_off + idx];
    p.ls = c.link_slot[L.cell_off + idx];
#pragma unroll
    for (int j = 0; j < 9; ++j) {
        int sx = px - CX[j], sy = py - CY[j];
        if (L.x_edge == EDGE_PERIODIC) sx = sx < 0 ? sx + L.nx : (sx >= L.nx ? sx - L.nx : sx);
        if (L.y_edge == EDGE_PERIODIC) sy = sy < 0 ? sy + L.ny : (sy >= L.ny ? sy - L.ny : sy);
        p.own[j] = c.base[j * L.plane + idx];
        p.pull[j] = c.base[j * L.plane + static_cast<int64_t>(sy) * L.pitch + sx];
    }
}

// Boundary-link data of one node: mask, schemes, force slots (int16 pairs) and q in
// registers, loaded with the record unconditionally; wall velocities (moving walls
// only) are read from the record where they are used.
template <typename R> struct LinkVals {
    uint32_t mask = 0, scheme = 0;
    uint32_t slots[5];
    R q[9];
    const LinkRec* rec = nullptr;
    __device__ __forceinline__ int slot(int i) const
    {
        return static_cast<int16_t>(static_cast<uint16_t>(slots[i >> 1] >> (16 * (i & 1))));
    }
};

template <typename R> __device__ __forceinline__ void load_link(const LinkRec& r, LinkVals<R>& v)
{
    v.rec = &r;
    v.mask = r.mask;
    v.scheme = r.scheme;
    const uint32_t* sl = reinterpret_cast<const uint32_t*>(r.slot);  // slot[9] and pad: 5 dwords
#pragma unroll
    for (int k = 0; k < 5; ++k) v.slots[k] = sl[k];
#pragma unroll
    for (int i = 0; i < 9; ++i) v.q[i] = static_cast<R>(r.q[i]);
}

// Post-stream, pre-boundary deviations at node (px,py), from the post-collision
// buffer (pull form of the streaming loops src/lbm.hpp:408-484).  lk: the node's
// boundary links (read when p.fl has CELL_LINKS).
template <int ST>
__device__ __forceinline__ void gather_general(const Ctx<ST>& c, int px, int py, const Pre<ST>& p,
                                               const LinkVals<typename Form<ST>::real>& lk,
                                               typename Form<ST>::real d[9])
{
    // unfused multiply-adds: the boundary values round the same way in every kernel
    // that inlines this (one or two steps per launch, any shape)
#pragma clang fp contract(off)
    using R = typename Form<ST>::real;
    const LatticeDev& L = c.L;
    if (p.fl & CELL_SOLID) {
        // solid nodes are never streamed into: f_next stays zero (src/lbm.hpp:400)
#pragma unroll
        for (int j = 0; j < 9; ++j) d[j] = -wgt<R>(j);
        return;
    }
    R own[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) own[i] = Form<ST>::dec(p.own[i], c.inv_scale, L.shift[i]);

    const uint32_t mask = (p.fl & CELL_LINKS) ? lk.mask : 0u;
    R rho_own = R(1);
    if (mask) {
        R s = R(0);
#pragma unroll
        for (int i = 0; i < 9; ++i) s += own[i];
        rho_own += s;
    }

    if (L.mei_static && mask && px >= 1 && px <= L.nx - 2 && py >= 1 && py <= L.ny - 2) {
        // every link Mei with a resting wall and the node off the edges: each slot is its
        // link's Mei value (the general expression below with u_w = 0) or the plain pull,
        // selected without branches
#pragma unroll
        for (int j = 0; j < 9; ++j) {
            const int i = OPP[j];
            const R q = lk.q[i];
            const R uwx = R(0), uwy = R(0);
            const R eu = R(CX[i]) * uwx + R(CY[i]) * uwy;
            const R uu = uwx * uwx + uwy * uwy;
            const R feq_dev =
                wgt<R>(i) * (rho_own * (R(1) + R(3) * eu + R(4.5) * eu * eu - R(1.5) * uu) - R(1));
            const R mei = q * feq_dev + (R(1) - q) * own[i] + (R(2) * q - R(1)) * wgt<R>(i) * rho_own * eu * R(3);
            const R pull = Form<ST>::dec(p.pull[j], c.inv_scale, L.shift[j]);
            d[j] = ((mask >> i) & 1u) ? mei : pull;
        }
        return;
    }

#pragma unroll
    for (int j = 0; j < 9; ++j) {
        const int i = OPP[j];  // push direction whose target is px - e_j
        if (mask & (1u << i)) {
            // boundary link (px,py) -> (px,py)+e_i solid: value for slot j = opp(i)
            const int scheme = (lk.scheme >> (2 * i)) & 3;
            const R q = lk.q[i];
            const bool moving = (lk.scheme & LINK_MOVING) != 0;
            const R uwx = moving ? static_cast<R>(lk.rec->uwx[i]) : R(0);
            const R uwy = moving ? static_cast<R>(lk.rec->uwy[i]) : R(0);
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
                        // the upstream node's stored value: the (unwrapped) source of direction i
                        const R up = Form<ST>::dec(p.pull[i], c.inv_scale, L.shift[i]);
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
        const int sx = px - CX[j], sy = py - CY[j];
        if ((sy < 0 || sy >= L.ny) && L.y_edge == EDGE_FREE_SLIP) {
            // specular reflection: the population of direction mirror(j) that node
            // (sx, py) emitted towards the edge comes back into slot j.  A solid node
            // there emits nothing, and this node's own population of direction i = opp(j)
            // meets that solid's mirror image: bounced back; beyond the open x edge it is
            // dropped (zero)
            if (sx < 0 || sx >= L.nx) {
                d[j] = -wgt<R>(j);
            } else {
                const int64_t e = static_cast<int64_t>(py) * L.pitch + sx;
                d[j] = (c.flags[L.cell_off + e] & CELL_SOLID)
                           ? own[i]
                           : Form<ST>::dec(c.base[MIRY[j] * L.plane + e], c.inv_scale, L.shift[MIRY[j]]);
            }
            continue;
        }
        bool wall = false, open = false;
        if ((sx < 0 || sx >= L.nx) && L.x_edge != EDGE_PERIODIC) (L.x_edge == EDGE_WALL ? wall : open) = true;
        if ((sy < 0 || sy >= L.ny) && L.y_edge != EDGE_PERIODIC) (L.y_edge == EDGE_WALL ? wall : open) = true;
        if (wall) {
            d[j] = own[i];  // bounced into the same node (src/lbm.hpp:418-419)
        } else if (open) {
            d[j] = -wgt<R>(j);  // dropped: f_next keeps its zero (src/lbm.hpp:470)
        } else {
            d[j] = Form<ST>::dec(p.pull[j], c.inv_scale, L.shift[j]);
        }
    }
}

// Zou/He velocity inlet on deviations (src/lbm.hpp:21-34 / 40-58); w1=w3, w5=w7=w6=w8
// so the closure keeps its form, and w0+w2+w4+2(w3+w6+w7) = 1.  one: 1, or the storage
// scale 2^k for deviations in stored units (every term then scales exactly by 2^k).
template <typename R> __device__ __forceinline__ void zou_he(R d[9], R u, R one = R(1))
{
#pragma clang fp contract(off)
    const R rho = (one + d[0] + d[2] + d[4] + R(2) * (d[3] + d[6] + d[7])) / (R(1) - u);
    d[1] = d[3] + R(2.0 / 3.0) * rho * u;
    d[5] = d[7] + R(1.0 / 6.0) * rho * u;
    d[8] = d[6] + R(1.0 / 6.0) * rho * u;
}

// Node state after streaming and the open boundaries (src/lbm.hpp:530-539).
// Boundary order: inlet, outlet copy, lid.  Solid nodes skip the inlet and the lid,
// which leave their zero populations unchanged.
// p holds the preloaded values of the gathered node: (x,y), or (nx-2,y) on the outlet
// column (gathered_x).
template <int ST> __device__ __forceinline__ int gathered_x(const LatticeDev& L, int x)
{
    return L.outlet && x == L.nx - 1 ? L.nx - 2 : x;  // enforce_outflow copies x = nx-2
}

// The velocity inlet at row y of a fluid node on x = 0: Zou/He with u_in, or the
// parabolic profile on rows y >= inlet_y0 (src/lbm.hpp:21-58).  one: as zou_he's (the
// equilibrium inlet's deviations are multiplied by it, an exact identity for 1).
template <typename R> __device__ __forceinline__ void inlet_closure(const LatticeDev& L, int y, R d[9], R one = R(1))
{
#pragma clang fp contract(off)
    if (L.inlet == INLET_EQUILIBRIUM) {
        // the node becomes feq(1, u_in, 0), stored as its deviation from w
        const R u = static_cast<R>(L.u_in);
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            const R eu = R(CX[i]) * u;
            d[i] = (wgt<R>(i) * (R(3) * eu + R(4.5) * eu * eu - R(1.5) * u * u)) * one;
        }
    } else if (L.inlet == INLET_ZOU_HE) {
        zou_he<R>(d, static_cast<R>(L.u_in), one);
    } else if (y >= L.inlet_y0) {
        const double yy = static_cast<double>(y - L.inlet_y0);
        const double h = static_cast<double>(L.inlet_h);
        zou_he<R>(d, static_cast<R>(L.u_in * 4.0 * yy * (h - yy) / (h * h)), one);
    }
}

template <int ST>
__device__ __forceinline__ void state_general(const Ctx<ST>& c, int x, int y, bool solid, const Pre<ST>& p,
                                              const LinkVals<typename Form<ST>::real>& lk,
                                              typename Form<ST>::real d[9])
{
    // unfused, as gather_general: the lid's equilibrium rounds the same way in every
    // kernel that inlines this
#pragma clang fp contract(off)
    using R = typename Form<ST>::real;
    const LatticeDev& L = c.L;
    const int gx = gathered_x<ST>(L, x);
    gather_general<ST>(c, gx, y, p, lk, d);
    if (gx == 0 && L.inlet != INLET_NONE) {
        const bool gsolid = (p.fl & CELL_SOLID) != 0;
        if (!gsolid) inlet_closure<R>(L, y, d);
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
// Links of the node a Pre was taken at (none unless it has CELL_LINKS).
template <int ST>
__device__ __forceinline__ void links_of(const Ctx<ST>& c, const Pre<ST>& p, LinkVals<typename Form<ST>::real>& lk)
{
    if (p.fl & CELL_LINKS) load_link(c.links[c.L.link_off + p.ls], lk);
}

// Momentum exchange of a node with boundary links on its post-boundary state d
// (src/lbm.hpp:542-581): the terms of the record's primary slot (that of its first
// counted link) summed in direction order and stored for the record; terms of any other
// slot go to the fallback accumulator.  ls_: the node's own links.
template <int ST>
__device__ __forceinline__ void node_forces(const LatticeDev& L, const Ctx<ST>& c, const ForceSink& fs,
                                            const LinkVals<typename Form<ST>::real>& ls_,
                                            const typename Form<ST>::real d[9])
{
#pragma clang fp contract(off)
    int ps = -1;
    double fx = 0.0, fy = 0.0;
#pragma unroll
    for (int i = 1; i < 9; ++i) {
        const int si = ls_.slot(i);
        if (!(ls_.mask & (1u << i)) || si < 0) continue;
        const double fb = WD[OPP[i]] + static_cast<double>(d[OPP[i]]);
        if (ps < 0) ps = si;
        if (si == ps) {
            fx += CX[i] * 2.0 * fb;
            fy += CY[i] * 2.0 * fb;
        } else {
            atomicAdd(fs.accp + 2 * (L.slot_base + si), CX[i] * 2.0 * fb);
            atomicAdd(fs.accp + 2 * (L.slot_base + si) + 1, CY[i] * 2.0 * fb);
        }
    }
    if (ps >= 0) {
        const int64_t r = ls_.rec - c.links;
        fs.nf[2 * r] = fx;
        fs.nf[2 * r + 1] = fy;
    }
}

// The general path of one node: boundary state from the preloaded gathered node (p,
// lg), momentum exchange on it with the node's own links (ls_), collision.
template <int ST>
__device__ __forceinline__ void node_general_d(const LatticeDev& L, const Ctx<ST>& c, const ForceSink& fs, int x,
                                               int y, uint8_t fl, const Pre<ST>& p,
                                               const LinkVals<typename Form<ST>::real>& lg,
                                               const LinkVals<typename Form<ST>::real>& ls_, uint16_t wraw,
                                               bool& bad, typename Form<ST>::real d[9])
{
    using R = typename Form<ST>::real;
    const bool solid = (fl & CELL_SOLID) != 0;
    state_general<ST>(c, x, y, solid, p, lg, d);
    if (fs.nf && (fl & CELL_LINKS) && L.forces) node_forces<ST>(L, c, fs, ls_, d);
    if (!solid) {
        R wd = R(0);
        if (L.les) wd = wraw == 0xffffu ? R(1e9) : static_cast<R>(wraw);
        collide<R>(d, L, wd, R(1));
        if (!(d[0] == d[0])) bad = true;
    }
}

template <int ST>
__device__ __forceinline__ void node_general(const LatticeDev& L, const StepArgs& a, const Ctx<ST>& c,
                                             typename Form<ST>::store* dst, const ForceSink& fs, int x, int y,
                                             uint8_t fl, const Pre<ST>& p,
                                             const LinkVals<typename Form<ST>::real>& lg,
                                             const LinkVals<typename Form<ST>::real>& ls_, uint16_t wraw,
                                             const Rounding& rd, bool& bad)
{
    using R = typename Form<ST>::real;
    R d[9];
    node_general_d<ST>(L, c, fs, x, y, fl, p, lg, ls_, wraw, bad, d);
    const int64_t idx = static_cast<int64_t>(y) * L.pitch + x;
    auto* o = dst + L.f_off + idx;
    uint32_t h = rd.on ? rd.node(L.cell_off + idx) : 0u;
#pragma unroll
    for (int j = 0; j < 9; ++j)
        store_val(o + j * L.plane, enc_node<ST>(d[j], a.scale, L.nshift[j], rd, h, j), a.options & LAUNCH_STREAM_STORES);
}

// Edge node number e of a lattice: rows 0 and ny-1, then columns 0 and nx-1 between them.
__device__ __forceinline__ void edge_position(const LatticeDev& L, int e, int& x, int& y)
{
    const int rows = L.ny == 1 ? 1 : 2;
    if (e < L.nx) {
        x = e;
        y = 0;
    } else if (rows == 2 && e < 2 * L.nx) {
        x = e - L.nx;
        y = L.ny - 1;
    } else {
        const int k = e - rows * L.nx, h = L.ny - 2;
        x = k < h ? 0 : L.nx - 1;
        y = (k < h ? k : k - h) + 1;
    }
}

// A general-path list node of lattice L: boundary-link record t (is_link; the record
// gives the position, so links, flags, stored values and sources load in two rounds),
// else edge node t (nodes with links are the records'; solid ones are written only when
// marked).  One call site of node_general keeps one copy of the general path in the
// kernel.
template <int ST>
__device__ __forceinline__ void list_node(const LatticeDev& L, const StepArgs& a, const Ctx<ST>& c,
                                          typename Form<ST>::store* dst, const ForceSink& fs, bool is_link,
                                          int t, const Rounding& rd, bool& bad)
{
    using R = typename Form<ST>::real;
    LinkVals<R> own_links;
    int x, y;
    if (is_link) {
        const LinkRec& rec = c.links[L.link_off + t];
        load_link(rec, own_links);
        const int32_t e = rec.cell;
        y = e / L.pitch;
        x = e - y * L.pitch;
    } else {
        edge_position(L, t, x, y);
    }
    const int gx = gathered_x<ST>(L, x);
    const int64_t idx = static_cast<int64_t>(y) * L.pitch + x;
    const uint8_t fl = a.flags[L.cell_off + idx];
    Pre<ST> p;
    preload<ST>(c, gx, y, p);
    const uint16_t wraw = L.les ? a.wdist[L.cell_off + idx] : uint16_t(0xffffu);
    if (!is_link && (fl & (CELL_LINKS | CELL_COLUMN))) return;
    if ((fl & CELL_SOLID) && !(fl & CELL_WRITE)) return;
    LinkVals<R> lg = own_links;
    if (!is_link || gx != x) links_of<ST>(c, p, lg);
    node_general<ST>(L, a, c, dst, fs, x, y, fl, p, lg, own_links, wraw, rd, bad);
}

// One node, bulk or general.  SKIP_SPECIAL: general nodes are left to the lists.
template <int ST, bool SKIP_SPECIAL>
__device__ __forceinline__ void cell_update(const LatticeDev& L, const StepArgs& a,
                                            const Ctx<ST>& c, typename Form<ST>::store* dst,
                                            const ForceSink& fs, int x, int y, const Rounding& rd, bool& bad)
{
    using S = typename Form<ST>::store;
    using R = typename Form<ST>::real;
    const int64_t idx = static_cast<int64_t>(y) * L.pitch + x;
    // The flag, the bulk pull values and the wall distance in one round trip: the pull
    // addresses of any node lie inside the population pool (which ends with a guard of
    // more than one row), and the asm below keeps the loads ahead of the branch.
    const uint8_t fl = a.flags[L.cell_off + idx];
    S raw[9];
    // on the outlet column the pulls of the copied node x - 1 (enforce_outflow)
    const S* b = c.base + idx - ((L.outlet && x == L.nx - 1) ? 1 : 0);
#pragma unroll
    for (int j = 0; j < 9; ++j) raw[j] = b[j * L.plane - (CY[j] * L.pitch + CX[j])];
    uint16_t wraw = 0xffffu;
    if (L.les) wraw = a.wdist[L.cell_off + idx];
    __asm__ volatile("" ::"v"(raw[0]), "v"(raw[1]), "v"(raw[2]), "v"(raw[3]), "v"(raw[4]), "v"(raw[5]),
                     "v"(raw[6]), "v"(raw[7]), "v"(raw[8]), "v"(wraw));
    const bool column = (fl & CELL_COLUMN) != 0;
    if ((fl & CELL_SPECIAL) && !column) {
        if (SKIP_SPECIAL) return;
        if ((fl & CELL_SOLID) && !(fl & CELL_WRITE)) return;
        const int gx = gathered_x<ST>(L, x);
        Pre<ST> p;
        preload<ST>(c, gx, y, p);
        LinkVals<R> lg;
        links_of<ST>(c, p, lg);
        LinkVals<R> ls_ = lg;
        if (gx != x && (fl & CELL_LINKS)) load_link(c.links[L.link_off + a.link_slot[L.cell_off + idx]], ls_);
        node_general<ST>(L, a, c, dst, fs, x, y, fl, p, lg, ls_, wraw, rd, bad);
        return;
    }
    R d[9];
#pragma unroll
    for (int j = 0; j < 9; ++j) d[j] = Form<ST>::dec(raw[j], a.inv_scale, L.shift[j]);
    // an inlet node: the closure replaces the three slots whose sources lie off the lattice
    if (column && x == 0) inlet_closure<R>(L, y, d);
    R wd = R(0);
    if (L.les) wd = wraw == 0xffffu ? R(1e9) : static_cast<R>(wraw);
    collide<R>(d, L, wd, R(1));
    if (!(d[0] == d[0])) bad = true;
    S* o = dst + L.f_off + idx;
    uint32_t h = rd.on ? rd.node(L.cell_off + idx) : 0u;
#pragma unroll
    for (int j = 0; j < 9; ++j)
        store_val(o + j * L.plane, enc_node<ST>(d[j], a.scale, L.nshift[j], rd, h, j), a.options & LAUNCH_STREAM_STORES);
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
                                             const Rounding& rd, bool& bad)
{
    using S = typename Form<ST>::store;
    using R = typename Form<ST>::real;
    using P = Pack<S, V>;
    const int64_t idx = static_cast<int64_t>(y) * L.pitch + x;
    P in[9];
#pragma unroll
    for (int j = 0; j < 9; ++j)
        in[j] = *reinterpret_cast<const P*>(c.base + j * L.plane + idx - CY[j] * L.pitch);
    // the wave's first and last lanes fetch the values from outside the wave in the
    // same round trip as the planes (interior row: every address is inside the pool)
    const int lane = __lane_id();
    S far[9];
    if (lane == 0) {
#pragma unroll
        for (int j = 0; j < 9; ++j)
            if (CX[j] == 1) far[j] = c.base[j * L.plane + idx - CY[j] * L.pitch - 1];
    }
    if (lane == 63) {
#pragma unroll
        for (int j = 0; j < 9; ++j)
            if (CX[j] == -1) far[j] = c.base[j * L.plane + idx - CY[j] * L.pitch + V];
    }
    Pack<uint16_t, V> wd4;
    if (L.les) wd4 = *reinterpret_cast<const Pack<uint16_t, V>*>(a.wdist + L.cell_off + idx);
    S edge[9];
#pragma unroll
    for (int j = 0; j < 9; ++j) {
        if (CX[j] == 1) edge[j] = shift_in<S>(in[j].v[V - 1], true);
        if (CX[j] == -1) edge[j] = shift_in<S>(in[j].v[0], false);
    }
    if (!bulk) return;
    if (lane == 0) {
#pragma unroll
        for (int j = 0; j < 9; ++j)
            if (CX[j] == 1) edge[j] = far[j];
    }
    if (lane == 63) {
#pragma unroll
        for (int j = 0; j < 9; ++j)
            if (CX[j] == -1) edge[j] = far[j];
    }
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
        uint32_t h = rd.on ? rd.node(L.cell_off + idx + k) : 0u;
#pragma unroll
        for (int j = 0; j < 9; ++j) {
            if constexpr (Form<ST>::shifted) d[j] += L.nshift_s[j];
            out[j].v[k] = enc_node_raw<ST>(d[j], rd, h, j);
        }
    }
#pragma unroll
    for (int j = 0; j < 9; ++j)
        store_val(reinterpret_cast<P*>(dst + L.f_off + j * L.plane + idx), out[j], a.options & LAUNCH_STREAM_STORES);
}

// Stochastic-rounding key of lattice l's step (its step count at the start of the batch
// + the launch's index in it + extra): the same in every launch shape and path.
__device__ __forceinline__ uint32_t step_key(const StepArgs& a, int l, int extra)
{
    const int64_t k = a.steps[l] + a.step_arg + extra;
    return mix32(a.sr_seed ^ mix32(static_cast<uint32_t>(k) ^ mix32(static_cast<uint32_t>(k >> 32) + 0x9e3779b9u)));
}

// One step of one lattice (SINGLE: descriptor, buffers and tile grid in the kernel
// arguments, nothing loaded before the populations) or of every lattice of an ensemble
// (tile table; per-lattice buffer parity from the device counters).  Forces go to this
// launch's accumulator block, copy blockIdx % kAccCopies; flush_kernel sums them.
template <int ST, int BX, int ROWS, int V, bool SINGLE>
__global__ __launch_bounds__(BX) void step_kernel(StepArgs a)
{
    using S = typename Form<ST>::store;

    // Workgroups are handed to the 8 XCDs round-robin; with LAUNCH_XCD_ORDER XCD x
    // takes one contiguous range of tiles, so lines shared by neighbouring tiles stay
    // in one L2.  The map is a bijection on [0, gridDim.x).
    int b = static_cast<int>(blockIdx.x);
    const int nb = static_cast<int>(gridDim.x);
    int lat_id, y0, x0;
    const S* src;
    S* dst;
    LatticeDev L;  // by value: scalar registers, loaded before any store
    if constexpr (SINGLE) {
        L = a.L0;
        lat_id = a.lattice0;
        src = static_cast<const S*>(a.src);
        dst = static_cast<S*>(a.dst);
        // every argument the first memory round needs, fetched from the kernel argument
        // segment in one group instead of between branches.  L.lattice and nan_flags keep
        // their registers live: a register of a group load the kernel does not read is
        // otherwise reused for a later argument, which must then wait for the whole group
        // (scalar loads return out of order) and takes a second round.
        __asm__ volatile("" ::"s"(src), "s"(dst), "s"(a.flags), "s"(a.wdist), "s"(a.acc), "s"(a.total_slots),
                         "s"(a.tiles_x), "s"(a.options), "s"(nb), "s"(L.nx), "s"(L.ny), "s"(L.pitch),
                         "s"(L.f_off), "s"(L.plane), "s"(L.cell_off), "s"(L.les), "s"(a.special),
                         "s"(a.n_special), "s"(a.special_blocks), "s"(a.n_links), "s"(a.n_edges), "s"(L.lattice),
                         "s"(a.nan_flags));
    }
    Rounding rd{a.sr != 0, 0u};
    if constexpr (SINGLE) {
        if (rd.on) rd.key = step_key(a, a.lattice0, 0);
        // The first blocks take the general-path nodes -- boundary-link records, edge
        // nodes, written solids: dispatched first, their longer load chains overlap the
        // bulk tiles, which skip those nodes.
        if (b < a.special_blocks) {
            double* accp = a.acc + static_cast<int64_t>(blockIdx.x % kAccCopies) * a.total_slots * 2;
            const Ctx<ST> c{L, src + L.f_off, a.flags, a.link_slot, a.links, a.inv_scale};
            const ForceSink fs{a.nf, accp};
            bool bad = false;
            // link nodes: kListLanes per wave, since their loads scatter and a wave's
            // memory requests issue one address at a time; edge nodes: every lane, since
            // consecutive edge nodes are consecutive along a row or a column
            constexpr int per_block = BX / 64 * kListLanes;
            const int gl = (a.n_links + per_block - 1) / per_block;
            const int ge = gl + (a.n_edges + BX - 1) / BX;
            if (b < ge) {
                // one call site: one copy of the general path in the kernel
                int t = -1;
                if (b < gl) {
                    const int lane = static_cast<int>(threadIdx.x) & 63;
                    const int k = b * per_block + (static_cast<int>(threadIdx.x) >> 6) * kListLanes + lane;
                    if (lane < kListLanes && k < a.n_links) t = k;
                } else {
                    const int k = a.n_links + (b - gl) * BX + static_cast<int>(threadIdx.x);
                    if (k < a.n_links + a.n_edges) t = k;
                }
                if (t >= 0) {
                    const bool is_link = t < a.n_links;
                    list_node<ST>(L, a, c, dst, fs, is_link, is_link ? t : t - a.n_links, rd, bad);
                }
            } else {
                // written solids (Bouzidi sources): zero populations, f - w = -w
                const int t = (b - ge) * BX + static_cast<int>(threadIdx.x);
                if (t < a.n_special) {
                    S* o = dst + L.f_off + a.special[t];
#pragma unroll
                    for (int j = 0; j < 9; ++j)
                        o[j * L.plane] = Form<ST>::enc(-static_cast<typename Form<ST>::real>(WD[j]), a.scale,
                                                       L.nshift[j]);
                }
            }
            if (bad) atomicOr(a.nan_flags + lat_id, 1);
            return;
        }
        b -= a.special_blocks;
    }
    if (a.options & LAUNCH_XCD_ORDER) {
        // over the tile blocks; (b & 7) still names one XCD's blocks after the offset
        const int nt = nb - (SINGLE ? a.special_blocks : 0);
        const int q = nt >> 3, r = nt & 7, xcd = b & 7;
        b = xcd * q + (xcd < r ? xcd : r) + (b >> 3);
    }
    if constexpr (SINGLE) {
        y0 = (b / a.tiles_x) * ROWS;
        x0 = (b % a.tiles_x) * (BX * V);
    } else {
        const Tile t = a.tiles[b];
        lat_id = t.lattice;
        y0 = t.y0;
        x0 = t.x0;
        L = a.lat[lat_id];
        if (rd.on) rd.key = step_key(a, lat_id, 0);
        const int sp = a.par[lat_id] ^ (a.flips & 1);
        src = static_cast<const S*>(sp ? a.buf1 : a.buf0);
        dst = static_cast<S*>(a.dst ? a.dst : (sp ? a.buf0 : a.buf1));
    }
    double* accp = a.acc + static_cast<int64_t>(blockIdx.x % kAccCopies) * a.total_slots * 2;

    const Ctx<ST> c{L, src + L.f_off, a.flags, a.link_slot, a.links, a.inv_scale};
    const ForceSink fs{a.nf, accp};
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
                if (x < L.pitch) group_update<ST, V>(L, a, c, dst, x, y, bulk, rd, bad);
                done = bulk;
            }
        }
        if (!done) {
            // one copy of the general path, looped (keeps the kernel small)
#pragma unroll 1
            for (int q = 0; q < V; ++q)
                if (x + q < L.nx) cell_update<ST, SINGLE>(L, a, c, dst, fs, x + q, y, rd, bad);
        }
    }
    if (bad) atomicOr(a.nan_flags + lat_id, 1);
}

// ------------------------------------------------------------------ two steps per launch

// Post-collision values of node (x,y) for the next state, read from the source buffer
// like cell_update; forces counted only when accp is non-null.  false: a solid node
// nothing reads (no value).
// A fluid node whose general path reduces to its pulls (wrapped around the periodic
// edges it lies on) and the collision: no boundary links, every edge it lies on periodic,
// and no inlet, outflow copy or lid on it.
__device__ __forceinline__ bool wrap_only(const LatticeDev& L, uint8_t fl, int x, int y)
{
    if (fl & (CELL_SOLID | CELL_LINKS | CELL_COLUMN)) return false;
    if ((x == 0 || x == L.nx - 1) && L.x_edge != EDGE_PERIODIC) return false;
    if ((y == 0 || y == L.ny - 1) && L.y_edge != EDGE_PERIODIC) return false;
    if (L.lid && y == L.ny - 1) return false;
    if (L.inlet != INLET_NONE && x == 0) return false;
    if (L.outlet && x == L.nx - 1) return false;
    return true;
}

// Bulk nodes take their pulls directly, the others the general path.  EDGES: so do the
// inlet and outlet column nodes (as the one-step kernel's bulk tiles take them) and the
// wrap-only nodes -- on small lattices, whose tiles often hold an edge, this keeps whole
// tiles off the general path; on large ones the extra code costs more than it saves.
//
// MODE: NV_ALL computes either; NV_PLAIN computes plain nodes and returns -1 for general
// ones (left for a later NV_GENERAL call, which takes the general path without the
// plain pulls).  Returns 1 when out holds the node's values, 0 for a solid node nothing
// reads.
enum { NV_ALL = 0, NV_PLAIN = 1, NV_GENERAL = 2 };

template <int ST, bool EDGES, int MODE>
__device__ __forceinline__ int node_values(const LatticeDev& L, const StepArgs& a, const Ctx<ST>& c,
                                           const ForceSink& fs, int x, int y, const Rounding& rd, bool& bad,
                                           typename Form<ST>::store out[9])
{
    using S = typename Form<ST>::store;
    using R = typename Form<ST>::real;
    const int64_t idx = static_cast<int64_t>(y) * L.pitch + x;
    const uint8_t fl = a.flags[L.cell_off + idx];
    S raw[9];
    if (MODE == NV_GENERAL) {
        // no plain pulls
    } else if (!EDGES || (x >= 1 && x <= L.nx - 2 && y >= 1 && y <= L.ny - 2)) {
        const S* b = c.base + idx;
#pragma unroll
        for (int j = 0; j < 9; ++j) raw[j] = b[j * L.plane - (CY[j] * L.pitch + CX[j])];
    } else {
        // edge nodes: the pulls of the gathered node (x - 1 on the outlet column),
        // wrapped on periodic edges (unwrapped sources of open edges stay inside the pool)
        const int px = gathered_x<ST>(L, x);
#pragma unroll
        for (int j = 0; j < 9; ++j) {
            int sx = px - CX[j], sy = y - CY[j];
            if (L.x_edge == EDGE_PERIODIC) sx = sx < 0 ? sx + L.nx : (sx >= L.nx ? sx - L.nx : sx);
            if (L.y_edge == EDGE_PERIODIC) sy = sy < 0 ? sy + L.ny : (sy >= L.ny ? sy - L.ny : sy);
            raw[j] = c.base[j * L.plane + static_cast<int64_t>(sy) * L.pitch + sx];
        }
    }
    uint16_t wraw = 0xffffu;
    if (L.les) wraw = a.wdist[L.cell_off + idx];
    if constexpr (MODE != NV_GENERAL)
        __asm__ volatile("" ::"v"(raw[0]), "v"(raw[1]), "v"(raw[2]), "v"(raw[3]), "v"(raw[4]), "v"(raw[5]),
                         "v"(raw[6]), "v"(raw[7]), "v"(raw[8]), "v"(wraw));
    R d[9];
    const bool column = EDGES && (fl & CELL_COLUMN) != 0;
    const bool general = (fl & CELL_SPECIAL) && !column && !(EDGES && wrap_only(L, fl, x, y));
    if (MODE == NV_GENERAL || (MODE == NV_ALL && general)) {
        if ((fl & CELL_SOLID) && !(fl & CELL_WRITE)) return 0;
        const int gx = gathered_x<ST>(L, x);
        Pre<ST> p;
        preload<ST>(c, gx, y, p);
        LinkVals<R> lg;
        links_of<ST>(c, p, lg);
        LinkVals<R> ls_ = lg;
        if (gx != x && (fl & CELL_LINKS)) load_link(c.links[L.link_off + a.link_slot[L.cell_off + idx]], ls_);
        node_general_d<ST>(L, c, fs, x, y, fl, p, lg, ls_, wraw, bad, d);
    } else {
        if (MODE == NV_PLAIN && general) return (fl & CELL_SOLID) && !(fl & CELL_WRITE) ? 0 : -1;
#pragma unroll
        for (int j = 0; j < 9; ++j) d[j] = Form<ST>::dec(raw[j], a.inv_scale, L.shift[j]);
        // an inlet node: the closure replaces the three slots whose sources lie off the lattice
        if (column && x == 0) inlet_closure<R>(L, y, d);
        R wd = R(0);
        if (L.les) wd = wraw == 0xffffu ? R(1e9) : static_cast<R>(wraw);
        collide<R>(d, L, wd, R(1));
        if (!(d[0] == d[0])) bad = true;
    }
    uint32_t h = rd.on ? rd.node(L.cell_off + idx) : 0u;
#pragma unroll
    for (int j = 0; j < 9; ++j) out[j] = enc_node<ST>(d[j], a.scale, L.nshift[j], rd, h, j);
    return 1;
}

// Steps k and k+1 of one lattice in one launch.  A block owns a 64 x TY tile: it
// computes C_k on the tile and a one-node ring around it from C_{k-1} in the source
// buffer (ring nodes wrap on periodic edges), keeps it in LDS in the stored form, then
// computes C_{k+1} on the tile from LDS and writes it to the destination.  Every
// arithmetic step is the one-step kernel's, so results are bit for bit the same; C_k
// itself is never written: after a batch that ends with a two-step launch the host
// recomputes it before anything reads the state (the state after step k+1 is gathered from
// C_k).  Forces of step k go to acc, of step k+1 to acc2.
//
// Ensemble launches (SINGLE false) take each block's lattice and tile origin from the
// two-step tile table and the buffers from the lattice's parity; a tile may start left
// of x = 0 (right-aligned rows, so that the outflow copy's source column lies inside the
// last tile), and nodes left of x = 0 are then not the tile's.  On their small lattices
// one tile in twelve holds boundary-link nodes, and a wave holding one runs the long
// general path beside the plain one: each step therefore queues its general-path nodes
// in LDS (per-wave segments, positions from ballots) and the whole block takes them
// after the plain nodes, a few dense waves instead of every wave of the tile.
template <int ST, int TY, bool SINGLE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(ST == ST_F64 ? 1 : 4))) void step2_kernel(
    StepArgs a)
{
    using S = typename Form<ST>::store;
    using R = typename Form<ST>::real;
    constexpr int TX = 64, WX = TX + 2, WY = TY + 2, P2 = TX * TY / 256;
    __shared__ S lds[9][WY][WX];
    // deferred general-path nodes (ensembles): per-wave segments of step-1 ring indices,
    // then of step-2 tile indices; SEG covers a wave's share of the larger step
    constexpr bool DEFER = !SINGLE;
    constexpr int SEG = 64 * ((WX * WY + 255) / 256);
    __shared__ int16_t queue[DEFER ? 4 * SEG : 1];
    __shared__ int qcnt[4];

    LatticeDev L;
    const S* src;
    S* dst;
    int lat_id;
    const int nb = static_cast<int>(gridDim.x);
    if constexpr (SINGLE) {
        L = a.L0;
        lat_id = a.lattice0;
        src = static_cast<const S*>(a.src);
        dst = static_cast<S*>(a.dst);
        __asm__ volatile("" ::"s"(src), "s"(dst), "s"(a.flags), "s"(a.wdist), "s"(a.acc), "s"(a.acc2),
                         "s"(a.total_slots), "s"(a.tiles_x), "s"(a.options), "s"(nb), "s"(L.nx), "s"(L.ny),
                         "s"(L.pitch), "s"(L.f_off), "s"(L.plane), "s"(L.cell_off), "s"(L.les));
    }
    int b = static_cast<int>(blockIdx.x);
    if (a.options & LAUNCH_XCD_ORDER) {
        const int q = nb >> 3, r = nb & 7, xcd = b & 7;
        b = xcd * q + (xcd < r ? xcd : r) + (b >> 3);
    }
    int x0, y0;
    if constexpr (SINGLE) {
        x0 = (b % a.tiles_x) * TX;
        y0 = (b / a.tiles_x) * TY;
    } else {
        const Tile t = a.tiles[b];
        lat_id = t.lattice;
        x0 = t.x0;
        y0 = t.y0;
        L = a.lat[lat_id];
        if (a.src) {
            // the other tiles of a one-lattice handle whose bulk tiles go to step2_bulk_kernel
            src = static_cast<const S*>(a.src);
            dst = static_cast<S*>(a.dst);
        } else {
            const int sp = a.par[lat_id] ^ (a.flips & 1);
            src = static_cast<const S*>(sp ? a.buf1 : a.buf0);
            dst = static_cast<S*>(a.dst ? a.dst : (sp ? a.buf0 : a.buf1));
        }
    }
    const int copy = static_cast<int>(blockIdx.x % kAccCopies);
    double* acc0 = a.acc + static_cast<int64_t>(copy) * a.total_slots * 2;
    double* acc1 = a.acc2 + static_cast<int64_t>(copy) * a.total_slots * 2;
    const Ctx<ST> c{L, src + L.f_off, a.flags, a.link_slot, a.links, a.inv_scale};
    const int tid = static_cast<int>(threadIdx.x);

    // the second step's per-node data, fetched with the first step's loads
    uint8_t fl2[P2];
    int32_t ls2[P2];
    uint16_t wd2[P2];
#pragma unroll
    for (int k = 0; k < P2; ++k) {
        const int n = tid + 256 * k, x = x0 + n % TX, y = y0 + n / TX;
        fl2[k] = 0;
        ls2[k] = 0;
        wd2[k] = 0xffffu;
        if (x < L.nx && y < L.ny && (SINGLE || x >= 0)) {
            const int64_t idx = L.cell_off + static_cast<int64_t>(y) * L.pitch + x;
            fl2[k] = a.flags[idx];
            ls2[k] = a.link_slot[idx];
            if (L.les) wd2[k] = a.wdist[idx];
        }
    }

    const ForceSink fs0{a.nf, acc0}, fs1{a.nf2, acc1}, none{nullptr, nullptr};
    bool bad = false;
    Rounding rd0{a.sr != 0, 0u}, rd1{a.sr != 0, 0u};
    if (rd0.on) {
        rd0.key = step_key(a, lat_id, 0);
        rd1.key = step_key(a, lat_id, 1);
    }
    // ensembles: ring index r -> LDS position and lattice node (wrapped); false: off a
    // non-periodic edge
    auto ring_node = [&](int r, int& lx, int& ly, int& gx, int& gy, bool& own) {
        lx = r % WX;
        ly = r / WX;
        const int ux = x0 - 1 + lx, uy = y0 - 1 + ly;  // unwrapped position
        // no tile node has a neighbour further than one node off the lattice (positions
        // there exist when the lattice is narrower or lower than the tile and its ring)
        if (ux < -1 || ux > L.nx || uy < -1 || uy > L.ny) return false;
        gx = ux;
        gy = uy;
        if (gx < 0 || gx >= L.nx) {
            if (L.x_edge != EDGE_PERIODIC) return false;
            gx = gx < 0 ? gx + L.nx : gx - L.nx;
        }
        if (gy < 0 || gy >= L.ny) {
            if (L.y_edge != EDGE_PERIODIC) return false;
            gy = gy < 0 ? gy + L.ny : gy - L.ny;
        }
        own = lx >= 1 && lx <= TX && ly >= 1 && ly <= TY && ux < L.nx && uy < L.ny && ux >= 0;
        return true;
    };
    const int lane = tid & 63, wave = tid >> 6;
    const uint64_t below = (uint64_t(1) << lane) - 1;
    // step k on the tile and its ring
    if constexpr (!DEFER) {
        for (int r = tid; r < WX * WY; r += 256) {
            const int lx = r % WX, ly = r / WX;
            const int ux = x0 - 1 + lx, uy = y0 - 1 + ly;  // unwrapped position
            // no tile node has a neighbour further than one node off the lattice
            if (ux > L.nx || uy > L.ny) continue;
            int gx = ux, gy = uy;
            if (gx < 0 || gx >= L.nx) {
                if (L.x_edge != EDGE_PERIODIC) continue;
                gx = gx < 0 ? gx + L.nx : gx - L.nx;
            }
            if (gy < 0 || gy >= L.ny) {
                if (L.y_edge != EDGE_PERIODIC) continue;
                gy = gy < 0 ? gy + L.ny : gy - L.ny;
            }
            const bool own = lx >= 1 && lx <= TX && ly >= 1 && ly <= TY && ux < L.nx && uy < L.ny;
            S out[9];
            if (node_values<ST, false, NV_ALL>(L, a, c, own ? fs0 : none, gx, gy, rd0, bad, out)) {
#pragma unroll
                for (int j = 0; j < 9; ++j) lds[j][ly][lx] = out[j];
            }
        }
    } else {
        int wcnt = 0;
        for (int r0 = 0; r0 < WX * WY; r0 += 256) {
            const int r = r0 + tid;
            bool defer = false;
            int lx, ly, gx, gy;
            bool own;
            if (r < WX * WY && ring_node(r, lx, ly, gx, gy, own)) {
                S out[9];
                const int st = node_values<ST, true, NV_PLAIN>(L, a, c, own ? fs0 : none, gx, gy, rd0, bad, out);
                if (st > 0) {
#pragma unroll
                    for (int j = 0; j < 9; ++j) lds[j][ly][lx] = out[j];
                }
                defer = st < 0;
            }
            const uint64_t m = __ballot(defer);
            if (defer) queue[wave * SEG + wcnt + __popcll(m & below)] = static_cast<int16_t>(r);
            wcnt += __popcll(m);
        }
        if (lane == 0) qcnt[wave] = wcnt;
        __syncthreads();
        const int c0 = qcnt[0], c1 = qcnt[1], c2 = qcnt[2], c3 = qcnt[3];
        for (int k = tid; k < c0 + c1 + c2 + c3; k += 256) {
            int i = k, w = 0;
            if (i >= c0) {
                i -= c0;
                w = 1;
                if (i >= c1) {
                    i -= c1;
                    w = 2;
                    if (i >= c2) {
                        i -= c2;
                        w = 3;
                    }
                }
            }
            int lx, ly, gx, gy;
            bool own;
            ring_node(queue[w * SEG + i], lx, ly, gx, gy, own);
            S out[9];
            if (node_values<ST, true, NV_GENERAL>(L, a, c, own ? fs0 : none, gx, gy, rd0, bad, out)) {
#pragma unroll
                for (int j = 0; j < 9; ++j) lds[j][ly][lx] = out[j];
            }
        }
    }
    __syncthreads();

    if constexpr (!DEFER) {
        // step k+1 on the tile, from LDS (one copy of the loop body: the general path is long)
#pragma unroll 1
        for (int k = 0; k < P2; ++k) {
            const int n = tid + 256 * k, tx = n % TX, ty = n / TX, x = x0 + tx, y = y0 + ty;
            if (x >= L.nx || y >= L.ny) continue;
            uint8_t fl = fl2[0];
            int32_t lsk = ls2[0];
            uint16_t wdk = wd2[0];
#pragma unroll
            for (int q = 1; q < P2; ++q)
                if (k == q) {
                    fl = fl2[q];
                    lsk = ls2[q];
                    wdk = wd2[q];
                }
            const int lx = tx + 1, ly = ty + 1;
            R d[9];
            if (fl & CELL_SPECIAL) {
                if ((fl & CELL_SOLID) && !(fl & CELL_WRITE)) continue;
                const int gx = gathered_x<ST>(L, x), glx = gx - x0 + 1;
                Pre<ST> p;
                if (gx == x) {
                    p.fl = fl;
                    p.ls = lsk;
                } else {
                    const int64_t gidx = L.cell_off + static_cast<int64_t>(y) * L.pitch + gx;
                    p.fl = a.flags[gidx];
                    p.ls = a.link_slot[gidx];
                }
#pragma unroll
                for (int j = 0; j < 9; ++j) {
                    p.own[j] = lds[j][ly][glx];
                    p.pull[j] = lds[j][ly - CY[j]][glx - CX[j]];
                }
                LinkVals<R> lg;
                links_of<ST>(c, p, lg);
                LinkVals<R> ls_ = lg;
                if (gx != x && (fl & CELL_LINKS)) load_link(c.links[L.link_off + lsk], ls_);
                node_general_d<ST>(L, c, fs1, x, y, fl, p, lg, ls_, wdk, bad, d);
            } else {
#pragma unroll
                for (int j = 0; j < 9; ++j)
                    d[j] = Form<ST>::dec(lds[j][ly - CY[j]][lx - CX[j]], a.inv_scale, L.shift[j]);
                R wd = R(0);
                if (L.les) wd = wdk == 0xffffu ? R(1e9) : static_cast<R>(wdk);
                collide<R>(d, L, wd, R(1));
                if (!(d[0] == d[0])) bad = true;
            }
            const int64_t idx = static_cast<int64_t>(y) * L.pitch + x;
            S* o = dst + L.f_off + idx;
            uint32_t h = rd1.on ? rd1.node(L.cell_off + idx) : 0u;
#pragma unroll
            for (int j = 0; j < 9; ++j)
                store_val(o + j * L.plane, enc_node<ST>(d[j], a.scale, L.nshift[j], rd1, h, j),
                          a.options & LAUNCH_STREAM_STORES);
        }
    } else {
        // step k+1 at tile node (x, y), from LDS: the general path
        auto general2 = [&](int x, int y, int lx, int ly, uint8_t fl, int32_t lsk, uint16_t wdk, R d[9]) {
            const int gx = gathered_x<ST>(L, x), glx = gx - x0 + 1;
            Pre<ST> p;
            if (gx == x) {
                p.fl = fl;
                p.ls = lsk;
            } else {
                const int64_t gidx = L.cell_off + static_cast<int64_t>(y) * L.pitch + gx;
                p.fl = a.flags[gidx];
                p.ls = a.link_slot[gidx];
            }
#pragma unroll
            for (int j = 0; j < 9; ++j) {
                p.own[j] = lds[j][ly][glx];
                p.pull[j] = lds[j][ly - CY[j]][glx - CX[j]];
            }
            LinkVals<R> lg;
            links_of<ST>(c, p, lg);
            LinkVals<R> ls_ = lg;
            if (gx != x && (fl & CELL_LINKS)) load_link(c.links[L.link_off + lsk], ls_);
            node_general_d<ST>(L, c, fs1, x, y, fl, p, lg, ls_, wdk, bad, d);
        };
        auto store2 = [&](int x, int y, const R d[9]) {
            const int64_t idx = static_cast<int64_t>(y) * L.pitch + x;
            S* o = dst + L.f_off + idx;
            uint32_t h = rd1.on ? rd1.node(L.cell_off + idx) : 0u;
#pragma unroll
            for (int j = 0; j < 9; ++j)
                store_val(o + j * L.plane, enc_node<ST>(d[j], a.scale, L.nshift[j], rd1, h, j),
                          a.options & LAUNCH_STREAM_STORES);
        };
        // step k+1 on the tile, from LDS: plain nodes now, general-path nodes queued
        int wcnt2 = 0;
#pragma unroll 1
        for (int k = 0; k < P2; ++k) {
            const int n = tid + 256 * k, tx = n % TX, ty = n / TX, x = x0 + tx, y = y0 + ty;
            bool defer = false;
            if (x < L.nx && y < L.ny && x >= 0) {
                uint8_t fl = fl2[0];
                uint16_t wdk = wd2[0];
#pragma unroll
                for (int q = 1; q < P2; ++q)
                    if (k == q) {
                        fl = fl2[q];
                        wdk = wd2[q];
                    }
                const int lx = tx + 1, ly = ty + 1;
                const bool column = (fl & CELL_COLUMN) != 0;
                if ((fl & CELL_SPECIAL) && !column && !wrap_only(L, fl, x, y)) {
                    defer = !((fl & CELL_SOLID) && !(fl & CELL_WRITE));
                } else {
                    // the ring holds the wrapped sources; the outlet column pulls those of x - 1
                    const int slx = (column && x == L.nx - 1 && L.outlet) ? lx - 1 : lx;
                    R d[9];
#pragma unroll
                    for (int j = 0; j < 9; ++j)
                        d[j] = Form<ST>::dec(lds[j][ly - CY[j]][slx - CX[j]], a.inv_scale, L.shift[j]);
                    if (column && x == 0) inlet_closure<R>(L, y, d);
                    R wd = R(0);
                    if (L.les) wd = wdk == 0xffffu ? R(1e9) : static_cast<R>(wdk);
                    collide<R>(d, L, wd, R(1));
                    if (!(d[0] == d[0])) bad = true;
                    store2(x, y, d);
                }
            }
            const uint64_t m = __ballot(defer);
            if (defer) queue[wave * SEG + wcnt2 + __popcll(m & below)] = static_cast<int16_t>(n);
            wcnt2 += __popcll(m);
        }
        // every thread read the step-1 counts before the barrier that opened step 2
        if (lane == 0) qcnt[wave] = wcnt2;
        __syncthreads();
        const int c0 = qcnt[0], c1 = qcnt[1], c2 = qcnt[2], c3 = qcnt[3];
        for (int k = tid; k < c0 + c1 + c2 + c3; k += 256) {
            int i = k, w = 0;
            if (i >= c0) {
                i -= c0;
                w = 1;
                if (i >= c1) {
                    i -= c1;
                    w = 2;
                    if (i >= c2) {
                        i -= c2;
                        w = 3;
                    }
                }
            }
            const int n = queue[w * SEG + i], tx = n % TX, ty = n / TX, x = x0 + tx, y = y0 + ty;
            const int64_t idx = L.cell_off + static_cast<int64_t>(y) * L.pitch + x;
            const uint8_t fl = a.flags[idx];
            const int32_t lsk = a.link_slot[idx];
            const uint16_t wdk = L.les ? a.wdist[idx] : uint16_t(0xffffu);
            R d[9];
            general2(x, y, tx + 1, ty + 1, fl, lsk, wdk, d);
            store2(x, y, d);
        }
    }
    if (bad) atomicOr(a.nan_flags + lat_id, 1);
}

// ------------------------------------------------------------------ two steps per launch: bulk tiles

// Row of ring position r of a 66-wide ring by a 24-bit multiply and a shift: floor(r * 993 /
// 2^16) = floor(r / 66) while r < 32768 (993 / 2^16 exceeds 1/66 by 4.6e-7, and
// 65/66 + 32768 * 4.6e-7 < 1).
__host__ __device__ constexpr int ring_row(int r) { return static_cast<int>((static_cast<uint32_t>(r) * 993u) >> 16); }

constexpr bool ring_row_exact(int n)
{
    for (int r = 0; r < n; ++r)
        if (ring_row(r) != r / 66) return false;
    return true;
}

// Value at a byte offset from a base pointer the whole wave shares: the offset is a 32-bit
// lane value, so the address is the scalar base plus the zero-extended offset.
template <typename T> __device__ __forceinline__ T load_at(const T* base, uint32_t byte_off)
{
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + byte_off);
}

// Stochastic rounding of x (the 16-bit forms' enc_sr_raw) into the low or high half of
// pk, keeping the other half.
template <int ST, bool HI> __device__ __forceinline__ uint32_t cvt_sr_half(uint32_t pk, float x, uint32_t rnd)
{
    if constexpr (ST == ST_F16S) {
        typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
        return __builtin_bit_cast(uint32_t,
                                  __builtin_amdgcn_cvt_sr_f16_f32(__builtin_bit_cast(half2_t, pk), x, rnd, HI));
    } else {
        typedef short short2_t __attribute__((ext_vector_type(2)));
        return __builtin_bit_cast(uint32_t,
                                  __builtin_amdgcn_cvt_sr_bf16_f32(__builtin_bit_cast(short2_t, pk), x, rnd, HI));
    }
}

// a and b rounded to nearest even into the low and high halves (the 16-bit forms).
template <int ST> __device__ __forceinline__ uint32_t cvt_pk_rne(float a, float b)
{
    typedef float float2_t __attribute__((ext_vector_type(2)));
    if constexpr (ST == ST_F16S) {
        typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
        return __builtin_bit_cast(uint32_t, __builtin_convertvector(float2_t{a, b}, half2_t));
    } else {
        typedef __bf16 bf162_t __attribute__((ext_vector_type(2)));
        return __builtin_bit_cast(uint32_t, __builtin_convertvector(float2_t{a, b}, bf162_t));
    }
}

// Threads of a bulk-kernel block: 512 when the tile's LDS (9 planes of 66 x (ty + 2) stored
// values) exceeds 24 KiB, so that the few blocks a CU holds still bring enough waves.
__host__ __device__ constexpr int bulk_threads(int bytes_per_value, int ty)
{
    return 9 * 66 * (ty + 2) * bytes_per_value > 24 * 1024 ? 512 : 256;
}

// Decode, collide, encode one node of a bulk or edge tile in place (stored units, as
// group_update; the shift additions are skipped without a shift, where they are exact
// identities).  inlet (EDGE): the node is an inlet node of row y.
template <int ST, bool SR, bool EDGE>
__device__ __forceinline__ void bulk_update(typename Form<ST>::store v[9], const LatticeDev& L,
                                            typename Form<ST>::real sc, bool les, uint16_t wraw, uint32_t key,
                                            uint32_t cell, bool inlet, int y, bool& bad)
{
    using S = typename Form<ST>::store;
    using R = typename Form<ST>::real;
    R d[9];
#pragma unroll
    for (int j = 0; j < 9; ++j) d[j] = Form<ST>::dec_raw(v[j]);
    if constexpr (Form<ST>::shifted) {
        if (L.shifted) {
#pragma unroll
            for (int j = 0; j < 9; ++j) d[j] += L.shift_s[j];
        }
    }
    if constexpr (EDGE) {
        if (inlet) inlet_closure<R>(L, y, d, sc);
    }
    R wd = R(0);
    if (les) wd = wraw == 0xffffu ? R(1e9) : static_cast<R>(wraw);
    collide<R>(d, L, wd, sc);
    if (!(d[0] == d[0])) bad = true;
    if constexpr (Form<ST>::shifted) {
        if (L.shifted) {
#pragma unroll
            for (int j = 0; j < 9; ++j) d[j] += L.nshift_s[j];
        }
    }
    if constexpr (SR) {
        // enc_node_raw's words and conversions, two populations per register (each
        // conversion writes one half and keeps the other)
        uint32_t h = mix32(key ^ (cell * 0x9e3779b1u));
#pragma unroll
        for (int j = 0; j < 9; j += 2) {
            // the pair's register starts as the first random word (dead after its use)
            const uint32_t r0 = next_rand(h, j);
            uint32_t pk = cvt_sr_half<ST, false>(r0, d[j], r0);
            if (j < 8) pk = cvt_sr_half<ST, true>(pk, d[j + 1], next_rand(h, j + 1));
            v[j] = static_cast<S>(pk);
            if (j < 8) v[j + 1] = static_cast<S>(pk >> 16);
        }
    } else if constexpr (Form<ST>::shifted) {
        // round to nearest even, two conversions per instruction (v_cvt_pk_f16_f32,
        // v_cvt_pk_bf16_f32): the bits of enc_raw's one at a time
#pragma unroll
        for (int j = 0; j < 8; j += 2) {
            const uint32_t pk = cvt_pk_rne<ST>(d[j], d[j + 1]);
            v[j] = static_cast<S>(pk);
            v[j + 1] = static_cast<S>(pk >> 16);
        }
        v[8] = Form<ST>::enc_raw(d[8]);
    } else {
#pragma unroll
        for (int j = 0; j < 9; ++j) v[j] = Form<ST>::enc_raw(d[j]);
    }
}

// Steps k and k+1 of the bulk tiles of a lattice: 64 x TY tiles whose ring of nodes
// (66 x (TY + 2)) lies inside the lattice and holds no general-path node -- no edge,
// boundary link, solid, inlet, outlet or lid node -- so every node of the launch is a plain
// pull and a collision.  The host lists them (a.bulk_tiles); the other tiles go to the
// ensemble form of step2_kernel through a table of their own, on a second stream.  ENS:
// the tiles belong to the lattices of an ensemble (descriptor and buffer parity per tile's
// lattice, as step2_kernel's ensemble form), else to the one lattice of a.L0 (buffers
// a.src, a.dst).  Without the general path the kernel keeps its values and pointers in
// registers: the pull addresses are per-plane scalar bases plus one 32-bit lane offset,
// the ring's row and column come from 24-bit integer arithmetic, and the
// stochastic-rounding choice is compiled in (SR).  Decoding, collision and encoding are
// group_update's, in stored units, so the results are bit for bit those of every other
// kernel.
//
// EDGE: the edge tiles instead (a.bulk_tiles holds them) -- 64 x TY tiles from a node of
// their lattice (cut off at its last column and row) whose ring holds, besides plain nodes,
// only inlet and outlet column nodes (CELL_COLUMN) and nodes whose general path reduces to
// wrapped pulls (wrap_only).  Ring positions wrap around periodic edges and are skipped off
// the others (only an inlet node's closed slots read there) and more than one node off the
// lattice (no tile node reads there); an inlet node closes its populations in stored units
// (inlet_closure scaled by 2^k: the bits of the closure in physical units, scaled
// exactly); an outlet node takes the pulls of the node before it (the outflow copy).  Pull
// addresses are whole node offsets within a plane, from the three source rows and columns
// of the node, wrapped.
template <int ST, int TY, bool SR, bool ENS, bool EDGE>
__global__ __launch_bounds__(bulk_threads(sizeof(typename Form<ST>::store), TY)) void step2_bulk_kernel(StepArgs a)
{
    using S = typename Form<ST>::store;
    using R = typename Form<ST>::real;
    constexpr int NT = bulk_threads(sizeof(S), TY);
    constexpr int TX = 64, WX = TX + 2, WY = TY + 2, NR = WX * WY, P1 = (NR + NT - 1) / NT, P2 = TX * TY / NT;
    static_assert(ring_row_exact(NR), "ring_row must be exact over the ring");
    static_assert(P2 * NT == TX * TY, "whole tiles of NT threads");
    __shared__ S lds[9][WY][WX];

    const int nb = static_cast<int>(gridDim.x);
    int b = static_cast<int>(blockIdx.x);
    if (a.options & LAUNCH_XCD_ORDER) {
        const int q = nb >> 3, r = nb & 7, xcd = b & 7;
        b = xcd * q + (xcd < r ? xcd : r) + (b >> 3);
    }
    const Tile t = a.bulk_tiles[b];
    const int lat_id = ENS ? t.lattice : a.lattice0;
    LatticeDev L;
    const S* base_src;
    S* base_dst;
    if constexpr (ENS) {
        L = a.lat[lat_id];
        const int sp = a.par[lat_id] ^ (a.flips & 1);
        base_src = static_cast<const S*>(sp ? a.buf1 : a.buf0);
        base_dst = static_cast<S*>(a.dst ? a.dst : (sp ? a.buf0 : a.buf1));
    } else {
        L = a.L0;
        base_src = static_cast<const S*>(a.src);
        base_dst = static_cast<S*>(a.dst);
    }
    // ring position (lx, ly) is lattice node (x0 - 1 + lx, y0 - 1 + ly): element ly * pitch
    // + lx from the ring origin (less than (TY + 2) pitch + 66, a 32-bit offset).  EDGE:
    // the origin is the plane's first node, offsets are node offsets in the plane
    const int64_t org = EDGE ? 0 : static_cast<int64_t>(t.y0 - 1) * L.pitch + (t.x0 - 1);
    const S* src = base_src + L.f_off + org;
    S* dst = base_dst + L.f_off + org;
    const uint16_t* wdb = a.wdist + L.cell_off + org;
    const uint32_t cell0 = static_cast<uint32_t>(L.cell_off + org);
    const int pitch = L.pitch;
    const bool nt = (a.options & LAUNCH_STREAM_STORES) != 0;
    const bool les = L.les != 0;
    const R sc = static_cast<R>(a.scale);
    uint32_t key0 = 0, key1 = 0;
    if constexpr (SR) {
        key0 = step_key(a, lat_id, 0);
        key1 = step_key(a, lat_id, 1);
    }
    const int tid = static_cast<int>(threadIdx.x);
    const int tx = tid & 63, ty0 = tid >> 6;
    constexpr int RS = NT / 64;  // tile rows per step-2 iteration
    // offset of the tile node at ring position (lx, ly) from the origin
    auto tile_off = [&](int lx, int ly) -> uint32_t {
        if constexpr (EDGE)
            return __umul24(t.y0 - 1 + ly, pitch) + static_cast<uint32_t>(t.x0 - 1 + lx);
        else
            return __umul24(ly, pitch) + static_cast<uint32_t>(lx);
    };
    // EDGE: lattice node (gx, gy) of ring position (lx, ly), wrapped around every edge and
    // kept inside the lattice; false off a non-periodic edge or more than one node off
    auto ring_xy = [&](int lx, int ly, int& gx, int& gy) {
        gx = t.x0 - 1 + lx;
        gy = t.y0 - 1 + ly;
        bool in = gx <= L.nx && gy <= L.ny;
        if (gx < 0 || gx >= L.nx) {
            in = in && L.x_edge == EDGE_PERIODIC;
            gx = gx < 0 ? gx + L.nx : gx - L.nx;
        }
        if (gy < 0 || gy >= L.ny) {
            in = in && L.y_edge == EDGE_PERIODIC;
            gy = gy < 0 ? gy + L.ny : gy - L.ny;
        }
        gx = min(gx, L.nx - 1);
        gy = min(gy, L.ny - 1);
        return in;
    };
    // EDGE: whether the tile node at ring position (lx, ly) lies in the lattice
    auto tile_node = [&](int lx, int ly) { return !EDGE || (t.x0 - 1 + lx < L.nx && t.y0 - 1 + ly < L.ny); };

    // the second step's wall distances, fetched with the first step's loads
    uint16_t wd2[P2];
#pragma unroll
    for (int k = 0; k < P2; ++k) {
        wd2[k] = 0xffffu;
        if (les && tile_node(tx + 1, ty0 + RS * k + 1)) wd2[k] = load_at(wdb, tile_off(tx + 1, ty0 + RS * k + 1) * 2u);
    }

    auto update = [&](S v[9], uint16_t wraw, uint32_t key, uint32_t cell, bool inlet, int y, bool& bad) {
        bulk_update<ST, SR, EDGE>(v, L, sc, les, wraw, key, cell, inlet, y, bad);
    };

    // pulls and wall distance of ring position r (r < NR), one register each (kept apart
    // until used: packing two 16-bit values into one register would wait for both loads)
    using W = typename std::conditional<sizeof(S) == 2, uint32_t, S>::type;
    auto fetch = [&](int r, W v[9], uint32_t& wraw) {
        const int ly = ring_row(r), lx = r - 66 * ly;
        if constexpr (EDGE) {
            int gx, gy;
            ring_xy(lx, ly, gx, gy);
            // the gathered column (x - 1 on the outlet column) and the source rows and
            // columns around it, wrapped (across an open edge too: an inlet node's sources
            // there are replaced by its closure, and the address stays in the plane)
            const int px = (L.outlet && gx == L.nx - 1) ? gx - 1 : gx;
            const uint32_t xl = px == 0 ? L.nx - 1 : px - 1, xr = px == L.nx - 1 ? 0 : px + 1;
            const uint32_t rd = __umul24(gy == 0 ? L.ny - 1 : gy - 1, pitch), r0 = __umul24(gy, pitch),
                           ru = __umul24(gy == L.ny - 1 ? 0 : gy + 1, pitch);
#pragma unroll
            for (int j = 0; j < 9; ++j) {
                const uint32_t row = CY[j] == 1 ? rd : (CY[j] == 0 ? r0 : ru);
                const uint32_t col = CX[j] == 1 ? xl : (CX[j] == 0 ? static_cast<uint32_t>(px) : xr);
                v[j] = load_at(src + j * L.plane, (row + col) * sizeof(S));
            }
            wraw = 0xffffu;
            if (les) wraw = load_at(wdb, (r0 + gx) * 2u);
        } else {
            const uint32_t o = __umul24(ly, pitch) + lx;  // pitch < 2^24
#pragma unroll
            for (int j = 0; j < 9; ++j) v[j] = load_at(src + j * L.plane - (CY[j] * pitch + CX[j]), o * sizeof(S));
            wraw = 0xffffu;
            if (les) wraw = load_at(wdb, o * 2u);
        }
    };

    bool bad = false;
    // step k on the tile and its ring, from the source buffer.  A thread's ring positions
    // are tid, tid + NT, ...; the loads of the next PF positions are in flight while one
    // is computed (16-bit loads move 128 bytes a wave: more of them must be outstanding to
    // cover the memory latency).  The last position of the ring stands in for those past
    // it: loaded, not used.
    constexpr int PF = sizeof(S) == 2 ? 2 : 1;
    W q[PF][9];
    uint32_t wq[PF];
#pragma unroll
    for (int p = 0; p < PF; ++p) fetch(tid + NT * p < NR ? tid + NT * p : NR - 1, q[p], wq[p]);
    // unrolled: the queue is renamed registers, not copies (a copy waits for its load)
#pragma unroll
    for (int k = 0; k < P1; ++k) {
        const int r = tid + NT * k;
        W nxt[9];
        uint32_t wnxt = 0xffffu;
        if (k + PF < P1) fetch(r + NT * PF < NR ? r + NT * PF : NR - 1, nxt, wnxt);
        if (r < NR) {
            const int ly = ring_row(r), lx = r - 66 * ly;
            int gx = 0, gy = 0;
            const bool in = EDGE ? ring_xy(lx, ly, gx, gy) : true;
            if (in) {
                S v[9];
#pragma unroll
                for (int j = 0; j < 9; ++j) v[j] = static_cast<S>(q[0][j]);
                const uint32_t cell = EDGE ? cell0 + __umul24(gy, pitch) + static_cast<uint32_t>(gx)
                                           : cell0 + __umul24(ly, pitch) + static_cast<uint32_t>(lx);
                update(v, static_cast<uint16_t>(wq[0]), key0, cell, EDGE && gx == 0 && L.inlet != INLET_NONE, gy, bad);
#pragma unroll
                for (int j = 0; j < 9; ++j) lds[j][ly][lx] = v[j];
            }
        }
#pragma unroll
        for (int p = 0; p + 1 < PF; ++p) {
#pragma unroll
            for (int j = 0; j < 9; ++j) q[p][j] = q[p + 1][j];
            wq[p] = wq[p + 1];
        }
#pragma unroll
        for (int j = 0; j < 9; ++j) q[PF - 1][j] = nxt[j];
        wq[PF - 1] = wnxt;
    }
    __syncthreads();
    // step k+1 on the tile, from LDS
#pragma unroll 1
    for (int k = 0; k < P2; ++k) {
        const int lx = tx + 1, ly = ty0 + RS * k + 1;
        if (!tile_node(lx, ly)) continue;
        const uint32_t o = tile_off(lx, ly);
        uint16_t wraw = wd2[0];
#pragma unroll
        for (int q = 1; q < P2; ++q)
            if (k == q) wraw = wd2[q];
        // EDGE: the outlet column pulls those of x - 1 (the ring holds them), an inlet node
        // closes its populations
        int slx = lx, y = 0;
        bool inlet = false;
        if constexpr (EDGE) {
            const int x = t.x0 + tx;
            y = t.y0 + ly - 1;
            if (L.outlet && x == L.nx - 1) slx = lx - 1;
            inlet = x == 0 && L.inlet != INLET_NONE;
        }
        S v[9];
#pragma unroll
        for (int j = 0; j < 9; ++j) v[j] = lds[j][ly - CY[j]][slx - CX[j]];
        update(v, wraw, key1, cell0 + o, inlet, y, bad);
#pragma unroll
        for (int j = 0; j < 9; ++j)
            store_val(reinterpret_cast<S*>(reinterpret_cast<char*>(dst + j * L.plane) + o * sizeof(S)), v[j], nt);
    }
    if (bad) atomicOr(a.nan_flags + lat_id, 1);
}

// Classes of two-step tiles (64 x ty from (x0, y0) of its lattice): out[t] = 1 (bulk) when
// the tile's ring of nodes lies inside the lattice and holds no general-path node (edge
// nodes are general-path nodes, so the ring's pulls stay inside the lattice too); 2 (edge)
// when the tile starts on a lattice node and every node of its ring, wrapped around
// periodic edges and skipped off the others and more than one node off the lattice, is
// plain, an inlet or outlet column node or a node whose general path reduces to wrapped
// pulls (and an outlet node is not the tile's first column); 3 (empty) when every node of the tile is
// a solid node nothing reads (no value to compute, no force: the tile can be left out);
// else 0.
__global__ __launch_bounds__(256) void tile_class_kernel(const Tile* tiles, const LatticeDev* lat,
                                                         const uint8_t* flags, int ty, uint8_t* out)
{
    const Tile t = tiles[blockIdx.x];
    const LatticeDev& L = lat[t.lattice];
    const int x0 = t.x0, y0 = t.y0;
    const bool inside = x0 >= 1 && x0 + 64 <= L.nx - 1 && y0 >= 1 && y0 + ty <= L.ny - 1;
    // edge tiles start on a lattice node (x0, y0 >= 0) and may be cut off at the last column
    // and row; an outlet node takes the ring column before it, which must lie in the ring
    const bool whole = x0 >= 0 && y0 >= 0 && !(L.outlet && x0 == L.nx - 1);
    int special = 0, general = 0, live = 0;
    for (int r = static_cast<int>(threadIdx.x); r < 64 * ty; r += 256) {
        const int x = x0 + (r & 63), y = y0 + (r >> 6);
        if (x >= 0 && x < L.nx && y < L.ny) {
            const uint8_t fl = flags[L.cell_off + static_cast<int64_t>(y) * L.pitch + x];
            if (!(fl & CELL_SOLID) || (fl & CELL_WRITE)) live = 1;
        }
    }
    if (whole) {
        const uint8_t* f = flags + L.cell_off;
        for (int r = static_cast<int>(threadIdx.x); r < 66 * (ty + 2); r += 256) {
            const int ly = r / 66, lx = r - 66 * ly;
            int gx = x0 - 1 + lx, gy = y0 - 1 + ly;
            bool in = gx <= L.nx && gy <= L.ny;
            if (gx < 0 || gx >= L.nx) {
                in = in && L.x_edge == EDGE_PERIODIC;
                gx = gx < 0 ? gx + L.nx : gx - L.nx;
            }
            if (gy < 0 || gy >= L.ny) {
                in = in && L.y_edge == EDGE_PERIODIC;
                gy = gy < 0 ? gy + L.ny : gy - L.ny;
            }
            if (!in) {
                special = 1;
                continue;
            }
            gx = min(gx, L.nx - 1);
            gy = min(gy, L.ny - 1);
            const uint8_t fl = f[static_cast<int64_t>(gy) * L.pitch + gx];
            if (fl & CELL_SPECIAL) {
                special = 1;
                if (!(fl & CELL_COLUMN) && !wrap_only(L, fl, gx, gy)) general = 1;
            }
        }
    }
    const int any = __syncthreads_or(special);
    const int gen = __syncthreads_or(general);
    const int alive = __syncthreads_or(live);
    if (threadIdx.x == 0) out[blockIdx.x] = !alive ? 3 : !whole ? 0 : (inside && !any ? 1 : (gen ? 0 : 2));
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
            Pre<ST> p;
            preload<ST>(c, gathered_x<ST>(L, x), y, p);
            LinkVals<R> lk;
            links_of<ST>(c, p, lk);
            state_general<ST>(c, x, y, solid, p, lk, d);
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
            for (int j = 0; j < 9; ++j) co[j * L.plane] = Form<ST>::enc(d[j], a.scale, L.nshift[j]);
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
    for (int j = 0; j < 9; ++j) st[j * L.plane] = Form<ST>::enc(d[j], a.scale, L.nshift[j]);
    if (!(a.flags[L.cell_off + idx] & CELL_SOLID)) {
        R wd = R(0);
        if (L.les) {
            const uint16_t w = a.wdist[L.cell_off + idx];
            wd = w == 0xffffu ? R(1e9) : static_cast<R>(w);
        }
        collide<R>(d, L, wd, R(1));
    }
#pragma unroll
    for (int j = 0; j < 9; ++j) co[j * L.plane] = Form<ST>::enc(d[j], a.scale, L.nshift[j]);
}

// ------------------------------------------------------------------ initial perturbation

// The cylinder drivers' shedding trigger (src/main.cpp:67-83) on the stored state: fluid
// nodes of columns [x0, x1) take feq(rho, u, v + dv) of their own moments, computed in
// FP64 in the reference's order (compute_macros, compute_equilibrium).
template <int ST>
__global__ __launch_bounds__(256) void perturb_kernel(PerturbArgs a)
{
#pragma clang fp contract(off)
    using S = typename Form<ST>::store;
    const LatticeDev L = a.lat[a.lattice];
    const int y = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    const int x = a.x0 + static_cast<int>(blockIdx.y);
    if (y >= L.ny || x >= a.x1) return;
    const int64_t idx = static_cast<int64_t>(y) * L.pitch + x;
    if (a.flags[L.cell_off + idx] & CELL_SOLID) return;
    S* st = static_cast<S*>(a.state) + L.f_off + idx;
    double f[9], rho = 0.0, mx = 0.0, my = 0.0;
#pragma unroll
    for (int j = 0; j < 9; ++j) {
        f[j] = WD[j] + static_cast<double>(Form<ST>::dec(st[j * L.plane], a.inv_scale, L.shift[j]));
        rho += f[j];
        mx += f[j] * CX[j];
        my += f[j] * CY[j];
    }
    const double u = mx / rho, v = my / rho + a.dv[static_cast<int64_t>(x - a.x0) * L.ny + y];
#pragma unroll
    for (int j = 0; j < 9; ++j) {
        const double eu = CX[j] * u + CY[j] * v;
        const double uu = u * u + v * v;
        const double feq = WD[j] * rho * (1.0 + 3.0 * eu + 4.5 * eu * eu - 1.5 * uu);
        st[j * L.plane] = Form<ST>::enc(static_cast<typename Form<ST>::real>(feq - WD[j]), a.scale, L.nshift[j]);
    }
}

// ------------------------------------------------------------------ obstacle rasterisation

// place_cylinder (src/lbm.hpp:931-944): solid where sqrt(dx^2 + dy^2) < r; place_polygon
// (src/lbm.hpp:951-964) with point_in_polygon's even-odd crossing rule
// (src/geometry.hpp:120-134) at the node's integer coordinates.  FP64, unfused, in the
// reference's operation order; cells only become solid (the reference never clears).
__global__ __launch_bounds__(256) void place_kernel(PlaceArgs a)
{
#pragma clang fp contract(off)
    const int x = a.x0 + static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    const int y = a.y0 + static_cast<int>(blockIdx.y);
    if (x >= a.x1 || y >= a.y1) return;
    const double px = static_cast<double>(x), py = static_cast<double>(y);
    bool in = false;
    if (a.kind == 0) {
        const double dx = px - a.cx, dy = py - a.cy;
        in = __builtin_sqrt(dx * dx + dy * dy) < a.r;
    } else {
        for (int i = 0, j = a.n - 1; i < a.n; j = i++) {
            const double xi = a.xy[2 * i], yi = a.xy[2 * i + 1];
            const double xj = a.xy[2 * j], yj = a.xy[2 * j + 1];
            if (((yi > py) != (yj > py)) && (px < (xj - xi) * (py - yi) / (yj - yi) + xi)) in = !in;
        }
    }
    if (in) a.mask[static_cast<int64_t>(y) * a.nx + x] = 1;
}

// ------------------------------------------------------------------ cell flags and link records

// Flags and link mask of node (x, y), as the lattice set-up loop of the step's streaming
// rules needs them: a solid node is SOLID | SPECIAL; a fluid node is SPECIAL on the
// lattice's edge, and LINKS | SPECIAL when a neighbour along e_i (i = 1..8, wrapped on
// periodic edges, absent off the others) is solid (src/lbm.hpp:457-484, 542-581).
__global__ __launch_bounds__(256) void geom_cells_kernel(GeomArgs g)
{
    const int x = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    const int y = static_cast<int>(blockIdx.y);
    bool fluid = false, link = false;
    if (x < g.nx) {
        const int64_t c = static_cast<int64_t>(y) * g.nx + x;
        uint8_t fl;
        uint32_t mask = 0;
        if (g.obs[c]) {
            fl = CELL_SOLID | CELL_SPECIAL;
        } else {
            fluid = true;
#pragma unroll
            for (int i = 1; i < 9; ++i) {
                int tx = x + CX[i], ty = y + CY[i];
                bool ok = true;
                if (tx < 0 || tx >= g.nx) {
                    ok = g.x_edge == EDGE_PERIODIC;
                    tx = tx < 0 ? tx + g.nx : tx - g.nx;
                }
                if (ty < 0 || ty >= g.ny) {
                    ok = ok && g.y_edge == EDGE_PERIODIC;
                    ty = ty < 0 ? ty + g.ny : ty - g.ny;
                }
                if (ok && g.obs[static_cast<int64_t>(ty) * g.nx + tx]) mask |= 1u << i;
            }
            const bool edge = x == 0 || x == g.nx - 1 || y == 0 || y == g.ny - 1;
            fl = edge ? CELL_SPECIAL : 0;
            if (mask) {
                fl |= CELL_SPECIAL | CELL_LINKS;
                link = true;
            }
        }
        g.flags[static_cast<int64_t>(y) * g.pitch + x] = fl;
        g.lmask[c] = static_cast<uint8_t>(mask >> 1);
    }
    const uint64_t fb = __ballot(fluid), lb = __ballot(link);
    if ((threadIdx.x & 63) == 0) {
        if (fb) atomicAdd(g.fluid, static_cast<unsigned long long>(__popcll(fb)));
        if (lb) atomicAdd(g.row_links + y, __popcll(lb));
    }
}

// The bulk tiles' inlet and outlet nodes (CELL_COLUMN): one thread per row 1..ny-2.
__global__ __launch_bounds__(256) void geom_columns_kernel(uint8_t* flags, int nx, int ny, int pitch,
                                                           int inlet_from, int outlet)
{
    const int y = 1 + static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    if (y > ny - 2) return;
    uint8_t* row = flags + static_cast<int64_t>(y) * pitch;
    constexpr uint8_t kBlocked = CELL_SOLID | CELL_LINKS;
    if (y >= inlet_from && !(row[0] & kBlocked)) row[0] |= CELL_COLUMN;
    if (outlet && !(row[nx - 1] & kBlocked) && !(row[nx - 2] & kBlocked)) row[nx - 1] |= CELL_COLUMN;
}

// Bouzidi reads the stored value of upstream nodes, which may be solid: every solid node
// is written, and the ones off the edges are counted per row for their list.
__global__ __launch_bounds__(256) void geom_write_kernel(uint8_t* flags, int nx, int ny, int pitch,
                                                         int32_t* row_count)
{
    const int x = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    const int y = static_cast<int>(blockIdx.y);
    bool listed = false;
    if (x < nx) {
        uint8_t& f = flags[static_cast<int64_t>(y) * pitch + x];
        if (f & CELL_SOLID) {
            f |= CELL_WRITE;
            listed = x >= 1 && x <= nx - 2 && y >= 1 && y <= ny - 2;
        }
    }
    const uint64_t b = __ballot(listed);
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(row_count + y, __popcll(b));
}

// One block per row: the row's written solids off the edges in x order, numbered from
// row_off[y] (ballot prefix per wave, the four waves' totals in LDS).
__global__ __launch_bounds__(256) void geom_solids_kernel(const uint8_t* flags, int nx, int ny, int pitch,
                                                          const int32_t* row_off, int32_t* out)
{
    __shared__ int wave_count[4];
    const int y = static_cast<int>(blockIdx.x);
    const int t = static_cast<int>(threadIdx.x), lane = t & 63, w = t >> 6;
    const bool inner_row = y >= 1 && y <= ny - 2;
    int base = row_off[y];
    for (int x0 = 0; x0 < nx; x0 += 256) {
        const int x = x0 + t;
        bool listed = false;
        if (inner_row && x >= 1 && x <= nx - 2) {
            const uint8_t f = flags[static_cast<int64_t>(y) * pitch + x];
            listed = (f & CELL_SOLID) && (f & CELL_WRITE);
        }
        const uint64_t b = __ballot(listed);
        const int below = __popcll(b & ((1ull << lane) - 1ull));
        if (lane == 0) wave_count[w] = __popcll(b);
        __syncthreads();
        int off = 0, total = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            off += k < w ? wave_count[k] : 0;
            total += wave_count[k];
        }
        if (listed) out[base + off + below] = y * pitch + x;
        base += total;
        __syncthreads();  // wave_count is rewritten by the next chunk
    }
}

// One block per row: the row's link cells in x order, numbered from row_off[y] with a
// ballot scan per wave and the four waves' totals in LDS.
__global__ __launch_bounds__(256) void geom_links_kernel(GeomArgs g)
{
    __shared__ int wave_links[4];
    const int y = static_cast<int>(blockIdx.x);
    const int t = static_cast<int>(threadIdx.x), lane = t & 63, w = t >> 6;
    int base = g.row_off[y];
    for (int x0 = 0; x0 < g.nx; x0 += 256) {
        const int x = x0 + t;
        const uint8_t m = x < g.nx ? g.lmask[static_cast<int64_t>(y) * g.nx + x] : uint8_t(0);
        const uint64_t b = __ballot(m != 0);
        const int below = __popcll(b & ((1ull << lane) - 1ull));
        if (lane == 0) wave_links[w] = __popcll(b);
        __syncthreads();
        int off = 0, total = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            off += k < w ? wave_links[k] : 0;
            total += wave_links[k];
        }
        if (m) {
            const int r = base + off + below;
            const int64_t p = static_cast<int64_t>(y) * g.pitch + x;
            g.rec_cell[r] = static_cast<int32_t>(p);
            g.rec_mask[r] = m;
            g.lslot[p] = r;
        }
        base += total;
        __syncthreads();  // wave_links is rewritten by the next chunk
    }
}

// ------------------------------------------------------------------ thermal populations

// Thermal kernel tiles: TW nodes wide (64 or 256), 256 / TW rows, one node per thread.
// The launch takes the width that leaves fewer idle lanes at the end of a row (256 on a
// tie): 800-wide rows idle 4% of the lanes in 64-wide tiles against 22% in 256-wide ones.

// Post-stream thermal populations (as deviations g - w) of node (x,y), pulled from the
// post-collision planes gs (src/lbm.hpp:489-528): y always periodic, x periodic on a
// periodic x edge; a source off the lattice or solid means the population bounced back
// into this node's opposite slot.  Solid nodes are never streamed into (zero).
template <int ST>
__device__ __forceinline__ void thermal_gather(const LatticeDev& L, const typename Form<ST>::store* gs,
                                               const uint8_t* flags, int x, int y, bool solid,
                                               typename Form<ST>::real gd[9])
{
    using R = typename Form<ST>::real;
    if (solid) {
#pragma unroll
        for (int j = 0; j < 9; ++j) gd[j] = -wgt<R>(j);
        return;
    }
    const int64_t idx = static_cast<int64_t>(y) * L.pitch + x;
#pragma unroll
    for (int j = 0; j < 9; ++j) {
        int sx = x - CX[j], sy = y - CY[j];
        sy = sy < 0 ? sy + L.ny : (sy >= L.ny ? sy - L.ny : sy);
        if (L.x_edge == EDGE_PERIODIC) sx = sx < 0 ? sx + L.nx : (sx >= L.nx ? sx - L.nx : sx);
        const int64_t sidx = static_cast<int64_t>(sy) * L.pitch + sx;
        const bool blocked = sx < 0 || sx >= L.nx || (flags[L.cell_off + sidx] & CELL_SOLID);
        gd[j] = Form<ST>::dec(blocked ? gs[OPP[j] * L.plane + idx] : gs[j * L.plane + sidx], 1.f, 0.f);
    }
}

// The thermal drivers' conditions on the state after a step (src/heated_cylinder.cpp:
// 155-164, src/thermal_cavity.cpp:153-188): rule k sets the thermal equilibrium at its
// temperature (src/thermal.hpp:24-30) with the node's velocity from the state d, or at
// rest; THERMAL_MIRROR then copies g[opp(i)] into g[i] in slot order.
template <typename R>
__device__ __forceinline__ void thermal_rules(const ThermalArgs& t, uint8_t code, const R d[9], R gd[9])
{
    const int k = (code & 15) - 1;
    if (k >= 0) {
        const R T = static_cast<R>(t.rule_t[k]);
        R u = R(0), v = R(0);
        if (!((t.rule_rest >> k) & 1u)) {
            // compute_macros on the state (src/lbm_types.hpp:278-288); sum w_i c_i = 0
            R s = R(0), mx = R(0), my = R(0);
#pragma unroll
            for (int i = 0; i < 9; ++i) {
                s += d[i];
                mx += R(CX[i]) * d[i];
                my += R(CY[i]) * d[i];
            }
            const R rho = R(1) + s;
            u = mx / rho;
            v = my / rho;
        }
        const R uu = u * u + v * v;
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            const R eu = R(CX[i]) * u + R(CY[i]) * v;
            gd[i] = wgt<R>(i) * T * (R(1) + R(3) * eu + R(4.5) * eu * eu - R(1.5) * uu) - wgt<R>(i);
        }
    }
    if (code & THERMAL_MIRROR) {
#pragma unroll
        for (int i = 0; i < 9; ++i) gd[i] = gd[OPP[i]];
    }
}

// Thermal collision and Boussinesq forcing of a fluid node (src/lbm.hpp:356-397), on the
// post-collision f deviations d and the pre-collision g deviations gd: the velocity is
// that of the post-collision f, T = sum g; g relaxes at omega_k to the thermal
// equilibrium (src/thermal.hpp:35-41); with beta > 0 and g_buoyancy != 0,
// F = (0, -beta (T - t_ref) g_buoyancy) (src/thermal.hpp:45-52) enters f as
// w_i (3 F.e_i + 9 (e_i.u)(e_i.F)).
template <typename R>
__device__ __forceinline__ void thermal_collide(const ThermalArgs& t, R d[9], R gd[9])
{
    R s = R(0), mx = R(0), my = R(0), dT = R(0);
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        s += d[i];
        mx += R(CX[i]) * d[i];
        my += R(CY[i]) * d[i];
        dT += gd[i];
    }
    const R rho = R(1) + s;
    const R u = mx / rho, v = my / rho;
    const R T = R(1) + dT;  // sum w_i = 1
    const R uu = u * u + v * v;
    const R om = static_cast<R>(t.omega_k);
    const bool buoy = t.beta > 0.0 && t.g_buoyancy != 0.0;
    const R fy = static_cast<R>(-t.beta) * (T - static_cast<R>(t.t_ref)) * static_cast<R>(t.g_buoyancy);
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        const R eu = R(CX[i]) * u + R(CY[i]) * v;
        // g_eq - w = w (T - 1) + w T (3 eu + 4.5 eu^2 - 1.5 u^2)
        const R geq = wgt<R>(i) * (dT + T * (R(3) * eu + R(4.5) * eu * eu - R(1.5) * uu));
        gd[i] -= om * (gd[i] - geq);
        if (buoy) {
            const R ef = R(CY[i]) * fy;
            d[i] += wgt<R>(i) * (R(3) * ef + R(9) * eu * ef);
        }
    }
}

// One lattice with thermal populations, one node per thread over TW-wide tiles of rows
// [t.y0, t.y1): the state of f and g from the plain pulls (interior nodes that are not
// special) or from the general path and thermal_gather, then the node conditions;
// then, by t.mode, the collisions
// and stores of a step (forces into the launch's record terms) or of the current state,
// or the thermal state / temperature as FP64.  With t.from_state the sources hold the
// state itself (uploaded, no step taken): no gather and no conditions.
template <int ST, int TW>
__global__ __launch_bounds__(256) void thermal_kernel(StepArgs a, ThermalArgs t)
{
    using S = typename Form<ST>::store;
    using R = typename Form<ST>::real;
    const LatticeDev& L = a.L0;
    // tiles row-major, each XCD taking one contiguous range of them (the blocks of a
    // launch go round-robin over the 8 XCDs): the rows a tile pulls from its neighbours
    // above and below are mostly in the same L2
    int b = static_cast<int>(blockIdx.x);
    const int nb = static_cast<int>(gridDim.x);
    const int q = nb >> 3, r = nb & 7, xcd = b & 7;
    b = xcd * q + (xcd < r ? xcd : r) + (b >> 3);
    const int tiles_x = (L.nx + TW - 1) / TW;
    const int x = (b % tiles_x) * TW + (static_cast<int>(threadIdx.x) % TW);
    const int y = t.y0 + (b / tiles_x) * (256 / TW) + (static_cast<int>(threadIdx.x) / TW);
    if (x >= L.nx || y >= t.y1) return;
    const int64_t idx = static_cast<int64_t>(y) * L.pitch + x;
    const S* fsrc = static_cast<const S*>(a.src) + L.f_off;
    const S* gsrc = static_cast<const S*>(t.gsrc) + L.f_off;
    // The flag and both bulk pulls in one round trip (as cell_update: every pull address
    // lies inside the population pool and its guard); an interior node that is not
    // special has nine in-lattice fluid neighbours, so both pulls are its state.
    const uint8_t fl = a.flags[L.cell_off + idx];
    S fr[9], gr[9];
#pragma unroll
    for (int j = 0; j < 9; ++j) {
        const int64_t o = j * L.plane + idx - (CY[j] * L.pitch + CX[j]);
        fr[j] = fsrc[o];
        gr[j] = gsrc[o];
    }
    __asm__ volatile("" ::"v"(fr[0]), "v"(fr[1]), "v"(fr[2]), "v"(fr[3]), "v"(fr[4]), "v"(fr[5]), "v"(fr[6]),
                     "v"(fr[7]), "v"(fr[8]), "v"(gr[0]), "v"(gr[1]), "v"(gr[2]), "v"(gr[3]), "v"(gr[4]),
                     "v"(gr[5]), "v"(gr[6]), "v"(gr[7]), "v"(gr[8]));
    const bool solid = (fl & CELL_SOLID) != 0;
    R d[9], gd[9];
    if (t.from_state) {
#pragma unroll
        for (int j = 0; j < 9; ++j) {
            d[j] = Form<ST>::dec(fsrc[j * L.plane + idx], 1.f, 0.f);
            gd[j] = Form<ST>::dec(gsrc[j * L.plane + idx], 1.f, 0.f);
        }
    } else if (!(fl & CELL_SPECIAL)) {
#pragma unroll
        for (int j = 0; j < 9; ++j) {
            d[j] = Form<ST>::dec(fr[j], 1.f, 0.f);
            gd[j] = Form<ST>::dec(gr[j], 1.f, 0.f);
        }
        if (t.code) thermal_rules<R>(t, t.code[L.cell_off + idx], d, gd);
    } else {
        const Ctx<ST> c{L, fsrc, a.flags, a.link_slot, a.links, a.inv_scale};
        const int gx = gathered_x<ST>(L, x);
        Pre<ST> p;
        preload<ST>(c, gx, y, p);
        LinkVals<R> lg;
        links_of<ST>(c, p, lg);
        state_general<ST>(c, x, y, solid, p, lg, d);
        if (t.mode == THERMAL_STEP && a.nf && (fl & CELL_LINKS) && L.forces) {
            LinkVals<R> ls_ = lg;
            if (gx != x) load_link(c.links[L.link_off + a.link_slot[L.cell_off + idx]], ls_);
            double* accp = a.acc + static_cast<int64_t>(blockIdx.x % kAccCopies) * a.total_slots * 2;
            node_forces<ST>(L, c, ForceSink{a.nf, accp}, ls_, d);
        }
        thermal_gather<ST>(L, gsrc, a.flags, x, y, solid, gd);
        if (t.code) thermal_rules<R>(t, t.code[L.cell_off + idx], d, gd);
    }
    if (t.mode == THERMAL_OUT_G) {
        double* o = t.out + (static_cast<int64_t>(y - t.y0) * L.nx + x) * 9;
#pragma unroll
        for (int j = 0; j < 9; ++j) o[j] = WD[j] + static_cast<double>(gd[j]);
        return;
    }
    if (t.mode == THERMAL_OUT_T) {
        // compute_temperature (src/thermal.hpp:55-60) in slot order
        double T = 0.0;
#pragma unroll
        for (int j = 0; j < 9; ++j) T += WD[j] + static_cast<double>(gd[j]);
        t.out[static_cast<int64_t>(y - t.y0) * L.nx + x] = T;
        return;
    }
    if (!solid) {
        R wd = R(0);
        if (L.les) {
            const uint16_t w = a.wdist[L.cell_off + idx];
            wd = w == 0xffffu ? R(1e9) : static_cast<R>(w);
        }
        collide<R>(d, L, wd, R(1));
        if (!(d[0] == d[0])) atomicOr(a.nan_flags + a.lattice0, 1);
        thermal_collide<R>(t, d, gd);
    }
    // as the one-step kernels: solid nodes keep a stored value only where it is read
    // (written solids); the post-upload collision stores every node
    const bool nt = (a.options & LAUNCH_STREAM_STORES) != 0;
    if (!solid || t.from_state || (fl & CELL_WRITE)) {
        S* o = static_cast<S*>(a.dst) + L.f_off + idx;
#pragma unroll
        for (int j = 0; j < 9; ++j) store_val(o + j * L.plane, Form<ST>::enc(d[j], 1.f, 0.f), nt);
    }
    if (!solid || t.from_state) {
        S* o = static_cast<S*>(t.gdst) + L.f_off + idx;
#pragma unroll
        for (int j = 0; j < 9; ++j) store_val(o + j * L.plane, Form<ST>::enc(gd[j], 1.f, 0.f), nt);
    }
}

// Rows of AoS FP64 thermal state (or 1.0 in every slot, the reference's initial value,
// src/lbm_types.hpp:215) into the state planes, as deviations from w.
template <int ST>
__global__ __launch_bounds__(256) void thermal_load_kernel(ThermalLoadArgs a)
{
    using S = typename Form<ST>::store;
    using R = typename Form<ST>::real;
    const LatticeDev L = a.lat[a.lattice];
    const int x = blockIdx.x * 256 + threadIdx.x;
    const int y = a.y0 + blockIdx.y;
    if (x >= L.nx || y >= a.y1) return;
    const int64_t idx = static_cast<int64_t>(y) * L.pitch + x;
    const double* in = a.in ? a.in + (static_cast<int64_t>(y - a.y0) * L.nx + x) * 9 : nullptr;
    S* o = static_cast<S*>(a.state_dst) + L.f_off + idx;
#pragma unroll
    for (int j = 0; j < 9; ++j) o[j * L.plane] = Form<ST>::enc(static_cast<R>((in ? in[j] : 1.0) - WD[j]), 1.f, 0.f);
}

// ------------------------------------------------------------------ wall distance

// The wall distance of src/lbm_types.hpp:231-264 -- the fixpoint of d = min(d, d_nbr + 1)
// over the 4 neighbours, solids 0, non-periodic -- is the 4-neighbour path length to the
// nearest solid node; a shortest path to the L1-nearest solid never needs to pass
// another solid first, so it equals the L1 distance transform
// d(x, y) = min over y' of (|y - y'| + h(x, y')), h the distance to the nearest solid in
// row y'.  Two passes, stored as uint16 saturating at 0xffff (no solid reachable, or
// 65535 and more).

constexpr uint32_t kFarWall = 0xffffu;

__device__ __forceinline__ uint32_t umin32(uint32_t a, uint32_t b) { return a < b ? a : b; }

// h of every node of row blockIdx.x: each of the 256 threads scans a contiguous chunk
// for its last and first solid, a scan over the chunks in LDS gives the nearest solid
// left and right of each chunk, and a second pass over the chunk writes min(x - left,
// right - x).
__global__ __launch_bounds__(256) void wall_rows_kernel(const uint8_t* flags, uint16_t* wd, int nx, int pitch)
{
    __shared__ int last[256], first[256];
    const int y = static_cast<int>(blockIdx.x);
    const int t = static_cast<int>(threadIdx.x);
    const int chunk = (nx + 255) / 256;
    const int x0 = min(nx, t * chunk), x1 = min(nx, x0 + chunk);
    const uint8_t* f = flags + static_cast<int64_t>(y) * pitch;
    int lo = -1, hi = -1;
    for (int x = x0; x < x1; ++x)
        if (f[x] & CELL_SOLID) {
            if (lo < 0) lo = x;
            hi = x;
        }
    last[t] = hi;                  // -1: none
    first[t] = lo < 0 ? nx : lo;   // nx: none
    __syncthreads();
    if (t == 0) {
        // exclusive scans: nearest solid left of chunk k (max), right of chunk k (min)
        int run = -1;
        for (int k = 0; k < 256; ++k) {
            const int v = last[k];
            last[k] = run;
            run = max(run, v);
        }
        run = nx;
        for (int k = 255; k >= 0; --k) {
            const int v = first[k];
            first[k] = run;
            run = min(run, v);
        }
    }
    __syncthreads();
    int left = last[t], right = first[t];
    uint16_t* o = wd + static_cast<int64_t>(y) * pitch;
    for (int x = x0; x < x1; ++x) {
        if (f[x] & CELL_SOLID) left = x;
        o[x] = static_cast<uint16_t>(left < 0 ? kFarWall : umin32(kFarWall, static_cast<uint32_t>(x - left)));
    }
    for (int x = x1 - 1; x >= x0; --x) {
        if (f[x] & CELL_SOLID) right = x;
        if (right < nx) o[x] = static_cast<uint16_t>(umin32(o[x], umin32(kFarWall, static_cast<uint32_t>(right - x))));
    }
}

// d(x, y) = min over y' of h(x, y') + |y - y'|: a forward and a backward min-plus scan down
// column x (one thread per column, so the rows are read coalesced), in place.
__global__ __launch_bounds__(256) void wall_cols_kernel(uint16_t* wd, int nx, int ny, int pitch)
{
    const int x = static_cast<int>(blockIdx.x) * 256 + static_cast<int>(threadIdx.x);
    if (x >= nx) return;
    uint16_t* c = wd + x;
    uint32_t d = c[0];
    for (int y = 1; y < ny; ++y) {
        uint16_t* p = c + static_cast<int64_t>(y) * pitch;
        d = umin32(*p, umin32(kFarWall, d + 1));
        *p = static_cast<uint16_t>(d);
    }
    for (int y = ny - 2; y >= 0; --y) {
        uint16_t* p = c + static_cast<int64_t>(y) * pitch;
        d = umin32(*p, umin32(kFarWall, d + 1));
        *p = static_cast<uint16_t>(d);
    }
}

// Advances the per-lattice step counters after a batch of launches; an odd number of
// launches (a two-step launch swaps the buffers once) leaves the state in the other buffer.
// No work: a split two-step launch runs it on the handle's stream between the fork and the
// bulk tiles' kernel, so that the kernels of the other streams, launched at the fork, reach
// the dispatcher first.
__global__ void hold_kernel() {}

__global__ void add_steps_kernel(int64_t* steps, int32_t* par, int32_t begin, int32_t count, int64_t inc,
                                 int32_t flips)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) {
        steps[begin + i] += inc;
        par[begin + i] ^= flips & 1;
    }
}

// Zeroes n doubles (a fill kernel rather than hipMemsetAsync: callers may capture it).
__global__ __launch_bounds__(256) void fill_zero_kernel(double* p, int64_t n)
{
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * 256)
        p[i] = 0.0;
}

// Moves the force totals of a batch of n launches to the history ring.  Block (i, s)
// sums, for launch i and force slot s, the terms its link records stored (records of
// the slot in record order, strided over the block, then a fixed LDS tree) and the
// kAccCopies fallback copies in copy order, clears those copies and writes history row
// (steps[lattice] + i) % R.  The order of every addition is fixed: the totals repeat bit
// for bit.  Runs before add_steps_kernel.
__global__ __launch_bounds__(256) void flush_kernel(FlushArgs f)
{
    __shared__ double sx[256], sy[256];
    const int pair = static_cast<int>(blockIdx.x);
    const int i = pair / f.slot_count;
    const int s = f.slot_begin + pair % f.slot_count;
    const int tid = static_cast<int>(threadIdx.x);
    const double* nfi = f.nf + static_cast<int64_t>(i) * f.n_recs * 2;
    double fx = 0.0, fy = 0.0;
    for (int r = f.slot_start[s] + tid; r < f.slot_start[s + 1]; r += 256) {
        const int64_t rec = f.slot_recs[r];
        fx += nfi[2 * rec];
        fy += nfi[2 * rec + 1];
    }
    sx[tid] = fx;
    sy[tid] = fy;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (tid < o) {
            sx[tid] += sx[tid + o];
            sy[tid] += sy[tid + o];
        }
        __syncthreads();
    }
    if (tid == 0) {
        double* blk = f.acc + static_cast<int64_t>(i) * kAccCopies * f.total_slots * 2;
        double gx = sx[0], gy = sy[0];
        for (int c = 0; c < kAccCopies; ++c) {
            double* q = blk + (static_cast<int64_t>(c) * f.total_slots + s) * 2;
            gx += q[0];
            gy += q[1];
            q[0] = 0.0;
            q[1] = 0.0;
        }
        const int64_t row = (f.steps[f.slot_lattice[s]] + i) % f.R;
        f.hist[(static_cast<int64_t>(s) * f.R + row) * 2] = gx;
        f.hist[(static_cast<int64_t>(s) * f.R + row) * 2 + 1] = gy;
    }
}

// A group of nodes moves at most 16 bytes per plane and thread (one dwordx4).
template <int ST, int V> constexpr bool group_fits() { return sizeof(typename Form<ST>::store) * V <= 16; }

template <int ST, int BX, int ROWS, int V>
hipError_t launch_step_t(bool single, int n_blocks, const StepArgs& a, hipStream_t s)
{
    if constexpr (!group_fits<ST, V>()) {
        return hipErrorInvalidValue;
    } else {
        if (single) hipLaunchKernelGGL((step_kernel<ST, BX, ROWS, V, true>), dim3(n_blocks), dim3(BX), 0, s, a);
        else hipLaunchKernelGGL((step_kernel<ST, BX, ROWS, V, false>), dim3(n_blocks), dim3(BX), 0, s, a);
        return hipGetLastError();
    }
}

// Launch shapes (threads per block, rows per block, nodes per thread) compiled in.
#define LBM_SHAPES(X)                                                                          \
    X(64, 1, 1) X(128, 1, 1) X(256, 1, 1) X(256, 2, 1) X(64, 1, 2) X(128, 1, 2) X(256, 1, 2) \
    X(64, 1, 4) X(128, 1, 4) X(256, 1, 4) X(64, 2, 4) X(128, 2, 4) X(64, 1, 8) X(128, 1, 8)

template <int ST>
hipError_t launch_step_st(int bx, int rows, int v, bool single, int n_blocks, const StepArgs& a, hipStream_t s)
{
#define LBM_SHAPE(B, RW, VV) \
    if (bx == B && rows == RW && v == VV) return launch_step_t<ST, B, RW, VV>(single, n_blocks, a, s);
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

hipError_t launch_step(int storage, int bx, int rows, int v, bool single, int n_blocks, const StepArgs& a,
                       hipStream_t s)
{
    switch (storage) {
    case ST_F64: return launch_step_st<ST_F64>(bx, rows, v, single, n_blocks, a, s);
    case ST_F32: return launch_step_st<ST_F32>(bx, rows, v, single, n_blocks, a, s);
    case ST_F16S: return launch_step_st<ST_F16S>(bx, rows, v, single, n_blocks, a, s);
    case ST_BF16S: return launch_step_st<ST_BF16S>(bx, rows, v, single, n_blocks, a, s);
    }
    return hipErrorInvalidValue;
}

int list_blocks(int block_x, int n_links, int n_edges)
{
    const int per_block = block_x / 64 * kListLanes;
    return (n_links + per_block - 1) / per_block + (n_edges + block_x - 1) / block_x;
}

template <int ST, bool SINGLE>
hipError_t launch_step2_st(int ty, int n_blocks, const StepArgs& a, hipStream_t s)
{
    if (ty == 4) {
        hipLaunchKernelGGL((step2_kernel<ST, 4, SINGLE>), dim3(n_blocks), dim3(256), 0, s, a);
    } else if (ty == 8) {
        hipLaunchKernelGGL((step2_kernel<ST, 8, SINGLE>), dim3(n_blocks), dim3(256), 0, s, a);
    } else if (ty == 16) {
        hipLaunchKernelGGL((step2_kernel<ST, 16, SINGLE>), dim3(n_blocks), dim3(256), 0, s, a);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <bool SINGLE>
hipError_t launch_step2_sg(int storage, int tile_rows, int n_blocks, const StepArgs& a, hipStream_t s)
{
    switch (storage) {
    case ST_F64: return launch_step2_st<ST_F64, SINGLE>(tile_rows, n_blocks, a, s);
    case ST_F32: return launch_step2_st<ST_F32, SINGLE>(tile_rows, n_blocks, a, s);
    case ST_F16S: return launch_step2_st<ST_F16S, SINGLE>(tile_rows, n_blocks, a, s);
    case ST_BF16S: return launch_step2_st<ST_BF16S, SINGLE>(tile_rows, n_blocks, a, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_step2(int storage, int tile_rows, bool single, int n_blocks, const StepArgs& a, hipStream_t s)
{
    return single ? launch_step2_sg<true>(storage, tile_rows, n_blocks, a, s)
                  : launch_step2_sg<false>(storage, tile_rows, n_blocks, a, s);
}

namespace {
template <int ST, int TY, bool ENS, bool EDGE> hipError_t launch_bulk_e(int n_blocks, const StepArgs& a, hipStream_t s)
{
    constexpr int nt = bulk_threads(sizeof(typename Form<ST>::store), TY);
    if (Form<ST>::shifted && a.sr)
        hipLaunchKernelGGL((step2_bulk_kernel<ST, TY, Form<ST>::shifted, ENS, EDGE>), dim3(n_blocks), dim3(nt), 0, s, a);
    else
        hipLaunchKernelGGL((step2_bulk_kernel<ST, TY, false, ENS, EDGE>), dim3(n_blocks), dim3(nt), 0, s, a);
    return hipGetLastError();
}

template <int ST, int TY, bool EDGE> hipError_t launch_bulk_t(bool single, int n_blocks, const StepArgs& a, hipStream_t s)
{
    return single ? launch_bulk_e<ST, TY, false, EDGE>(n_blocks, a, s) : launch_bulk_e<ST, TY, true, EDGE>(n_blocks, a, s);
}

template <int ST>
hipError_t launch_bulk_st(int ty, bool single, bool edge, int n_blocks, const StepArgs& a, hipStream_t s)
{
    switch (ty) {
    case 8: return edge ? launch_bulk_t<ST, 8, true>(single, n_blocks, a, s) : launch_bulk_t<ST, 8, false>(single, n_blocks, a, s);
    case 16:
        return edge ? launch_bulk_t<ST, 16, true>(single, n_blocks, a, s) : launch_bulk_t<ST, 16, false>(single, n_blocks, a, s);
    case 32:
        return edge ? launch_bulk_t<ST, 32, true>(single, n_blocks, a, s) : launch_bulk_t<ST, 32, false>(single, n_blocks, a, s);
    }
    return hipErrorInvalidValue;
}
}  // namespace

bool bulk_supported(int storage, int tile_rows)
{
    return storage != ST_F64 && (tile_rows == 8 || tile_rows == 16 || tile_rows == 32);
}

hipError_t launch_step2_bulk(int storage, int tile_rows, bool single, bool edge, int n_blocks, const StepArgs& a,
                             hipStream_t s)
{
    if (n_blocks <= 0) return hipSuccess;
    switch (storage) {
    case ST_F32: return launch_bulk_st<ST_F32>(tile_rows, single, edge, n_blocks, a, s);
    case ST_F16S: return launch_bulk_st<ST_F16S>(tile_rows, single, edge, n_blocks, a, s);
    case ST_BF16S: return launch_bulk_st<ST_BF16S>(tile_rows, single, edge, n_blocks, a, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_hold(hipStream_t s)
{
    hipLaunchKernelGGL(hold_kernel, dim3(1), dim3(64), 0, s);
    return hipGetLastError();
}

hipError_t launch_tile_class(const Tile* tiles, int n_tiles, const LatticeDev* lat, const uint8_t* flags, int ty,
                             uint8_t* out, hipStream_t s)
{
    if (n_tiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(tile_class_kernel, dim3(static_cast<unsigned>(n_tiles)), dim3(256), 0, s, tiles, lat, flags,
                       ty, out);
    return hipGetLastError();
}

hipError_t launch_flush(const FlushArgs& f, hipStream_t s)
{
    const int n = f.n * f.slot_count;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(flush_kernel, dim3(n), dim3(256), 0, s, f);
    return hipGetLastError();
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

hipError_t launch_wall_distance(const uint8_t* flags, uint16_t* wdist, int nx, int ny, int pitch, hipStream_t s)
{
    hipLaunchKernelGGL(wall_rows_kernel, dim3(ny), dim3(256), 0, s, flags, wdist, nx, pitch);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(wall_cols_kernel, dim3((nx + 255) / 256), dim3(256), 0, s, wdist, nx, ny, pitch);
    return hipGetLastError();
}

hipError_t launch_geom_cells(const GeomArgs& g, hipStream_t s)
{
    hipLaunchKernelGGL(geom_cells_kernel, dim3((g.nx + 255) / 256, g.ny), dim3(256), 0, s, g);
    return hipGetLastError();
}

hipError_t launch_geom_columns(uint8_t* flags, int nx, int ny, int pitch, int inlet_from, int outlet,
                               hipStream_t s)
{
    if (ny < 3) return hipSuccess;
    hipLaunchKernelGGL(geom_columns_kernel, dim3((ny - 2 + 255) / 256), dim3(256), 0, s, flags, nx, ny, pitch,
                       inlet_from, outlet);
    return hipGetLastError();
}

hipError_t launch_geom_write(uint8_t* flags, int nx, int ny, int pitch, int32_t* row_count, hipStream_t s)
{
    hipLaunchKernelGGL(geom_write_kernel, dim3((nx + 255) / 256, ny), dim3(256), 0, s, flags, nx, ny, pitch,
                       row_count);
    return hipGetLastError();
}

hipError_t launch_geom_solids(const uint8_t* flags, int nx, int ny, int pitch, const int32_t* row_off,
                              int32_t* out, hipStream_t s)
{
    hipLaunchKernelGGL(geom_solids_kernel, dim3(ny), dim3(256), 0, s, flags, nx, ny, pitch, row_off, out);
    return hipGetLastError();
}

hipError_t launch_geom_links(const GeomArgs& g, hipStream_t s)
{
    hipLaunchKernelGGL(geom_links_kernel, dim3(g.ny), dim3(256), 0, s, g);
    return hipGetLastError();
}

hipError_t launch_place(const PlaceArgs& a, hipStream_t s)
{
    if (a.x1 <= a.x0 || a.y1 <= a.y0) return hipSuccess;
    hipLaunchKernelGGL(place_kernel, dim3((a.x1 - a.x0 + 255) / 256, a.y1 - a.y0), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_perturb(int storage, const PerturbArgs& a, int ny, hipStream_t s)
{
    dim3 grid((ny + 255) / 256, a.x1 - a.x0);
    switch (storage) {
    case ST_F64: hipLaunchKernelGGL(perturb_kernel<ST_F64>, grid, dim3(256), 0, s, a); break;
    case ST_F32: hipLaunchKernelGGL(perturb_kernel<ST_F32>, grid, dim3(256), 0, s, a); break;
    case ST_F16S: hipLaunchKernelGGL(perturb_kernel<ST_F16S>, grid, dim3(256), 0, s, a); break;
    case ST_BF16S: hipLaunchKernelGGL(perturb_kernel<ST_BF16S>, grid, dim3(256), 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int TW> hipError_t launch_thermal_w(int storage, const StepArgs& a, const ThermalArgs& t, hipStream_t s)
{
    constexpr int rows = 256 / TW;
    dim3 grid(((a.L0.nx + TW - 1) / TW) * ((t.y1 - t.y0 + rows - 1) / rows));
    switch (storage) {
    case ST_F64: hipLaunchKernelGGL((thermal_kernel<ST_F64, TW>), grid, dim3(256), 0, s, a, t); break;
    case ST_F32: hipLaunchKernelGGL((thermal_kernel<ST_F32, TW>), grid, dim3(256), 0, s, a, t); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_thermal(int storage, const StepArgs& a, const ThermalArgs& t, hipStream_t s)
{
    const int nx = a.L0.nx;
    const int idle64 = (nx + 63) / 64 * 64 - nx, idle256 = (nx + 255) / 256 * 256 - nx;
    return idle64 < idle256 ? launch_thermal_w<64>(storage, a, t, s) : launch_thermal_w<256>(storage, a, t, s);
}

hipError_t launch_thermal_load(int storage, const ThermalLoadArgs& a, int nx, hipStream_t s)
{
    dim3 grid((nx + 255) / 256, a.y1 - a.y0);
    switch (storage) {
    case ST_F64: hipLaunchKernelGGL(thermal_load_kernel<ST_F64>, grid, dim3(256), 0, s, a); break;
    case ST_F32: hipLaunchKernelGGL(thermal_load_kernel<ST_F32>, grid, dim3(256), 0, s, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_fill_zero(double* p, int64_t n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const int64_t want = (n + 255) / 256;
    const int64_t blocks = want < 2048 ? want : 2048;
    hipLaunchKernelGGL(fill_zero_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

hipError_t launch_add_steps(int64_t* steps, int32_t* par, int32_t begin, int32_t count, int64_t inc,
                            int32_t flips, hipStream_t s)
{
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(add_steps_kernel, dim3((count + 255) / 256), dim3(256), 0, s, steps, par, begin, count,
                       inc, flips);
    return hipGetLastError();
}

}  // namespace lbm
