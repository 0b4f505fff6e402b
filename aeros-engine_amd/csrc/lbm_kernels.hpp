// HIP kernels of the MI355X (gfx950) lattice Boltzmann solver.
//
// One fused pull kernel per lattice step: every thread owns one cell, pulls
// the 19 post-collision populations its neighbours stored in the previous
// step (half-way bounce-back from its own opposite population where the
// source is solid, wrap / copy-in-place on the domain shell, equilibrium on
// the inlet, the convective rule on the outlet), collides them (BGK/TRT, Guo
// forcing, Smagorinsky with Van Driest damping from the previous step's
// velocity), then advances the D3Q7 temperature lattice with the fresh
// velocity and writes everything back.  Populations live in HBM as
// structure-of-arrays f[q][cell] so a wave's 64 lanes touch one contiguous
// 256-byte (FP32) row per direction.
//
// Arithmetic follows the reference's stepLBMCPU / thermalStep operation by
// operation (aeros/src/lbm.cpp:907-1709, aeros/src/thermal_lbm.cpp:199-355):
// built with -ffp-contract=off, correctly rounded division and square root,
// so the FP32 path reproduces the reference's float results bit for bit.
// The diagnostics of every step are reduced in a fixed order (per-thread,
// then an LDS tree, then the last block folds the block partials in index
// order), so results do not depend on scheduling.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include <type_traits>

#include "lbm_common.h"

namespace aeros {

// ------------------------------------------------------------ lattice tables
// D3Q19 in the reference's order (aeros/src/lbm.cpp:92-109).
__device__ __host__ constexpr int cx_of(int q) {
    return q == 1 || q == 7 || q == 9 || q == 11 || q == 13 ? 1 : (q == 2 || q == 8 || q == 10 || q == 12 || q == 14 ? -1 : 0);
}
__device__ __host__ constexpr int cy_of(int q) {
    return q == 3 || q == 7 || q == 10 || q == 15 || q == 17 ? 1 : (q == 4 || q == 8 || q == 9 || q == 16 || q == 18 ? -1 : 0);
}
__device__ __host__ constexpr int cz_of(int q) {
    return q == 5 || q == 11 || q == 14 || q == 15 || q == 18 ? 1 : (q == 6 || q == 12 || q == 13 || q == 16 || q == 17 ? -1 : 0);
}
__device__ __host__ constexpr int opp_of(int q) { return q == 0 ? 0 : (q & 1 ? q + 1 : q - 1); }

template <class R> __device__ __host__ inline R w_of(int q) {
    return q == 0 ? R(1) / R(3) : (q <= 6 ? R(1) / R(18) : R(1) / R(36));
}
template <class R> __device__ __host__ inline R w7_of(int q) { return q == 0 ? R(0.25) : R(0.125); }

// ------------------------------------------------------------ storage
struct bf16_t { uint16_t bits; };

// Every solver array lives in device global memory.  Accesses go through
// global-address-space references, so pointers that reach a kernel through
// memory (served mode's per-case tables, which are generic pointers) still
// use global, not flat, loads and stores.
#define AEROS_GLOBAL __attribute__((address_space(1)))
template <class T> using gref = AEROS_GLOBAL T&;
template <class T> using gcref = const AEROS_GLOBAL T&;

// Element i of an array through a 32-bit byte offset: with a uniform base the
// load or store uses the scalar-base + 32-bit vector-offset address form.
template <class T> __device__ inline gcref<T> at(const T* base, uint32_t i) {
    const AEROS_GLOBAL char* b = (const AEROS_GLOBAL char*)base;
    return *(const AEROS_GLOBAL T*)(b + i * (uint32_t)sizeof(T));
}
template <class T> __device__ inline gref<T> at(T* base, uint32_t i) {
    AEROS_GLOBAL char* b = (AEROS_GLOBAL char*)base;
    return *(AEROS_GLOBAL T*)(b + i * (uint32_t)sizeof(T));
}

// Population loads and stores.  Stores are non-temporal: a lattice written in
// one step is read back only in the next, after the rest of the lattice has
// streamed through the caches, so the stores should not displace lines that
// are still to be read (measured 4-6 % faster steps).
template <class T> __device__ inline void st_nt(AEROS_GLOBAL T& dst, T v) { __builtin_nontemporal_store(v, &dst); }

template <class S, class R> struct Pop;
// Each storage: M, the element as memory holds it; dec / enc between it and
// the arithmetic type; ld / st of element i of an array.
template <> struct Pop<float, float> {
    using M = float;
    __device__ static float dec(float s, float) { return s; }
    __device__ static float enc(float v, float) { return v; }
    __device__ static float ld(const float* p, uint32_t i, float) { return at(p, i); }
    __device__ static void st(float* p, uint32_t i, float v, float) { st_nt(at(p, i), v); }
};
// float deviations f - w with double arithmetic: the stored value's spacing
// follows the deviation, not w, so the small per-step changes of a slow
// decay are kept (a float f near w = 1/18 has a spacing of 3.7e-9)
template <> struct Pop<float, double> {
    using M = float;
    __device__ static double dec(float s, double w) { return w + (double)s; }
    __device__ static float enc(double v, double w) { return (float)(v - w); }
    __device__ static double ld(const float* p, uint32_t i, double w) { return w + (double)at(p, i); }
    __device__ static void st(float* p, uint32_t i, double v, double w) { st_nt(at(p, i), (float)(v - w)); }
};
template <> struct Pop<double, double> {
    __device__ static double ld(const double* p, uint32_t i, double) { return at(p, i); }
    __device__ static void st(double* p, uint32_t i, double v, double) { st_nt(at(p, i), v); }
};
// h + w for a half h, and the half nearest to v - w: each one mixed-precision
// multiply-add (h * 1 + w in FP32; v * 1 - w rounded once to half), where
// the compiler would otherwise spend a conversion and an add on each.
__device__ inline float half_plus(_Float16 h, float w) {
    float r;
    asm("v_fma_mix_f32 %0, %1, 1.0, %2 op_sel_hi:[1,0,0]" : "=v"(r) : "v"(h), "v"(w));
    return r;
}
// The same h + w as a conversion (exact) and one rounded add, for the
// free-slip kernels: with the opaque mixed-precision form in their
// reflection path the compiler failed to unroll the direction loop and the
// step of a 39x35x35 lattice took 36 us instead of 11.
__device__ inline float half_plus_plain(_Float16 h, float w) { return (float)h + w; }
__device__ inline _Float16 half_minus(float v, float w) {
    _Float16 r;
    asm("v_fma_mixlo_f16 %0, %1, 1.0, -%2" : "=v"(r) : "v"(v), "v"(w));
    return r;
}
// dec serves the byte-addressed step kernels, ld the element-indexed ones
// (the free-slip kernels and the set-up and read-back kernels).
template <> struct Pop<_Float16, float> {   // deviation from w_q, IEEE half, round to nearest even
    using M = _Float16;
    __device__ static float dec(_Float16 h, float w) { return half_plus(h, w); }
    __device__ static _Float16 enc(float v, float w) { return half_minus(v, w); }
    __device__ static float ld(const _Float16* p, uint32_t i, float w) { return half_plus_plain(at(p, i), w); }
    __device__ static void st(_Float16* p, uint32_t i, float v, float w) { st_nt(at(p, i), half_minus(v, w)); }
};
__device__ inline float bf16_to_f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ inline uint16_t f_to_bf16(float x) {   // round to nearest even (one conversion on gfx950)
    return __builtin_bit_cast(uint16_t, (__bf16)x);
}
template <> struct Pop<bf16_t, float> {     // deviation from w_q, bfloat16
    using M = uint16_t;
    __device__ static float dec(uint16_t b, float w) { return w + bf16_to_f(b); }
    __device__ static uint16_t enc(float v, float w) { return f_to_bf16(v - w); }
    __device__ static float ld(const bf16_t* p, uint32_t i, float w) { return w + bf16_to_f(at(p, i).bits); }
    __device__ static void st(bf16_t* p, uint32_t i, float v, float w) { st_nt(at(p, i).bits, f_to_bf16(v - w)); }
};
// Equilibrium-shifted 16-bit storage: the stored value is f - e_q, where e_q
// is the inflow equilibrium (passed where the storages above take w_q), so
// the free stream stores values near zero.  The half form is scaled by 2^12
// (deviations down to 1.5e-8 stay normal; |f - e| <= 11 keeps it below the
// half range): stored = half(float(f - e) * 4096), read = h * 2^-12 + e, each
// rounded once as written.
struct h16e_t { _Float16 h; };
struct b16e_t { uint16_t bits; };
// h * 2^-12 is exact in FP32, so the one rounding is the add's (measured
// 1.7 % faster at 33.5M cells than the mixed-precision multiply-add, and 3x
// faster in the free-slip kernels, as for half_plus_plain)
__device__ inline float half_scaled_plus(_Float16 h, float e) { return (float)h * 0x1p-12f + e; }
__device__ inline _Float16 half_scaled(float d) {
    _Float16 r;
    asm("v_fma_mixlo_f16 %0, %1, %2, 0" : "=v"(r) : "v"(d), "v"(4096.0f));
    return r;
}
template <> struct Pop<h16e_t, float> {
    using M = _Float16;
    __device__ static float dec(_Float16 h, float e) { return half_scaled_plus(h, e); }
    __device__ static _Float16 enc(float v, float e) { return half_scaled(v - e); }
    __device__ static float ld(const h16e_t* p, uint32_t i, float e) { return half_scaled_plus(at(p, i).h, e); }
    __device__ static void st(h16e_t* p, uint32_t i, float v, float e) { st_nt(at(p, i).h, half_scaled(v - e)); }
};
template <> struct Pop<b16e_t, float> {
    using M = uint16_t;
    __device__ static float dec(uint16_t b, float e) { return e + bf16_to_f(b); }
    __device__ static uint16_t enc(float v, float e) { return f_to_bf16(v - e); }
    __device__ static float ld(const b16e_t* p, uint32_t i, float e) { return e + bf16_to_f(at(p, i).bits); }
    __device__ static void st(b16e_t* p, uint32_t i, float v, float e) { st_nt(at(p, i).bits, f_to_bf16(v - e)); }
};
template <class S> constexpr bool kEqShift = std::is_same<S, h16e_t>::value || std::is_same<S, b16e_t>::value;
// The value a stored population of direction q is taken relative to.
template <class S, class R> __device__ __forceinline__ R sshift(const StepParams<R>& P, int q) {
    if constexpr (kEqShift<S>) return P.st_shift[q];
    else return w_of<R>(q);
}

// A lattice seen by the step kernel: direction planes of n cells, element
// plane + idx with plane = q n (less the pull shift) >= 0.  Element types up
// to 4 bytes use one base and a 32-bit byte offset (the host guarantees
// 19 n sizeof(S) < 2^32), which keeps the 19 directions off the scalar
// register file; doubles use a per-direction base.
// The byte offset of element idx, hidden from the optimiser: the element
// address is then the plane's uniform byte offset plus this one per-cell value
// (one vector add per access), where it would otherwise form (plane + idx) *
// size with an add and a shift per access.  Fields are the same bits either
// way; measured at 33.5M cells (TRT + LES): FP32 1.185 -> 1.171 ms, and with
// 5 waves per SIMD asked for the 16-bit kernels (kStepWaves below; the
// uniform plane offsets cost them registers, and without the request they
// fall to 4 waves) FP16 0.804 -> 0.764 ms, BF16 0.811 -> 0.787 ms, FP16
// shifted by the inflow equilibrium 0.893 -> 0.793 ms.
__device__ __forceinline__ uint32_t cell_bytes(uint32_t idx, uint32_t size) {
    uint32_t b = idx * size;
    asm("" : "+v"(b));
    return b;
}
template <class S, bool BYTE = true> struct Lat {
    S* p;
    __device__ Lat(const void* base) : p(static_cast<S*>(const_cast<void*>(base))) {}
    template <class R> __device__ R ld(uint32_t idx, uint32_t plane, R w) const {
        if constexpr (sizeof(S) <= 4 && !BYTE) {
            return Pop<S, R>::ld(p, plane + idx, w);
        } else if constexpr (sizeof(S) <= 4) {
            using M = typename Pop<S, R>::M;
            const uint32_t off = plane * (uint32_t)sizeof(S) + cell_bytes(idx, (uint32_t)sizeof(S));
            return Pop<S, R>::dec(*(const AEROS_GLOBAL M*)((const AEROS_GLOBAL char*)p + off), w);
        } else {
            return Pop<S, R>::ld(p + plane, idx, w);
        }
    }
    template <class R> __device__ void st(uint32_t idx, uint32_t plane, R v, R w) const {
        if constexpr (sizeof(S) <= 4 && !BYTE) {
            Pop<S, R>::st(p, plane + idx, v, w);
        } else if constexpr (sizeof(S) <= 4) {
            using M = typename Pop<S, R>::M;
            const uint32_t off = plane * (uint32_t)sizeof(S) + cell_bytes(idx, (uint32_t)sizeof(S));
            st_nt(*(AEROS_GLOBAL M*)((AEROS_GLOBAL char*)p + off), Pop<S, R>::enc(v, w));
        } else {
            Pop<S, R>::st(p + plane, idx, v, w);
        }
    }
};

template <class R> __device__ inline bool fin(R x) { return __builtin_isfinite(x); }

// Faces whose shell cells lose the source of direction q (source outside the
// domain along that face's normal).
__device__ __host__ constexpr uint32_t out_faces(int q) {
    return (cx_of(q) > 0 ? M_X0 : 0u) | (cx_of(q) < 0 ? M_X1 : 0u) | (cy_of(q) > 0 ? M_Y0 : 0u) |
           (cy_of(q) < 0 ? M_Y1 : 0u) | (cz_of(q) > 0 ? M_Z0 : 0u) | (cz_of(q) < 0 ? M_Z1 : 0u);
}

// Index of the direction with velocity (x, y, z); -1 if there is none.
__device__ __host__ constexpr int dir_of(int x, int y, int z) {
    for (int q = 0; q < Q; ++q)
        if (cx_of(q) == x && cy_of(q) == y && cz_of(q) == z) return q;
    return -1;
}
// Direction q with its components along the axes in `axes` (bit a = axis a)
// negated: the direction after a specular reflection off those walls.
__device__ __host__ constexpr int mirror_of(int q, int axes) {
    return dir_of(axes & 1 ? -cx_of(q) : cx_of(q), axes & 2 ? -cy_of(q) : cy_of(q), axes & 4 ? -cz_of(q) : cz_of(q));
}

// Source of the pull of direction q into cell (x, y, z).  Beyond a free-slip
// face (a specular wall half-way outside it) the source is reflected: the
// population left the cell's mirror image across the wall, moving in the
// mirrored direction dq.  Beyond a periodic face it wraps.  Returns false
// when the source lies beyond any other face: the cell then keeps its own
// population (aeros/src/lbm.cpp:1392-1394).
__device__ __host__ inline bool pull_source(int x, int y, int z, int q, int nx, int ny, int nz, int periodic_x,
                                            int periodic_z, uint32_t slip, int& sx, int& sy, int& sz, int& dq) {
    sx = x - cx_of(q);
    sy = y - cy_of(q);
    sz = z - cz_of(q);
    int axes = 0;
    if ((sx < 0 && (slip & M_X0)) || (sx >= nx && (slip & M_X1))) { sx = x; axes |= 1; }
    if ((sy < 0 && (slip & M_Y0)) || (sy >= ny && (slip & M_Y1))) { sy = y; axes |= 2; }
    if ((sz < 0 && (slip & M_Z0)) || (sz >= nz && (slip & M_Z1))) { sz = z; axes |= 4; }
    if (periodic_x) { if (sx < 0) sx += nx; else if (sx >= nx) sx -= nx; }
    if (periodic_z) { if (sz < 0) sz += nz; else if (sz >= nz) sz -= nz; }
    dq = mirror_of(q, axes);
    return sx >= 0 && sx < nx && sy >= 0 && sy < ny && sz >= 0 && sz < nz;
}

// Cell index -> coordinates without an integer division.
template <class R> __device__ inline void cell_xyz(const StepParams<R>& P, uint32_t c, int& x, int& y, int& z) {
    const uint32_t t = (uint32_t)__builtin_fma((double)c, P.inv_nx, P.half_inv_nx);   // c / nx
    x = (int)(c - t * (uint32_t)P.nx);
    const uint32_t zz = (uint32_t)__builtin_fma((double)t, P.inv_ny, P.half_inv_ny); // t / ny
    y = (int)(t - zz * (uint32_t)P.ny);
    z = (int)zz;
}

// Division, reciprocal and square root of the step kernel.  The FP32 and
// FP64 paths round them correctly, as the reference does; the 16-bit
// storage paths use the hardware's approximate reciprocal and square root
// (1 ulp), far below the storage's own rounding.
template <class S, class R> __device__ inline R k_div(R a, R b) {
    if constexpr (sizeof(S) == 2) return a * __builtin_amdgcn_rcpf(b);
    else return a / b;
}
template <class S, class R> __device__ inline R k_sqrt(R x) {
    if constexpr (sizeof(S) == 2) return __builtin_amdgcn_sqrtf(x);
    else return __builtin_sqrt(x);
}

// ------------------------------------------------------------ equilibria
// equilibriumKernel (aeros/src/lbm.cpp:141-144).  cu is formed as the
// reference does, (cx*ux + cy*uy) + cz*uz with cx,cy,cz in {-1,0,1}; the
// products by 0 and 1 are exact, so the sum of the signed nonzero terms in
// the same order gives the same bits.
template <class R, int q> __device__ inline R cdotu(R ux, R uy, R uz) {
    constexpr int a = cx_of(q), b = cy_of(q), c = cz_of(q);
    R s = R(0);
    bool have = false;
    if (a != 0) { s = a > 0 ? ux : -ux; have = true; }
    if (b != 0) { const R t = b > 0 ? uy : -uy; s = have ? s + t : t; have = true; }
    if (c != 0) { const R t = c > 0 ? uz : -uz; s = have ? s + t : t; have = true; }
    return s;
}
template <class R, int q> __device__ inline R feq_k(R rho, R ux, R uy, R uz, R usqr) {
    const R cu = cdotu<R, q>(ux, uy, uz);
    return w_of<R>(q) * rho * (R(1) + cu * R(3) + cu * cu * R(4.5) - usqr * R(1.5));
}
// Same value as feq_k for a runtime direction (used only on rare fix-up paths).
template <class R> __device__ inline R feq_rt(int q, R rho, R ux, R uy, R uz, R usqr) {
    const int a = cx_of(q), b = cy_of(q), c = cz_of(q);
    R s = R(0);
    bool have = false;
    if (a != 0) { s = a > 0 ? ux : -ux; have = true; }
    if (b != 0) { const R t = b > 0 ? uy : -uy; s = have ? s + t : t; have = true; }
    if (c != 0) { const R t = c > 0 ? uz : -uz; s = have ? s + t : t; have = true; }
    return w_of<R>(q) * rho * (R(1) + s * R(3) + s * s * R(4.5) - usqr * R(1.5));
}

// computeEquilibriumFast (aeros/src/lbm.cpp:146-184), guarded form
template <class R> __device__ inline R feq_guarded(int q, R rho, R ux, R uy, R uz, R usqr) {
    if (!fin(rho)) rho = R(1);
    if (usqr > R(0.09)) {
        const R s = __builtin_sqrt(R(0.09) / usqr);
        ux *= s; uy *= s; uz *= s;
        usqr = ux * ux + uy * uy + uz * uz;
    }
    R cu = (R)cx_of(q) * ux + (R)cy_of(q) * uy + (R)cz_of(q) * uz;
    if (!fin(cu)) cu = R(0);
    R f = w_of<R>(q) * rho * (R(1) + cu * R(3) + cu * cu * R(4.5) - usqr * R(1.5));
    if (!fin(f)) f = w_of<R>(q) * rho;
    return f;
}

// Guo source term 3 w [(c - u_e) + 3 (c.u_e) c] . F (aeros/src/lbm.cpp:1213-1217)
template <class R, int q> __device__ inline R guo(R ex, R ey, R ez, R Fx, R Fy, R Fz) {
    const R cx = (R)cx_of(q), cy = (R)cy_of(q), cz = (R)cz_of(q);
    const R cu = cx * ex + cy * ey + cz * ez;
    return R(3) * w_of<R>(q) * ((cx - ex + R(3) * cu * cx) * Fx + (cy - ey + R(3) * cu * cy) * Fy +
                                (cz - ez + R(3) * cu * cz) * Fz);
}

// ------------------------------------------------------------ reductions
// The diagnostics of a step are reduced without any cross-block fence:
//  * every block writes one Partial (plain stores) to the step's half of a
//    double-buffered partials array, and adds its rho sum and NaN count into
//    the step's set of decision accumulators (int64 atomics spread over
//    kAccSlots words, so the result does not depend on their order);
//  * the next launch (each block, from the accumulators) decides whether the
//    step broke the loop (NaN storm / divergence), and its block 0 folds the
//    partials in block order into the device Status; after the last step of a
//    call a one-block kernel does the same.
// Kernel boundaries order these accesses, so no fence is needed.
struct Partial {
    int64_t rho_fx;          // sum of rho in 2^-32 fixed point (exact for float rho in [0.5, 1.5])
    double ke, tke;
    float maxv;
    int32_t nan, neg, bad, rho_out, rho_nf, ulim;
};

constexpr int kAccSlots = 16;              // atomic words per accumulated quantity
constexpr int kAccSet = 2 * kAccSlots;     // rho_fx[kAccSlots], nan[kAccSlots]
constexpr int kAccSets = 3;                // step k adds into set k % 3, reads set (k - 1) % 3, clears (k + 1) % 3

// Device-resident scalar state, updated as the reference updates its globals
// at the end of every step (aeros/src/lbm.cpp:1313-1319, 1554-1580).
struct Status {
    int32_t step;            // lbmCurrentStep
    int32_t abort;           // 0, 1 = NaN storm (step not counted), 2 = divergence (step counted)
    int32_t abort_step;      // reference step number at which the loop broke
    int32_t converged;
    float avg_rho, avg_ke, max_u_lb, max_u_world, tke, convergence, prev_ke;
    int32_t pad0;
    int64_t neg, bad, rho_out, rho_nf, ulim;
};

constexpr int BLOCK = 256;
constexpr int WAVES = BLOCK / 64;

// rho in 2^-32 fixed point.  A float in [0.5, 1.5] is a multiple of 2^-24,
// so the conversion is exact and any summation order gives the same integer.
__device__ inline int64_t rho_fixed(float r) {
    const uint32_t u = __float_as_uint(r);
    const uint64_t m = (u & 0x7fffffu) | 0x800000u;      // 24-bit significand
    const int e = (int)((u >> 23) & 0xffu) - 127;        // -1 or 0 for [0.5, 2)
    return (int64_t)(m << (9 + e));                      // m 2^(e-23) 2^32
}
__device__ inline int64_t rho_fixed(double r) { return (int64_t)__builtin_rint(r * 4294967296.0); }

// Whether step k broke the loop, from its decision accumulators: 1 = NaN
// storm, 2 = divergence of the fluid-mean density, else 0.
template <class R>
__device__ inline int step_decision(const StepParams<R>& P, int64_t rho_fx, int64_t nan, float* avg_rho) {
    const float a = (float)((double)rho_fx * 0x1p-32 / (double)P.fluid_cells);
    if (avg_rho) *avg_rho = a;
    if (nan > P.nan_limit) return 1;
    if (!__builtin_isfinite(a) || a > 1.2f || a < 0.8f) return 2;
    return 0;
}

// Totals of one accumulator set (uniform loads, summed in a fixed order).
__device__ inline void acc_totals(const int64_t* set, int64_t& rho_fx, int64_t& nan) {
    int64_t a = 0, b = 0;
#pragma unroll
    for (int i = 0; i < kAccSlots; ++i) { a += set[i]; b += set[kAccSlots + i]; }
    rho_fx = a;
    nan = b;
}

template <class T> __device__ inline T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
__device__ inline float wave_max(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}

template <class T> __device__ inline void lds_tree_sum(T* s, int tid) {
    for (int off = BLOCK / 2; off > 0; off >>= 1) {
        __syncthreads();
        if (tid < off) s[tid] += s[tid + off];
    }
    __syncthreads();
}

struct RedLds {
    Partial w[WAVES];
    Partial fold[BLOCK];
    int flag;
};

// Block sum of the per-thread partials: xor butterflies inside each wave
// (every lane ends with the same value), then the waves in index order.
// The counters are almost always zero and are reduced only when some lane
// of the wave has one.
__device__ inline Partial block_reduce(RedLds& L, int tid, const Partial& p) {
    Partial r;
    r.rho_fx = wave_sum(p.rho_fx);
    r.ke = wave_sum(p.ke);
    r.tke = wave_sum(p.tke);
    r.maxv = wave_max(p.maxv);
    const bool any = __ballot((p.nan | p.neg | p.bad | p.rho_out | p.rho_nf | p.ulim) != 0) != 0;
    if (any) {
        r.nan = wave_sum(p.nan); r.neg = wave_sum(p.neg); r.bad = wave_sum(p.bad);
        r.rho_out = wave_sum(p.rho_out); r.rho_nf = wave_sum(p.rho_nf); r.ulim = wave_sum(p.ulim);
    } else {
        r.nan = r.neg = r.bad = r.rho_out = r.rho_nf = r.ulim = 0;
    }
    const int wid = tid >> 6;
    if ((tid & 63) == 0) L.w[wid] = r;
    __syncthreads();
    Partial t = L.w[0];
#pragma unroll
    for (int k = 1; k < WAVES; ++k) {
        const Partial& q = L.w[k];
        t.rho_fx += q.rho_fx; t.ke += q.ke; t.tke += q.tke; t.maxv = fmaxf(t.maxv, q.maxv);
        t.nan += q.nan; t.neg += q.neg; t.bad += q.bad; t.rho_out += q.rho_out; t.rho_nf += q.rho_nf;
        t.ulim += q.ulim;
    }
    return t;
}

// Fold the nb partials of one step in index order (fixed assignment of
// partials to threads, then a fixed tree) and apply the reference's
// end-of-step logic to the Status (aeros/src/lbm.cpp:1313-1319, 1554-1580).
// `dec` is the step's decision from its accumulators, which every block of
// the following launch computes identically.
template <class R>
__device__ void fold_step(const StepParams<R>& P, const Partial* parts, int nb, Status* st, RedLds& L, int tid,
                          int dec, float avg_rho) {
    double ke = 0, tke = 0;
    float m = 0.0f;
    int64_t c[5] = {0, 0, 0, 0, 0};
    for (int k = tid; k < nb; k += BLOCK) {
        const Partial& q = parts[k];
        ke += q.ke; tke += q.tke; m = fmaxf(m, q.maxv);
        c[0] += q.neg; c[1] += q.bad; c[2] += q.rho_out; c[3] += q.rho_nf; c[4] += q.ulim;
    }
    // the int64 counters travel in the fixed-point slot pairs of the LDS partials
    Partial& mine = L.fold[tid];
    mine.ke = ke; mine.tke = tke; mine.maxv = m; mine.rho_fx = c[0];
    __syncthreads();
    for (int off = BLOCK / 2; off > 0; off >>= 1) {
        if (tid < off) {
            Partial& a = L.fold[tid];
            const Partial& b = L.fold[tid + off];
            a.ke += b.ke; a.tke += b.tke; a.maxv = fmaxf(a.maxv, b.maxv); a.rho_fx += b.rho_fx;
        }
        __syncthreads();
    }
    const double ske = L.fold[0].ke, stk = L.fold[0].tke;
    const float mv = L.fold[0].maxv;
    int64_t tot[5];
    tot[0] = L.fold[0].rho_fx;
    __syncthreads();
    for (int r = 1; r < 5; ++r) {
        L.fold[tid].rho_fx = c[r];
        __syncthreads();
        for (int off = BLOCK / 2; off > 0; off >>= 1) {
            if (tid < off) L.fold[tid].rho_fx += L.fold[tid + off].rho_fx;
            __syncthreads();
        }
        tot[r] = L.fold[0].rho_fx;
        __syncthreads();
    }
    if (tid == 0) {
        if (dec == 1) {                      // NaN storm: break before the stream
            st->abort = 1;
            st->abort_step = st->step + 1;
        } else {
            st->step += 1;
            const float avg_ke = (float)(ske / (double)P.fluid_cells);
            st->avg_rho = avg_rho;
            st->avg_ke = avg_ke;
            st->max_u_lb = mv;
            st->tke = (float)(stk / (double)P.fluid_cells);
            st->max_u_world = mv * P.vel_scale;
            st->neg = tot[0]; st->bad = tot[1]; st->rho_out = tot[2]; st->rho_nf = tot[3]; st->ulim = tot[4];
            st->convergence = fabsf(avg_ke - st->prev_ke);
            st->prev_ke = avg_ke;
            st->converged = st->convergence < P.conv_threshold;
            if (dec == 2) {
                st->abort = 2;
                st->abort_step = st->step;
            }
        }
    }
}

// ------------------------------------------------------------ the step kernel
struct StepBuffers {
    const void* f_in;
    void* f_out;
    const uint32_t* mask;
    const void* u_prev;        // Smagorinsky stencil input: previous step's velocity, AoS (x,y,z[,pad])
    void* u_next;
    void* fld[6];              // rho, ux, uy, uz, tke, strain (type R)
    const void* g_in;          // thermal populations (type R)
    void* g_out;
    const void* b_in;          // buoyancy from the previous thermal step
    void* b_out;
    void* t_out;               // temperature
    Partial* parts;            // 2 x nblk: the partials of steps of even / odd index
    int64_t* acc;              // kAccSets x kAccSet decision accumulators
    Status* status;
};

// Velocity stencil storage: 3 components, padded to 4 for half so that one
// 8-byte load fetches a neighbour.
template <class U> struct VelAos { static constexpr int stride = 3; };
template <> struct VelAos<_Float16> { static constexpr int stride = 4; };

// Flags of one launch: the state holds pre-collision populations (collide
// only), write the output fields, re-run of an aborted step (no abort check,
// no reduction), the step's index k since the lattice was (re)initialised
// (selects the partials half and the accumulator sets), and whether step
// k - 1 is still to be folded (every launch of a call but the first).
struct StepFlags {
    int first, write_fields, dry, k, pending;
};

// Population addresses as byte offsets (BYTE) except in the free-slip
// kernels: there the uniform plane offsets exhaust the scalar registers, the
// compiler no longer unrolls the reflection's direction loop, and the step
// of a 39x35x35 lattice took 37 us instead of 11 (FP32).
template <bool SLIP>
constexpr bool kByteAddr = !SLIP;

// One lattice step over blocks blk of nblk (the whole grid for a single
// lattice, the lattice's share of the launch in served mode).
template <class R, class S, class U, bool LES, bool THERM, bool FORCE, bool SLIP = false, bool BYTE = true>
__device__ __forceinline__ void step_grid(const StepParams<R>& P, const StepBuffers& B, const StepFlags F,
                                          const uint32_t blk, const uint32_t nblk, RedLds& L) {
    Status* st = B.status;
    if (!F.dry) {
        if (__hip_atomic_load(&st->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;   // an earlier step broke the loop
        if (F.pending) {
            // did step k - 1 break the loop?  Every block decides the same from
            // the accumulators; block 0 also folds that step's partials.
            int64_t rfx, nan;
            acc_totals(B.acc + (size_t)((F.k + kAccSets - 1) % kAccSets) * kAccSet, rfx, nan);
            float avg_rho;
            const int dec = step_decision(P, rfx, nan, &avg_rho);
            if (blk == 0)
                fold_step<R>(P, B.parts + (size_t)((F.k - 1) & 1) * nblk, (int)nblk, st, L, threadIdx.x, dec, avg_rho);
            if (dec) return;
        }
        if (blk == 0 && threadIdx.x < kAccSet) B.acc[(size_t)((F.k + 1) % kAccSets) * kAccSet + threadIdx.x] = 0;
    }

    const uint32_t n = (uint32_t)P.n;
    const Lat<S, BYTE> lin(B.f_in), lout(B.f_out);
    const int nx = P.nx, ny = P.ny, nz = P.nz;
    const uint32_t sy = (uint32_t)nx, sz = (uint32_t)nx * (uint32_t)ny;
    const int tid = threadIdx.x;
    constexpr int US = VelAos<U>::stride;

    Partial acc{0, 0.0, 0.0, 0.0f, 0, 0, 0, 0, 0, 0};
    R maxu2 = R(0);

    // The block owns one contiguous run of cells, `per` = a whole number of
    // 256-cell rows, and walks it a row at a time.  Each iteration issues the
    // next row's mask load first, then the Smagorinsky stencil and the
    // populations back to back, so a cell costs one memory round trip.
    const uint32_t per = ((n + nblk * BLOCK - 1) / (nblk * BLOCK)) * BLOCK;
    const uint32_t cbeg = blk * per;
    const uint32_t cend = cbeg + per < n ? cbeg + per : n;
    uint32_t cell = cbeg + tid;
    uint32_t m_next = cell < cend ? at(B.mask, cell) : 0u;
    for (; cell < cend; cell += BLOCK) {
        const uint32_t m = m_next;
        // The population planes' offsets q n: for the two BF16 storages and
        // the thermal kernels recomputed per cell from a copy of n the
        // optimiser cannot hoist, so that they do not hold 38 scalar registers
        // across the loop (which spilled into vector lanes).  At 33.5M cells
        // BF16 3.4 % and shifted BF16 5.6 % faster, the thermal lattice about
        // 1 %; FP16 and FP32 were slower so.
        uint32_t n_it = n;
        if constexpr (std::is_same<S, bf16_t>::value || std::is_same<S, b16e_t>::value || THERM)
            asm volatile("" : "+s"(n_it));
        if (cell + BLOCK < cend) m_next = at(B.mask, cell + BLOCK);
        if (m & M_SOLID) continue;   // solid/ground: fields are constant (rho 1, u 0), never stored
        const bool inlet = m & P.inlet_face, outlet = m & P.outlet_face;

        // ---------------- previous-step velocity around the cell for the
        // Smagorinsky strain (aeros/src/lbm.cpp:979-1036).  With 4- and 8-byte
        // storage these loads join the population loads, so a cell costs one
        // memory round trip; with 16-bit storage they follow the moments,
        // which holds fewer registers and keeps one more wave per SIMD, worth
        // more there than the round trip.
        constexpr bool EARLY_U = sizeof(S) >= 4;
        const bool strain = LES && !inlet && !outlet && (m & M_STRAIN);
        R xpx = 0, xpy = 0, xpz = 0, xmx = 0, xmy = 0, xmz = 0, ypx = 0, ypy = 0, ypz = 0;
        R ymx = 0, ymy = 0, ymz = 0, zpx = 0, zpy = 0, zpz = 0, zmx = 0, zmy = 0, zmz = 0;
        auto load_u = [&]() {
            const U* u0 = static_cast<const U*>(B.u_prev);
            auto ld3 = [&](uint32_t c, R& a, R& b, R& d) {
                const AEROS_GLOBAL U* p = &at(u0, c * (uint32_t)US);
                a = (R)p[0]; b = (R)p[1]; d = (R)p[2];
            };
            ld3(cell + 1, xpx, xpy, xpz); ld3(cell - 1, xmx, xmy, xmz);
            ld3(cell + sy, ypx, ypy, ypz); ld3(cell - sy, ymx, ymy, ymz);
            ld3(cell + sz, zpx, zpy, zpz); ld3(cell - sz, zmx, zmy, zmz);
        };
        // strain rate |S| = sqrt(2 S:S) by central differences
        auto strain_rate = [&]() -> R {
            const R h = R(0.5);
            const R dux_dx = (xpx - xmx) * h, duy_dy = (ypy - ymy) * h, duz_dz = (zpz - zmz) * h;
            const R dux_dy = (ypx - ymx) * h, dux_dz = (zpx - zmx) * h;
            const R duy_dx = (xpy - xmy) * h, duy_dz = (zpy - zmy) * h;
            const R duz_dx = (xpz - xmz) * h, duz_dy = (ypz - ymz) * h;
            const R sxy = h * (dux_dy + duy_dx), sxz = h * (dux_dz + duz_dx), syz = h * (duy_dz + duz_dy);
            const R s2 = dux_dx * dux_dx + duy_dy * duy_dy + duz_dz * duz_dz + R(2) * (sxy * sxy + sxz * sxz + syz * syz);
            return k_sqrt<S, R>(R(2) * s2);
        };
        if (EARLY_U && strain) load_u();
        // the shifted temperature populations and the buoyancy of an interior
        // cell join the population loads with every storage type (solid
        // sources are replaced after the collision); the thermal kernels keep
        // 4 waves per SIMD either way
        R gq[Q7], bq = R(0);
        const bool early_g = THERM && !(m & M_FACES);
        if (early_g) {
            const R* g = static_cast<const R*>(B.g_in);
#pragma unroll
            for (int i = 0; i < Q7; ++i) gq[i] = at(g + (size_t)i * n - (cx_of(i) + cy_of(i) * nx + cz_of(i) * (int)sz), cell);
            bq = at(static_cast<const R*>(B.b_in), cell);
        }

        // ---------------- pull S(x, q) from the previous post-collision state
        R f[Q];
        if (F.first) {
#pragma unroll
            for (int q = 0; q < Q; ++q) f[q] = lin.ld(cell, (uint32_t)q * n_it, sshift<S>(P, q));
        } else if (!(m & M_FACES)) {
            // interior: the source of q is cell - c_q; a uniform shifted base per direction
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const int off = cx_of(q) + cy_of(q) * nx + cz_of(q) * (int)sz;
                f[q] = lin.ld(cell, (uint32_t)q * n_it - (uint32_t)off, sshift<S>(P, q));
            }
            if (m & M_LINKS) {   // half-way bounce-back: the cell's own opposite population
#pragma unroll
                for (int q = 1; q < Q; ++q)
                    if ((m >> q) & 1u) f[q] = lin.ld(cell, (uint32_t)opp_of(q) * n_it, sshift<S>(P, opp_of(q)));
            }
        } else if (m & M_PULL_INLET) {
#pragma unroll
            for (int q = 0; q < Q; ++q) f[q] = P.feq_in[q];
        } else if (m & M_PULL_OUTLET) {
            // f_out + c (f_in - f_out) + 0.15 c (f_in - f_in2) on the planes N-1, N-2, N-3
            // (aeros/src/lbm.cpp:1465-1544)
            const uint32_t ci = (uint32_t)((int64_t)cell - P.out_d1), ci2 = (uint32_t)((int64_t)cell - P.out_d2);
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const R w = sshift<S>(P, q);
                const R fo = lin.ld(cell, (uint32_t)q * n_it, w);
                const R fi = lin.ld(ci, (uint32_t)q * n_it, w);
                const R fi2 = lin.ld(ci2, (uint32_t)q * n_it, w);
                f[q] = fo + P.conv_c * (fi - fo) + P.conv_k2 * (fi - fi2);
            }
        } else if (!(m & (SLIP ? (P.wrap_faces | P.slip_faces) : P.wrap_faces))) {
            // domain shell with no periodic wrap or free-slip wall: a source outside the domain
            // is replaced by the cell's own population (aeros/src/lbm.cpp:1392-1394);
            // in-domain sources pull (or bounce back) as in the interior
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const int off = cx_of(q) + cy_of(q) * nx + cz_of(q) * (int)sz;
                const uint32_t src = (m & out_faces(q)) ? cell : cell - (uint32_t)off;
                f[q] = lin.ld(src, (uint32_t)q * n_it, sshift<S>(P, q));
            }
            if (m & M_LINKS) {
#pragma unroll
                for (int q = 1; q < Q; ++q)
                    if ((m >> q) & 1u) f[q] = lin.ld(cell, (uint32_t)opp_of(q) * n_it, sshift<S>(P, opp_of(q)));
            }
        } else {
            // domain shell on a periodic or free-slip face: a source beyond a
            // periodic face wraps to the opposite face (aeros/src/lbm.cpp:1384-1391),
            // one beyond a free-slip face is read at the cell's mirror image in
            // the mirrored direction (specular reflection), one beyond any
            // other face is replaced by the cell's own population; in-domain
            // sources pull (or bounce back) as in the interior
            if constexpr (SLIP) {
                const uint32_t special = P.wrap_faces | P.slip_faces;
#pragma unroll
                for (int q = 0; q < Q; ++q) {
                    const int off = cx_of(q) + cy_of(q) * nx + cz_of(q) * (int)sz;
                    const uint32_t of = m & out_faces(q), wr = of & P.wrap_faces, rf = of & P.slip_faces;
                    const bool out = (of & ~special) != 0;
                    const bool rx = (rf & (M_X0 | M_X1)) != 0, ry = (rf & (M_Y0 | M_Y1)) != 0,
                               rz = (rf & (M_Z0 | M_Z1)) != 0;
                    const uint32_t src = cell - (uint32_t)off + ((wr & (M_X0 | M_X1)) ? (uint32_t)(cx_of(q) * nx) : 0u) +
                                         ((wr & (M_Z0 | M_Z1)) ? (uint32_t)(cz_of(q) * nz) * sz : 0u) +
                                         (rx ? (uint32_t)cx_of(q) : 0u) + (ry ? (uint32_t)(cy_of(q) * nx) : 0u) +
                                         (rz ? (uint32_t)cz_of(q) * sz : 0u);
                    // a D3Q19 velocity has at most two non-zero components, and
                    // only those can be reflected: two reflections give opp(q)
                    const int dq = ((int)rx + (int)ry + (int)rz >= 2)
                                       ? opp_of(q)
                                       : (rx ? mirror_of(q, 1) : (ry ? mirror_of(q, 2) : (rz ? mirror_of(q, 4) : q)));
                    f[q] = lin.ld(out ? cell : src, (uint32_t)(out ? q : dq) * n_it, sshift<S>(P, out ? q : dq));
                }
            } else {
#pragma unroll
                for (int q = 0; q < Q; ++q) {
                    const int off = cx_of(q) + cy_of(q) * nx + cz_of(q) * (int)sz;
                    const uint32_t xf = out_faces(q) & (M_X0 | M_X1), zf = out_faces(q) & (M_Z0 | M_Z1);
                    const bool out = (m & out_faces(q) & ~P.wrap_faces) != 0;
                    const uint32_t wrap = ((m & xf) ? (uint32_t)(cx_of(q) * nx) : 0u) +
                                          ((m & zf) ? (uint32_t)(cz_of(q) * nz) * sz : 0u);
                    const uint32_t src = out ? cell : cell - (uint32_t)off + wrap;
                    f[q] = lin.ld(src, (uint32_t)q * n_it, sshift<S>(P, q));
                }
            }
            if (m & M_LINKS) {   // the link bits refer to the wrapped or reflected source
#pragma unroll
                for (int q = 1; q < Q; ++q)
                    if ((m >> q) & 1u) f[q] = lin.ld(cell, (uint32_t)opp_of(q) * n_it, sshift<S>(P, opp_of(q)));
            }
        }

        R smag = R(0);
        if (EARLY_U && strain) smag = strain_rate();

        // ---------------- moments with sanitising (aeros/src/lbm.cpp:1061-1089)
        // Every population in [-1, 10] (which also excludes NaN/Inf) is the
        // common case; only otherwise are populations replaced and counted.
        R rho = R(0), ux = R(0), uy = R(0), uz = R(0);
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            // rho += f; u += f c with c in {-1,0,1}: the products by 0 only add
            // signed zeros, so they are left out
            const R fi = f[q];
            rho += fi;
            if (cx_of(q) > 0) ux += fi; else if (cx_of(q) < 0) ux -= fi;
            if (cy_of(q) > 0) uy += fi; else if (cy_of(q) < 0) uy -= fi;
            if (cz_of(q) > 0) uz += fi; else if (cz_of(q) < 0) uz -= fi;
        }
        bool clean;
        if constexpr (sizeof(R) == 4) {
            // NaN-propagating minimum/maximum (three-operand on gfx950): a NaN
            // fails both bounds, an infinity one of them
            R lo = f[0], hi = f[0];
#pragma unroll
            for (int q = 1; q < Q; ++q) {
                lo = __builtin_elementwise_minimum(lo, f[q]);
                hi = __builtin_elementwise_maximum(hi, f[q]);
            }
            clean = lo >= R(-1) && hi <= R(10);
        } else {
            clean = true;
#pragma unroll
            for (int q = 0; q < Q; ++q) clean = clean && (f[q] >= R(-1)) && (f[q] <= R(10));
        }
        int nanc = 0, badc = 0;
        if (!clean) {
            rho = ux = uy = uz = R(0);
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                R fi = f[q];
                if (!fin(fi)) { fi = w_of<R>(q); nanc++; badc++; }
                if (fi < R(-1) || fi > R(10)) { fi = w_of<R>(q); badc++; }
                rho += fi;
                if (cx_of(q) > 0) ux += fi; else if (cx_of(q) < 0) ux -= fi;
                if (cy_of(q) > 0) uy += fi; else if (cy_of(q) < 0) uy -= fi;
                if (cz_of(q) > 0) uz += fi; else if (cz_of(q) < 0) uz -= fi;
            }
        }
        if (!fin(rho)) { acc.rho_nf++; nanc++; rho = R(1); }
        else if (rho < R(0.5) || rho > R(1.5)) { acc.rho_out++; nanc++; rho = R(1); }
        {
            const R inv = k_div<S, R>(R(1), rho);
            ux *= inv; uy *= inv; uz *= inv;
        }
        if (!fin(ux)) { ux = uy = uz = R(0); nanc++; }
        {
            const R ut = ux * ux + uy * uy + uz * uz;
            if (ut > R(0.09)) { const R s = __builtin_sqrt(R(0.09) / ut); ux *= s; uy *= s; uz *= s; }
        }
        if (inlet && !P.periodic_x) { rho = R(1); ux = P.in_u[0]; uy = P.in_u[1]; uz = P.in_u[2]; }

        // ---------------- Smagorinsky eddy viscosity (aeros/src/lbm.cpp:1126-1145)
        R tau_eff = P.tau0;
        if (LES && !inlet && !outlet) {
            if (!EARLY_U && strain) {
                load_u();
                smag = strain_rate();
            }
            if (fin(smag) && smag < R(10) && smag > R(1e-6)) {
                R cs = P.cs;
                const int vd = (int)((m & M_VD_MASK) >> M_VD_SHIFT);
                if (vd <= 5) cs *= P.vd[vd];
                const R nut = cs * cs * smag;
                tau_eff = P.tau0 + k_div<S, R>(nut, R(1) / R(3));
                if (tau_eff > R(1.5)) tau_eff = R(1.5);
                if (tau_eff < R(0.51)) tau_eff = R(0.51);
            }
        }

        R usqr = ux * ux + uy * uy + uz * uz;
        if (usqr > R(0.09)) {
            acc.ulim++;
            const R s = __builtin_sqrt(R(0.09) / usqr);
            ux *= s; uy *= s; uz *= s;
            usqr = ux * ux + uy * uy + uz * uz;
        }
        const bool base_tau = tau_eff == P.tau0;
        const R inv_tau = base_tau ? P.inv_tau0 : k_div<S, R>(R(1), tau_eff);

        // ---------------- Guo forcing (aeros/src/lbm.cpp:1161-1176)
        R Fx = R(0), Fy = R(0), Fz = R(0);
        if (FORCE) { Fx = P.fx; Fy = P.fy; Fz = P.fz; }
        if (THERM) Fy += early_g ? bq : at(static_cast<const R*>(B.b_in), cell);
        const bool has_force = (FORCE || THERM) && (Fx != R(0) || Fy != R(0) || Fz != R(0));
        R ex = ux, ey = uy, ez = uz, es = usqr;
        if (has_force) {
            const R hf = k_div<S, R>(R(0.5), rho);
            ex += Fx * hf; ey += Fy * hf; ez += Fz * hf;
            es = ex * ex + ey * ey + ez * ez;
        }
        R s_minus = inv_tau;
        if (P.trt && base_tau) s_minus = P.s_minus0;
        else if (P.trt) {   // magic parameter 3/16 (aeros/src/lbm.cpp:1196-1202)
            const R mm = inv_tau - R(0.5);
            const R den = (__builtin_fabs(mm) > R(1e-6)) ? (k_div<S, R>(R(0.1875), mm) + R(0.5)) : R(1);
            s_minus = k_div<S, R>(R(1), den > R(0.05) ? den : R(0.05));
        }

        // ---------------- pairwise TRT/BGK relaxation (aeros/src/lbm.cpp:1204-1264)
        // A non-finite input population relaxes from its equilibrium; inputs
        // are finite on the clean path, so the replacement runs only otherwise.
        if (!clean) {
#pragma unroll
            for (int q = 0; q < Q; ++q)
                if (!fin(f[q])) f[q] = feq_rt<R>(q, rho, ex, ey, ez, es);
        }
        const R kp = R(1) - R(0.5) * inv_tau, km = R(1) - R(0.5) * s_minus;
        if constexpr (sizeof(S) == 2) {
            // 16-bit storage (no counterpart in the reference): the same
            // relaxation, folded and fused.  With d+ = (f_i + f_o)/2 - e+ and
            // d- = (f_i - f_o)/2 - e-, where e+ = w rho (1 - 1.5 u.u + 4.5 (c.u)^2)
            // and e- = 3 w rho c.u are the even and odd equilibrium parts,
            //   f_i' = f_i - s+ d+ - s- d-,   f_o' = f_o - s+ d+ + s- d-.
            // The FP32 and FP64 paths keep the reference's operation order.
            // (Working on the stored deviations directly saves the weight add
            // and subtract per population but raised register use and cost
            // one wave per SIMD: measured 15 % slower.)
            const R base = __builtin_fmaf(R(-1.5), es, R(1));
            const R wr0 = w_of<R>(0) * rho, wr1 = w_of<R>(1) * rho, wr2 = w_of<R>(7) * rho;
            f[0] = __builtin_fmaf(-inv_tau, f[0] - wr0 * base, f[0]);
            if (has_force) f[0] = __builtin_fmaf(kp, guo<R, 0>(ex, ey, ez, Fx, Fy, Fz), f[0]);
#define AEROS_PAIR_FUSED(QI)                                                                       \
            {                                                                                       \
                constexpr int qi = QI, qo = QI + 1;                                                 \
                const R cu = cdotu<R, qi>(ex, ey, ez);                                              \
                const R wr = qi <= 6 ? wr1 : wr2;                                                   \
                const R fi = f[qi], fo = f[qo];                                                     \
                const R sym = __builtin_fmaf(R(4.5) * cu, cu, base);                                \
                const R dp = __builtin_fmaf(R(0.5), fi + fo, -(wr * sym));                          \
                const R dm = __builtin_fmaf(R(0.5), fi - fo, -(wr * (R(3) * cu)));                  \
                R ni = __builtin_fmaf(-inv_tau, dp, __builtin_fmaf(-s_minus, dm, fi));              \
                R no = __builtin_fmaf(-inv_tau, dp, __builtin_fmaf(s_minus, dm, fo));               \
                if (has_force) {                                                                    \
                    const R si = guo<R, qi>(ex, ey, ez, Fx, Fy, Fz);                                \
                    const R so = guo<R, qo>(ex, ey, ez, Fx, Fy, Fz);                                \
                    const R sp = R(0.5) * (si + so), sm = R(0.5) * (si - so);                       \
                    ni = __builtin_fmaf(kp, sp, __builtin_fmaf(km, sm, ni));                        \
                    no = __builtin_fmaf(kp, sp, __builtin_fmaf(-km, sm, no));                       \
                }                                                                                   \
                f[qi] = ni; f[qo] = no;                                                             \
            }
            AEROS_PAIR_FUSED(1) AEROS_PAIR_FUSED(3) AEROS_PAIR_FUSED(5) AEROS_PAIR_FUSED(7) AEROS_PAIR_FUSED(9)
            AEROS_PAIR_FUSED(11) AEROS_PAIR_FUSED(13) AEROS_PAIR_FUSED(15) AEROS_PAIR_FUSED(17)
#undef AEROS_PAIR_FUSED
        } else {
        {
            const R feq0 = feq_k<R, 0>(rho, ex, ey, ez, es);
            const R fi = f[0];
            R fn = fi - (fi - feq0) * inv_tau;
            if (has_force) fn = fn + kp * guo<R, 0>(ex, ey, ez, Fx, Fy, Fz);
            f[0] = fn;
        }
#define AEROS_PAIR(QI)                                                                             \
        {                                                                                           \
            constexpr int qi = QI, qo = QI + 1;                                                     \
            const R feq = feq_k<R, qi>(rho, ex, ey, ez, es);                                        \
            const R feqo = feq_k<R, qo>(rho, ex, ey, ez, es);                                       \
            const R fi = f[qi], fo = f[qo];                                                         \
            const R fp = R(0.5) * (fi + fo), fm = R(0.5) * (fi - fo);                               \
            const R ep = R(0.5) * (feq + feqo), em = R(0.5) * (feq - feqo);                         \
            const R fpn = fp - inv_tau * (fp - ep);                                                 \
            const R fmn = fm - s_minus * (fm - em);                                                 \
            R ni, no;                                                                               \
            if (has_force) {                                                                        \
                const R si = guo<R, qi>(ex, ey, ez, Fx, Fy, Fz);                                    \
                const R so = guo<R, qo>(ex, ey, ez, Fx, Fy, Fz);                                    \
                const R sp = R(0.5) * (si + so), sm = R(0.5) * (si - so);                           \
                ni = fpn + fmn + kp * sp + km * sm;                                                 \
                no = fpn - fmn + kp * sp - km * sm;                                                 \
            } else {                                                                                \
                ni = fpn + fmn;                                                                     \
                no = fpn - fmn;                                                                     \
            }                                                                                       \
            f[qi] = ni; f[qo] = no;                                                                 \
        }
        AEROS_PAIR(1) AEROS_PAIR(3) AEROS_PAIR(5) AEROS_PAIR(7) AEROS_PAIR(9)
        AEROS_PAIR(11) AEROS_PAIR(13) AEROS_PAIR(15) AEROS_PAIR(17)
#undef AEROS_PAIR
        }
        // non-finite results fall back to equilibrium (with bounded
        // parameters a clean cell cannot produce one)
        if (!clean || !P.bounded) {
            bool fin_all = true;
#pragma unroll
            for (int q = 0; q < Q; ++q) fin_all = fin_all && fin(f[q]);
            if (!fin_all) {
#pragma unroll
                for (int q = 0; q < Q; ++q)
                    if (!fin(f[q])) f[q] = feq_rt<R>(q, rho, ex, ey, ez, es);
            }
        }
        {   // negative populations, counted per wave (ballot + popcount on the
            // scalar unit) and credited to the wave's first active lane
            const uint64_t live = __ballot(1);
            int wn = 0;
#pragma unroll
            for (int q = 0; q < Q; ++q) wn += __popcll(__ballot(f[q] < R(0)));
            if (__lane_id() == (unsigned)(__ffsll((unsigned long long)live) - 1)) acc.neg += wn;
        }

#pragma unroll
        for (int q = 0; q < Q; ++q) lout.st(cell, (uint32_t)q * n_it, f[q], sshift<S>(P, q));

        // ---------------- velocity and diagnostics (aeros/src/lbm.cpp:1266-1305)
        R vx = ux, vy = uy, vz = uz;   // the stored velocity includes F/(2 rho)
        if (has_force) {
            const R hi = k_div<S, R>(R(0.5), rho);
            vx = ux + Fx * hi; vy = uy + Fy * hi; vz = uz + Fz * hi;
        }
        if (LES) {
            AEROS_GLOBAL U* un = &at(static_cast<U*>(B.u_next), cell * (uint32_t)US);
            un[0] = (U)vx; un[1] = (U)vy; un[2] = (U)vz;
            if (US == 4) un[3] = (U)0;
        }
        const R tl = (P.cs * smag) * (P.cs * smag);
        R tk = tl * P.vel_scale2;
        if (!fin(tk) || tk < R(0)) tk = R(0);
        if (tk > R(1.0e6)) tk = R(1.0e6);
        if (F.write_fields) {
            at(static_cast<R*>(B.fld[0]), cell) = rho;
            at(static_cast<R*>(B.fld[1]), cell) = vx;
            at(static_cast<R*>(B.fld[2]), cell) = vy;
            at(static_cast<R*>(B.fld[3]), cell) = vz;
            at(static_cast<R*>(B.fld[4]), cell) = tk;
            at(static_cast<R*>(B.fld[5]), cell) = smag;
        }
        maxu2 = usqr > maxu2 ? usqr : maxu2;   // max |u|^2; the sqrt is taken once per thread
        acc.rho_fx += rho_fixed(rho);
        acc.ke += (double)usqr;
        acc.tke += (double)tk;
        acc.nan += nanc; acc.bad += badc;

        // ---------------- D3Q7 temperature: pull, then BGK with the new velocity
        // (aeros/src/thermal_lbm.cpp:205-336).  x and z always wrap for this
        // lattice; a solid source bounces back, anti-bounce-back with the wall
        // temperature on the rows y <= 0 and y >= Ny-1.
        if (THERM) {
            const R* g = static_cast<const R*>(B.g_in);
            R* go = static_cast<R*>(B.g_out);
            R gi[Q7];
            if (!(m & M_FACES)) {
#pragma unroll
                for (int i = 0; i < Q7; ++i) {
                    const int off = cx_of(i) + cy_of(i) * nx + cz_of(i) * (int)sz;
                    if ((m >> i) & 1u) {
                        const R go_ = at(g + (size_t)opp_of(i) * n, cell);
                        int cx_, cy_, cz_;
                        cell_xyz(P, cell, cx_, cy_, cz_);
                        const int wy = cy_ - cy_of(i);
                        gi[i] = (P.t_fixed && (wy <= 0 || wy >= ny - 1))
                                    ? (-go_ + R(2) * w7_of<R>(i) * (wy <= 0 ? P.t_hot : P.t_cold)) : go_;
                    } else {
                        gi[i] = early_g ? gq[i] : at(g + (size_t)i * n - off, cell);
                    }
                }
            } else {
                // x and z always wrap for this lattice; a source outside in y
                // bounces back (anti-bounce-back with the wall temperature on
                // the rows y = 0 and y = Ny-1), as does a solid source
                int x, y, z;
                cell_xyz(P, cell, x, y, z);
#pragma unroll
                for (int i = 0; i < Q7; ++i) {
                    const int off = cx_of(i) + cy_of(i) * nx + cz_of(i) * (int)sz;
                    const uint32_t xf = out_faces(i) & (M_X0 | M_X1), yf = out_faces(i) & (M_Y0 | M_Y1),
                                       zf = out_faces(i) & (M_Z0 | M_Z1);
                    bool bounce;
                    int wy;
                    uint32_t src = 0;
                    if (m & yf) {
                        bounce = true;
                        wy = cy_of(i) > 0 ? 0 : ny - 1;
                    } else {
                        src = cell - (uint32_t)off + ((m & xf) ? (uint32_t)(cx_of(i) * nx) : 0u) +
                              ((m & zf) ? (uint32_t)(cz_of(i) * nz) * sz : 0u);
                        bounce = (at(B.mask, src) & M_SOLID) != 0;
                        wy = y - cy_of(i);
                    }
                    if (bounce) {
                        const R go_ = at(g + (size_t)opp_of(i) * n, cell);
                        gi[i] = (P.t_fixed && (wy <= 0 || wy >= ny - 1))
                                    ? (-go_ + R(2) * w7_of<R>(i) * (wy <= 0 ? P.t_hot : P.t_cold)) : go_;
                    } else {
                        gi[i] = at(g + (size_t)i * n, src);
                    }
                }
                (void)x; (void)z;
            }
            R T = R(0);
#pragma unroll
            for (int i = 0; i < Q7; ++i) T += gi[i];
            if (!fin(T)) T = P.t_ref;
            T = (P.t_lo > T ? P.t_lo : T);
            T = (P.t_hi < T ? P.t_hi : T);
#pragma unroll
            for (int i = 0; i < Q7; ++i) {
                const R cu = (R)cx_of(i) * vx + (R)cy_of(i) * vy + (R)cz_of(i) * vz;
                const R e = w7_of<R>(i) * T * (R(1) + R(4) * cu);
                at(go + (size_t)i * n, cell) = gi[i] - (gi[i] - e) * P.t_inv_tau;
            }
            at(static_cast<R*>(B.t_out), cell) = T;
            at(static_cast<R*>(B.b_out), cell) = P.t_g * (T - P.t_ref);
        }
    }

    // ---------------- this step's partial and decision accumulators
    if (F.dry) return;
    // max sqrt(usqr) == sqrt(max usqr): correctly rounded sqrt and the cast
    // to float are monotone, so the block maximum is unchanged.
    acc.maxv = (float)__builtin_sqrt(maxu2);
    const Partial bp = block_reduce(L, tid, acc);
    if (tid == 0) {
        B.parts[(size_t)(F.k & 1) * nblk + blk] = bp;
        int64_t* set = B.acc + (size_t)(F.k % kAccSets) * kAccSet;
        atomicAdd((unsigned long long*)&set[blk % kAccSlots], (unsigned long long)bp.rho_fx);
        if (bp.nan) atomicAdd((unsigned long long*)&set[kAccSlots + blk % kAccSlots], (unsigned long long)(int64_t)bp.nan);
    }
}

// After the last launch of a call: fold that step (index k) as the next
// launch would.  One block.
template <class R>
__device__ __forceinline__ void fold_last(const StepParams<R>& P, const StepBuffers& B, int k, uint32_t nblk, RedLds& L) {
    Status* st = B.status;
    if (__hip_atomic_load(&st->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
    int64_t rfx, nan;
    acc_totals(B.acc + (size_t)(k % kAccSets) * kAccSet, rfx, nan);
    float avg_rho;
    const int dec = step_decision(P, rfx, nan, &avg_rho);
    fold_step<R>(P, B.parts + (size_t)(k & 1) * nblk, (int)nblk, st, L, threadIdx.x, dec, avg_rho);
}
template <class R>
__global__ __launch_bounds__(BLOCK) void k_fold(StepParams<R> P, StepBuffers B, int k, uint32_t nblk) {
    __shared__ RedLds L;
    fold_last<R>(P, B, k, nblk, L);
}

// SLIP: some domain face carries a free-slip wall (a separate instance, so
// that the reflection's addressing costs the other lattices no registers).
// Waves per SIMD asked of the compiler: at least 5 for the 16-bit storages
// without the thermal lattice (their natural allocation lands a few registers
// above 96, i.e. 4 waves; asked, the plain LES kernels fit 80-90 and run 6 or
// 5), except where the request makes the compiler spill (a body force; free
// slip without LES); no request otherwise.
template <class S, bool LES, bool THERM, bool FORCE, bool SLIP>
constexpr int kStepWaves = (sizeof(S) == 2 && !THERM && !FORCE && (!SLIP || LES)) ? 5 : 1;

template <class R, class S, class U, bool LES, bool THERM, bool FORCE, bool SLIP = false>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(kStepWaves<S, LES, THERM, FORCE, SLIP>)))
void k_step(StepParams<R> P, StepBuffers B) {
    __shared__ RedLds L;
    step_grid<R, S, U, LES, THERM, FORCE, SLIP, kByteAddr<SLIP>>(
        P, B, StepFlags{P.first, P.write_fields, P.dry, P.k, P.pending}, blockIdx.x, gridDim.x, L);
}

// ------------------------------------------------------------ served mode
// Many lattices advance in one launch: block b works on lattice map[b].ci as
// its block map[b].blk of cases[ci].nblocks.  A lattice's block count depends
// on its size only, so its arithmetic, reductions included, is the same alone
// and in any mix.
template <class R> struct BatchCase {
    StepParams<R> P;          // first / write_fields / dry are set per launch
    StepBuffers buf[2];       // by parity of the step's output (f_in of a first step is f_out)
    int32_t start;            // step index before the round
    int32_t pre;              // buffers hold pre-collision populations
    uint32_t nblocks;
    uint32_t pad;
};
struct BlockMap {
    uint32_t ci, blk;
};

template <class R, class S, class U, bool LES, bool THERM, bool FORCE, bool SLIP>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(kStepWaves<S, LES, THERM, FORCE, SLIP>)))
void k_step_batch(const BatchCase<R>* __restrict__ cases,
                                                      const BlockMap* __restrict__ map, int i, int nsteps) {
    __shared__ RedLds L;
    // block-uniform, read-only: scalar loads, so the case's parameters sit in
    // scalar registers as a kernel argument's would
    const BlockMap bm = map[blockIdx.x];
    const BatchCase<R>& C = cases[__builtin_amdgcn_readfirstlane(bm.ci)];
    const int k = C.start + i;
    const StepFlags F{(C.pre && i == 0) ? 1 : 0, (i == nsteps - 1) ? 1 : 0, 0, k, i > 0 ? 1 : 0};
    StepBuffers B = C.buf[k & 1];
    if (F.first) B.f_in = B.f_out;
    step_grid<R, S, U, LES, THERM, FORCE, SLIP, kByteAddr<SLIP>>(C.P, B, F, bm.blk, C.nblocks, L);
}
// The end of a served tick: one block per case folds its last step and
// copies the case's status to out[case], so that one copy brings every
// case's status to the host.
template <class R>
__global__ __launch_bounds__(BLOCK) void k_fold_batch(const BatchCase<R>* cases, int nsteps, Status* out) {
    __shared__ RedLds L;
    const BatchCase<R>& C = cases[blockIdx.x];
    const int k = C.start + nsteps - 1;
    fold_last<R>(C.P, C.buf[k & 1], k, C.nblocks, L);
    if (threadIdx.x == 0) out[blockIdx.x] = *C.buf[k & 1].status;   // thread 0 wrote the fold's result
}

// ------------------------------------------------------------ other kernels
// Fill both lattices: buffer A gets the initial equilibria, buffer B zeros
// (the reference's f / fNext after initLBM, aeros/src/lbm.cpp:581-582, 690-739).
template <class R, class S>
__global__ void k_init(S* fa, S* fb, const uint32_t* mask, const float* u3, int64_t n, R u0, R u1, R u2,
                       StepParams<R> P) {
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
        const bool sol = mask[c] & M_SOLID;
        R a = R(0), b = R(0), d = R(0);
        if (!sol) {
            if (u3) { a = (R)u3[c * 3]; b = (R)u3[c * 3 + 1]; d = (R)u3[c * 3 + 2]; }
            else { a = u0; b = u1; d = u2; }
        }
        R us = a * a + b * b + d * d;
        if (us > R(0.09)) {
            const R s = __builtin_sqrt(R(0.09) / us);
            a *= s; b *= s; d *= s; us = a * a + b * b + d * d;
        }
        for (int q = 0; q < Q; ++q) {
            Pop<S, R>::st(fa + (size_t)q * n, (uint32_t)c, feq_guarded<R>(q, R(1), a, b, d, us), sshift<S>(P, q));
            Pop<S, R>::st(fb + (size_t)q * n, (uint32_t)c, R(0), sshift<S>(P, q));
        }
    }
}

// Divergence reset of the fluid cells (aeros/src/lbm.cpp:1583-1600).
template <class R, class S, class U>
__global__ void k_reset(S* f, const uint32_t* mask, int64_t n, StepParams<R> P, U* uaos, R* rho, R* fux, R* fuy,
                        R* fuz) {
    constexpr int US = VelAos<U>::stride;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
        if (mask[c] & M_SOLID) continue;
        for (int q = 0; q < Q; ++q) Pop<S, R>::st(f + (size_t)q * n, (uint32_t)c, P.feq_in[q], sshift<S>(P, q));
        if (uaos) {
            U* u = uaos + (size_t)c * US;
            u[0] = (U)P.in_u[0]; u[1] = (U)P.in_u[1]; u[2] = (U)P.in_u[2];
        }
        rho[c] = R(1); fux[c] = P.in_u[0]; fuy[c] = P.in_u[1]; fuz[c] = P.in_u[2];
    }
}

// Per-cell link masks from the solid/ground/wall-distance bytes.
__global__ void k_build_mask(uint32_t* mask, const uint8_t* solid, const uint8_t* ground, const uint8_t* wall, int nx,
                             int ny, int nz, int axis, int sign, int periodic_x, int periodic_z, int conv_outlet,
                             uint32_t slip) {
    const int64_t n = (int64_t)nx * ny * nz;
    const int64_t sy = nx, sz = (int64_t)nx * ny;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
        const int x = (int)(c % nx), y = (int)((c / nx) % ny), z = (int)(c / sz);
        uint32_t m = 0;
        const bool sol = solid[c] || ground[c];
        if (sol) m |= M_SOLID;
        const int p[3] = {x, y, z}, N[3] = {nx, ny, nz};
        const bool inl = sign > 0 ? p[axis] == 0 : p[axis] == N[axis] - 1;
        const bool outl = sign > 0 ? p[axis] == N[axis] - 1 : p[axis] == 0;
        if (x == 0) m |= M_X0;
        if (x == nx - 1) m |= M_X1;
        if (y == 0) m |= M_Y0;
        if (y == ny - 1) m |= M_Y1;
        if (z == 0) m |= M_Z0;
        if (z == nz - 1) m |= M_Z1;
        const bool shell = (m & M_FACES) != 0;
        if (!sol) {
            if (inl && !periodic_x) m |= M_PULL_INLET;
            else if (outl && conv_outlet && !periodic_x) m |= M_PULL_OUTLET;
            for (int q = 0; q < Q; ++q) {
                int sx, syy, szz, dq;
                if (!pull_source(x, y, z, q, nx, ny, nz, periodic_x, periodic_z, slip, sx, syy, szz, dq)) continue;
                const int64_t s = ((int64_t)szz * ny + syy) * nx + sx;
                if (solid[s] || ground[s]) m |= 1u << q;
            }
            // strain defined: interior, off the inlet/outlet planes, no solid
            // 6-neighbour (ground neighbours count as fluid here, as in
            // aeros/src/lbm.cpp:1016-1022)
            if (!inl && !outl && !shell && !solid[c - 1] && !solid[c + 1] && !solid[c - sy] && !solid[c + sy] &&
                !solid[c - sz] && !solid[c + sz])
                m |= M_STRAIN;
        }
        const int wd = wall[c] <= 5 ? wall[c] : 6;
        m |= (uint32_t)wd << M_VD_SHIFT;
        mask[c] = m;
    }
}

// Post-collision -> post-stream: materialise the populations the reference
// holds after the same number of steps (its f).  Fluid cells pull exactly as
// the step kernel does; solid cells show the untouched buffer of that parity.
template <class R, class S>
__global__ void k_pull_only(const S* fin_, const S* fsolid, R* out, const uint32_t* mask, StepParams<R> P) {
    const int64_t n = P.n;
    const int nx = P.nx, ny = P.ny, nz = P.nz;
    const int64_t sy = nx, sz = (int64_t)nx * ny;
    for (int64_t cell = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; cell < n; cell += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t m = mask[cell];
        if (m & M_SOLID) {
            for (int q = 0; q < Q; ++q) out[q * n + cell] = Pop<S, R>::ld(fsolid + (size_t)q * n, (uint32_t)cell, sshift<S>(P, q));
            continue;
        }
        if (m & M_PULL_INLET) {
            for (int q = 0; q < Q; ++q) out[q * n + cell] = P.feq_in[q];
            continue;
        }
        if (m & M_PULL_OUTLET) {
            const int64_t ci = cell - P.out_d1, ci2 = cell - P.out_d2;
            for (int q = 0; q < Q; ++q) {
                const R w = sshift<S>(P, q);
                const S* b = fin_ + (size_t)q * n;
                const R fo = Pop<S, R>::ld(b, (uint32_t)cell, w), fi = Pop<S, R>::ld(b, (uint32_t)ci, w);
                const R fi2 = Pop<S, R>::ld(b, (uint32_t)ci2, w);
                out[q * n + cell] = fo + P.conv_c * (fi - fo) + P.conv_k2 * (fi - fi2);
            }
            continue;
        }
        const int x = (int)(cell % nx), y = (int)((cell / nx) % ny), z = (int)(cell / sz);
        for (int q = 0; q < Q; ++q) {
            int sx, syy, szz, mq;
            const bool in = pull_source(x, y, z, q, nx, ny, nz, P.periodic_x, P.periodic_z, P.slip_faces, sx, syy, szz, mq);
            const bool bb = in && ((m >> q) & 1u);
            const int64_t src = (!in || bb) ? cell : ((int64_t)szz * ny + syy) * nx + sx;
            const int dq = bb ? opp_of(q) : (in ? mq : q);
            out[q * n + cell] = Pop<S, R>::ld(fin_ + (size_t)dq * n, (uint32_t)src, sshift<S>(P, dq));
        }
        (void)sy;
    }
}

// Expand storage to R (populations of a pre-collision state).
template <class R, class S>
__global__ void k_widen(const S* in, R* out, int64_t n, StepParams<R> P) {
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x)
        for (int q = 0; q < Q; ++q) out[q * n + c] = Pop<S, R>::ld(in + (size_t)q * n, (uint32_t)c, sshift<S>(P, q));
}

// thermalUpload / thermalResetField populations: g = g_eq(T, 0) = w T.
template <class R>
__global__ void k_thermal_eq(R* g, const R* T, int64_t n) {
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
        const R t = T[c];
        for (int i = 0; i < Q7; ++i) g[i * n + c] = w7_of<R>(i) * t * (R(1) + R(4) * R(0));
    }
}

// World velocity and pressure post-pass (aeros/src/lbm.cpp:1602-1660).
template <class R>
__global__ void k_post(const uint32_t* mask, const R* rho, const R* ux, const R* uy, const R* uz, R* wx, R* wy,
                       R* wz, R* p, int64_t n, R scale, R rho0, unsigned long long* bad) {
    unsigned long long nb = 0;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
        if (mask[c] & M_SOLID) { wx[c] = wy[c] = wz[c] = p[c] = R(0); continue; }
        R vx = ux[c];
        if (!fin(vx)) vx = R(0);
        wx[c] = vx * scale; wy[c] = uy[c] * scale; wz[c] = uz[c] * scale;
        R r = rho[c];
        if (!fin(r)) r = R(1);
        const R pl = (R(1) / R(3)) * (r - R(1));
        R pp = pl * scale * scale * rho0;
        if (!fin(pp)) { nb++; pp = R(0); }
        p[c] = pp;
    }
    if (nb) atomicAdd(bad, nb);
}

// Sampled pressure range and rho residual (aeros/src/lbm.cpp:1662-1705):
// cells 0, stride, 2 stride, ...  Each block reduces its grid-stride share
// of the samples into one partial; k_post_fold folds the partials in block
// order, so the result does not depend on scheduling.
struct PostPart {
    float mn, mx;
    double ss;
    int cnt, pad;
};
constexpr int kPostBlocks = 512;   // at most this many partials
template <class R>
__global__ void k_post_scalars(const R* p, const R* rho, float* prev, int64_t n, int64_t stride, int valid,
                               PostPart* parts) {
    __shared__ float smin[BLOCK], smax[BLOCK];
    __shared__ double ss[BLOCK];
    __shared__ int sn[BLOCK];
    const int tid = threadIdx.x;
    float mn = 1e30f, mx = -1e30f;
    double acc = 0.0;
    int cnt = 0;
    const int64_t ns = (n + stride - 1) / stride;
    for (int64_t k = (int64_t)blockIdx.x * BLOCK + tid; k < ns; k += (int64_t)gridDim.x * BLOCK) {
        const int64_t c = k * stride;
        const float pv = (float)p[c];
        if (__builtin_isfinite(pv)) { mn = fminf(mn, pv); mx = fmaxf(mx, pv); }
        const float r = (float)rho[c];
        if (valid && __builtin_isfinite(r)) { const float d = r - prev[k]; acc += (double)d * (double)d; cnt++; }
        prev[k] = __builtin_isfinite(r) ? r : 1.0f;
    }
    smin[tid] = mn; smax[tid] = mx; ss[tid] = acc; sn[tid] = cnt;
    for (int off = BLOCK / 2; off > 0; off >>= 1) {
        __syncthreads();
        if (tid < off) {
            smin[tid] = fminf(smin[tid], smin[tid + off]);
            smax[tid] = fmaxf(smax[tid], smax[tid + off]);
            ss[tid] += ss[tid + off];
            sn[tid] += sn[tid + off];
        }
    }
    __syncthreads();
    if (tid == 0) parts[blockIdx.x] = PostPart{smin[0], smax[0], ss[0], sn[0], 0};
}
__global__ void k_post_fold(const PostPart* parts, int nparts, int valid, float* out /* pmin, pmax, resid, nsampled */) {
    __shared__ float smin[BLOCK], smax[BLOCK];
    __shared__ double ss[BLOCK];
    __shared__ int sn[BLOCK];
    const int tid = threadIdx.x;
    float mn = 1e30f, mx = -1e30f;
    double acc = 0.0;
    int cnt = 0;
    for (int k = tid; k < nparts; k += BLOCK) {
        mn = fminf(mn, parts[k].mn); mx = fmaxf(mx, parts[k].mx);
        acc += parts[k].ss; cnt += parts[k].cnt;
    }
    smin[tid] = mn; smax[tid] = mx; ss[tid] = acc; sn[tid] = cnt;
    for (int off = BLOCK / 2; off > 0; off >>= 1) {
        __syncthreads();
        if (tid < off) {
            smin[tid] = fminf(smin[tid], smin[tid + off]);
            smax[tid] = fmaxf(smax[tid], smax[tid + off]);
            ss[tid] += ss[tid + off];
            sn[tid] += sn[tid + off];
        }
    }
    __syncthreads();
    if (tid == 0) {
        const float a = smin[0], b = smax[0];
        out[0] = (a <= b && __builtin_isfinite(a)) ? a : 0.0f;
        out[1] = (a <= b && __builtin_isfinite(b)) ? b : 0.0f;
        out[2] = (valid && sn[0] > 0) ? (float)__builtin_sqrt(ss[0] / (double)sn[0]) : 0.0f;
        out[3] = (float)sn[0];
    }
}

// Control-volume momentum balance (aeros/src/lbm.cpp:2183-2232): each face
// cell's float term is formed as the reference forms it and accumulated in
// double; per-block partials are folded in block order by the host.
template <class R>
__device__ __forceinline__ void cv_force_block(const uint8_t* solid, const R* rho, const R* ux, const R* uy,
                                               const R* uz, int nx, int ny, int nz, int x0, int x1, int y0, int y1,
                                               int z0, int z1, double* partial) {
    __shared__ double s[3][BLOCK];
    const int tid = threadIdx.x;
    const int64_t nyz = (int64_t)(z1 - z0 + 1) * (y1 - y0 + 1), nzx = (int64_t)(z1 - z0 + 1) * (x1 - x0 + 1),
                  nyx = (int64_t)(y1 - y0 + 1) * (x1 - x0 + 1);
    const int64_t total = 2 * (nyz + nzx + nyx);
    double fx = 0, fy = 0, fz = 0;
    for (int64_t k = (int64_t)blockIdx.x * BLOCK + tid; k < total; k += (int64_t)gridDim.x * BLOCK) {
        int x, y, z;
        R nxv = 0, nyv = 0, nzv = 0;
        int64_t r = k;
        if (r < 2 * nyz) {
            const int64_t pair = r / 2; const int side = (int)(r % 2);
            z = z0 + (int)(pair / (y1 - y0 + 1)); y = y0 + (int)(pair % (y1 - y0 + 1));
            x = side ? x1 : x0; nxv = side ? R(1) : R(-1);
        } else if ((r -= 2 * nyz) < 2 * nzx) {
            const int64_t pair = r / 2; const int side = (int)(r % 2);
            z = z0 + (int)(pair / (x1 - x0 + 1)); x = x0 + (int)(pair % (x1 - x0 + 1));
            y = side ? y1 : y0; nyv = side ? R(1) : R(-1);
        } else {
            r -= 2 * nzx;
            const int64_t pair = r / 2; const int side = (int)(r % 2);
            y = y0 + (int)(pair / (x1 - x0 + 1)); x = x0 + (int)(pair % (x1 - x0 + 1));
            z = side ? z1 : z0; nzv = side ? R(1) : R(-1);
        }
        const int64_t c = ((int64_t)z * ny + y) * nx + x;
        if (solid[c]) continue;
        const R rr = rho[c], vx = ux[c], vy = uy[c], vz = uz[c];
        const R pp = (R(1) / R(3)) * (rr - R(1));
        const R un = vx * nxv + vy * nyv + vz * nzv;
        fx += (double)(pp * nxv + rr * vx * un);
        fy += (double)(pp * nyv + rr * vy * un);
        fz += (double)(pp * nzv + rr * vz * un);
    }
    s[0][tid] = fx; s[1][tid] = fy; s[2][tid] = fz;
    for (int k = 0; k < 3; ++k) lds_tree_sum(s[k], tid);
    if (tid == 0) { partial[blockIdx.x * 3] = s[0][0]; partial[blockIdx.x * 3 + 1] = s[1][0]; partial[blockIdx.x * 3 + 2] = s[2][0]; }
}
template <class R>
__global__ void k_cv_force(const uint8_t* solid, const R* rho, const R* ux, const R* uy, const R* uz, int nx, int ny,
                           int nz, int x0, int x1, int y0, int y1, int z0, int z1, double* partial) {
    cv_force_block<R>(solid, rho, ux, uy, uz, nx, ny, nz, x0, x1, y0, y1, z0, z1, partial);
}

// The control-volume force of several lattices in one launch: blockIdx.y
// selects the lattice, and each lattice's 256 blocks partition its faces
// exactly as k_cv_force's grid of 256 does, so the partials (and their sum
// in block order on the host) are the same bits as a k_cv_force call.
constexpr int kCvBlocks = 256;   // blocks of a control-volume force launch (the partition of its faces)
template <class R> struct CvCase {
    const uint8_t* solid;
    const R *rho, *ux, *uy, *uz;
    int nx, ny, nz, x0, x1, y0, y1, z0, z1;
};
template <class R>
__global__ void k_cv_force_batch(const CvCase<R>* cases, double* partial) {
    const CvCase<R>& c = cases[blockIdx.y];
    cv_force_block<R>(c.solid, c.rho, c.ux, c.uy, c.uz, c.nx, c.ny, c.nz, c.x0, c.x1, c.y0, c.y1, c.z0, c.z1,
                      partial + (size_t)blockIdx.y * gridDim.x * 3);
}

// Exact checksum of the rho field: in the float modes every rho lies in
// [0.5, 1.5] and is an integer multiple of 2^-24, so the sum is accumulated
// exactly in 64-bit integers (order independent, equal to any exact or
// sequential double sum).  The u-sum and the FP64 mode use a fixed-order
// double tree.
template <class R>
__global__ void k_checksums(const R* rho, const R* ux, const R* uy, const R* uz, int64_t n, double* part,
                            unsigned long long* exact, int exact_ok) {
    __shared__ double s[2][BLOCK];
    const int tid = threadIdx.x;
    double a = 0, b = 0;
    unsigned long long e = 0;
    for (int64_t c = (int64_t)blockIdx.x * BLOCK + tid; c < n; c += (int64_t)gridDim.x * BLOCK) {
        a += (double)rho[c];
        b += (double)(ux[c] + uy[c] + uz[c]);
        if (exact_ok) e += (unsigned long long)((double)rho[c] * 16777216.0);
    }
    s[0][tid] = a; s[1][tid] = b;
    lds_tree_sum(s[0], tid);
    lds_tree_sum(s[1], tid);
    if (tid == 0) { part[blockIdx.x * 2] = s[0][0]; part[blockIdx.x * 2 + 1] = s[1][0]; }
    if (exact_ok && e) atomicAdd(exact, e);
}

// Projected area: number of (a, b) index pairs across the flow axis with at
// least one solid cell along it (aeros/src/lbm.cpp:2236-2256).
__global__ void k_proj_area(const uint8_t* solid, int nx, int ny, int nz, int axis, unsigned long long* count) {
    const int na = axis == 0 ? ny : nx, nb = axis == 2 ? ny : nz, nl = axis == 0 ? nx : (axis == 1 ? ny : nz);
    const int64_t np = (int64_t)na * nb;
    unsigned long long cnt = 0;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < np; k += (int64_t)gridDim.x * blockDim.x) {
        const int a = (int)(k % na), b = (int)(k / na);
        bool hit = false;
        for (int l = 0; l < nl && !hit; ++l) {
            int x, y, z;
            if (axis == 0) { x = l; y = a; z = b; }
            else if (axis == 1) { x = a; y = l; z = b; }
            else { x = a; y = b; z = l; }
            hit = solid[((int64_t)z * ny + y) * nx + x] != 0;
        }
        cnt += hit;
    }
    if (cnt) atomicAdd(count, cnt);
}

// lbmEquilibriumMoments (aeros/src/lbm.cpp:249-266) on the device.
__global__ void k_eq_moments(float rho, float ux, float uy, float uz, float* out, int* ok) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (!__builtin_isfinite(rho) || !__builtin_isfinite(ux) || !__builtin_isfinite(uy) || !__builtin_isfinite(uz)) {
        *ok = 0;
        return;
    }
    const float us = ux * ux + uy * uy + uz * uz;
    float m = 0, px = 0, py = 0, pz = 0;
    for (int q = 0; q < Q; ++q) {
        const float fe = feq_guarded<float>(q, rho, ux, uy, uz, us);
        if (!__builtin_isfinite(fe)) { *ok = 0; return; }
        m += fe; px += fe * (float)cx_of(q); py += fe * (float)cy_of(q); pz += fe * (float)cz_of(q);
    }
    out[0] = m; out[1] = px; out[2] = py; out[3] = pz;
    *ok = 1;
}

// Mean temperature over the fluid rows and the Nusselt sum at the first
// fluid row (aeros/src/thermal_lbm.cpp:67-80, 338-351); block partials.
template <class R>
__global__ void k_thermal_diag(const R* T, const uint32_t* mask, int nx, int ny, int nz, int ylo, int yhi, R hot,
                               double* part) {
    __shared__ double s[3][BLOCK];
    const int tid = threadIdx.x;
    const int64_t rows = (int64_t)(yhi - ylo + 1) * nx;
    const int64_t m = rows * nz;
    double st = 0, sq = 0, sn = 0;
    for (int64_t k = (int64_t)blockIdx.x * BLOCK + tid; k < m; k += (int64_t)gridDim.x * BLOCK) {
        const int z = (int)(k / rows), r = (int)(k % rows), y = ylo + r / nx, x = r % nx;
        const int64_t c = ((int64_t)z * ny + y) * nx + x;
        if (mask[c] & M_SOLID) continue;
        st += (double)T[c];
        if (y == ylo) { sq += (double)(hot - T[c]) / 0.5; sn += 1.0; }
    }
    s[0][tid] = st; s[1][tid] = sq; s[2][tid] = sn;
    for (int k = 0; k < 3; ++k) lds_tree_sum(s[k], tid);
    if (tid == 0) for (int k = 0; k < 3; ++k) part[blockIdx.x * 3 + k] = s[k][0];
}

__global__ void k_probe(int* d) { d[threadIdx.x] += 1; }

// ------------------------------------------------------------ post-processing on the device
// computeLBMVorticityAndQ (aeros/src/lbm.cpp:2005-2093): central differences
// of the world velocity over the interior cells, in float as the reference
// computes them; |omega| (capped at 1000), Q = (|Omega|^2 - |S|^2)/2 (capped
// at +-1e6) and sqrt(2 S:S), which overwrites the strain field as in the
// reference.  Solid and ground cells get 0 vorticity and Q; the shell keeps
// its values.
template <class R>
__global__ void k_vort_q(const uint32_t* mask, const R* uxw, const R* uyw, const R* uzw, R* vort, R* qc, R* strain,
                         int nx, int ny, int nz, float icx, float icy, float icz) {
    const int64_t n = (int64_t)nx * ny * nz, sy = nx, sz = (int64_t)nx * ny;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
        const int x = (int)(c % nx), y = (int)((c / nx) % ny), z = (int)(c / sz);
        if (x < 1 || x > nx - 2 || y < 1 || y > ny - 2 || z < 1 || z > nz - 2) continue;
        if (mask[c] & M_SOLID) { vort[c] = R(0); qc[c] = R(0); continue; }
        auto d = [&](const R* f, int64_t o, float ic) { return ((float)f[c + o] - (float)f[c - o]) * 0.5f * ic; };
        const float dux_dx = d(uxw, 1, icx), dux_dy = d(uxw, sy, icy), dux_dz = d(uxw, sz, icz);
        const float duy_dx = d(uyw, 1, icx), duy_dy = d(uyw, sy, icy), duy_dz = d(uyw, sz, icz);
        const float duz_dx = d(uzw, 1, icx), duz_dy = d(uzw, sy, icy), duz_dz = d(uzw, sz, icz);
        if (!__builtin_isfinite(dux_dx)) { vort[c] = R(0); qc[c] = R(0); continue; }
        const float wx = duz_dy - duy_dz, wy = dux_dz - duz_dx, wz = duy_dx - dux_dy;
        float vm = __builtin_sqrtf(wx * wx + wy * wy + wz * wz);
        if (!__builtin_isfinite(vm)) vm = 0.0f;
        if (vm > 1000.0f) vm = 1000.0f;
        vort[c] = (R)vm;
        const float sxy = 0.5f * (dux_dy + duy_dx), sxz = 0.5f * (dux_dz + duz_dx), syz = 0.5f * (duy_dz + duz_dy);
        const float oxy = 0.5f * (dux_dy - duy_dx), oxz = 0.5f * (dux_dz - duz_dx), oyz = 0.5f * (duy_dz - duz_dy);
        const float s2 = dux_dx * dux_dx + duy_dy * duy_dy + duz_dz * duz_dz + 2.0f * (sxy * sxy + sxz * sxz + syz * syz);
        const float o2 = 2.0f * (oxy * oxy + oxz * oxz + oyz * oyz);
        float qv = 0.5f * (o2 - s2);
        if (!__builtin_isfinite(qv)) qv = 0.0f;
        if (qv > 1e6f) qv = 1e6f;
        if (qv < -1e6f) qv = -1e6f;
        qc[c] = (R)qv;
        float sm = __builtin_sqrtf(2.0f * s2);
        if (!__builtin_isfinite(sm)) sm = 0.0f;
        strain[c] = (R)sm;
    }
}

// The lattice as the world-space samplers see it (lbmMin*, lbmCellSize*, lbmN*).
struct SampleGrid {
    float minx, miny, minz, maxx, maxy, maxz, csx, csy, csz;
    int nx, ny, nz;
};
__device__ inline float clamp01(float t) { return t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t); }

// sampleFieldTrilinear with solid cells skipped (aeros/src/lbm.cpp:1860-1914).
template <class R>
__device__ bool sample_field(const R* field, const uint32_t* mask, const SampleGrid& G, float px, float py, float pz,
                             float* out) {
    *out = 0.0f;
    if (!__builtin_isfinite(px) || !__builtin_isfinite(py) || !__builtin_isfinite(pz)) return false;
    if (G.csx < 1e-9f || G.csy < 1e-9f || G.csz < 1e-9f) return false;
    if (G.nx < 2 || G.ny < 2 || G.nz < 2) return false;
    const float fx = (px - G.minx) / G.csx, fy = (py - G.miny) / G.csy, fz = (pz - G.minz) / G.csz;
    const int ix = (int)__builtin_floorf(fx), iy = (int)__builtin_floorf(fy), iz = (int)__builtin_floorf(fz);
    if (ix < 0 || iy < 0 || iz < 0) return false;
    if (ix > G.nx - 2 || iy > G.ny - 2 || iz > G.nz - 2) return false;
    const float tx = clamp01(fx - (float)ix), ty = clamp01(fy - (float)iy), tz = clamp01(fz - (float)iz);
    const float wx[2] = {1.0f - tx, tx}, wy[2] = {1.0f - ty, ty}, wz[2] = {1.0f - tz, tz};
    float acc = 0.0f, wsum = 0.0f;
    for (int dz = 0; dz <= 1; ++dz)
        for (int dy = 0; dy <= 1; ++dy)
            for (int dx = 0; dx <= 1; ++dx) {
                const int64_t c = ((int64_t)(iz + dz) * G.ny + (iy + dy)) * G.nx + (ix + dx);
                if (mask[c] & M_SOLID) continue;
                const float v = (float)field[c];
                if (!__builtin_isfinite(v)) continue;
                const float w = wx[dx] * wy[dy] * wz[dz];
                acc += v * w;
                wsum += w;
            }
    if (wsum < 1e-6f) return false;
    acc /= wsum;
    if (!__builtin_isfinite(acc)) return false;
    *out = acc;
    return true;
}

// Scalar samplers: mode 0 = value or 0 (getLBMPressureWorld, getLBMVorticityWorld,
// getLBMQWorld, getLBMTKEWorld, getLBMStrainWorld), mode 1 = density clamped
// to [0.8, 1.2] or 1 (getLBMDensityWorld, aeros/src/lbm.cpp:1942-1999).
template <class R>
__global__ void k_sample(const R* field, const uint32_t* mask, SampleGrid G, const float* pos, int64_t np, int mode,
                         float* out, uint8_t* valid) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < np; i += (int64_t)gridDim.x * blockDim.x) {
        float v = 0.0f;
        const bool ok = sample_field(field, mask, G, pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], &v);
        if (mode == 1) v = ok ? (v < 0.8f ? 0.8f : (v > 1.2f ? 1.2f : v)) : 1.0f;
        else if (!ok) v = 0.0f;
        out[i] = v;
        if (valid) valid[i] = ok ? 1 : 0;
    }
}

// getLBMVelocityLB (aeros/src/lbm.cpp:1781-1843): trilinear over the fluid
// corners only, weights renormalised.
template <class R>
__device__ void velocity_lb(const R* ux, const R* uy, const R* uz, const uint32_t* mask, const SampleGrid& G, float px,
                            float py, float pz, float v[3]) {
    v[0] = v[1] = v[2] = 0.0f;
    // (the reference tests x only; a non-finite y or z floors to INT_MIN on
    // its x86 build and returns 0 the same way)
    if (!__builtin_isfinite(px) || !__builtin_isfinite(py) || !__builtin_isfinite(pz)) return;
    const float fx = (px - G.minx) / G.csx, fy = (py - G.miny) / G.csy, fz = (pz - G.minz) / G.csz;
    const int ix = (int)__builtin_floorf(fx), iy = (int)__builtin_floorf(fy), iz = (int)__builtin_floorf(fz);
    const float tx = clamp01(fx - (float)ix), ty = clamp01(fy - (float)iy), tz = clamp01(fz - (float)iz);
    if (ix < 0 || ix >= G.nx - 1 || iy < 0 || iy >= G.ny - 1 || iz < 0 || iz >= G.nz - 1) return;
    const float w[8] = {(1 - tx) * (1 - ty) * (1 - tz), tx * (1 - ty) * (1 - tz), (1 - tx) * ty * (1 - tz), tx * ty * (1 - tz),
                        (1 - tx) * (1 - ty) * tz,       tx * (1 - ty) * tz,       (1 - tx) * ty * tz,       tx * ty * tz};
    float sv[8][3], sw[8], tot = 0.0f;
    int cnt = 0;
    for (int k = 0; k < 8; ++k) {   // corner order 000, 100, 010, 110, 001, 101, 011, 111
        const int x = ix + (k & 1), y = iy + ((k >> 1) & 1), z = iz + ((k >> 2) & 1);
        const int64_t c = ((int64_t)z * G.ny + y) * G.nx + x;
        if (mask[c] & M_SOLID) continue;
        const float a = (float)ux[c], b = (float)uy[c], d = (float)uz[c];
        if (!__builtin_isfinite(a)) continue;
        sv[cnt][0] = a; sv[cnt][1] = b; sv[cnt][2] = d;
        sw[cnt] = w[k];
        tot += w[k];
        cnt++;
    }
    if (cnt == 0) return;
    if (tot < 1e-6f) { v[0] = sv[0][0]; v[1] = sv[0][1]; v[2] = sv[0][2]; return; }
    float r[3] = {0.0f, 0.0f, 0.0f};
    for (int i = 0; i < cnt; ++i) {
        const float k = sw[i] / tot;
        r[0] += sv[i][0] * k; r[1] += sv[i][1] * k; r[2] += sv[i][2] * k;
    }
    if (__builtin_isfinite(r[0])) { v[0] = r[0]; v[1] = r[1]; v[2] = r[2]; }
}

// world = 0: getLBMVelocityLB; world = 1: getLBMVelocityWorld (aeros/src/lbm.cpp:
// 1845-1869): the free stream outside the lattice box, 0 inside a solid cell
// and where |u_LB|^2 < 1e-12, else u_LB flowSpeed/U0, capped at 2 flowSpeed.
template <class R>
__global__ void k_sample_velocity(const R* ux, const R* uy, const R* uz, const uint32_t* mask, SampleGrid G,
                                  const float* pos, int64_t np, int world, float scale, float vmax, float fsx, float fsy,
                                  float fsz, float* out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < np; i += (int64_t)gridDim.x * blockDim.x) {
        const float px = pos[3 * i], py = pos[3 * i + 1], pz = pos[3 * i + 2];
        float v[3];
        if (!world) {
            velocity_lb(ux, uy, uz, mask, G, px, py, pz, v);
        } else if (px < G.minx || px > G.maxx || py < G.miny || py > G.maxy || pz < G.minz || pz > G.maxz) {
            v[0] = fsx; v[1] = fsy; v[2] = fsz;
        } else {
            // isLBMSolidWorld (aeros/src/lbm.cpp:1954-1969): nearest cell by floor
            bool solid = false;
            if (__builtin_isfinite(px) && __builtin_isfinite(py) && __builtin_isfinite(pz) && G.csx >= 1e-9f &&
                G.csy >= 1e-9f && G.csz >= 1e-9f) {
                const int ix = (int)__builtin_floorf((px - G.minx) / G.csx), iy = (int)__builtin_floorf((py - G.miny) / G.csy),
                          iz = (int)__builtin_floorf((pz - G.minz) / G.csz);
                if (ix >= 0 && ix < G.nx && iy >= 0 && iy < G.ny && iz >= 0 && iz < G.nz)
                    solid = (mask[((int64_t)iz * G.ny + iy) * G.nx + ix] & M_SOLID) != 0;
            }
            v[0] = v[1] = v[2] = 0.0f;
            if (!solid) {
                float l[3];
                velocity_lb(ux, uy, uz, mask, G, px, py, pz, l);
                const float m2 = l[0] * l[0] + l[1] * l[1] + l[2] * l[2];
                if (!(m2 < 1e-12f)) {
                    float w0 = l[0] * scale, w1 = l[1] * scale, w2 = l[2] * scale;
                    if (!__builtin_isfinite(w0)) {
                        w0 = fsx; w1 = fsy; w2 = fsz;
                    } else {
                        const float mag = __builtin_sqrtf(w0 * w0 + w1 * w1 + w2 * w2);
                        if (mag > vmax) { const float k = vmax / mag; w0 *= k; w1 *= k; w2 *= k; }
                    }
                    v[0] = w0; v[1] = w1; v[2] = w2;
                }
            }
        }
        out[3 * i] = v[0]; out[3 * i + 1] = v[1]; out[3 * i + 2] = v[2];
    }
}

// ------------------------------------------------------------ voxelizer
// The three-ray majority vote of the set-mask producer (aeros/src/mesh_bvh.h:
// 62-185, aeros/src/voxel_grid.cpp:337-351) on the device: one thread per
// voxel casts its three jittered axis rays through the host-built BVH with
// the same float slab test and double-precision Moller-Trumbore test as the
// host voxelizer, so the mask is the same bit for bit.  Hits are counted
// per triangle, so the traversal order does not matter.
__device__ inline bool vox_box(const float o[3], const float inv[3], const VoxNode& nd) {
    float t0 = -__builtin_inff(), t1 = __builtin_inff();
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        float a = (nd.lo[k] - o[k]) * inv[k], b = (nd.hi[k] - o[k]) * inv[k];
        if (a > b) { const float t = a; a = b; b = t; }
        if (__builtin_isnan(a) || __builtin_isnan(b)) continue;   // ray in the slab plane
        t0 = t0 < a ? a : t0;
        t1 = b < t1 ? b : t1;
        if (t1 < t0) return false;
    }
    return t1 >= 0.0f;
}
__device__ inline int vox_tri(const float o[3], const float d[3], const VoxTri& t) {
    const double e1[3] = {(double)t.b[0] - t.a[0], (double)t.b[1] - t.a[1], (double)t.b[2] - t.a[2]};
    const double e2[3] = {(double)t.c[0] - t.a[0], (double)t.c[1] - t.a[1], (double)t.c[2] - t.a[2]};
    const double dd[3] = {d[0], d[1], d[2]};
    const double p[3] = {dd[1] * e2[2] - dd[2] * e2[1], dd[2] * e2[0] - dd[0] * e2[2], dd[0] * e2[1] - dd[1] * e2[0]};
    const double det = e1[0] * p[0] + e1[1] * p[1] + e1[2] * p[2];
    const double scale = __builtin_fabs(e1[0] * e2[1] - e1[1] * e2[0]) + __builtin_fabs(e1[1] * e2[2] - e1[2] * e2[1]) +
                         __builtin_fabs(e1[2] * e2[0] - e1[0] * e2[2]) + 1e-30;
    if (__builtin_fabs(det) < 1e-12 * scale) return 0;
    const double inv = 1.0 / det;
    const double sv[3] = {(double)o[0] - t.a[0], (double)o[1] - t.a[1], (double)o[2] - t.a[2]};
    const double u = (sv[0] * p[0] + sv[1] * p[1] + sv[2] * p[2]) * inv;
    if (u < 0.0 || u > 1.0) return 0;
    const double q[3] = {sv[1] * e1[2] - sv[2] * e1[1], sv[2] * e1[0] - sv[0] * e1[2], sv[0] * e1[1] - sv[1] * e1[0]};
    const double v = (dd[0] * q[0] + dd[1] * q[1] + dd[2] * q[2]) * inv;
    if (v < 0.0 || u + v > 1.0) return 0;
    const double tt = (e2[0] * q[0] + e2[1] * q[1] + e2[2] * q[2]) * inv;
    return (tt > 1e-7 && tt < 1e30) ? 1 : 0;
}
// Crossings of the ray o + t e_axis, t > 1e-7.
__device__ inline int vox_crossings(const VoxNode* nodes, const VoxTri* tris, const float o[3], int axis) {
    const float d[3] = {axis == 0 ? 1.0f : 0.0f, axis == 1 ? 1.0f : 0.0f, axis == 2 ? 1.0f : 0.0f};
    const float inv[3] = {axis == 0 ? 1.0f : __builtin_inff(), axis == 1 ? 1.0f : __builtin_inff(),
                          axis == 2 ? 1.0f : __builtin_inff()};
    if (!vox_box(o, inv, nodes[0])) return 0;
    constexpr int kStack = 64;   // the BVH is at most 48 levels deep, so at most 49 entries are pending
    int stack[kStack];
    int sp = 0, hits = 0;
    stack[sp++] = 0;
    while (sp) {
        const VoxNode& nd = nodes[stack[--sp]];
        if (nd.count) {
            for (int i = nd.first; i < nd.first + nd.count; ++i) hits += vox_tri(o, d, tris[i]);
        } else {
            if (sp < kStack && vox_box(o, inv, nodes[nd.left])) stack[sp++] = nd.left;
            if (sp < kStack && vox_box(o, inv, nodes[nd.right])) stack[sp++] = nd.right;
        }
    }
    return hits;
}
__global__ __launch_bounds__(256) void k_voxelize(const VoxNode* nodes, const VoxTri* tris, int nx, int ny, int64_t n,
                                                  float mnx, float mny, float mnz, float csx, float csy, float csz,
                                                  float jx, float jy, float jz, int32_t* out) {
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(c % nx);
        const int64_t r = c / nx;   // k ny + j
        const int j = (int)(r % ny), k = (int)(r / ny);
        const float pz = mnz + ((float)k + 0.5f) * csz, py = mny + ((float)j + 0.5f) * csy;
        const float px = mnx + ((float)i + 0.5f) * csx;
        const float ox[3] = {px, py + jy, pz + jz};
        const float oy[3] = {px + jz, py, pz + jx};
        const float oz[3] = {px + jy, py + jx, pz};
        int votes = 0;
        votes += vox_crossings(nodes, tris, ox, 0) & 1;
        votes += vox_crossings(nodes, tris, oy, 1) & 1;
        votes += vox_crossings(nodes, tris, oz, 2) & 1;
        out[c] = votes >= 2 ? 1 : 0;
    }
}

// A field of R values as float (the asynchronous frames publish floats).
template <class R> __global__ void k_to_f32(const R* in, float* out, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = (float)in[i];
}

// Fill a typed buffer with a value (graph-capturable, unlike hipMemset).
template <class T> __global__ void k_fill(T* p, int64_t n, T v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

// ------------------------------------------------------------ launchers
template <class R, class S, class U, bool LES, bool THERM, bool FORCE, bool SLIP = false>
static hipError_t launch_step_t(const StepParams<R>& P, const StepBuffers& B, int grid, hipStream_t s) {
    hipLaunchKernelGGL((k_step<R, S, U, LES, THERM, FORCE, SLIP>), dim3(grid), dim3(BLOCK), 0, s, P, B);
    return hipGetLastError();
}

template <class R, class S, class U>
hipError_t launch_step(const StepParams<R>& P, const StepBuffers& B, int grid, hipStream_t s) {
    const bool force = P.fx != R(0) || P.fy != R(0) || P.fz != R(0);
    if (P.slip_faces) {
        if (P.thermal) {
            if (P.les) return force ? launch_step_t<R, S, U, true, true, true, true>(P, B, grid, s)
                                    : launch_step_t<R, S, U, true, true, false, true>(P, B, grid, s);
            return force ? launch_step_t<R, S, U, false, true, true, true>(P, B, grid, s)
                         : launch_step_t<R, S, U, false, true, false, true>(P, B, grid, s);
        }
        if (P.les) return force ? launch_step_t<R, S, U, true, false, true, true>(P, B, grid, s)
                                : launch_step_t<R, S, U, true, false, false, true>(P, B, grid, s);
        return force ? launch_step_t<R, S, U, false, false, true, true>(P, B, grid, s)
                     : launch_step_t<R, S, U, false, false, false, true>(P, B, grid, s);
    }
    if (P.les) {
        if (P.thermal) return force ? launch_step_t<R, S, U, true, true, true>(P, B, grid, s)
                                    : launch_step_t<R, S, U, true, true, false>(P, B, grid, s);
        return force ? launch_step_t<R, S, U, true, false, true>(P, B, grid, s)
                     : launch_step_t<R, S, U, true, false, false>(P, B, grid, s);
    }
    if (P.thermal) return force ? launch_step_t<R, S, U, false, true, true>(P, B, grid, s)
                                : launch_step_t<R, S, U, false, true, false>(P, B, grid, s);
    return force ? launch_step_t<R, S, U, false, false, true>(P, B, grid, s)
                 : launch_step_t<R, S, U, false, false, false>(P, B, grid, s);
}

// Served mode: one launch over the cases of one (LES, thermal, free-slip)
// group; served cases carry no body force.
template <class R, class S, class U, bool LES, bool THERM, bool SLIP>
static hipError_t launch_batch_t(const BatchCase<R>* cases, const BlockMap* map, int nblocks, int i, int nsteps,
                                 hipStream_t s) {
    hipLaunchKernelGGL((k_step_batch<R, S, U, LES, THERM, false, SLIP>), dim3(nblocks), dim3(BLOCK), 0, s, cases, map,
                       i, nsteps);
    return hipGetLastError();
}
template <class R, class S, class U, bool SLIP>
static hipError_t launch_batch_s(bool les, bool thermal, const BatchCase<R>* cases, const BlockMap* map, int nblocks,
                                 int i, int nsteps, hipStream_t s) {
    if (les) return thermal ? launch_batch_t<R, S, U, true, true, SLIP>(cases, map, nblocks, i, nsteps, s)
                            : launch_batch_t<R, S, U, true, false, SLIP>(cases, map, nblocks, i, nsteps, s);
    return thermal ? launch_batch_t<R, S, U, false, true, SLIP>(cases, map, nblocks, i, nsteps, s)
                   : launch_batch_t<R, S, U, false, false, SLIP>(cases, map, nblocks, i, nsteps, s);
}
template <class R, class S, class U>
hipError_t launch_batch(bool les, bool thermal, bool slip, const BatchCase<R>* cases, const BlockMap* map, int nblocks,
                        int i, int nsteps, hipStream_t s) {
    return slip ? launch_batch_s<R, S, U, true>(les, thermal, cases, map, nblocks, i, nsteps, s)
                : launch_batch_s<R, S, U, false>(les, thermal, cases, map, nblocks, i, nsteps, s);
}

}  // namespace aeros
