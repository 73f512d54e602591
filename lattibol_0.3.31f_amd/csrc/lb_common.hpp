// lb_common.hpp -- device-side numerics shared by every LBM kernel.
//
// Everything here is compiled with -ffp-contract=off: each float/double
// operation rounds exactly as written, so results match a host evaluation of
// the same expressions bit for bit.  Expression order follows the reference
// kernels (src/collision.cu), cited per function.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "../../include/lattibol_amd.h"

#define LB_DEV __device__ __forceinline__

namespace lb {

// ---------------------------------------------------------------------------
// Velocity sets.  D2Q9 storage rows grow downward (streaming.cu:48-54) while
// the moments use physical y up (collision.cu:274): physical c_y = -row c_y.
// D3Q19 directions per collision.cu:463-478 / streaming.cu:116-161.
// ---------------------------------------------------------------------------
// Tables as constexpr functions so device code can evaluate them with
// compile-time indices (every loop over directions is unrolled by
// static_for below).
template <int Q> struct Vel;
template <> struct Vel<9> {
    static constexpr __host__ __device__ int cx(int i) {
        return (i == 1 || i == 5 || i == 8) ? 1 : ((i == 3 || i == 6 || i == 7) ? -1 : 0);
    }
    static constexpr __host__ __device__ int cy(int i) {   // storage rows
        return (i == 4 || i == 7 || i == 8) ? 1 : ((i == 2 || i == 5 || i == 6) ? -1 : 0);
    }
    static constexpr __host__ __device__ int cz(int) { return 0; }
    static constexpr __host__ __device__ int opp(int i) {
        return i == 0 ? 0 : (i <= 4 ? ((i + 1) % 4 + 1) : ((i - 5 + 2) % 4 + 5));
    }
    static constexpr __host__ __device__ int wc(int i) { return i == 0 ? 0 : (i < 5 ? 1 : 2); }
    static constexpr __host__ __device__ int pcy(int i) { return -cy(i); }   // physical y
};
template <> struct Vel<19> {
    static constexpr __host__ __device__ int cx(int i) {
        return (i == 1 || (i >= 7 && i <= 10)) ? 1 : ((i == 2 || (i >= 11 && i <= 14)) ? -1 : 0);
    }
    static constexpr __host__ __device__ int cy(int i) {
        return (i == 3 || i == 7 || i == 11 || i == 15 || i == 16) ? 1
             : ((i == 4 || i == 8 || i == 12 || i == 17 || i == 18) ? -1 : 0);
    }
    static constexpr __host__ __device__ int cz(int i) {
        return (i == 5 || i == 9 || i == 13 || i == 15 || i == 17) ? 1
             : ((i == 6 || i == 10 || i == 14 || i == 16 || i == 18) ? -1 : 0);
    }
    static constexpr __host__ __device__ int opp(int i) {
        return i == 0 ? 0 : (i <= 6 ? (i % 2 ? i + 1 : i - 1)
             : i == 7 ? 12 : i == 8 ? 11 : i == 9 ? 14 : i == 10 ? 13
             : i == 11 ? 8 : i == 12 ? 7 : i == 13 ? 10 : i == 14 ? 9 : 33 - i);
    }
    static constexpr __host__ __device__ int wc(int i) { return i == 0 ? 0 : (i < 7 ? 1 : 2); }
    static constexpr __host__ __device__ int pcy(int i) { return cy(i); }
};
static_assert(Vel<9>::opp(1) == 3 && Vel<9>::opp(2) == 4 && Vel<9>::opp(3) == 1 && Vel<9>::opp(4) == 2, "opp9");
static_assert(Vel<9>::opp(5) == 7 && Vel<9>::opp(6) == 8 && Vel<9>::opp(7) == 5 && Vel<9>::opp(8) == 6, "opp9");
static_assert(Vel<19>::opp(7) == 12 && Vel<19>::opp(8) == 11 && Vel<19>::opp(9) == 14 && Vel<19>::opp(10) == 13, "opp19");
static_assert(Vel<19>::opp(15) == 18 && Vel<19>::opp(16) == 17 && Vel<19>::opp(1) == 2 && Vel<19>::opp(6) == 5, "opp19");

// Position of a direction inside the five +z (5,9,13,15,17) or five -z
// (6,10,14,16,18) populations: the halo-plane slot.
constexpr __host__ __device__ int zslot(int i) {
    return (i == 5 || i == 6) ? 0 : (i == 9 || i == 10) ? 1 : (i == 13 || i == 14) ? 2
         : (i == 15 || i == 16) ? 3 : 4;
}

template <int B, int E, class F> LB_DEV void static_for(F &&f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

// Lattice weights as float members assigned from double quotients
// (simulation_d2q9.cu:56-58, simulation_d3q19.cu:54-56); exact double
// quotients for FP64 runs.
template <int Q, class T> LB_DEV T weight(int wclass);
template <> LB_DEV float weight<9, float>(int c) {
    return c == 0 ? (float)(16. / 36.) : (c == 1 ? (float)(4. / 36.) : (float)(1. / 36.));
}
template <> LB_DEV float weight<19, float>(int c) {
    return c == 0 ? (float)(12. / 36.) : (c == 1 ? (float)(2. / 36.) : (float)(1. / 36.));
}
template <> LB_DEV double weight<9, double>(int c) {
    return c == 0 ? 16. / 36. : (c == 1 ? 4. / 36. : 1. / 36.);
}
template <> LB_DEV double weight<19, double>(int c) {
    return c == 0 ? 12. / 36. : (c == 1 ? 2. / 36. : 1. / 36.);
}

// ---------------------------------------------------------------------------
// Storage policies.  Shifted 16-bit storage keeps f_i - w_i (float weight).
// ---------------------------------------------------------------------------
// Round to nearest even on the bit pattern (the oracle's formula), NaN kept
// quiet.  Device stores use the conversion instruction below; this form
// remains as its check (lb_check_bf16_rounding).
LB_DEV uint16_t f32_to_bf16_bits(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
// v_cvt_pk_bf16_f32 (gfx950): IEEE round to nearest even with FP32
// denormals preserved, one instruction instead of five; the same bits as
// f32_to_bf16_bits for every non-NaN input (checked over all 2^32 patterns).
// The empty asm pins f as a rounded FP32 value (no fused producer).
LB_DEV uint16_t f32_to_bf16(float f) {
    __asm__("" : "+v"(f));
    return __builtin_bit_cast(uint16_t, (__bf16)f);
}
LB_DEV float bf16_to_f32(uint16_t h) { return __builtin_bit_cast(float, (uint32_t)h << 16); }
// The empty asm pins f as a rounded FP32 register value.  Without it the
// compiler folds a producing multiply or FMA and this conversion into one
// v_fma_mixlo_f16 (a single rounding straight to FP16), which differs from
// FP32-then-FP16 rounding whenever the FP32 result sits on an FP16 tie.
LB_DEV uint16_t f32_to_f16(float f) {
    __asm__("" : "+v"(f));
    return __builtin_bit_cast(uint16_t, (_Float16)f);
}
LB_DEV float f16_to_f32(uint16_t h) { return (float)__builtin_bit_cast(_Float16, h); }

// Stochastic rounding of 16-bit stores.  Random bits per (step, node, slot):
// the SplitMix64 finaliser of step * golden + node gives the node's word;
// Marsaglia's xorshift32 (13, 17, 5) run from that word with its low bit set
// gives word m + 1 for slots 2m and 2m + 1, high 16 bits to the even slot,
// low 16 bits to the odd one (a counter-based stream: the same bits on any
// device, launch shape or thread count; shifts and exclusive ors only, ten
// words per node, no multiply).  The 16 bits sit in the top half of the word
// the rounding functions read.  The value rounds up in magnitude with probability
// (distance to the lower neighbour) / spacing, so the stored value is
// unbiased (exactly for BF16 and FP16 normals, whose dropped fractions have
// 16 and 13 bits; FP16 subnormals round up with that probability quantised
// to 2^-16) and increments below half a spacing survive.
// The node's word (64-bit arithmetic, once per node) and each slot's bits
// from it; sr_bits is the two together.
LB_DEV uint32_t sr_seed(uint32_t step, uint32_t id) {
    uint64_t z = 0x5EEDull + (uint64_t)step * 0x9E3779B97F4A7C15ull + (uint64_t)id;
    z ^= z >> 30; z *= 0xBF58476D1CE4E5B9ull; z ^= z >> 27; z *= 0x94D049BB133111EBull; z ^= z >> 31;
    return (uint32_t)z;
}
// (slot is a compile-time constant at every call: the chains of the 19
// slots share their common prefix.)
LB_DEV uint32_t sr_mix(uint32_t seed, int slot) {
    uint32_t x = seed | 1u;
    for (int m = 0; m <= (slot >> 1); m++) {
        x ^= x << 13; x ^= x >> 17; x ^= x << 5;
    }
    return (slot & 1) ? x << 16 : x & 0xffff0000u;
}
LB_DEV uint32_t sr_bits(uint32_t step, uint32_t id, int slot) { return sr_mix(sr_seed(step, id), slot); }
// sr_mix for a conversion that reads only the top 16 bits (BF16): the even
// slot's word is passed unmasked, the same result without the mask.
LB_DEV uint32_t sr_mix_hi(uint32_t seed, int slot) {
    uint32_t x = seed | 1u;
    for (int m = 0; m <= (slot >> 1); m++) {
        x ^= x << 13; x ^= x >> 17; x ^= x << 5;
    }
    return (slot & 1) ? x << 16 : x;
}
// BF16: add the top 16 bits of r below the kept bits and truncate;
// infinities and NaN (quieted) as in round to nearest even.  The formula
// on the bit pattern; the stores use the instruction below.
LB_DEV uint16_t f32_to_bf16_sr_bits(float f, uint32_t r) {
    const uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7f800000u) == 0x7f800000u) return f32_to_bf16_bits(f);
    return (uint16_t)((u + (r >> 16)) >> 16);
}
// v_cvt_sr_bf16_f32 (gfx950) rounds exactly so: one instruction per value
// (checked against f32_to_bf16_sr_bits over every FP32 pattern,
// lb_check_sr_rounding).
typedef __bf16 lb_bf16x2 __attribute__((ext_vector_type(2)));
typedef _Float16 lb_f16x2 __attribute__((ext_vector_type(2)));
LB_DEV uint16_t f32_to_bf16_sr(float f, uint32_t r) {
    const lb_bf16x2 v = __builtin_amdgcn_cvt_sr_bf16_f32(lb_bf16x2{}, f, r, false);
    return (uint16_t)(__builtin_bit_cast(uint32_t, v) & 0xffffu);
}
// FP16: the same with 13 dropped bits in the normal range (the truncated
// value is exact in FP16); below 2^-14 the spacing is 2^-24:
// q = floor(|v| 2^24 + u), u = r23 / 2^23 from 23 random bits.  With
// |v| = m 2^(e-150) (m the 24-bit significand, e >= 1 the biased exponent)
// |v| 2^24 = m 2^-k, k = 126 - e >= 14, and q = (m 2^(23-k) + r23) >> 23
// exactly: for k > 23 the fraction dropped from m 2^(23-k) cannot carry,
// since m 2^(23-k) + r23 crosses a multiple of 2^23 only through its
// integer part.  Integer arithmetic only.
// The normal-range result is formed for every lane; the subnormal and
// overflow results only when a lane of the wave needs them (a wave-uniform
// branch, no divergence), then one is selected per lane.  For k <= 23 the sum
// splits as (m >> k) + (((m mod 2^k) 2^(23-k) + r23) >> 23), whose second
// term stays below 2^24, so 32-bit words suffice; k is clamped to [14, 47]
// (at 47 and above q is 0 either way).
// HW = false: the normal-range result from the bit pattern (the form the
// instruction is checked against); HW = true: from v_cvt_sr_f16_f32
// (gfx950), which equals it on 2^-14 <= |v| < 65520 (checked over every
// FP32 pattern, lb_check_sr_rounding); the instruction's subnormal and
// overflow results differ (it truncates after adding the same 13 bits), so
// those lanes keep the integer forms.
template <bool HW = true> LB_DEV uint16_t f32_to_f16_sr(float f, uint32_t r) {
    const uint32_t x = __builtin_bit_cast(uint32_t, f);
    const uint32_t sign = (x >> 16) & 0x8000u, ax = x & 0x7fffffffu;
    uint16_t hn;
    if constexpr (HW) {
        const lb_f16x2 v = __builtin_amdgcn_cvt_sr_f16_f32(lb_f16x2{}, f, r, false);
        hn = (uint16_t)(__builtin_bit_cast(uint32_t, v) & 0xffffu);
    } else {
        const uint32_t t = (ax + (r >> 19)) & ~0x1fffu;
        hn = (uint16_t)(sign | f32_to_f16(__builtin_bit_cast(float, t)));
    }
    const bool rare = ax >= 0x477ff000u || ax < 0x38800000u;
    if (__builtin_amdgcn_ballot_w64(rare) == 0) return hn;
    const uint32_t e = ax >> 23, m = (ax & 0x7fffffu) | (e ? 0x800000u : 0u);
    const int k0 = 126 - (int)(e ? e : 1u);
    const int k = k0 < 14 ? 14 : (k0 > 47 ? 47 : k0);
    const uint32_t r23 = r >> 9;
    const uint32_t ql = (m >> (k & 31)) + ((((m & ((1u << (k & 31)) - 1u)) << ((23 - k) & 31)) + r23) >> 23);
    const uint32_t qh = ((m >> ((k - 23) & 31)) + r23) >> 23;
    const uint16_t hs = (uint16_t)(sign | (k <= 23 ? ql : qh));
    const uint16_t ho = f32_to_f16(f);
    return ax >= 0x477ff000u ? ho : (ax < 0x38800000u ? hs : hn);
}

// DEV = false: the population itself is stored and the arithmetic follows the
// reference expressions.  DEV = true: the deviation s_i = f_i - w_i is stored
// and the whole collision runs on deviations (moments, equilibria, boundary
// closures), which keeps FP32 rounding noise relative to the deviation, not
// to w_i (see the deviation-form functions below).
struct StF32 {
    typedef float raw; typedef float T;
    static constexpr bool DEV = false;
    static LB_DEV T ld(raw v, float) { return v; }
    static LB_DEV raw st(T f, float) { return f; }
};
struct StF64 {
    typedef double raw; typedef double T;
    static constexpr bool DEV = false;
    static LB_DEV T ld(raw v, float) { return v; }
    static LB_DEV raw st(T f, float) { return f; }
};
struct StF32S {
    typedef float raw; typedef float T;
    static constexpr bool DEV = true;
    static LB_DEV T ld(raw v, float) { return v; }
    static LB_DEV raw st(T s, float) { return s; }
};
struct StF16 {
    typedef uint16_t raw; typedef float T;
    static constexpr bool DEV = true;
    static LB_DEV T ld(raw v, float) { return f16_to_f32(v); }
    static LB_DEV raw st(T s, float) { return f32_to_f16(s); }
};
struct StBF16 {
    typedef uint16_t raw; typedef float T;
    static constexpr bool DEV = true;
    static LB_DEV T ld(raw v, float) { return bf16_to_f32(v); }
    static LB_DEV raw st(T s, float) { return f32_to_bf16(s); }
};

// ---------------------------------------------------------------------------
// exp: Cody-Waite reduction with fdlibm's split ln2, degree-13 Taylor
// polynomial on |r| <= ln2/2, exact 2^k scaling.  The same IEEE operation
// sequence as the host evaluation, hence bit-identical.
// ---------------------------------------------------------------------------
LB_DEV double exp_cw(double x) {
    const double INV_LN2 = 1.44269504088896338700e+00;
    const double LN2_HI = 6.93147180369123816490e-01;
    const double LN2_LO = 1.90821492927058770002e-10;
    if (x != x) return x;
    if (x > 709.0) return __builtin_huge_val();
    if (x < -708.0) return 0.0;
    double kd = __builtin_rint(x * INV_LN2);
    double r = (x - kd * LN2_HI) - kd * LN2_LO;
    double p = 1.0 / 6227020800.0;
    p = p * r + 1.0 / 479001600.0;
    p = p * r + 1.0 / 39916800.0;
    p = p * r + 1.0 / 3628800.0;
    p = p * r + 1.0 / 362880.0;
    p = p * r + 1.0 / 40320.0;
    p = p * r + 1.0 / 5040.0;
    p = p * r + 1.0 / 720.0;
    p = p * r + 1.0 / 120.0;
    p = p * r + 1.0 / 24.0;
    p = p * r + 1.0 / 6.0;
    p = p * r + 0.5;
    p = p * r + 1.0;
    p = p * r + 1.0;
    int k = (int)kd;
    uint64_t bits = (uint64_t)(int64_t)(k + 1023) << 52;
    return p * __builtin_bit_cast(double, bits);
}

// exp in single precision: Cody-Waite reduction x = k ln2 + r with ln2 split
// into a 15-bit head (k * head exact for |k| < 2^9) and a tail, a degree-7
// Taylor polynomial in Horner form with fused multiply-adds on |r| <= ln2/2
// (truncation below 1e-8 relative), and scaling by 2^k through the exponent
// field.  Below -87 the result is 0, above 88 +inf, NaN stays NaN; within
// [-87, 88] it is within 2 ulp of exp.  Selects instead of branches, so a
// wave runs one straight-line sequence.  The host oracle evaluates the same
// IEEE operations.
LB_DEV float exp_f(float x) {
    const float xc = x < -87.0f ? -87.0f : (x > 88.0f ? 88.0f : x);
    const float kf = __builtin_rintf(xc * 0x1.715476p+0f);        // 1 / ln2
    float r = __builtin_fmaf(-kf, 0x1.62e4p-1f, xc);                // ln2 head
    r = __builtin_fmaf(-kf, 0x1.7f7d1cp-20f, r);                    // ln2 tail
    float p = 1.0f / 5040.0f;
    p = __builtin_fmaf(p, r, 1.0f / 720.0f);
    p = __builtin_fmaf(p, r, 1.0f / 120.0f);
    p = __builtin_fmaf(p, r, 1.0f / 24.0f);
    p = __builtin_fmaf(p, r, 1.0f / 6.0f);
    p = __builtin_fmaf(p, r, 0.5f);
    p = __builtin_fmaf(p, r, 1.0f);
    p = __builtin_fmaf(p, r, 1.0f);
    const float v = p * __builtin_bit_cast(float, (uint32_t)((int)kf + 127) << 23);
    const float w = x < -87.0f ? 0.0f : (x > 88.0f ? __builtin_huge_valf() : v);
    return x != x ? x : w;
}

// psi = rho0 (1 - exp(-rho/rho0)): FP32 collisions evaluate it in single
// precision (exp_f), FP64 collisions in double (exp_cw).
template <class T> LB_DEV T sc_psi(T rho, float rho0);
template <> LB_DEV float sc_psi<float>(float rho, float rho0) {
    if (rho0 == 1.0f) return 1.0f - exp_f(-rho);
    return rho0 * (1.0f - exp_f(-(rho / rho0)));
}
template <> LB_DEV double sc_psi<double>(double rho, float rho0) {
    // rho0 = 1 (a uniform branch) skips the division and product, which are
    // exact there: the same bits either way
    if (rho0 == 1.0f) return 1.0 - exp_cw(-rho);
    double r = rho / (double)rho0;
    return (double)rho0 * (1.0 - exp_cw(-r));
}

// ---------------------------------------------------------------------------
// BGK and equilibria with the reference expression order.
// D2Q9: collision.cu:299-330.  D3Q19: collision.cu:481-541.
// ---------------------------------------------------------------------------
template <class T>
LB_DEV void bgk9(T *f, T rho, T ux, T uy, T ti, T w0, T w1, T w5) {
    T uxs = ux * ux, uys = uy * uy;
    T usq = uxs + uys;
    T a0 = w0 * rho, a1 = w1 * rho, a5 = w5 * rho;
    T s5 = +ux + uy;
    T s6 = -ux + uy;
    const T one = T(1), three = T(3), c45 = T(4.5), c15 = T(1.5);
    f[0] = f[0] - ti * (f[0] - a0 * (one - c15 * usq));
    f[1] = f[1] - ti * (f[1] - a1 * (one + three * ux + c45 * uxs - c15 * usq));
    f[2] = f[2] - ti * (f[2] - a1 * (one + three * uy + c45 * uys - c15 * usq));
    f[3] = f[3] - ti * (f[3] - a1 * (one - three * ux + c45 * uxs - c15 * usq));
    f[4] = f[4] - ti * (f[4] - a1 * (one - three * uy + c45 * uys - c15 * usq));
    f[5] = f[5] - ti * (f[5] - a5 * (one + three * s5 + c45 * s5 * s5 - c15 * usq));
    f[6] = f[6] - ti * (f[6] - a5 * (one + three * s6 + c45 * s6 * s6 - c15 * usq));
    f[7] = f[7] - ti * (f[7] - a5 * (one - three * s5 + c45 * s5 * s5 - c15 * usq));
    f[8] = f[8] - ti * (f[8] - a5 * (one - three * s6 + c45 * s6 * s6 - c15 * usq));
}
template <class T>
LB_DEV void feq9(T *f, T rho, T ux, T uy, T w0, T w1, T w5) {
    T uxs = ux * ux, uys = uy * uy;
    T usq = uxs + uys;
    T a0 = w0 * rho, a1 = w1 * rho, a5 = w5 * rho;
    T s5 = +ux + uy;
    T s6 = -ux + uy;
    const T one = T(1), three = T(3), c45 = T(4.5), c15 = T(1.5);
    f[0] = a0 * (one - c15 * usq);
    f[1] = a1 * (one + three * ux + c45 * uxs - c15 * usq);
    f[2] = a1 * (one + three * uy + c45 * uys - c15 * usq);
    f[3] = a1 * (one - three * ux + c45 * uxs - c15 * usq);
    f[4] = a1 * (one - three * uy + c45 * uys - c15 * usq);
    f[5] = a5 * (one + three * s5 + c45 * s5 * s5 - c15 * usq);
    f[6] = a5 * (one + three * s6 + c45 * s6 * s6 - c15 * usq);
    f[7] = a5 * (one - three * s5 + c45 * s5 * s5 - c15 * usq);
    f[8] = a5 * (one - three * s6 + c45 * s6 * s6 - c15 * usq);
}
template <class T>
LB_DEV void bgk19(T *f, T rho, T ux, T uy, T uz, T ti, T w0, T w1, T w7) {
    T uxsq = ux * ux, uysq = uy * uy, uzsq = uz * uz;
    T usq = uxsq + uysq + uzsq;
    T a0 = w0 * rho, a1 = w1 * rho, a7 = w7 * rho;
    T s7 = +ux + uy, s8 = +ux - uy, s9 = +ux + uz, s10 = +ux - uz, s15 = +uy + uz, s16 = +uy - uz;
    const T one = T(1), three = T(3), c45 = T(4.5), c15 = T(1.5);
    f[0] = f[0] - ti * (f[0] - a0 * (one - c15 * usq));
    f[1] = f[1] - ti * (f[1] - a1 * (one + three * ux + c45 * uxsq - c15 * usq));
    f[2] = f[2] - ti * (f[2] - a1 * (one - three * ux + c45 * uxsq - c15 * usq));
    f[3] = f[3] - ti * (f[3] - a1 * (one + three * uy + c45 * uysq - c15 * usq));
    f[4] = f[4] - ti * (f[4] - a1 * (one - three * uy + c45 * uysq - c15 * usq));
    f[5] = f[5] - ti * (f[5] - a1 * (one + three * uz + c45 * uzsq - c15 * usq));
    f[6] = f[6] - ti * (f[6] - a1 * (one - three * uz + c45 * uzsq - c15 * usq));
    f[7] = f[7] - ti * (f[7] - a7 * (one + three * s7 + c45 * s7 * s7 - c15 * usq));
    f[8] = f[8] - ti * (f[8] - a7 * (one + three * s8 + c45 * s8 * s8 - c15 * usq));
    f[9] = f[9] - ti * (f[9] - a7 * (one + three * s9 + c45 * s9 * s9 - c15 * usq));
    f[10] = f[10] - ti * (f[10] - a7 * (one + three * s10 + c45 * s10 * s10 - c15 * usq));
    f[11] = f[11] - ti * (f[11] - a7 * (one - three * s8 + c45 * s8 * s8 - c15 * usq));
    f[12] = f[12] - ti * (f[12] - a7 * (one - three * s7 + c45 * s7 * s7 - c15 * usq));
    f[13] = f[13] - ti * (f[13] - a7 * (one - three * s10 + c45 * s10 * s10 - c15 * usq));
    f[14] = f[14] - ti * (f[14] - a7 * (one - three * s9 + c45 * s9 * s9 - c15 * usq));
    f[15] = f[15] - ti * (f[15] - a7 * (one + three * s15 + c45 * s15 * s15 - c15 * usq));
    f[16] = f[16] - ti * (f[16] - a7 * (one + three * s16 + c45 * s16 * s16 - c15 * usq));
    f[17] = f[17] - ti * (f[17] - a7 * (one - three * s16 + c45 * s16 * s16 - c15 * usq));
    f[18] = f[18] - ti * (f[18] - a7 * (one - three * s15 + c45 * s15 * s15 - c15 * usq));
}
template <class T>
LB_DEV void feq19(T *f, T rho, T ux, T uy, T uz, T w0, T w1, T w7) {
    T uxsq = ux * ux, uysq = uy * uy, uzsq = uz * uz;
    T usq = uxsq + uysq + uzsq;
    T a0 = w0 * rho, a1 = w1 * rho, a7 = w7 * rho;
    T s7 = +ux + uy, s8 = +ux - uy, s9 = +ux + uz, s10 = +ux - uz, s15 = +uy + uz, s16 = +uy - uz;
    const T one = T(1), three = T(3), c45 = T(4.5), c15 = T(1.5);
    f[0] = a0 * (one - c15 * usq);
    f[1] = a1 * (one + three * ux + c45 * uxsq - c15 * usq);
    f[2] = a1 * (one - three * ux + c45 * uxsq - c15 * usq);
    f[3] = a1 * (one + three * uy + c45 * uysq - c15 * usq);
    f[4] = a1 * (one - three * uy + c45 * uysq - c15 * usq);
    f[5] = a1 * (one + three * uz + c45 * uzsq - c15 * usq);
    f[6] = a1 * (one - three * uz + c45 * uzsq - c15 * usq);
    f[7] = a7 * (one + three * s7 + c45 * s7 * s7 - c15 * usq);
    f[8] = a7 * (one + three * s8 + c45 * s8 * s8 - c15 * usq);
    f[9] = a7 * (one + three * s9 + c45 * s9 * s9 - c15 * usq);
    f[10] = a7 * (one + three * s10 + c45 * s10 * s10 - c15 * usq);
    f[11] = a7 * (one - three * s8 + c45 * s8 * s8 - c15 * usq);
    f[12] = a7 * (one - three * s7 + c45 * s7 * s7 - c15 * usq);
    f[13] = a7 * (one - three * s10 + c45 * s10 * s10 - c15 * usq);
    f[14] = a7 * (one - three * s9 + c45 * s9 * s9 - c15 * usq);
    f[15] = a7 * (one + three * s15 + c45 * s15 * s15 - c15 * usq);
    f[16] = a7 * (one + three * s16 + c45 * s16 * s16 - c15 * usq);
    f[17] = a7 * (one - three * s16 + c45 * s16 * s16 - c15 * usq);
    f[18] = a7 * (one - three * s15 + c45 * s15 * s15 - c15 * usq);
}

// Moments: collision.cu:265-275 (D2Q9, physical uy) and :454-478 (D3Q19).
template <class T> LB_DEV void mom9(const T *f, T &rho, T &ux, T &uy) {
    rho = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8];
    ux = (f[1] - f[3] + f[5] - f[6] - f[7] + f[8]) / rho;
    uy = (f[2] - f[4] + f[5] + f[6] - f[7] - f[8]) / rho;
}
template <class T> LB_DEV void mom19(const T *f, T &rho, T &ux, T &uy, T &uz) {
    rho = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8] + f[9] + f[10] + f[11] +
          f[12] + f[13] + f[14] + f[15] + f[16] + f[17] + f[18];
    ux = (f[1] + f[7] + f[8] + f[9] + f[10] - f[2] - f[11] - f[12] - f[13] - f[14]) / rho;
    uy = (f[3] + f[7] + f[11] + f[15] + f[16] - f[4] - f[8] - f[12] - f[17] - f[18]) / rho;
    uz = (f[5] + f[9] + f[13] + f[15] + f[17] - f[6] - f[10] - f[14] - f[16] - f[18]) / rho;
}

// Zou-He with the reference's double promotions (collision.cu:211-262 D2Q9,
// :413-451 D3Q19 incl. the f11-for-f17 z correction and the outlet's reuse
// of the freshly written f11).
template <class T> LB_DEV void zouhe9_in(T *f, float ux_in) {
    T rho = (T)(((double)(f[0] + f[2] + f[4]) + 2. * (double)(f[3] + f[7] + f[6])) /
                (1. - (double)ux_in));
    T ru = rho * (T)ux_in;
    f[1] = (T)((double)f[3] + (2. / 3.) * (double)ru);
    f[5] = (T)((double)f[7] + (1. / 6.) * (double)ru - 0.5 * (double)(f[2] - f[4]));
    f[8] = (T)((double)f[6] + (1. / 6.) * (double)ru - 0.5 * (double)(f[4] - f[2]));
}
template <class T> LB_DEV void zouhe9_out(T *f, float rho_out) {
    T ux = (T)(-1. + ((double)(f[0] + f[2] + f[4]) + 2. * (double)(f[1] + f[5] + f[8])) /
                         (double)rho_out);
    T ru = (T)rho_out * ux;
    f[3] = (T)((double)f[1] - (2. / 3.) * (double)ru);
    f[7] = (T)((double)f[5] - (1. / 6.) * (double)ru + 0.5 * (double)(f[2] - f[4]));
    f[6] = (T)((double)f[8] - (1. / 6.) * (double)ru + 0.5 * (double)(f[4] - f[2]));
}
template <class T> LB_DEV void zouhe19_in(T *f, float ux_in) {
    T rho = (T)(((double)(f[0] + f[3] + f[4] + f[5] + f[6] + f[15] + f[16] + f[17] + f[18]) +
                 2. * (double)(f[2] + f[11] + f[12] + f[13] + f[14])) /
                (1. - (double)ux_in));
    T ru = rho * (T)ux_in;
    double ny = (double)(f[3] + f[15] + f[16] - f[4] - f[17] - f[18]);
    double nz = (double)(f[5] + f[11] + f[15] - f[6] - f[16] - f[18]);
    f[1] = (T)((double)f[2] + (1. / 3.) * (double)ru);
    f[8] = (T)((double)f[11] + (1. / 6.) * (double)ru + 0.5 * ny);
    f[7] = (T)((double)f[12] + (1. / 6.) * (double)ru - 0.5 * ny);
    f[9] = (T)((double)f[14] + (1. / 6.) * (double)ru - 0.5 * nz);
    f[10] = (T)((double)f[13] + (1. / 6.) * (double)ru + 0.5 * nz);
}
template <class T> LB_DEV void zouhe19_out(T *f, float rho_out) {
    T ux = (T)(-1 + ((double)(f[0] + f[3] + f[4] + f[5] + f[6] + f[15] + f[16] + f[17] + f[18]) +
                     2. * (double)(f[1] + f[7] + f[8] + f[9] + f[10])) /
                        (double)rho_out);
    T ru = (T)rho_out * ux;
    double ny = (double)(f[3] + f[15] + f[16] - f[4] - f[17] - f[18]);
    f[2] = (T)((double)f[1] - (1. / 3.) * (double)ru);
    f[11] = (T)((double)f[8] - (1. / 6.) * (double)ru - 0.5 * ny);
    f[12] = (T)((double)f[7] - (1. / 6.) * (double)ru + 0.5 * ny);
    double nz = (double)(f[5] + f[11] + f[15] - f[6] - f[16] - f[18]);
    f[14] = (T)((double)f[9] - (1. / 6.) * (double)ru + 0.5 * nz);
    f[13] = (T)((double)f[10] - (1. / 6.) * (double)ru - 0.5 * nz);
}

// ---------------------------------------------------------------------------
// Deviation form (s_i = f_i - w_i).  With drho = rho - 1:
//   drho = sum s_i,  u = (sum s_i c_i) * (1 / rho)  (the w_i c_i cancel pairwise),
//   seq_i = w_i fma(rho, g_i, drho),  g_i = 3 c.u + 4.5 (c.u)^2 - 1.5 u^2,
//   s_i <- fma(1/tau, seq_i - s_i, s_i),
// i.e. the BGK of the reference written for deviations, in an operation
// sequence chosen for instruction count: one reciprocal for the three
// velocity components, and per opposite pair (c, -c) one a = 4.5 (c.u)^2 -
// 1.5 u^2 (a fused multiply-add) shared by g = a + 3 c.u and g = a - 3 c.u.
// Explicit fmas are exact-rounded on both the GPU and the host, so the
// oracle evaluates the same sequence bit for bit.
// ---------------------------------------------------------------------------
LB_DEV float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
LB_DEV double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <class T> LB_DEV void geq_pair(T cu, T r15, T &gp, T &gm) {
    const T t = T(3) * cu;
    const T a = fma_t(T(4.5) * cu, cu, -r15);
    gp = a + t;
    gm = a - t;
}
template <class T> LB_DEV void geq9(T *g, T ux, T uy) {
    const T r15 = T(1.5) * fma_t(ux, ux, uy * uy);
    g[0] = -r15;
    geq_pair<T>(ux, r15, g[1], g[3]);
    geq_pair<T>(uy, r15, g[2], g[4]);
    geq_pair<T>(ux + uy, r15, g[5], g[7]);
    geq_pair<T>(uy - ux, r15, g[6], g[8]);
}
template <class T> LB_DEV void geq19(T *g, T ux, T uy, T uz) {
    const T r15 = T(1.5) * fma_t(ux, ux, fma_t(uy, uy, uz * uz));
    g[0] = -r15;
    geq_pair<T>(ux, r15, g[1], g[2]);
    geq_pair<T>(uy, r15, g[3], g[4]);
    geq_pair<T>(uz, r15, g[5], g[6]);
    geq_pair<T>(ux + uy, r15, g[7], g[12]);
    geq_pair<T>(ux - uy, r15, g[8], g[11]);
    geq_pair<T>(ux + uz, r15, g[9], g[14]);
    geq_pair<T>(ux - uz, r15, g[10], g[13]);
    geq_pair<T>(uy + uz, r15, g[15], g[18]);
    geq_pair<T>(uy - uz, r15, g[16], g[17]);
}

// Zou-He in deviation form (the known-population weights sum to one, the
// paired weights cancel); z correction with f17 (the reference's Q1 slip is
// kept only in the population form).
template <class T> LB_DEV void zouhe9_in_dev(T *s, T ux_in) {
    T S = (s[0] + s[2] + s[4]) + T(2) * (s[3] + s[7] + s[6]);
    T drho = (S + ux_in) / (T(1) - ux_in);
    T ru = (T(1) + drho) * ux_in;
    s[1] = s[3] + (T(2) / T(3)) * ru;
    s[5] = s[7] + (T(1) / T(6)) * ru - T(0.5) * (s[2] - s[4]);
    s[8] = s[6] + (T(1) / T(6)) * ru - T(0.5) * (s[4] - s[2]);
}
template <class T> LB_DEV void zouhe9_out_dev(T *s, T rho_out) {
    T S = (s[0] + s[2] + s[4]) + T(2) * (s[1] + s[5] + s[8]);
    T ru = (T(1) + S) - rho_out;
    s[3] = s[1] - (T(2) / T(3)) * ru;
    s[7] = s[5] - (T(1) / T(6)) * ru + T(0.5) * (s[2] - s[4]);
    s[6] = s[8] - (T(1) / T(6)) * ru + T(0.5) * (s[4] - s[2]);
}
template <class T> LB_DEV void zouhe19_in_dev(T *s, T ux_in) {
    T S = (s[0] + s[3] + s[4] + s[5] + s[6] + s[15] + s[16] + s[17] + s[18]) +
          T(2) * (s[2] + s[11] + s[12] + s[13] + s[14]);
    T drho = (S + ux_in) / (T(1) - ux_in);
    T ru = (T(1) + drho) * ux_in;
    T ny = s[3] + s[15] + s[16] - s[4] - s[17] - s[18];
    T nz = s[5] + s[15] + s[17] - s[6] - s[16] - s[18];
    s[1] = s[2] + (T(1) / T(3)) * ru;
    s[8] = s[11] + (T(1) / T(6)) * ru + T(0.5) * ny;
    s[7] = s[12] + (T(1) / T(6)) * ru - T(0.5) * ny;
    s[9] = s[14] + (T(1) / T(6)) * ru - T(0.5) * nz;
    s[10] = s[13] + (T(1) / T(6)) * ru + T(0.5) * nz;
}
template <class T> LB_DEV void zouhe19_out_dev(T *s, T rho_out) {
    T S = (s[0] + s[3] + s[4] + s[5] + s[6] + s[15] + s[16] + s[17] + s[18]) +
          T(2) * (s[1] + s[7] + s[8] + s[9] + s[10]);
    T ru = (T(1) + S) - rho_out;
    T ny = s[3] + s[15] + s[16] - s[4] - s[17] - s[18];
    T nz = s[5] + s[15] + s[17] - s[6] - s[16] - s[18];
    s[2] = s[1] - (T(1) / T(3)) * ru;
    s[11] = s[8] - (T(1) / T(6)) * ru - T(0.5) * ny;
    s[12] = s[7] - (T(1) / T(6)) * ru + T(0.5) * ny;
    s[14] = s[9] - (T(1) / T(6)) * ru + T(0.5) * nz;
    s[13] = s[10] - (T(1) / T(6)) * ru - T(0.5) * nz;
}

// ---------------------------------------------------------------------------
// FP64 wavefront reduction (64 lanes) via cross-lane shuffles, then LDS.
// ---------------------------------------------------------------------------
LB_DEV double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

}  // namespace lb

// Sticky error state for the void reference entry points (lb_runtime.cpp).
extern "C" void lb_set_error(int code);
int lb_hip_err(hipError_t e);
