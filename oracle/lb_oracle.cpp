/*
 * lb_oracle.cpp -- CPU oracle for the lattibol LBM hot path.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker; never linked by the product.
 * Built with g++ -O2 -ffp-contract=off (no FMA contraction, IEEE division) so
 * every float/double operation rounds exactly as written.
 *
 * Reference files restated here (paths relative to the reference tree):
 *   src/streaming.cu:24-169        in-place swap streaming D2Q9 / D3Q19
 *   src/collision.cu:35-166        D2Q9 Shan-Chen force + adhesion + BGK
 *   src/collision.cu:168-341       D2Q9 swap-back, Zou-He, moments, forcing, BGK
 *   src/collision.cu:357-552       D3Q19 swap-back, Zou-He (with the f11/f17
 *                                  slip in the z correction), moments, BGK
 *   src/get_border.cu:26-98        halo pack;  src/apply_border.cu:25-97 unpack
 *   src/copy_border.cu:26-81       ring copy of the packed halos
 *   src/read_value.cu:27-51        point probe of the plot buffer
 *   src/simulation_d2q9.cu:52-421  host initialisation (equilibrium, Shan-Chen,
 *                                  solid_bound);  :423-560 step loop
 *   src/simulation_d3q19.cu:50-319 host initialisation; :321-436 step loop
 *   src/input.cpp:168-171          tau_inv = 1/(3 nu + 0.5)
 *
 * How each function is pinned (tests/test_oracle_*.py):
 *   orc_exp                 vs math.exp on 20k points (<= 2 ulp) and exp(0)=1,
 *                           exp(ln2)=2 exactly-to-rounding.
 *   orc_expf                single-precision exp of the FP32 Shan-Chen psi:
 *                           vs math.exp rounded to float on 30k points over
 *                           [-87, 88] (<= 2 ulp), exp(0)=1, 0 / inf / NaN edges;
 *                           psi in single vs double precision within 2 ulp.
 *   orc_splitmix64_next     published SplitMix64 outputs for seed 0.
 *   orc_f32_to_f16          numpy astype(float16) (round-to-nearest-even) on
 *                           normals, subnormals, ties, overflow.
 *   orc_f32_to_bf16         torch CPU bfloat16 conversion on the same set.
 *   orc_ref_streaming_*     permutation property: a delta in slot i moves to
 *                           x + c_i; sums of every population preserved.
 *   orc_ref_collision_*     equilibrium is a fixed point; per-node mass and
 *                           momentum conserved by BGK; Zou-He reproduces the
 *                           prescribed inlet velocity / outlet density.
 *   orc_ref_collision_d2q9  Shan-Chen / adhesion pass (G < 0), read back as
 *                           the post-collision momentum of rest populations
 *                           (tests/test_oracle_shanchen_d2q9.py): psi store =
 *                           4 exp(-200/rho); per-node force = the textbook
 *                           -G psi(x) sum w psi(x+e) e by array shifts; total
 *                           momentum conserved on a periodic box; uniform psi
 *                           beside solids gives (G psi - Gads) psi sum_solid w e
 *                           in closed form; a flat interface settles at the
 *                           reference's own coexistence pair 85.7 / 524.39
 *                           (simulation_d2q9.cu:150-153) and near the
 *                           mechanical-stability integral.  Planted slips (a
 *                           stencil sign, a transposed neighbour, an adhesion
 *                           sign, the psi constant) each fail these checks.
 *   orc_ref_*_border        get->copy->apply with k subdomains equals the
 *                           single-domain run (single phase, no Zou-He).
 *   orc_ref_init_*          Sigma feq = rho, Sigma feq c = rho u (FP64 check).
 *   orc_ref_run_*           composition of the above, plus committed golden
 *                           fixtures (tests/golden) that freeze its output.
 *   orc_pull_step           bit-identical to orc_ref_run_* on fluid nodes for
 *                           periodic / full-way bounce-back / Zou-He decks;
 *                           half-way walls: plane Couette profile u=U(y+1/2)/H
 *                           and body-force Poiseuille profile; moving lid term
 *                           by the Couette test; equilibrium inlet fixed point;
 *                           Shan-Chen D3Q19: flat-interface coexistence
 *                           densities vs the mechanical-stability integral,
 *                           and mass conservation.
 *   orc_pull_reduce         vs math.fsum of the decoded populations.
 *   orc_residual            vs a numpy evaluation of the same norm ratio.
 */
#include "lb_oracle.h"

#include <cmath>
#include <cstring>
#include <cstdlib>
#include <vector>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

/* ------------------------------------------------------------------------
 * Velocity sets.  D2Q9 storage rows grow downward (streaming.cu:48-54); the
 * physical y axis used by the collision moments points up (collision.cu:274),
 * so physical c_y = -storage c_y.  D3Q19 has no flip (collision.cu:463-478).
 * ---------------------------------------------------------------------- */
const int C9[9][3] = {{0, 0, 0}, {1, 0, 0},  {0, -1, 0}, {-1, 0, 0}, {0, 1, 0},
                      {1, -1, 0}, {-1, -1, 0}, {-1, 1, 0}, {1, 1, 0}};
const int OPP9[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
const int C19[19][3] = {{0, 0, 0},  {1, 0, 0},  {-1, 0, 0}, {0, 1, 0},  {0, -1, 0},
                        {0, 0, 1},  {0, 0, -1}, {1, 1, 0},  {1, -1, 0}, {1, 0, 1},
                        {1, 0, -1}, {-1, 1, 0}, {-1, -1, 0}, {-1, 0, 1}, {-1, 0, -1},
                        {0, 1, 1},  {0, 1, -1}, {0, -1, 1}, {0, -1, -1}};
const int OPP19[19] = {0, 2, 1, 4, 3, 6, 5, 12, 11, 14, 13, 8, 7, 10, 9, 18, 17, 16, 15};

/* Lattice weights as the reference passes them: float members assigned from
 * double quotients (simulation_d2q9.cu:56-58, simulation_d3q19.cu:54-56). */
const float W9F[3] = {(float)(16. / 36.), (float)(4. / 36.), (float)(1. / 36.)};
const float W19F[3] = {(float)(12. / 36.), (float)(2. / 36.), (float)(1. / 36.)};

inline int wclass9(int i) { return i == 0 ? 0 : (i < 5 ? 1 : 2); }
inline int wclass19(int i) { return i == 0 ? 0 : (i < 7 ? 1 : 2); }

int g_threads = 1;

/* ------------------------------------------------------------------------
 * exp: Cody-Waite reduction x = k ln2 + r with the two-part ln2 of fdlibm's
 * e_exp.c, then a degree-13 Taylor polynomial in Horner form on
 * |r| <= ln2/2 (truncation < 5e-18 relative), then scaling by 2^k built from
 * the exponent bits.  The GPU path evaluates the same sequence of IEEE
 * operations, so the two agree bit for bit.
 * ---------------------------------------------------------------------- */
double exp_cw(double x) {
    const double INV_LN2 = 1.44269504088896338700e+00;
    const double LN2_HI = 6.93147180369123816490e-01;
    const double LN2_LO = 1.90821492927058770002e-10;
    if (x != x) return x;
    if (x > 709.0) return HUGE_VAL;
    if (x < -708.0) return 0.0;
    double kd = std::nearbyint(x * INV_LN2);
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
    double scale;
    std::memcpy(&scale, &bits, 8);
    return p * scale;
}

/* ------------------------------------------------------------------------
 * exp in single precision: Cody-Waite reduction with ln2 split into a 15-bit
 * head (0x1.62e4p-1) and a tail (0x1.7f7d1cp-20), degree-7 Taylor polynomial
 * in Horner form with fused multiply-adds on |r| <= ln2/2, scaling by 2^k
 * through the exponent field; 0 below -87, +inf above 88, NaN kept.  The GPU
 * path evaluates the same IEEE single-precision operations.
 * ---------------------------------------------------------------------- */
float exp_f(float x) {
    const float xc = x < -87.0f ? -87.0f : (x > 88.0f ? 88.0f : x);
    const float kf = std::nearbyintf(xc * 0x1.715476p+0f);
    float r = std::fmaf(-kf, 0x1.62e4p-1f, xc);
    r = std::fmaf(-kf, 0x1.7f7d1cp-20f, r);
    float p = 1.0f / 5040.0f;
    p = std::fmaf(p, r, 1.0f / 720.0f);
    p = std::fmaf(p, r, 1.0f / 120.0f);
    p = std::fmaf(p, r, 1.0f / 24.0f);
    p = std::fmaf(p, r, 1.0f / 6.0f);
    p = std::fmaf(p, r, 0.5f);
    p = std::fmaf(p, r, 1.0f);
    p = std::fmaf(p, r, 1.0f);
    const uint32_t sb = (uint32_t)((int)kf + 127) << 23;
    float sc;
    std::memcpy(&sc, &sb, 4);
    const float v = p * sc;
    const float w = x < -87.0f ? 0.0f : (x > 88.0f ? HUGE_VALF : v);
    return x != x ? x : w;
}

inline uint32_t fbits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline float bitsf(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

/* float -> IEEE half, round to nearest even, subnormals kept. */
uint16_t f2h(float f) {
    uint32_t x = fbits(f);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t ax = x & 0x7fffffffu;
    if (ax >= 0x7f800000u) return (uint16_t)(sign | (ax > 0x7f800000u ? 0x7e00u : 0x7c00u));
    if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);  /* >= 65520 -> inf */
    if (ax < 0x38800000u) {                                     /* below 2^-14 */
        float t = bitsf(ax) * 16777216.0f;                      /* exact: |f| * 2^24 */
        uint32_t m = (uint32_t)std::nearbyint(t);               /* RNE, 0..1024 */
        return (uint16_t)(sign | m);
    }
    uint32_t e = (ax >> 23) - 127 + 15;
    uint32_t m = ax & 0x7fffffu;
    uint32_t h = (e << 10) | (m >> 13);
    uint32_t rem = m & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
    return (uint16_t)(sign | h);
}

float h2f(uint16_t h) {
    uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    uint32_t e = (h >> 10) & 0x1fu;
    uint32_t m = h & 0x3ffu;
    if (e == 0) {
        float v = (float)m * (1.0f / 16777216.0f);  /* exact */
        return sign ? -v : v;
    }
    if (e == 31) return bitsf(sign | 0x7f800000u | (m << 13));
    return bitsf(sign | ((e - 15 + 127) << 23) | (m << 13));
}

uint16_t f2bf(float f) {
    uint32_t u = fbits(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
float bf2f(uint16_t h) { return bitsf((uint32_t)h << 16); }

/* ------------------------------------------------------------------------
 * Stochastic rounding of the 16-bit stores (round_mode 1).  Random bits:
 * one SplitMix64 finaliser per (step, node), then a xorshift32 (13, 17, 5)
 * sequence from it with the low bit set, one word per pair of population
 * slots (high 16 bits for the even slot, low 16 for the odd one, placed in
 * the top half of the word) -- a counter-based stream, so
 * every run, thread count and device produces the same bits.  v rounds up
 * (in magnitude) to the next representable value with probability equal to
 * its distance from the lower one over the spacing (for FP16 subnormals
 * quantised to 2^-16 by the 16 random bits): the expected stored value is v,
 * so increments below half a spacing are not lost (nearest-even rounding
 * drops them every step).
 * ---------------------------------------------------------------------- */
uint32_t sr_bits(uint32_t step, uint32_t id, int slot) {
    uint64_t z = 0x5EEDull + (uint64_t)step * 0x9E3779B97F4A7C15ull + (uint64_t)id;
    z ^= z >> 30; z *= 0xBF58476D1CE4E5B9ull; z ^= z >> 27; z *= 0x94D049BB133111EBull; z ^= z >> 31;
    uint32_t x = (uint32_t)z | 1u;                    /* xorshift32 (13, 17, 5) from the node word */
    for (int m = 0; m <= (slot >> 1); m++) { x ^= x << 13; x ^= x >> 17; x ^= x << 5; }
    return (slot & 1) ? x << 16 : x & 0xffff0000u;   /* 16 bits per slot, one word per two slots */
}
/* BF16: add 16 random bits below the kept ones, truncate. */
uint16_t f2bf_sr(float f, uint32_t r) {
    uint32_t u = fbits(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return f2bf(f);   /* inf / NaN */
    return (uint16_t)((u + (r >> 16)) >> 16);
}
/* FP16: normal range as for BF16 with 13 dropped bits (the truncated value is
 * exact in FP16); below 2^-14 the spacing is 2^-24: floor(|v| 2^24 + u),
 * u in [0, 1) from 23 random bits, exact in double. */
uint16_t f2h_sr(float f, uint32_t r) {
    const uint32_t x = fbits(f);
    const uint32_t sign = (x >> 16) & 0x8000u, ax = x & 0x7fffffffu;
    if (ax >= 0x477ff000u) return f2h(f);                     /* overflow, inf, NaN */
    if (ax < 0x38800000u) {
        const double q = (double)bitsf(ax) * 16777216.0 + (double)(r >> 9) * (1.0 / 8388608.0);
        return (uint16_t)(sign | (uint32_t)q);
    }
    const uint32_t t = (ax + (r >> 19)) & ~0x1fffu;
    return (uint16_t)(sign | f2h(bitsf(t)));
}

/* ------------------------------------------------------------------------
 * Storage policies for the pull specification.  T is the compute type.
 * Shifted storage keeps f_i - w_i (w_i the float lattice weight).
 * ---------------------------------------------------------------------- */
struct StF32 { typedef float raw; typedef float T; static const bool DEV = false;
    static T ld(raw v, float) { return v; } static raw st(T f, float) { return f; } };
struct StF64 { typedef double raw; typedef double T; static const bool DEV = false;
    static T ld(raw v, float) { return v; } static raw st(T f, float) { return f; } };
/* deviation storages: values are s_i = f_i - w_i and stay deviations */
struct StF32S { typedef float raw; typedef float T; static const bool DEV = true;
    static T ld(raw v, float) { return v; } static raw st(T s, float) { return s; } };
struct StF16 { typedef uint16_t raw; typedef float T; static const bool DEV = true;
    static T ld(raw v, float) { return h2f(v); } static raw st(T s, float) { return f2h(s); } };
struct StBF16 { typedef uint16_t raw; typedef float T; static const bool DEV = true;
    static T ld(raw v, float) { return bf2f(v); } static raw st(T s, float) { return f2bf(s); } };

/* Store of one population: nearest-even, or stochastic for 16-bit storage. */
template <class S> typename S::raw store_pop(typename S::T v, float w, int mode, uint32_t step, uint32_t id, int slot) {
    (void)mode; (void)step; (void)id; (void)slot;
    return S::st(v, w);
}
template <> uint16_t store_pop<StF16>(float v, float w, int mode, uint32_t step, uint32_t id, int slot) {
    return mode ? f2h_sr(v, sr_bits(step, id, slot)) : StF16::st(v, w);
}
template <> uint16_t store_pop<StBF16>(float v, float w, int mode, uint32_t step, uint32_t id, int slot) {
    return mode ? f2bf_sr(v, sr_bits(step, id, slot)) : StBF16::st(v, w);
}

/* Lattice weights in the compute type: float weights for float compute
 * (as the reference), exact double quotients for the FP64 accuracy run. */
template <class T> struct Wt;
template <> struct Wt<float> {
    static float w9(int c) { return W9F[c]; }
    static float w19(int c) { return W19F[c]; }
};
template <> struct Wt<double> {
    static double w9(int c) { return c == 0 ? 16. / 36. : (c == 1 ? 4. / 36. : 1. / 36.); }
    static double w19(int c) { return c == 0 ? 12. / 36. : (c == 1 ? 2. / 36. : 1. / 36.); }
};

/* ------------------------------------------------------------------------
 * BGK collision with the reference expression order.
 * D2Q9: collision.cu:299-330 (ueq = velocity after forcing shifts).
 * ---------------------------------------------------------------------- */
template <class T>
inline void bgk9(T *f, T rho, T ux, T uy, T ti, T w0, T w1, T w5) {
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

/* Equilibrium with the same expressions as bgk9 (feq_i = a_i * (...)). */
template <class T>
inline void feq9(T *f, T rho, T ux, T uy, T w0, T w1, T w5) {
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

/* D3Q19: collision.cu:481-541. */
template <class T>
inline void bgk19(T *f, T rho, T ux, T uy, T uz, T ti, T w0, T w1, T w7) {
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
inline void feq19(T *f, T rho, T ux, T uy, T uz, T w0, T w1, T w7) {
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

/* Moments, summation order of collision.cu:265-275 (D2Q9, physical uy). */
template <class T>
inline void mom9(const T *f, T &rho, T &ux, T &uy) {
    rho = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8];
    ux = (f[1] - f[3] + f[5] - f[6] - f[7] + f[8]) / rho;
    uy = (f[2] - f[4] + f[5] + f[6] - f[7] - f[8]) / rho;
}
/* collision.cu:454-478 (D3Q19). */
template <class T>
inline void mom19(const T *f, T &rho, T &ux, T &uy, T &uz) {
    rho = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8] + f[9] + f[10] + f[11] +
          f[12] + f[13] + f[14] + f[15] + f[16] + f[17] + f[18];
    ux = (f[1] + f[7] + f[8] + f[9] + f[10] - f[2] - f[11] - f[12] - f[13] - f[14]) / rho;
    uy = (f[3] + f[7] + f[11] + f[15] + f[16] - f[4] - f[8] - f[12] - f[17] - f[18]) / rho;
    uz = (f[5] + f[9] + f[13] + f[15] + f[17] - f[6] - f[10] - f[14] - f[16] - f[18]) / rho;
}

/* Zou-He, D2Q9 (collision.cu:211-222 velocity inlet, 252-262 pressure
 * outlet).  The reference promotes to double through its double literals;
 * sums of floats stay float until they meet a double.  For T=double every
 * operation is double. */
template <class T>
inline void zouhe9_in(T *f, float ux_in) {
    T rho = (T)(((double)(f[0] + f[2] + f[4]) + 2. * (double)(f[3] + f[7] + f[6])) /
                (1. - (double)ux_in));
    T ru = rho * (T)ux_in;
    f[1] = (T)((double)f[3] + (2. / 3.) * (double)ru);
    f[5] = (T)((double)f[7] + (1. / 6.) * (double)ru - 0.5 * (double)(f[2] - f[4]));
    f[8] = (T)((double)f[6] + (1. / 6.) * (double)ru - 0.5 * (double)(f[4] - f[2]));
}
template <class T>
inline void zouhe9_out(T *f, float rho_out) {
    T ux = (T)(-1. + ((double)(f[0] + f[2] + f[4]) + 2. * (double)(f[1] + f[5] + f[8])) /
                         (double)rho_out);
    T ru = (T)rho_out * ux;
    f[3] = (T)((double)f[1] - (2. / 3.) * (double)ru);
    f[7] = (T)((double)f[5] - (1. / 6.) * (double)ru + 0.5 * (double)(f[2] - f[4]));
    f[6] = (T)((double)f[8] - (1. / 6.) * (double)ru + 0.5 * (double)(f[4] - f[2]));
}
/* D3Q19 (collision.cu:413-451), including the reference's use of f11 where
 * f17 belongs in the z correction and the outlet's reuse of the freshly
 * written f11 in the f14/f13 lines. */
template <class T>
inline void zouhe19_in(T *f, float ux_in) {
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
template <class T>
inline void zouhe19_out(T *f, float rho_out) {
    T ux = (T)(-1 + ((double)(f[0] + f[3] + f[4] + f[5] + f[6] + f[15] + f[16] + f[17] + f[18]) +
                     2. * (double)(f[1] + f[7] + f[8] + f[9] + f[10])) /
                        (double)rho_out);
    T ru = (T)rho_out * ux;
    double ny = (double)(f[3] + f[15] + f[16] - f[4] - f[17] - f[18]);
    f[2] = (T)((double)f[1] - (1. / 3.) * (double)ru);
    f[11] = (T)((double)f[8] - (1. / 6.) * (double)ru - 0.5 * ny);
    f[12] = (T)((double)f[7] - (1. / 6.) * (double)ru + 0.5 * ny);
    double nz = (double)(f[5] + f[11] + f[15] - f[6] - f[16] - f[18]); /* new f11 */
    f[14] = (T)((double)f[9] - (1. / 6.) * (double)ru + 0.5 * nz);
    f[13] = (T)((double)f[10] - (1. / 6.) * (double)ru - 0.5 * nz);
}

/* ---- deviation form: s_i = f_i - w_i, drho = rho - 1 ---------------------
 * drho = sum s_i, u = (sum s_i c_i) * (1 / rho), seq_i = w_i fma(rho, g_i,
 * drho) with g_i = 3 c.u + 4.5 (c.u)^2 - 1.5 u^2 evaluated per opposite pair
 * (a = fma(4.5 c.u, c.u, -1.5 u^2), g = a +- 3 c.u), s_i <- fma(1/tau,
 * seq_i - s_i, s_i): the kernels' operation sequence (lb_common.hpp,
 * lb_pull.hpp), with std::fma exact-rounded like the GPU's fma.
 * Zou-He closures follow because the known weights sum to one and paired
 * weights cancel; the D3Q19 z correction uses f17. */
template <class T> void geq_pair(T cu, T r15, T &gp, T &gm) {
    const T t = T(3) * cu;
    const T a = std::fma(T(4.5) * cu, cu, -r15);
    gp = a + t;
    gm = a - t;
}
template <class T> void geq9(T *g, T ux, T uy) {
    const T r15 = T(1.5) * std::fma(ux, ux, uy * uy);
    g[0] = -r15;
    geq_pair<T>(ux, r15, g[1], g[3]);
    geq_pair<T>(uy, r15, g[2], g[4]);
    geq_pair<T>(ux + uy, r15, g[5], g[7]);
    geq_pair<T>(uy - ux, r15, g[6], g[8]);
}
template <class T> void geq19(T *g, T ux, T uy, T uz) {
    const T r15 = T(1.5) * std::fma(ux, ux, std::fma(uy, uy, uz * uz));
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
template <class T> void zouhe9_in_dev(T *s, T u) {
    T S = (s[0] + s[2] + s[4]) + T(2) * (s[3] + s[7] + s[6]);
    T drho = (S + u) / (T(1) - u);
    T ru = (T(1) + drho) * u;
    s[1] = s[3] + (T(2) / T(3)) * ru;
    s[5] = s[7] + (T(1) / T(6)) * ru - T(0.5) * (s[2] - s[4]);
    s[8] = s[6] + (T(1) / T(6)) * ru - T(0.5) * (s[4] - s[2]);
}
template <class T> void zouhe9_out_dev(T *s, T rho_out) {
    T S = (s[0] + s[2] + s[4]) + T(2) * (s[1] + s[5] + s[8]);
    T ru = (T(1) + S) - rho_out;
    s[3] = s[1] - (T(2) / T(3)) * ru;
    s[7] = s[5] - (T(1) / T(6)) * ru + T(0.5) * (s[2] - s[4]);
    s[6] = s[8] - (T(1) / T(6)) * ru + T(0.5) * (s[4] - s[2]);
}
template <class T> void zouhe19_in_dev(T *s, T u) {
    T S = (s[0] + s[3] + s[4] + s[5] + s[6] + s[15] + s[16] + s[17] + s[18]) +
          T(2) * (s[2] + s[11] + s[12] + s[13] + s[14]);
    T drho = (S + u) / (T(1) - u);
    T ru = (T(1) + drho) * u;
    T ny = s[3] + s[15] + s[16] - s[4] - s[17] - s[18];
    T nz = s[5] + s[15] + s[17] - s[6] - s[16] - s[18];
    s[1] = s[2] + (T(1) / T(3)) * ru;
    s[8] = s[11] + (T(1) / T(6)) * ru + T(0.5) * ny;
    s[7] = s[12] + (T(1) / T(6)) * ru - T(0.5) * ny;
    s[9] = s[14] + (T(1) / T(6)) * ru - T(0.5) * nz;
    s[10] = s[13] + (T(1) / T(6)) * ru + T(0.5) * nz;
}
template <class T> void zouhe19_out_dev(T *s, T rho_out) {
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

inline int wrapi(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

/* ========================================================================
 * Pull-scheme specification.
 * ====================================================================== */
template <int Q> struct Lat;
template <> struct Lat<9> {
    static int c(int i, int a) { return C9[i][a]; }
    static int opp(int i) { return OPP9[i]; }
    static int wc(int i) { return wclass9(i); }
    static int cphys(int i, int a) { return a == 1 ? -C9[i][1] : C9[i][a]; }
};
template <> struct Lat<19> {
    static int c(int i, int a) { return C19[i][a]; }
    static int opp(int i) { return OPP19[i]; }
    static int wc(int i) { return wclass19(i); }
    static int cphys(int i, int a) { return C19[i][a]; }
};

template <int Q, class T> inline T wq(int i) {
    return Q == 9 ? Wt<T>::w9(wclass9(i)) : Wt<T>::w19(wclass19(i));
}

struct Geo {
    int n[3];
    int64_t N;
    Geo(const orc_cfg *c) { n[0] = c->nx; n[1] = c->ny; n[2] = c->nz; N = (int64_t)n[0] * n[1] * n[2]; }
    int64_t idx(int x, int y, int z) const { return x + (int64_t)n[0] * (y + (int64_t)n[1] * z); }
};

/* Which face's inlet plane node p lies on, or -1. */
inline int inlet_face(const orc_cfg *cfg, const Geo &g, const int p[3], int dims) {
    for (int a = 0; a < dims; a++) {
        if (cfg->bc[2 * a] == ORC_BC_INLET_EQ && p[a] == 0) return 2 * a;
        if (cfg->bc[2 * a + 1] == ORC_BC_INLET_EQ && p[a] == g.n[a] - 1) return 2 * a + 1;
    }
    return -1;
}

/* Resolve the source of population i for a pull at p.  Returns false when the
 * link crosses a wall face (face index in *wface). */
template <int Q>
inline bool resolve_src(const orc_cfg *cfg, const Geo &g, const int p[3], int i, int dims,
                        int s[3], int *wface) {
    for (int a = 0; a < 3; a++) s[a] = p[a] - Lat<Q>::c(i, a);
    for (int a = 0; a < dims; a++) {
        int n = g.n[a];
        if (s[a] >= 0 && s[a] < n) continue;
        int face = 2 * a + (s[a] >= n ? 1 : 0);
        switch (cfg->bc[face]) {
        case ORC_BC_PERIODIC: s[a] = wrapi(s[a], n); break;
        case ORC_BC_OUTFLOW:
        case ORC_BC_INLET_EQ: s[a] = s[a] < 0 ? 0 : n - 1; break;
        default: *wface = face; return false;
        }
    }
    return true;
}

template <int Q, class S>
void pull_node_pop(const orc_cfg *cfg, const Geo &g, const typename S::raw *A, const uint8_t *solid,
                   const int p[3], int64_t id, int dims, typename S::T *f) {
    typedef typename S::T T;
    for (int i = 0; i < Q; i++) {
        int s[3], wface = -1;
        bool ok = resolve_src<Q>(cfg, g, p, i, dims, s, &wface);
        int64_t sid = ok ? g.idx(s[0], s[1], s[2]) : -1;
        bool bounce = !ok || (!cfg->bb_full && solid[sid]);
        if (bounce) {
            int j = Lat<Q>::opp(i);
            T v = S::ld(A[(int64_t)j * g.N + id], (float)wq<Q, float>(j));
            if (wface >= 0 && cfg->bc[wface] == ORC_BC_MOVING_WALL) {
                /* half-way moving wall: f_i = f*_opp(i) + 6 w_i rho_w (c_i . u_w), rho_w = 1 */
                T cu = (T)Lat<Q>::cphys(i, 0) * (T)cfg->bc_u[wface][0] +
                       (T)Lat<Q>::cphys(i, 1) * (T)cfg->bc_u[wface][1] +
                       (T)Lat<Q>::cphys(i, 2) * (T)cfg->bc_u[wface][2];
                v = v + (T(6) * wq<Q, T>(i)) * cu;
            }
            f[i] = v;
        } else {
            f[i] = S::ld(A[(int64_t)i * g.N + sid], (float)wq<Q, float>(i));
        }
    }
}

template <int Q, class T>
inline void moments(const T *f, T &rho, T &ux, T &uy, T &uz) {
    if (Q == 9) { mom9(f, rho, ux, uy); uz = T(0); }
    else mom19(f, rho, ux, uy, uz);
}

/* Deviation-form moments and equilibrium (storages with S::DEV). */
template <int Q, class T>
inline void moments_dev(const T *s, T &drho, T &ux, T &uy, T &uz) {
    if (Q == 9) {
        drho = s[0] + s[1] + s[2] + s[3] + s[4] + s[5] + s[6] + s[7] + s[8];
        T inv = T(1) / (T(1) + drho);
        ux = (s[1] - s[3] + s[5] - s[6] - s[7] + s[8]) * inv;
        uy = (s[2] - s[4] + s[5] + s[6] - s[7] - s[8]) * inv;
        uz = T(0);
    } else {
        drho = s[0] + s[1] + s[2] + s[3] + s[4] + s[5] + s[6] + s[7] + s[8] + s[9] + s[10] + s[11] + s[12] +
               s[13] + s[14] + s[15] + s[16] + s[17] + s[18];
        T inv = T(1) / (T(1) + drho);
        ux = (s[1] + s[7] + s[8] + s[9] + s[10] - s[2] - s[11] - s[12] - s[13] - s[14]) * inv;
        uy = (s[3] + s[7] + s[11] + s[15] + s[16] - s[4] - s[8] - s[12] - s[17] - s[18]) * inv;
        uz = (s[5] + s[9] + s[13] + s[15] + s[17] - s[6] - s[10] - s[14] - s[16] - s[18]) * inv;
    }
}
template <int Q, class T>
inline void feq_dev(T *seq, T drho, T ux, T uy, T uz) {
    T g[19];
    if (Q == 9) geq9<T>(g, ux, uy);
    else geq19<T>(g, ux, uy, uz);
    T rho = T(1) + drho;
    for (int i = 0; i < Q; i++) seq[i] = wq<Q, T>(i) * std::fma(rho, g[i], drho);
}
template <int Q, class S>
inline void zou_he_any(const orc_cfg *cfg, int x, int nx, typename S::T *f) {
    typedef typename S::T T;
    if (!cfg->zou_he) return;
    if (x == 0) {
        if (S::DEV) { if (Q == 9) zouhe9_in_dev<T>(f, (T)cfg->zh_ux_in); else zouhe19_in_dev<T>(f, (T)cfg->zh_ux_in); }
        else if (Q == 9) zouhe9_in(f, cfg->zh_ux_in);
        else zouhe19_in(f, cfg->zh_ux_in);
    } else if (x == nx - 1) {
        if (S::DEV) { if (Q == 9) zouhe9_out_dev<T>(f, (T)cfg->zh_rho_out); else zouhe19_out_dev<T>(f, (T)cfg->zh_rho_out); }
        else if (Q == 9) zouhe9_out(f, cfg->zh_rho_out);
        else zouhe19_out(f, cfg->zh_rho_out);
    }
}

/* Nodes away from every face whose neighbours are all fluid pull with plain
 * index offsets (the same values resolve_src would find). */
template <int Q>
std::vector<uint8_t> simple_nodes(const orc_cfg *cfg, const Geo &g, const uint8_t *solid, int dims) {
    std::vector<uint8_t> m(g.N, 0);
    for (int z = 0; z < g.n[2]; z++)
        for (int y = 0; y < g.n[1]; y++)
            for (int x = 0; x < g.n[0]; x++) {
                if (x == 0 || x == g.n[0] - 1 || y == 0 || y == g.n[1] - 1) continue;
                if (dims == 3 && (z == 0 || z == g.n[2] - 1)) continue;
                int64_t id = g.idx(x, y, z);
                bool ok = true;
                if (!cfg->bb_full)
                    for (int i = 1; i < Q && ok; i++)
                        ok = !solid[g.idx(x - Lat<Q>::c(i, 0), y - Lat<Q>::c(i, 1), z - Lat<Q>::c(i, 2))];
                m[id] = ok;
            }
    return m;
}

template <int Q, class S>
inline void pull_simple(const Geo &g, const typename S::raw *A, int64_t id, typename S::T *f) {
    for (int i = 0; i < Q; i++) {
        int64_t off = Lat<Q>::c(i, 0) + (int64_t)g.n[0] * (Lat<Q>::c(i, 1) + (int64_t)g.n[1] * Lat<Q>::c(i, 2));
        f[i] = S::ld(A[(int64_t)i * g.N + id - off], (float)wq<Q, float>(i));
    }
}

template <int Q, class S>
typename S::T node_density(const orc_cfg *cfg, const Geo &g, const typename S::raw *A,
                           const uint8_t *solid, const int p[3], int64_t id, int dims) {
    typedef typename S::T T;
    T f[19];
    int inf = inlet_face(cfg, g, p, dims);
    if (inf >= 0) return (T)cfg->bc_rho[inf];
    pull_node_pop<Q, S>(cfg, g, A, solid, p, id, dims, f);
    zou_he_any<Q, S>(cfg, p[0], g.n[0], f);
    T rho, ux, uy, uz;
    if (S::DEV) {
        T drho;
        moments_dev<Q, T>(f, drho, ux, uy, uz);
        rho = T(1) + drho;
    } else {
        moments<Q, T>(f, rho, ux, uy, uz);
    }
    return rho;
}

/* psi = rho0 * (1 - exp(-rho/rho0)): single precision (exp_f) for FP32
 * collisions, double (exp_cw) for FP64 ones. */
template <class T> inline T sc_psi(T rho, float rho0);
template <> inline float sc_psi<float>(float rho, float rho0) {
    if (rho0 == 1.0f) return 1.0f - exp_f(-rho);
    return rho0 * (1.0f - exp_f(-(rho / rho0)));
}
template <> inline double sc_psi<double>(double rho, float rho0) {
    double r = rho / (double)rho0;
    return (double)rho0 * (1.0 - exp_cw(-r));
}

template <int Q, class S>
void pull_step_impl(const orc_cfg *cfg, const typename S::raw *A, typename S::raw *B,
                    const uint8_t *solid, typename S::T *macros, const uint8_t *simple) {
    typedef typename S::T T;
    Geo g(cfg);
    const int dims = Q == 9 ? 2 : 3;
    std::vector<uint8_t> own;
    if (!simple) { own = simple_nodes<Q>(cfg, g, solid, dims); simple = own.data(); }
    const T ti = (T)cfg->tau_inv;
    const bool sc = cfg->sc_G != 0.0f;
    std::vector<T> psi;
    if (sc) {
        psi.resize(g.N);
#pragma omp parallel for schedule(static) num_threads(g_threads)
        for (int z = 0; z < g.n[2]; z++)
            for (int y = 0; y < g.n[1]; y++)
                for (int x = 0; x < g.n[0]; x++) {
                    int p[3] = {x, y, z};
                    int64_t id = g.idx(x, y, z);
                    if (solid[id]) { psi[id] = (T)cfg->sc_psi_solid; continue; }
                    psi[id] = sc_psi<T>(node_density<Q, S>(cfg, g, A, solid, p, id, dims), cfg->sc_rho0);
                }
    }
    // Rows are processed in bursts: the pulls of the simple nodes are staged
    // per population (one contiguous stream at a time) and the results are
    // written back per population, which keeps the CPU's prefetchers fed.
    // The arithmetic per node is unchanged.
    const int nx = g.n[0];
    int64_t offs[19];
    for (int i = 0; i < Q; i++)
        offs[i] = Lat<Q>::c(i, 0) + (int64_t)g.n[0] * (Lat<Q>::c(i, 1) + (int64_t)g.n[1] * Lat<Q>::c(i, 2));
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (int z = 0; z < g.n[2]; z++) {
        std::vector<T> rin((size_t)Q * nx);
        std::vector<typename S::raw> rout((size_t)Q * nx);
        std::vector<uint8_t> wrote(nx);
        for (int y = 0; y < g.n[1]; y++) {
            const int64_t rowbase = g.idx(0, y, z);
            for (int i = 0; i < Q; i++) {
                const typename S::raw *src = A + (int64_t)i * g.N + rowbase - offs[i];
                const float w = (float)wq<Q, float>(i);
                for (int x = 0; x < nx; x++)
                    if (simple[rowbase + x]) rin[(size_t)i * nx + x] = S::ld(src[x], w);
            }
            std::fill(wrote.begin(), wrote.end(), 0);
            for (int x = 0; x < nx; x++) {
                int p[3] = {x, y, z};
                int64_t id = g.idx(x, y, z);
                T f[19];
                if (solid[id]) {
                    if (cfg->bb_full) {
                        pull_node_pop<Q, S>(cfg, g, A, solid, p, id, dims, f);
                        for (int i = 0; i < Q; i++) {
                            int j = Lat<Q>::opp(i);
                            rout[(size_t)j * nx + x] = store_pop<S>(f[i], (float)wq<Q, float>(j), cfg->round_mode,
                                                                    cfg->step, (uint32_t)id, j);
                        }
                        wrote[x] = 1;
                    }
                    if (macros) for (int k = 0; k < 4; k++) macros[k * g.N + id] = T(0);
                    continue;
                }
                int inf = inlet_face(cfg, g, p, dims);
                if (inf >= 0) {
                    T rho = (T)cfg->bc_rho[inf], ux = (T)cfg->bc_u[inf][0], uy = (T)cfg->bc_u[inf][1],
                      uz = (T)cfg->bc_u[inf][2];
                    if (S::DEV) feq_dev<Q, T>(f, rho - T(1), ux, uy, uz);
                    else if (Q == 9) feq9<T>(f, rho, ux, uy, wq<9, T>(0), wq<9, T>(1), wq<9, T>(5));
                    else feq19<T>(f, rho, ux, uy, uz, wq<19, T>(0), wq<19, T>(1), wq<19, T>(7));
                    for (int i = 0; i < Q; i++)
                        rout[(size_t)i * nx + x] = store_pop<S>(f[i], (float)wq<Q, float>(i), cfg->round_mode,
                                                                cfg->step, (uint32_t)id, i);
                    wrote[x] = 1;
                    if (macros) {
                        macros[id] = rho; macros[g.N + id] = ux; macros[2 * g.N + id] = uy; macros[3 * g.N + id] = uz;
                    }
                    continue;
                }
                if (simple[id]) for (int i = 0; i < Q; i++) f[i] = rin[(size_t)i * nx + x];
                else pull_node_pop<Q, S>(cfg, g, A, solid, p, id, dims, f);
                zou_he_any<Q, S>(cfg, x, g.n[0], f);
                T rho, ux, uy, uz, drho = T(0);
                if (S::DEV) {
                    moments_dev<Q, T>(f, drho, ux, uy, uz);
                    rho = T(1) + drho;
                } else {
                    moments<Q, T>(f, rho, ux, uy, uz);
                }
                if (macros) {
                    macros[id] = rho; macros[g.N + id] = ux; macros[2 * g.N + id] = uy; macros[3 * g.N + id] = uz;
                }
                T ex = ux, ey = uy, ez = uz;
                if (cfg->force[0] != 0.0f) ex = ux + (T)cfg->force[0] / (rho * ti);
                if (cfg->force[1] != 0.0f) ey = uy + (T)cfg->force[1] / (rho * ti);
                if (cfg->force[2] != 0.0f) ez = uz + (T)cfg->force[2] / (rho * ti);
                if (sc) {
                    T IF[3] = {T(0), T(0), T(0)};
                    for (int i = 1; i < Q; i++) {
                        int q[3];
                        for (int a = 0; a < 3; a++) q[a] = p[a] + Lat<Q>::c(i, a);
                        for (int a = 0; a < dims; a++) {
                            if (q[a] >= 0 && q[a] < g.n[a]) continue;
                            int face = 2 * a + (q[a] >= g.n[a] ? 1 : 0);
                            q[a] = cfg->bc[face] == ORC_BC_PERIODIC ? wrapi(q[a], g.n[a]) : (q[a] < 0 ? 0 : g.n[a] - 1);
                        }
                        T wp = wq<Q, T>(i) * psi[g.idx(q[0], q[1], q[2])];
                        for (int a = 0; a < dims; a++) {
                            int ca = Lat<Q>::cphys(i, a);
                            if (ca > 0) IF[a] = IF[a] + wp;
                            else if (ca < 0) IF[a] = IF[a] - wp;
                        }
                    }
                    T gp = (T)(-cfg->sc_G) * psi[id];
                    /* Shan-Chen velocity shift: one division, then
                     * u_eq += (-G psi / (tau^-1 rho)) F_sum with fused
                     * multiply-adds (the GPU sweeps' sequence) */
                    const T kf = gp / (ti * rho);
                    ex = std::fma(kf, IF[0], ex);
                    ey = std::fma(kf, IF[1], ey);
                    if (Q == 19) ez = std::fma(kf, IF[2], ez);
                }
                if (S::DEV) {
                    T eq[19];
                    feq_dev<Q, T>(eq, drho, ex, ey, ez);
                    if (cfg->rest_closure) {
                        T sum = eq[1];
                        for (int i = 2; i < Q; i++) sum = sum + eq[i];
                        eq[0] = drho - sum;
                    }
                    for (int i = 0; i < Q; i++) f[i] = std::fma(ti, eq[i] - f[i], f[i]);
                } else if (cfg->rest_closure) {
                    /* equilibria first (same roundings as the BGK expressions),
                       rest population closes the mass balance */
                    T eq[19];
                    if (Q == 9) feq9<T>(eq, rho, ex, ey, wq<9, T>(0), wq<9, T>(1), wq<9, T>(5));
                    else feq19<T>(eq, rho, ex, ey, ez, wq<19, T>(0), wq<19, T>(1), wq<19, T>(7));
                    T sum = eq[1];
                    for (int i = 2; i < Q; i++) sum = sum + eq[i];
                    eq[0] = rho - sum;
                    for (int i = 0; i < Q; i++) f[i] = f[i] - ti * (f[i] - eq[i]);
                } else if (Q == 9) bgk9<T>(f, rho, ex, ey, ti, wq<9, T>(0), wq<9, T>(1), wq<9, T>(5));
                else bgk19<T>(f, rho, ex, ey, ez, ti, wq<19, T>(0), wq<19, T>(1), wq<19, T>(7));
                for (int i = 0; i < Q; i++)
                    rout[(size_t)i * nx + x] = store_pop<S>(f[i], (float)wq<Q, float>(i), cfg->round_mode, cfg->step,
                                                            (uint32_t)id, i);
                wrote[x] = 1;
            }
            for (int i = 0; i < Q; i++) {
                typename S::raw *dst = B + (int64_t)i * g.N + rowbase;
                for (int x = 0; x < nx; x++)
                    if (wrote[x]) dst[x] = rout[(size_t)i * nx + x];
            }
        }
    }
}

template <int Q, class S>
void init_eq_impl(const orc_cfg *cfg, const double *rho, const double *ux, const double *uy,
                  const double *uz, typename S::raw *A) {
    typedef typename S::T T;
    Geo g(cfg);
    for (int64_t id = 0; id < g.N; id++) {
        T f[19];
        T r = (T)rho[id], a = (T)ux[id], b = (T)uy[id], c = uz ? (T)uz[id] : T(0);
        if (S::DEV) feq_dev<Q, T>(f, r - T(1), a, b, c);
        else if (Q == 9) feq9<T>(f, r, a, b, wq<9, T>(0), wq<9, T>(1), wq<9, T>(5));
        else feq19<T>(f, r, a, b, c, wq<19, T>(0), wq<19, T>(1), wq<19, T>(7));
        for (int i = 0; i < Q; i++) A[(int64_t)i * g.N + id] = S::st(f[i], (float)wq<Q, float>(i));
    }
}

template <int Q, class S>
void decode_impl(const orc_cfg *cfg, const typename S::raw *A, double *out) {
    Geo g(cfg);
    for (int i = 0; i < Q; i++)
        for (int64_t id = 0; id < g.N; id++)
            out[(int64_t)i * g.N + id] = (double)S::ld(A[(int64_t)i * g.N + id], (float)wq<Q, float>(i)) +
                                         (S::DEV ? (double)wq<Q, float>(i) : 0.0);
}

template <int Q, class S>
void reduce_impl(const orc_cfg *cfg, const typename S::raw *A, const uint8_t *solid, double out[4]) {
    Geo g(cfg);
    long double m = 0, px = 0, py = 0, pz = 0;
    for (int64_t id = 0; id < g.N; id++) {
        if (solid[id]) continue;
        for (int i = 0; i < Q; i++) {
            long double v = (long double)S::ld(A[(int64_t)i * g.N + id], (float)wq<Q, float>(i)) +
                            (S::DEV ? (long double)wq<Q, float>(i) : 0.0L);
            m += v;
            px += v * Lat<Q>::cphys(i, 0);
            py += v * Lat<Q>::cphys(i, 1);
            pz += v * Lat<Q>::cphys(i, 2);
        }
    }
    out[0] = (double)m; out[1] = (double)px; out[2] = (double)py; out[3] = (double)pz;
}

}  // namespace

/* Dispatch (lattice, storage) to a template instantiation. */
#define ORC_DISPATCH(cfg, CALL)                                                     \
    do {                                                                            \
        if ((cfg)->lattice == 9) {                                                  \
            switch ((cfg)->storage) {                                               \
            case ORC_STORE_F32: { const int Q = 9; typedef StF32 S; CALL; } break;  \
            case ORC_STORE_F32S: { const int Q = 9; typedef StF32S S; CALL; } break;\
            case ORC_STORE_F16: { const int Q = 9; typedef StF16 S; CALL; } break;  \
            case ORC_STORE_BF16: { const int Q = 9; typedef StBF16 S; CALL; } break;\
            default: { const int Q = 9; typedef StF64 S; CALL; } break;             \
            }                                                                       \
        } else {                                                                    \
            switch ((cfg)->storage) {                                               \
            case ORC_STORE_F32: { const int Q = 19; typedef StF32 S; CALL; } break; \
            case ORC_STORE_F32S: { const int Q = 19; typedef StF32S S; CALL; } break;\
            case ORC_STORE_F16: { const int Q = 19; typedef StF16 S; CALL; } break; \
            case ORC_STORE_BF16: { const int Q = 19; typedef StBF16 S; CALL; } break;\
            default: { const int Q = 19; typedef StF64 S; CALL; } break;            \
            }                                                                       \
        }                                                                           \
    } while (0)

extern "C" {

double orc_exp(double x) { return exp_cw(x); }
float orc_expf(float x) { return exp_f(x); }
float orc_psi_f32(float rho, float rho0) { return sc_psi<float>(rho, rho0); }
double orc_psi_f64(double rho, float rho0) { return sc_psi<double>(rho, rho0); }

/* SplitMix64 (Steele, Lea, Flood 2014): state += golden gamma, then the
 * variant-13 finaliser. */
uint64_t orc_splitmix64_next(uint64_t *state) {
    uint64_t z = (*state += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

/* out[k] = (float)(base + amp * (2u - 1)), u = top 53 bits of the k-th
 * SplitMix64 output of stream `seed`, scaled to [0,1). */
void orc_uniform_noise(uint64_t seed, int64_t n, double base, double amp, float *out) {
    uint64_t st = seed;
    for (int64_t k = 0; k < n; k++) {
        uint64_t v = orc_splitmix64_next(&st);
        double u = (double)(v >> 11) * (1.0 / 9007199254740992.0);
        out[k] = (float)(base + amp * (2.0 * u - 1.0));
    }
}

uint16_t orc_f32_to_f16(float f) { return f2h(f); }
float orc_f16_to_f32(uint16_t h) { return h2f(h); }
uint16_t orc_f32_to_bf16(float f) { return f2bf(f); }
uint32_t orc_sr_bits(uint32_t step, uint32_t id, int slot) { return sr_bits(step, id, slot); }
uint16_t orc_f32_to_f16_sr(float f, uint32_t r) { return f2h_sr(f, r); }
uint16_t orc_f32_to_bf16_sr(float f, uint32_t r) { return f2bf_sr(f, r); }
float orc_bf16_to_f32(uint16_t h) { return bf2f(h); }
void orc_f32_to_f16_array(const float *in, uint16_t *out, int64_t n) { for (int64_t k = 0; k < n; k++) out[k] = f2h(in[k]); }
void orc_f32_to_bf16_array(const float *in, uint16_t *out, int64_t n) { for (int64_t k = 0; k < n; k++) out[k] = f2bf(in[k]); }
void orc_set_threads(int n) { g_threads = n < 1 ? 1 : n; }
int orc_get_threads(void) { return g_threads; }

/* ===================================================================== *
 * Reference transcription.
 * ===================================================================== */

/* streaming.cu:24-74.  Each non-solid_bound node swaps the slot pairs of its
 * "positive half" directions with its periodic neighbours. */
void orc_ref_streaming_d2q9(float *f, const uint8_t *sb, int W, int H) {
    const int64_t N = (int64_t)W * H;
    float *F[9];
    for (int i = 0; i < 9; i++) F[i] = f + i * N;
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            int64_t o0 = x + (int64_t)y * W;
            if (sb[o0]) continue;
            int xn = x > 0 ? x - 1 : W - 1;
            int xp = x < W - 1 ? x + 1 : 0;
            int yp = y > 0 ? y - 1 : H - 1;
            int64_t o1 = xp + (int64_t)y * W, o2 = x + (int64_t)yp * W;
            int64_t o3 = xp + (int64_t)yp * W, o4 = xn + (int64_t)yp * W;
            std::swap(F[1][o0], F[3][o1]);
            std::swap(F[2][o0], F[4][o2]);
            std::swap(F[5][o0], F[7][o3]);
            std::swap(F[6][o0], F[8][o4]);
        }
}

/* streaming.cu:84-163. */
void orc_ref_streaming_d3q19(float *f, const uint8_t *sb, int W, int H, int D) {
    const int64_t N = (int64_t)W * H * D;
    float *F[19];
    for (int i = 0; i < 19; i++) F[i] = f + i * N;
    for (int z = 0; z < D; z++)
        for (int y = 0; y < H; y++)
            for (int x = 0; x < W; x++) {
                auto id = [&](int a, int b, int c) { return a + (int64_t)W * (b + (int64_t)H * c); };
                int64_t o0 = id(x, y, z);
                if (sb[o0]) continue;
                int xp = x < W - 1 ? x + 1 : 0;
                int yn = y > 0 ? y - 1 : H - 1, yp = y < H - 1 ? y + 1 : 0;
                int zn = z > 0 ? z - 1 : D - 1, zp = z < D - 1 ? z + 1 : 0;
                std::swap(F[1][o0], F[2][id(xp, y, z)]);
                std::swap(F[3][o0], F[4][id(x, yp, z)]);
                std::swap(F[5][o0], F[6][id(x, y, zp)]);
                std::swap(F[7][o0], F[12][id(xp, yp, z)]);
                std::swap(F[8][o0], F[11][id(xp, yn, z)]);
                std::swap(F[9][o0], F[14][id(xp, y, zp)]);
                std::swap(F[10][o0], F[13][id(xp, y, zn)]);
                std::swap(F[15][o0], F[18][id(x, yp, zp)]);
                std::swap(F[16][o0], F[17][id(x, yp, zn)]);
            }
}

/* collision.cu:168-341 (collision_d2q9_kernel) followed, when G < 0, by
 * collision.cu:35-166 (multiphase_collision_d2q9_kernel); both launched by
 * collision.cu:343-355. */
void orc_ref_collision_d2q9(float *plot, float *f, const uint8_t *solid, float *uxa, float *uya,
                            float *rhoa, float *psi, float *ueqx, float *ueqy, int W, int H,
                            int dev, int ndev, float w0, float w1, float w5, float ti,
                            float ux_in, float uy_in, float rho_out, float Fx, float Fy,
                            float G, float Gads, int zou_he) {
    (void)uy_in;
    const int64_t N = (int64_t)W * H;
    float *F[9];
    for (int i = 0; i < 9; i++) F[i] = f + i * N;
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            int64_t o = x + (int64_t)y * W;
            if (!solid[o]) {
                float v[9];
                for (int i = 0; i < 9; i++) v[i] = F[i][o];
                /* local swap restores natural slots (collision.cu:193-208) */
                std::swap(v[1], v[3]); std::swap(v[2], v[4]);
                std::swap(v[5], v[7]); std::swap(v[6], v[8]);
                if (zou_he) {
                    if (x == 0 && dev == 0) zouhe9_in(v, ux_in);
                    else if (x == W - 1 && dev == ndev - 1) zouhe9_out(v, rho_out);
                }
                float rho, ux, uy;
                mom9(v, rho, ux, uy);
                float ex = ux, ey = uy;
                if (Fx != 0) ex = ux + Fx / (rho * ti);
                if (Fy != 0) ey = uy + Fy / (rho * ti);
                uxa[o] = ux; uya[o] = uy; rhoa[o] = rho;
                if (G != 0) {
                    psi[o] = (float)(4. * exp_cw(-200. / (double)rho));
                    ueqx[o] = ex; ueqy[o] = ey;
                } else {
                    plot[o * 4 + 0] = ux; plot[o * 4 + 1] = uy; plot[o * 4 + 2] = 0.f; plot[o * 4 + 3] = rho;
                    bgk9<float>(v, rho, ex, ey, ti, w0, w1, w5);
                }
                for (int i = 0; i < 9; i++) F[i][o] = v[i];
            } else {
                plot[o * 4 + 0] = 0; plot[o * 4 + 1] = 0; plot[o * 4 + 2] = 0; plot[o * 4 + 3] = 0;
                plot[(o + N) * 4] = 0;
                uxa[o] = 0; uya[o] = 0; rhoa[o] = 0;
            }
        }
    if (!(G < 0)) return;
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            int64_t o = x + (int64_t)y * W;
            if (solid[o]) continue;
            int xn = x > 0 ? x - 1 : W - 1, xp = x < W - 1 ? x + 1 : 0;
            int yn = y > 0 ? y - 1 : H - 1, yp = y < H - 1 ? y + 1 : 0;
            int64_t m[9] = {o,
                            xp + (int64_t)y * W, x + (int64_t)yp * W, xn + (int64_t)y * W,
                            x + (int64_t)yn * W, xp + (int64_t)yp * W, xn + (int64_t)yp * W,
                            xn + (int64_t)yn * W, xp + (int64_t)yn * W};
            float IFx = 0, IFy = 0;
            IFx = IFx + w1 * psi[m[1]];
            IFx = IFx - w1 * psi[m[3]];
            IFx = IFx + w5 * psi[m[5]];
            IFx = IFx - w5 * psi[m[6]];
            IFx = IFx - w5 * psi[m[7]];
            IFx = IFx + w5 * psi[m[8]];
            IFy = IFy - w1 * psi[m[2]];
            IFy = IFy + w1 * psi[m[4]];
            IFy = IFy - w5 * psi[m[5]];
            IFy = IFy - w5 * psi[m[6]];
            IFy = IFy + w5 * psi[m[7]];
            IFy = IFy + w5 * psi[m[8]];
            IFx = -G * psi[o] * IFx;
            IFy = -G * psi[o] * IFy;
            float ex = ueqx[o] + (IFx / (ti * rhoa[o]));
            float ey = ueqy[o] + (IFy / (ti * rhoa[o]));
            if (Gads != 0) {
                float Fsx = 0, Fsy = 0;
                if (solid[m[1]]) Fsx = Fsx + w1;
                if (solid[m[2]]) Fsy = Fsy - w1;
                if (solid[m[3]]) Fsx = Fsx - w1;
                if (solid[m[4]]) Fsy = Fsy + w1;
                if (solid[m[5]]) { Fsx = Fsx + w5; Fsy = Fsy - w5; }
                if (solid[m[6]]) { Fsx = Fsx - w5; Fsy = Fsy - w5; }
                if (solid[m[7]]) { Fsx = Fsx - w5; Fsy = Fsy + w5; }
                if (solid[m[8]]) { Fsx = Fsx + w5; Fsy = Fsy + w5; }
                Fsx = -Gads * psi[o] * Fsx;
                Fsy = -Gads * psi[o] * Fsy;
                ex = ex + (Fsx / (ti * rhoa[o]));
                ey = ey + (Fsy / (ti * rhoa[o]));
            }
            plot[o * 4 + 0] = uxa[o]; plot[o * 4 + 1] = uya[o]; plot[o * 4 + 2] = 0.f; plot[o * 4 + 3] = rhoa[o];
            float v[9];
            for (int i = 0; i < 9; i++) v[i] = F[i][o];
            bgk9<float>(v, rhoa[o], ex, ey, ti, w0, w1, w5);
            for (int i = 0; i < 9; i++) F[i][o] = v[i];
        }
}

/* collision.cu:357-552. */
void orc_ref_collision_d3q19(float *plot, float *f, const uint8_t *solid, float *uxa, float *uya,
                             float *uza, float *rhoa, int W, int H, int D, int dev, int ndev,
                             float w0, float w1, float w7, float ti, float ux_in, float uy_in,
                             float uz_in, float rho_out, int zou_he) {
    (void)uy_in; (void)uz_in;
    const int64_t N = (int64_t)W * H * D;
    float *F[19];
    for (int i = 0; i < 19; i++) F[i] = f + i * N;
    for (int z = 0; z < D; z++)
        for (int y = 0; y < H; y++)
            for (int x = 0; x < W; x++) {
                int64_t o = x + (int64_t)W * (y + (int64_t)H * z);
                if (!solid[o]) {
                    float v[19];
                    for (int i = 0; i < 19; i++) v[i] = F[i][o];
                    std::swap(v[1], v[2]); std::swap(v[3], v[4]); std::swap(v[5], v[6]);
                    std::swap(v[7], v[12]); std::swap(v[8], v[11]); std::swap(v[9], v[14]);
                    std::swap(v[10], v[13]); std::swap(v[15], v[18]); std::swap(v[16], v[17]);
                    if (zou_he) {
                        if (x == 0 && dev == 0) zouhe19_in(v, ux_in);
                        else if (x == W - 1 && dev == ndev - 1) zouhe19_out(v, rho_out);
                    }
                    float rho, ux, uy, uz;
                    mom19(v, rho, ux, uy, uz);
                    uxa[o] = ux; uya[o] = uy; uza[o] = uz; rhoa[o] = rho;
                    plot[o * 4 + 0] = ux; plot[o * 4 + 1] = uy; plot[o * 4 + 2] = uz; plot[o * 4 + 3] = rho;
                    bgk19<float>(v, rho, ux, uy, uz, ti, w0, w1, w7);
                    for (int i = 0; i < 19; i++) F[i][o] = v[i];
                } else {
                    plot[o * 4 + 0] = 0; plot[o * 4 + 1] = 0; plot[o * 4 + 2] = 0; plot[o * 4 + 3] = 0;
                    uxa[o] = 0; uya[o] = 0; uza[o] = 0; rhoa[o] = 0;
                }
            }
}

/* get_border.cu:26-50: x=W-1 column of f1,f5,f8 and x=0 column of f3,f6,f7.
 * border6 = [f1 | f5 | f8 | f3 | f6 | f7], H floats each. */
void orc_ref_get_border_d2q9(const float *f, float *b, int W, int H) {
    const int64_t N = (int64_t)W * H;
    const int sr[3] = {1, 5, 8}, sl[3] = {3, 6, 7};
    for (int y = 0; y < H; y++) {
        for (int k = 0; k < 3; k++) b[k * H + y] = f[sr[k] * N + (W - 1) + (int64_t)y * W];
        for (int k = 0; k < 3; k++) b[(3 + k) * H + y] = f[sl[k] * N + (int64_t)y * W];
    }
}
/* apply_border.cu:25-48 (same buffer order). */
void orc_ref_apply_border_d2q9(float *f, const float *b, int W, int H) {
    const int64_t N = (int64_t)W * H;
    const int sr[3] = {1, 5, 8}, sl[3] = {3, 6, 7};
    for (int y = 0; y < H; y++) {
        for (int k = 0; k < 3; k++) f[sl[k] * N + (int64_t)y * W] = b[(3 + k) * H + y];
        for (int k = 0; k < 3; k++) f[sr[k] * N + (W - 1) + (int64_t)y * W] = b[k * H + y];
    }
}
/* get_border.cu:59-98, border10 = [1,7,8,9,10 | 2,11,12,13,14], index y+z*H. */
void orc_ref_get_border_d3q19(const float *f, float *b, int W, int H, int D) {
    const int64_t N = (int64_t)W * H * D, P = (int64_t)H * D;
    const int sr[5] = {1, 7, 8, 9, 10}, sl[5] = {2, 11, 12, 13, 14};
    for (int z = 0; z < D; z++)
        for (int y = 0; y < H; y++) {
            int64_t bd = y + (int64_t)z * H;
            int64_t oR = (W - 1) + (int64_t)W * (y + (int64_t)H * z), oL = (int64_t)W * (y + (int64_t)H * z);
            for (int k = 0; k < 5; k++) b[k * P + bd] = f[sr[k] * N + oR];
            for (int k = 0; k < 5; k++) b[(5 + k) * P + bd] = f[sl[k] * N + oL];
        }
}
/* apply_border.cu:58-97. */
void orc_ref_apply_border_d3q19(float *f, const float *b, int W, int H, int D) {
    const int64_t N = (int64_t)W * H * D, P = (int64_t)H * D;
    const int sr[5] = {1, 7, 8, 9, 10}, sl[5] = {2, 11, 12, 13, 14};
    for (int z = 0; z < D; z++)
        for (int y = 0; y < H; y++) {
            int64_t bd = y + (int64_t)z * H;
            int64_t oR = (W - 1) + (int64_t)W * (y + (int64_t)H * z), oL = (int64_t)W * (y + (int64_t)H * z);
            for (int k = 0; k < 5; k++) f[sl[k] * N + oL] = b[(5 + k) * P + bd];
            for (int k = 0; k < 5; k++) f[sr[k] * N + oR] = b[k * P + bd];
        }
}

/* read_value.cu:27-34 / 44-51. */
void orc_ref_read_value(const float *plot, int W, int H, int x, int y, int z, float out[4]) {
    int64_t o = x + (int64_t)y * W + (int64_t)z * W * H;
    for (int k = 0; k < 4; k++) out[k] = plot[o * 4 + k];
}

/* simulation_d2q9.cu:318-334: interior columns only, periodic in y. */
void orc_ref_solid_bound_d2q9(const uint8_t *s, uint8_t *sb, int W, int H) {
    std::memset(sb, 0, (size_t)W * H);
    for (int y = 0; y < H; y++)
        for (int x = 1; x < W - 1; x++) {
            int xn = x - 1, xp = x + 1;
            int yn = y > 0 ? y - 1 : H - 1, yp = y < H - 1 ? y + 1 : 0;
            auto S = [&](int a, int b) { return s[a + (int64_t)b * W] != 0; };
            sb[x + (int64_t)y * W] = S(x, y) && S(xp, y) && S(x, yp) && S(xn, y) && S(x, yn) &&
                                     S(xp, yp) && S(xn, yp) && S(xn, yn) && S(xp, yn);
        }
}
/* simulation_d3q19.cu:166-204. */
void orc_ref_solid_bound_d3q19(const uint8_t *s, uint8_t *sb, int W, int H, int D) {
    std::memset(sb, 0, (size_t)W * H * D);
    for (int z = 0; z < D; z++)
        for (int y = 0; y < H; y++)
            for (int x = 1; x < W - 1; x++) {
                int xn = x - 1, xp = x + 1;
                int yn = y > 0 ? y - 1 : H - 1, yp = y < H - 1 ? y + 1 : 0;
                int zn = z > 0 ? z - 1 : D - 1, zp = z < D - 1 ? z + 1 : 0;
                auto S = [&](int a, int b, int c) { return s[a + (int64_t)W * (b + (int64_t)H * c)] != 0; };
                bool all = S(x, y, z) && S(xp, y, z) && S(xn, y, z) && S(x, yp, z) && S(x, yn, z) &&
                           S(x, y, zp) && S(x, y, zn) && S(xp, yp, z) && S(xp, yn, z) && S(xp, y, zp) &&
                           S(xp, y, zn) && S(xn, yp, z) && S(xn, yn, z) && S(xn, y, zp) && S(xn, y, zn) &&
                           S(x, yp, zp) && S(x, yp, zn) && S(x, yn, zp) && S(x, yn, zn);
                sb[x + (int64_t)W * (y + (int64_t)H * z)] = all;
            }
}

/* simulation_d2q9.cu:52-421.  Equilibrium with the reference's mixed
 * float/double expressions.  Shan-Chen branch: uniform rho, psi in double,
 * forcing-shifted ueq (with the initialiser's own IFy sign and its sequential
 * reuse of ueq_x[0]); solid nodes get psi = 0 and the rest equilibrium
 * w_i*rho0 (the reference leaves them unset). */
int orc_ref_init_d2q9(const uint8_t *all_solids, int Wfull, int H, int ndev, float ux0,
                      float uy0, float rho0, float tauInv, float Fx, float Fy, float G,
                      float Gads, float *f, float *ux, float *uy, float *rho, float *psi,
                      float *ueq_x, float *ueq_y, uint8_t *solid, uint8_t *solid_bound) {
    if (G != 0 && ndev != 1) return -1;
    const int W = Wfull / ndev;
    const int64_t N = (int64_t)W * H;
    const float w0 = W9F[0], w1 = W9F[1], w2 = W9F[2];
    std::vector<float> F(9 * N), UX(N), UY(N), RHO(N), PSI(N, 0.f), EX(N, 0.f), EY(N, 0.f);
    UX[0] = ux0; UY[0] = uy0; RHO[0] = rho0;
    EX[0] = ux0; EY[0] = uy0;
    if (Fx != 0) EX[0] = UX[0] + Fx / (tauInv * RHO[0]);
    if (Fy != 0) EY[0] = UY[0] + Fy / (tauInv * RHO[0]);
    auto feq_ref = [&](float *o, int64_t k, float r, float ex, float ey) {
        float sq = ex * ex + ey * ey;
        o[0 * N + k] = (float)((double)(w0 * r) * (1. - 1.5 * (double)sq));
        o[1 * N + k] = (float)((double)(w1 * r) * (1. + 3. * (double)(+ex) + 4.5 * (double)(+ex) * (double)(+ex) - 1.5 * (double)sq));
        o[2 * N + k] = (float)((double)(w1 * r) * (1. + 3. * (double)(+ey) + 4.5 * (double)(+ey) * (double)(+ey) - 1.5 * (double)sq));
        o[3 * N + k] = (float)((double)(w1 * r) * (1. + 3. * (double)(-ex) + 4.5 * (double)(-ex) * (double)(-ex) - 1.5 * (double)sq));
        o[4 * N + k] = (float)((double)(w1 * r) * (1. + 3. * (double)(-ey) + 4.5 * (double)(-ey) * (double)(-ey) - 1.5 * (double)sq));
        float s5 = ex + ey, s6 = -ex + ey, s7 = -ex - ey, s8 = ex - ey;
        o[5 * N + k] = (float)((double)(w2 * r) * (1. + 3. * (double)s5 + 4.5 * (double)s5 * (double)s5 - 1.5 * (double)sq));
        o[6 * N + k] = (float)((double)(w2 * r) * (1. + 3. * (double)s6 + 4.5 * (double)s6 * (double)s6 - 1.5 * (double)sq));
        o[7 * N + k] = (float)((double)(w2 * r) * (1. + 3. * (double)s7 + 4.5 * (double)s7 * (double)s7 - 1.5 * (double)sq));
        o[8 * N + k] = (float)((double)(w2 * r) * (1. + 3. * (double)s8 + 4.5 * (double)s8 * (double)s8 - 1.5 * (double)sq));
    };
    if (G != 0) {
        /* ndev == 1: W == Wfull */
        for (int y = 0; y < H; y++)
            for (int x = 0; x < W; x++) {
                int64_t o = (int64_t)y * W + x;
                if (!all_solids[o]) {
                    RHO[o] = rho0;
                    PSI[o] = (float)(4. * exp_cw(-200. / (double)RHO[o]));
                    UX[o] = UX[0]; UY[o] = UY[0];
                }
            }
        for (int y = 0; y < H; y++) {
            int yp = y < H - 1 ? y + 1 : 0, yn = y > 0 ? y - 1 : H - 1;
            for (int x = 0; x < W; x++) {
                int xp = x < W - 1 ? x + 1 : 0, xn = x > 0 ? x - 1 : W - 1;
                int64_t o[9];
                o[0] = (int64_t)y * W + x;
                if (all_solids[o[0]]) continue;
                o[1] = (int64_t)y * W + xp; o[2] = (int64_t)yp * W + x; o[3] = (int64_t)y * W + xn;
                o[4] = (int64_t)yn * W + x; o[5] = (int64_t)yp * W + xp; o[6] = (int64_t)yp * W + xn;
                o[7] = (int64_t)yn * W + xn; o[8] = (int64_t)yn * W + xp;
                float IFx = 0, IFy = 0;
                IFx = IFx + w1 * PSI[o[1]]; IFx = IFx - w1 * PSI[o[3]];
                IFx = IFx + w2 * PSI[o[5]]; IFx = IFx - w2 * PSI[o[6]];
                IFx = IFx - w2 * PSI[o[7]]; IFx = IFx + w2 * PSI[o[8]];
                IFy = IFy + w1 * PSI[o[2]]; IFy = IFy - w1 * PSI[o[4]];
                IFy = IFy + w2 * PSI[o[5]]; IFy = IFy + w2 * PSI[o[6]];
                IFy = IFy - w2 * PSI[o[7]]; IFy = IFy - w2 * PSI[o[8]];
                IFx = -G * PSI[o[0]] * IFx;
                IFy = -G * PSI[o[0]] * IFy;
                EX[o[0]] = EX[0] + (IFx / (tauInv * RHO[o[0]]));
                EY[o[0]] = EY[0] + (IFy / (tauInv * RHO[o[0]]));
                float Fsx = 0, Fsy = 0;
                if (all_solids[o[1]]) Fsx = Fsx + w1;
                if (all_solids[o[2]]) Fsy = Fsy - w1;
                if (all_solids[o[3]]) Fsx = Fsx - w1;
                if (all_solids[o[4]]) Fsy = Fsy + w1;
                if (all_solids[o[5]]) { Fsx = Fsx + w2; Fsy = Fsy - w2; }
                if (all_solids[o[6]]) { Fsx = Fsx - w2; Fsy = Fsy - w2; }
                if (all_solids[o[7]]) { Fsx = Fsx - w2; Fsy = Fsy + w2; }
                if (all_solids[o[8]]) { Fsx = Fsx + w2; Fsy = Fsy + w2; }
                Fsx = -Gads * PSI[o[0]] * Fsx;
                Fsy = -Gads * PSI[o[0]] * Fsy;
                EX[o[0]] = EX[o[0]] + (Fsx / (tauInv * RHO[o[0]]));
                EY[o[0]] = EY[o[0]] + (Fsy / (tauInv * RHO[o[0]]));
                feq_ref(F.data(), o[0], RHO[o[0]], EX[o[0]], EY[o[0]]);
            }
        }
        for (int64_t o = 0; o < N; o++)
            if (all_solids[o]) {
                RHO[o] = 0; UX[o] = 0; UY[o] = 0; PSI[o] = 0;
                for (int i = 0; i < 9; i++) F[i * N + o] = W9F[wclass9(i)] * rho0;
            }
    } else {
        feq_ref(F.data(), 0, RHO[0], EX[0], EY[0]);
        for (int64_t i = 0; i < N; i++) {
            UX[i] = UX[0]; UY[i] = UY[0]; RHO[i] = RHO[0];
            for (int q = 0; q < 9; q++) F[q * N + i] = F[q * N];
        }
    }
    for (int d = 0; d < ndev; d++) {
        uint8_t *sd = solid + d * N;
        for (int y = 0; y < H; y++)
            for (int x = 0; x < W; x++)
                sd[x + (int64_t)y * W] = all_solids[(x + d * W) + (int64_t)y * W * ndev];
        orc_ref_solid_bound_d2q9(sd, solid_bound + d * N, W, H);
        std::memcpy(f + d * 9 * N, F.data(), 9 * N * sizeof(float));
        std::memcpy(ux + d * N, UX.data(), N * 4); std::memcpy(uy + d * N, UY.data(), N * 4);
        std::memcpy(rho + d * N, RHO.data(), N * 4); std::memcpy(psi + d * N, PSI.data(), N * 4);
        std::memcpy(ueq_x + d * N, EX.data(), N * 4); std::memcpy(ueq_y + d * N, EY.data(), N * 4);
    }
    return 0;
}

/* simulation_d3q19.cu:50-319 (uz is filled everywhere here; the reference
 * leaves uz[i>0] unset, an output-only field overwritten at step 1). */
int orc_ref_init_d3q19(const uint8_t *all_solids, int Wfull, int H, int D, int ndev, float ux0,
                       float uy0, float uz0, float rho0, float *f, float *ux, float *uy,
                       float *uz, float *rho, uint8_t *solid, uint8_t *solid_bound) {
    const int W = Wfull / ndev;
    const int64_t N = (int64_t)W * H * D;
    const float w0 = W19F[0], w1 = W19F[1], w2 = W19F[2];
    const float a = ux0, b = uy0, c = uz0;
    float sq = a * a + b * b + c * c;
    auto e = [&](float wr, float cu) {
        return (float)((double)wr * (1. + 3. * (double)cu + 4.5 * (double)cu * (double)cu - 1.5 * (double)sq));
    };
    float q[19];
    q[0] = (float)((double)(w0 * rho0) * (1. - 1.5 * (double)sq));
    q[1] = e(w1 * rho0, +a); q[2] = e(w1 * rho0, -a);
    q[3] = e(w1 * rho0, +b); q[4] = e(w1 * rho0, -b);
    q[5] = e(w1 * rho0, +c); q[6] = e(w1 * rho0, -c);
    q[7] = e(w2 * rho0, +a + b); q[8] = e(w2 * rho0, +a - b);
    q[9] = e(w2 * rho0, +a + c); q[10] = e(w2 * rho0, +a - c);
    q[11] = e(w2 * rho0, -a + b); q[12] = e(w2 * rho0, -a - b);
    q[13] = e(w2 * rho0, -a + c); q[14] = e(w2 * rho0, -a - c);
    q[15] = e(w2 * rho0, +b + c); q[16] = e(w2 * rho0, +b - c);
    q[17] = e(w2 * rho0, -b + c); q[18] = e(w2 * rho0, -b - c);
    for (int d = 0; d < ndev; d++) {
        for (int i = 0; i < 19; i++)
            for (int64_t k = 0; k < N; k++) f[(int64_t)d * 19 * N + i * N + k] = q[i];
        for (int64_t k = 0; k < N; k++) {
            ux[d * N + k] = a; uy[d * N + k] = b; uz[d * N + k] = c; rho[d * N + k] = rho0;
        }
        uint8_t *sd = solid + d * N;
        for (int z = 0; z < D; z++)
            for (int y = 0; y < H; y++)
                for (int x = 0; x < W; x++)
                    sd[x + (int64_t)y * W + (int64_t)z * W * H] =
                        all_solids[(x + d * W) + (int64_t)y * W * ndev + (int64_t)z * W * ndev * H];
        orc_ref_solid_bound_d3q19(sd, solid_bound + d * N, W, H, D);
    }
    return 0;
}

/* simulation_d2q9.cu:462-554: stream every subdomain; with more than one,
 * pack / ring-copy / unpack the x-edge halos; collide; probe subdomain 0. */
void orc_ref_run_d2q9(int ndev, int W, int H, float *f, const uint8_t *solid,
                      const uint8_t *sb, float *ux, float *uy, float *rho, float *psi,
                      float *ueq_x, float *ueq_y, float *plot, float ux_in, float uy_in,
                      float rho_out, float ti, float Fx, float Fy, float G, float Gads,
                      int steps, int cx, int cy, float value_cur[4], int zou_he) {
    const int64_t N = (int64_t)W * H;
    std::vector<float> border(ndev * 6 * (size_t)H), nb(ndev * 6 * (size_t)H);
    for (int s = 0; s < steps; s++) {
        for (int d = 0; d < ndev; d++) orc_ref_streaming_d2q9(f + d * 9 * N, sb + d * N, W, H);
        if (ndev > 1) {
            for (int d = 0; d < ndev; d++) orc_ref_get_border_d2q9(f + d * 9 * N, &border[d * 6 * H], W, H);
            for (int d = 0; d < ndev; d++) {
                int left = d == 0 ? ndev - 1 : d - 1, right = d == ndev - 1 ? 0 : d + 1;
                /* nb[f1,f5,f8] <- right's border[f1,f5,f8]; nb[f3,f6,f7] <- left's (copy_border.cu:26-47) */
                std::memcpy(&nb[d * 6 * H], &border[right * 6 * H], 3 * H * sizeof(float));
                std::memcpy(&nb[d * 6 * H + 3 * H], &border[left * 6 * H + 3 * H], 3 * H * sizeof(float));
            }
            for (int d = 0; d < ndev; d++) orc_ref_apply_border_d2q9(f + d * 9 * N, &nb[d * 6 * H], W, H);
        }
        for (int d = 0; d < ndev; d++)
            orc_ref_collision_d2q9(plot + d * 8 * N, f + d * 9 * N, solid + d * N, ux + d * N, uy + d * N,
                                   rho + d * N, psi + d * N, ueq_x + d * N, ueq_y + d * N, W, H, d, ndev,
                                   W9F[0], W9F[1], W9F[2], ti, ux_in, uy_in, rho_out, Fx, Fy, G, Gads, zou_he);
        orc_ref_read_value(plot, W, H, cx, cy, 0, value_cur);
    }
}

/* simulation_d3q19.cu:352-430. */
void orc_ref_run_d3q19(int ndev, int W, int H, int D, float *f, const uint8_t *solid,
                       const uint8_t *sb, float *ux, float *uy, float *uz, float *rho,
                       float *plot, float ux_in, float uy_in, float uz_in, float rho_out,
                       float ti, int steps, int cx, int cy, int cz, float value_cur[4], int zou_he) {
    const int64_t N = (int64_t)W * H * D, P = (int64_t)H * D;
    std::vector<float> border(ndev * 10 * (size_t)P), nb(ndev * 10 * (size_t)P);
    for (int s = 0; s < steps; s++) {
        for (int d = 0; d < ndev; d++) orc_ref_streaming_d3q19(f + d * 19 * N, sb + d * N, W, H, D);
        if (ndev > 1) {
            for (int d = 0; d < ndev; d++) orc_ref_get_border_d3q19(f + d * 19 * N, &border[d * 10 * P], W, H, D);
            for (int d = 0; d < ndev; d++) {
                int left = d == 0 ? ndev - 1 : d - 1, right = d == ndev - 1 ? 0 : d + 1;
                std::memcpy(&nb[d * 10 * P], &border[right * 10 * P], 5 * P * sizeof(float));
                std::memcpy(&nb[d * 10 * P + 5 * P], &border[left * 10 * P + 5 * P], 5 * P * sizeof(float));
            }
            for (int d = 0; d < ndev; d++) orc_ref_apply_border_d3q19(f + d * 19 * N, &nb[d * 10 * P], W, H, D);
        }
        for (int d = 0; d < ndev; d++)
            orc_ref_collision_d3q19(plot + d * 4 * N, f + d * 19 * N, solid + d * N, ux + d * N, uy + d * N,
                                    uz + d * N, rho + d * N, W, H, D, d, ndev, W19F[0], W19F[1], W19F[2], ti,
                                    ux_in, uy_in, uz_in, rho_out, zou_he);
        orc_ref_read_value(plot, W, H, cx, cy, cz, value_cur);
    }
}

/* ===================================================================== */
void orc_pull_init_eq(const orc_cfg *cfg, const double *rho, const double *ux, const double *uy,
                      const double *uz, void *A) {
    ORC_DISPATCH(cfg, (init_eq_impl<Q, S>(cfg, rho, ux, uy, uz, (S::raw *)A)));
}

void orc_pull_step(const orc_cfg *cfg, const void *A, void *B, const uint8_t *solid, void *macros) {
    ORC_DISPATCH(cfg, (pull_step_impl<Q, S>(cfg, (const S::raw *)A, (S::raw *)B, solid, (S::T *)macros, nullptr)));
}

void orc_pull_run(orc_cfg *cfg, void *A, void *B, const uint8_t *solid, int steps, void *macros) {
    size_t esz = (cfg->storage == ORC_STORE_F32 || cfg->storage == ORC_STORE_F32S) ? 4 : (cfg->storage == ORC_STORE_F64 ? 8 : 2);
    size_t bytes = esz * (size_t)cfg->lattice * cfg->nx * cfg->ny * cfg->nz;
    void *a = A, *b = B;
    std::vector<uint8_t> simple;
    {
        Geo g(cfg);
        simple = cfg->lattice == 9 ? simple_nodes<9>(cfg, g, solid, 2) : simple_nodes<19>(cfg, g, solid, 3);
    }
    const uint8_t *sp = simple.data();
    for (int s = 0; s < steps; s++) {
        void *mac = s == steps - 1 ? macros : nullptr;
        ORC_DISPATCH(cfg, (pull_step_impl<Q, S>(cfg, (const S::raw *)a, (S::raw *)b, solid, (S::T *)mac, sp)));
        cfg->step++;
        std::swap(a, b);
    }
    if (a != A) std::memcpy(A, a, bytes);
}

void orc_pull_decode(const orc_cfg *cfg, const void *A, double *out) {
    ORC_DISPATCH(cfg, (decode_impl<Q, S>(cfg, (const S::raw *)A, out)));
}

void orc_pull_reduce(const orc_cfg *cfg, const void *A, const uint8_t *solid, double out[4]) {
    ORC_DISPATCH(cfg, (reduce_impl<Q, S>(cfg, (const S::raw *)A, solid, out)));
}

double orc_residual(int ncomp, int64_t n, const float *u_now, const float *u_prev, const uint8_t *solid) {
    long double num = 0, den = 0;
    for (int64_t k = 0; k < n; k++) {
        if (solid && solid[k]) continue;
        for (int c = 0; c < ncomp; c++) {
            long double a = u_now[c * n + k], b = u_prev[c * n + k];
            num += (a - b) * (a - b);
            den += a * a;
        }
    }
    return den > 0 ? (double)std::sqrt(num / den) : 0.0;
}

}  // extern "C"
