// One-pass Shan-Chen step for fully periodic D3Q19 lattices without solids,
// inlets, Zou-He faces or halo planes (a multiphase box such as spinodal
// decomposition): the sweep of one time step in one launch, with psi in LDS.
//
// A workgroup of NW waves owns whole x-rows: BY rows of the lattice plus one
// frame row on each side (periodic in y), over a chunk of zc z-planes, and
// marches up through the chunk.  Rows are cut into segments of 64 nodes (one
// wave each).  Wave w holds the interior segments s = w + j NW (j < SPI, row
// 1 + s / nseg) and at most one frame segment f = w (row 0 or BY + 1; the
// autotuner may move the frame rows to the last waves, see below), so there
// is no frame column at all: the x-wrap stays inside the row (LDS rows
// carry one padding element per side holding the psi of the opposite end).
//
// Iteration k of the march:
//   A   issue the population loads of plane k (the wave's interior and frame
//       segments), one register per value, no wait;
//   B1  sweep plane k - 2 from the populations the wave loaded two iterations
//       before and kept in registers: moments, the Shan-Chen force from the
//       psi of planes k - 3 .. k - 1 in LDS, BGK; the values to store stay
//       in registers (two 16-bit values per register);
//   C   wait for the loads of A, psi of plane k into LDS slot k & 3, and the
//       interior populations packed for their sweep;
//   B2  issue the stores of plane k - 2;
//   one workgroup barrier.
// The loads of A are in flight while B1 computes.  The stores come after C:
// vector memory counts loads and stores together, so stores issued between A
// and C would make C wait for them too (measured: issuing them at the start
// of the next iteration instead, beside the loads, is 8 % slower).  Every
// population crosses HBM once
// per step (the frame rows and the planes just outside the chunk are loaded
// again by the workgroups that own them, mostly from L2).  Three planes of
// populations live in registers per interior segment; the 16-bit storages
// keep the two older ones with two populations per register.  Four LDS
// slots suffice: the slot written at iteration k was last read by the sweep
// of iteration k - 1, which every wave finished before that iteration's
// barrier.
//
// Every psi and every swept node runs the arithmetic of the two-pass step
// (psi_of_node / step_body through the same moment, force, collision and
// rounding functions, in the same order), so populations, psi and macros are
// bit-identical to the density pass + sweep and to the oracle.
//
// Addressing: a segment's row and plane are uniform per wave, so each of the
// 19 loads of a node is a buffer load with a scalar byte offset (population
// slot, source row and plane) and one of three per-lane x offsets (x - c_x,
// wrapped); stores use the node's x as the per-lane offset.  The host admits
// the form only when the population buffer fits a 32-bit byte offset.
#include <hip/hip_runtime.h>

#include "lb_pull.hpp"

namespace lb {
namespace {

LB_DEV uint16_t scr_ld16(rsrc_t r, uint32_t vo, uint32_t so) { return __builtin_amdgcn_raw_buffer_load_b16(r, vo, so, 0); }
LB_DEV float scr_ld32(rsrc_t r, uint32_t vo, uint32_t so) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, vo, so, 0));
}
// AUX 2: non-temporal stores (nothing in the step reads them back)
template <int AUX> LB_DEV void scr_st(uint16_t v, rsrc_t r, uint32_t vo, uint32_t so) {
    __builtin_amdgcn_raw_buffer_store_b16(v, r, vo, so, AUX);
}
template <int AUX> LB_DEV void scr_st(float v, rsrc_t r, uint32_t vo, uint32_t so) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, vo, so, AUX);
}

// Wrapped coordinate v in [-1, n] (one step outside at most).
LB_DEV int wrap1(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

// The 19 raw populations of one node, held in registers: two 16-bit values
// per register, or one FP32 value.
template <class S, bool P = sizeof(typename S::raw) == 2> struct Held;
template <class S> struct Held<S, true> {
    typedef unsigned short us2 __attribute__((ext_vector_type(2)));
    uint32_t w[10];
    template <int I> LB_DEV uint16_t get() const {
        return (I & 1) ? (uint16_t)(w[I >> 1] >> 16) : (uint16_t)(w[I >> 1] & 0xffffu);
    }
    // loads slot i into its half (odd slots high)
    template <class F> LB_DEV void fill(F &&ld) {
        static_for<0, 9>([&](auto M) {
            constexpr int m = decltype(M)::value;
            us2 v;
            v.x = ld(std::integral_constant<int, 2 * m>{});
            v.y = ld(std::integral_constant<int, 2 * m + 1>{});
            w[m] = __builtin_bit_cast(uint32_t, v);
        });
        w[9] = (uint32_t)ld(std::integral_constant<int, 18>{});
    }
};
template <class S> struct Held<S, false> {
    float w[19];
    template <int I> LB_DEV float get() const { return w[I]; }
    template <class F> LB_DEV void fill(F &&ld) {
        static_for<0, 19>([&](auto I) { w[decltype(I)::value] = ld(I); });
    }
};

// The 19 raw populations of one node as loaded, one register each: the loads
// stay in flight until the values are used (packing them into Held at load
// time would wait for every load at once).
// (16-bit values widened to 32 bits: an array of 16-bit elements is kept as
// packed pairs, and inserting a loaded value into a pair waits for it.)
template <class S> struct Raw {
    typedef typename std::conditional<sizeof(typename S::raw) == 2, uint32_t, typename S::raw>::type E;
    E v[19];
};

// Marks the loaded values as used here (an empty asm that reads and rewrites
// them, ordered after the preceding memory operations): the conversions and
// the packing after it cannot be hoisted next to the loads, where they would
// wait for them before the sweep has run.
template <class S> LB_DEV void scr_pin(Raw<S> &r) {
    auto &v = r.v;
    __asm__ volatile("" : "+v"(v[0]), "+v"(v[1]), "+v"(v[2]), "+v"(v[3]), "+v"(v[4]), "+v"(v[5]), "+v"(v[6]),
                          "+v"(v[7]), "+v"(v[8]), "+v"(v[9]), "+v"(v[10]), "+v"(v[11]), "+v"(v[12]), "+v"(v[13]),
                          "+v"(v[14]), "+v"(v[15]), "+v"(v[16]), "+v"(v[17]), "+v"(v[18])
                        :
                        : "memory");
}

template <class S> LB_DEV void scr_pack(const Raw<S> &r, Held<S> &h) {
    if constexpr (sizeof(typename S::raw) == 2) {
        static_for<0, 9>([&](auto M) {
            constexpr int m = decltype(M)::value;
            h.w[m] = r.v[2 * m] | (r.v[2 * m + 1] << 16);
        });
        h.w[9] = r.v[18];
    } else {
        static_for<0, 19>([&](auto I) { h.w[decltype(I)::value] = r.v[decltype(I)::value]; });
    }
}

// Loads the populations of the node at (x, y, z) after streaming: slot i
// comes from (x - c_x, y - c_y, z - c_z), periodic.  po: byte offsets of the
// population slots (i ps sizeof); ysb / zsb: byte offsets of the source rows
// / planes for c = +1, 0, -1 (index 1 - c); xb: per-lane byte offsets of the
// source columns, same indexing.  The nine row + plane sums are formed once,
// so each load costs one scalar add.
template <class S>
LB_DEV void scr_pull_raw(rsrc_t rs, const uint32_t po[19], const uint32_t ysb[3], const uint32_t zsb[3],
                         const uint32_t xb[3], Raw<S> &r) {
    typedef Vel<19> L;
    constexpr uint32_t ES = sizeof(typename S::raw);
    uint32_t rz[3][3];
    for (int u = 0; u < 3; u++)
        for (int v = 0; v < 3; v++) rz[u][v] = ysb[u] + zsb[v];
    static_for<0, 19>([&](auto I) {
        constexpr int i = decltype(I)::value;
        constexpr int cx = L::cx(i), cy = L::cy(i), cz = L::cz(i);
        const uint32_t so = po[i] + rz[1 - cy][1 - cz];
        if constexpr (ES == 2) r.v[i] = scr_ld16(rs, xb[1 - cx], so);
        else r.v[i] = scr_ld32(rs, xb[1 - cx], so);
    });
}

// Populations in compute form, as pull_node converts them.
template <class S> LB_DEV void scr_unpack(const Held<S> &h, typename S::T *f) {
    static_for<0, 19>([&](auto I) {
        constexpr int i = decltype(I)::value;
        f[i] = S::ld(h.template get<i>(), weight<19, float>(Vel<19>::wc(i)));
    });
}

// Byte offsets of the planes z - c_z for c_z = +1, 0, -1.
template <class S> LB_DEV void scr_planes(const StepArgs<S> &a, int z, uint32_t zsb[3]) {
    const uint32_t bz = a.sz * (uint32_t)sizeof(typename S::raw);
    zsb[0] = (uint32_t)wrap1(z - 1, a.nz) * bz;
    zsb[1] = (uint32_t)z * bz;
    zsb[2] = (uint32_t)wrap1(z + 1, a.nz) * bz;
}

// psi of a node from its pulled populations (psi_of_node's arithmetic).
template <class S> LB_DEV typename S::T scr_psi(const StepArgs<S> &a, const Raw<S> &r) {
    typedef typename S::T T;
    T f[19];
    static_for<0, 19>([&](auto I) {
        constexpr int i = decltype(I)::value;
        f[i] = S::ld((typename S::raw)r.v[i], weight<19, float>(Vel<19>::wc(i)));
    });
    T m0, ux, uy, uz, rho;
    if constexpr (S::DEV) {
        moments_dev<19, T>(f, m0, ux, uy, uz);
        rho = T(1) + m0;
    } else {
        moments<19, T>(f, m0, ux, uy, uz);
        rho = m0;
    }
    return sc_psi<T>(rho, a.scrho0);
}

// A row segment: row r of the (BY + 2)-row window, lattice row y, the lane's
// x, and the byte offsets of its source rows and columns.
struct Seg {
    int r, y, x;
    bool on;        // the segment exists in this tile (uniform per wave)
    bool lane;      // the lane's x lies in the lattice
    uint32_t ysb[3], xb[3];
};
template <class S> LB_DEV void scr_seg(const StepArgs<S> &a, Seg &g, int r, int xs, int y0, bool on) {
    g.r = r;
    g.on = on;
    g.y = wrap1(y0 - 1 + r, a.ny);
    g.x = xs * 64 + (int)(threadIdx.x & 63);
    g.lane = g.x < a.nx;
    for (int c = 0; c < 3; c++) {
        g.ysb[c] = (uint32_t)wrap1(g.y + c - 1, a.ny) * a.sy * (uint32_t)sizeof(typename S::raw);   // row y - c_y, c_y = 1 - c
        g.xb[c] = (uint32_t)wrap1(g.x + c - 1, a.nx) * (uint32_t)sizeof(typename S::raw);   // column x - c_x
    }
}

// Waves per SIMD each workgroup shape is compiled for: two workgroups of 6
// waves, or one of 8, 12 or 16, per CU.
constexpr int scr_min_waves(int nw) { return nw == 6 ? 3 : nw / 4; }

template <class S, bool MACROS, int NW, int SPI>
__global__ __launch_bounds__(NW * 64, scr_min_waves(NW)) void k_step_scr(const StepArgs<S> *__restrict__ pa, uint32_t dstep) {
    constexpr int Q = 19;
    typedef typename S::T T;
    typedef Vel<19> L;
    constexpr uint32_t ES = sizeof(typename S::raw);
    extern __shared__ float scr_lds[];
    T *psl = (T *)scr_lds;
    const StepArgs<S> &a = *pa;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int nseg = a.ntx, BY = a.by, W = nx + 2;
    const int slot = (BY + 2) * W;
    const int b = tile_of_block(a);
    const int ty = b % a.nty, zch = b / a.nty;
    const int y0 = ty * BY, bye = min(BY, ny - y0);
    const int z0 = zch * a.zc, z1 = min(nz, z0 + a.zc);
    const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);

    Seg si[SPI], sf;
    static_for<0, SPI>([&](auto J) {
        constexpr int j = decltype(J)::value;
        const int s = w + j * NW;
        scr_seg(a, si[j], 1 + s / nseg, s % nseg, y0, s < bye * nseg);
    });
    // the wave's frame segment: f = w, or with the frame rows on the last
    // waves (frame_rev) f = NW - 1 - w for the 16-bit storages and, for FP32,
    // f = w - foff with foff the largest multiple of nseg not above
    // NW - 2 nseg, which keeps the frame segment in the wave's own column
    // w % nseg (the x offsets of its interior segment serve both: FP32 has
    // no registers to spare at 16 waves)
    int f = w, fcol = w % nseg;
    if (a.frame_rev) {
        if constexpr (ES == 2) {
            f = NW - 1 - w;
            fcol = f % nseg;
        } else {
            f = w - (NW - 2 * nseg) / nseg * nseg;
        }
    }
    scr_seg(a, sf, f < nseg ? 0 : bye + 1, fcol, y0, f >= 0 && f < 2 * nseg);
    uint32_t po[19];                // byte offsets of the population slots
    for (int i = 0; i < Q; i++) po[i] = (uint32_t)i * a.ps * ES;
    const rsrc_t rsa = make_rsrc(a.A, (uint64_t)Q * a.N * ES);
    const rsrc_t rsb = make_rsrc(a.B, (uint64_t)Q * a.N * ES);
    uint32_t step = 0;
    if (a.round_mode) step = ldm(&a.ctr[0]) + dstep;

    Held<S> h1[SPI], h2[SPI];   // populations of planes k - 1 and k - 2 (interior segments)
    Held<S> res[SPI];               // values to store of the plane being swept
    float mac[SPI][4];
    // B2: the stores of plane q
    auto store_plane = [&](int q) {
        const uint32_t qe = (uint32_t)q * a.sz * ES;
        static_for<0, SPI>([&](auto J) {
            constexpr int j = decltype(J)::value;
            const Seg &g = si[j];
            if (!(g.on && g.lane)) return;
            const uint32_t vo = (uint32_t)g.x * ES;
            const uint32_t rowe = (uint32_t)g.y * a.sy * ES + qe;
            static_for<0, Q>([&](auto I) {
                constexpr int i = decltype(I)::value;
                // (non-temporal: 216 against 226 us for FP32 at 5 rows x 32
                // planes, BF16 / FP16 stochastic 151 / 174 against 155 / 178)
                scr_st<2>(res[j].template get<i>(), rsb, vo, po[i] + rowe);
            });
            if (MACROS) {
                const uint32_t id = (uint32_t)g.x + (uint32_t)nx * ((uint32_t)g.y + (uint32_t)ny * (uint32_t)q);
                store_macros(a, id, mac[j][0], mac[j][1], mac[j][2], mac[j][3]);
            }
        });
    };
    for (int k = z0 - 1; k <= z1 + 1; k++) {
        // A: loads of plane k for its psi (and, interior segments, its sweep)
        Raw<S> di[SPI], df;
        const bool dens = k <= z1;
        uint32_t zsk[3];
        scr_planes(a, wrap1(k, nz), zsk);
        if (dens) {
            static_for<0, SPI>([&](auto J) {
                constexpr int j = decltype(J)::value;
                if (si[j].on && si[j].lane) scr_pull_raw<S>(rsa, po, si[j].ysb, zsk, si[j].xb, di[j]);
            });
            if (sf.on && sf.lane) scr_pull_raw<S>(rsa, po, sf.ysb, zsk, sf.xb, df);
        }
        // B1: sweep plane q = k - 2 with the psi of planes q - 1, q, q + 1;
        // the values to store wait in registers for B2
        const int q = k - 2;
        const bool sw = q >= z0;
        if (sw) {
            const T *pzm = psl + ((q + 3) & 3) * slot, *pz0 = psl + (q & 3) * slot,
                    *pzp = psl + ((q + 1) & 3) * slot;
            static_for<0, SPI>([&](auto J) {
                constexpr int j = decltype(J)::value;
                const Seg &g = si[j];
                if (!(g.on && g.lane)) return;
                T f[19];
                scr_unpack<S>(h2[j], f);
                T m0, ux, uy, uz, rho, drho = T(0);
                if constexpr (S::DEV) {
                    moments_dev<Q, T>(f, m0, ux, uy, uz);
                    drho = m0;
                    rho = T(1) + drho;
                } else {
                    moments<Q, T>(f, m0, ux, uy, uz);
                    rho = m0;
                }
                const T ti = (T)a.ti;
                T ex = ux, ey = uy, ez = uz;
                if (a.force[0] != 0.0f) ex = ux + (T)a.force[0] / (rho * ti);
                if (a.force[1] != 0.0f) ey = uy + (T)a.force[1] / (rho * ti);
                if (a.force[2] != 0.0f) ez = uz + (T)a.force[2] / (rho * ti);
                const int c = g.r * W + g.x + 1;
                T IF[3] = {T(0), T(0), T(0)};
                static_for<1, Q>([&](auto I) {
                    constexpr int i = decltype(I)::value;
                    constexpr int cx = L::cx(i), cy = L::cy(i), cz = L::cz(i);
                    const T *pl = cz < 0 ? pzm : (cz > 0 ? pzp : pz0);
                    T wp = weight<Q, T>(L::wc(i)) * pl[c + cx + cy * W];
                    if constexpr (cx > 0) IF[0] = IF[0] + wp; else if constexpr (cx < 0) IF[0] = IF[0] - wp;
                    if constexpr (L::pcy(i) > 0) IF[1] = IF[1] + wp; else if constexpr (L::pcy(i) < 0) IF[1] = IF[1] - wp;
                    if constexpr (cz > 0) IF[2] = IF[2] + wp; else if constexpr (cz < 0) IF[2] = IF[2] - wp;
                });
                T gp = (T)(-a.scG) * pz0[c];
                const T kf = gp / (ti * rho);
                ex = fma_t(kf, IF[0], ex);
                ey = fma_t(kf, IF[1], ey);
                ez = fma_t(kf, IF[2], ez);
                if constexpr (S::DEV) collide_dev<Q, T>(f, drho, ex, ey, ez, ti, a.rest_closure != 0);
                else collide<Q, T>(f, rho, ex, ey, ez, ti, a.rest_closure != 0);
                if (MACROS) { mac[j][0] = (float)rho; mac[j][1] = (float)ux; mac[j][2] = (float)uy; mac[j][3] = (float)uz; }
                NodePos p;
                p.x = g.x; p.y = g.y; p.z = q;
                p.id = (uint32_t)g.x + (uint32_t)nx * ((uint32_t)g.y + (uint32_t)ny * (uint32_t)q);
                p.step = step;
                p.seed = a.round_mode ? sr_seed(step, p.id) : 0u;
                auto put = [&](auto sr) {
                    if constexpr (decltype(sr)::value && std::is_same<S, StBF16>::value) {
                        // stochastic BF16: the words of the stream for slot
                        // pairs (2m, 2m + 1) as store_val_m draws them, each
                        // pair converted straight into its packed register
                        // (the instruction's high-half form, no packing step)
                        uint32_t x = p.seed | 1u;
                        static_for<0, 9>([&](auto M) {
                            constexpr int m = decltype(M)::value;
                            x ^= x << 13; x ^= x >> 17; x ^= x << 5;
                            lb_bf16x2 v = __builtin_amdgcn_cvt_sr_bf16_f32(lb_bf16x2{}, f[2 * m], x, false);
                            v = __builtin_amdgcn_cvt_sr_bf16_f32(v, f[2 * m + 1], x << 16, true);
                            res[j].w[m] = __builtin_bit_cast(uint32_t, v);
                        });
                        x ^= x << 13; x ^= x >> 17; x ^= x << 5;
                        res[j].w[9] = __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_sr_bf16_f32(lb_bf16x2{}, f[18], x, false));
                    } else if constexpr (decltype(sr)::value && std::is_same<S, StF16>::value) {
                        // stochastic FP16, by pairs: the instruction's results
                        // straight into the packed register unless a lane of
                        // the wave needs the subnormal or overflow form for
                        // either value; then both through f32_to_f16_sr's
                        // integer forms (equal to the instruction on the
                        // other lanes), one wave-wide vote per pair
                        auto rare = [](float v) {
                            const uint32_t a = __builtin_bit_cast(uint32_t, v) & 0x7fffffffu;
                            return a - 0x38800000u >= 0x477ff000u - 0x38800000u;   // |v| < 2^-14 or >= 65520
                        };
                        uint32_t x = p.seed | 1u;
                        static_for<0, 10>([&](auto M) {
                            constexpr int m = decltype(M)::value;
                            x ^= x << 13; x ^= x >> 17; x ^= x << 5;
                            const uint32_t r0 = x & 0xffff0000u, r1 = x << 16;
                            const float v0 = f[2 * m];
                            lb_f16x2 h = __builtin_amdgcn_cvt_sr_f16_f32(lb_f16x2{}, v0, r0, false);
                            if constexpr (m < 9) {
                                const float v1 = f[2 * m + 1];
                                h = __builtin_amdgcn_cvt_sr_f16_f32(h, v1, r1, true);
                                uint32_t wd = __builtin_bit_cast(uint32_t, h);
                                if (__builtin_amdgcn_ballot_w64(rare(v0) || rare(v1)))
                                    wd = (uint32_t)f32_to_f16_sr<false>(v0, r0) | ((uint32_t)f32_to_f16_sr<false>(v1, r1) << 16);
                                res[j].w[m] = wd;
                            } else {
                                uint32_t wd = __builtin_bit_cast(uint32_t, h);
                                if (__builtin_amdgcn_ballot_w64(rare(v0))) wd = (uint32_t)f32_to_f16_sr<false>(v0, r0);
                                res[j].w[m] = wd;
                            }
                        });
                    } else {
                        res[j].fill([&](auto I) {
                            constexpr int i = decltype(I)::value;
                            return store_val_m<S, decltype(sr)::value>(p, f[i], weight<Q, float>(L::wc(i)), i);
                        });
                    }
                };
                if constexpr (kRoundable<S>) {
                    if (a.round_mode) put(std::true_type{});
                    else put(std::false_type{});
                } else {
                    put(std::false_type{});
                }
            });
        }
        // C: psi of plane k into slot k & 3 (and the rows' padding elements);
        // the interior populations packed for their sweep two iterations on
        if (dens) {
            static_for<0, SPI>([&](auto J) {
                constexpr int j = decltype(J)::value;
                if (si[j].on && si[j].lane) scr_pin<S>(di[j]);
            });
            if (sf.on && sf.lane) scr_pin<S>(df);
            T *ps = psl + ((k + 4) & 3) * slot;
            auto put_psi = [&](const Seg &g, const Raw<S> &r) {
                const T v = scr_psi<S>(a, r);
                const int row = g.r * W;
                ps[row + g.x + 1] = v;
                if (g.x == nx - 1) ps[row] = v;
                if (g.x == 0) ps[row + nx + 1] = v;
            };
            static_for<0, SPI>([&](auto J) {
                constexpr int j = decltype(J)::value;
                if (si[j].on && si[j].lane) put_psi(si[j], di[j]);
            });
            if (sf.on && sf.lane) put_psi(sf, df);
        }
        // B2: the stores of plane q, issued after C has used the loads of A
        if (sw) store_plane(q);
        static_for<0, SPI>([&](auto J) {
            constexpr int j = decltype(J)::value;
            h2[j] = h1[j];
            if (dens) scr_pack<S>(di[j], h1[j]);
        });
        __syncthreads();
    }
}

template <class S, int NW, int SPI>
int launch_v(bool macros, unsigned nb, size_t lds, hipStream_t st, const StepArgs<S> *a, uint32_t dstep) {
    if (macros) hipLaunchKernelGGL((k_step_scr<S, true, NW, SPI>), dim3(nb), dim3(NW * 64), lds, st, a, dstep);
    else hipLaunchKernelGGL((k_step_scr<S, false, NW, SPI>), dim3(nb), dim3(NW * 64), lds, st, a, dstep);
    return (int)hipGetLastError();
}

template <class S>
int launch_s(bool macros, int nw, int spi, unsigned nb, size_t lds, hipStream_t st, const void *pa, uint32_t dstep) {
    const StepArgs<S> *a = (const StepArgs<S> *)pa;
    switch (nw * 4 + spi) {
    case 6 * 4 + 1: return launch_v<S, 6, 1>(macros, nb, lds, st, a, dstep);
    case 8 * 4 + 1: return launch_v<S, 8, 1>(macros, nb, lds, st, a, dstep);
    case 8 * 4 + 2: return launch_v<S, 8, 2>(macros, nb, lds, st, a, dstep);
    case 12 * 4 + 1: return launch_v<S, 12, 1>(macros, nb, lds, st, a, dstep);
    case 16 * 4 + 1: return launch_v<S, 16, 1>(macros, nb, lds, st, a, dstep);
    default: break;
    }
    if constexpr (std::is_same<S, StBF16>::value) {
        switch (nw * 4 + spi) {
        case 6 * 4 + 2: return launch_v<S, 6, 2>(macros, nb, lds, st, a, dstep);
        case 8 * 4 + 3: return launch_v<S, 8, 3>(macros, nb, lds, st, a, dstep);
        case 12 * 4 + 2: return launch_v<S, 12, 2>(macros, nb, lds, st, a, dstep);
        default: break;
        }
    }
    return LB_ERR_UNSUPPORTED;
}

}  // namespace

// Interior segments per wave a workgroup of nw waves holds: three planes of
// populations per segment live in registers (one value per register as
// loaded, the two older planes of the 16-bit storages packed two per
// register), and each instance fits the register budget of scr_min_waves(nw)
// waves per SIMD without spilling (checked at build time with
// -Rpass-analysis=kernel-resource-usage).
int scr_max_spi(int nw, int storage) {
    const bool bf = storage == LB_STORE_BF16;
    if (nw == 6) return bf ? 2 : 1;
    if (nw == 8) return bf ? 3 : 2;
    if (nw == 12) return bf ? 2 : 1;
    if (nw == 16) return 1;
    return 0;
}

// Largest tile (rows) for nw waves: the interior segments fit nw x the
// segments per wave, the two frame rows one segment per wave, the four psi
// planes 64 KB of LDS.
int scr_max_rows(int nw, int nx, int storage) {
    const int nseg = (nx + 63) / 64;
    if (2 * nseg > nw) return 0;
    int by = nw * scr_max_spi(nw, storage) / nseg;
    while (by >= 1 && scr_lds_bytes(nx, by) > 65536) by--;
    return by;
}

size_t scr_lds_bytes(int nx, int by) { return (size_t)4 * (by + 2) * (nx + 2) * sizeof(float); }

int launch_step_scr(int storage, bool macros, int nw, int by, int nx, unsigned nblocks, hipStream_t stream,
                    const void *args, uint32_t dstep) {
    const int nseg = (nx + 63) / 64;
    const int spi = (by * nseg + nw - 1) / nw;
    const size_t lds = scr_lds_bytes(nx, by);
    switch (storage) {
    case LB_STORE_F32: return launch_s<StF32>(macros, nw, spi, nblocks, lds, stream, args, dstep);
    case LB_STORE_F16: return launch_s<StF16>(macros, nw, spi, nblocks, lds, stream, args, dstep);
    case LB_STORE_BF16: return launch_s<StBF16>(macros, nw, spi, nblocks, lds, stream, args, dstep);
    default: return LB_ERR_UNSUPPORTED;
    }
}

}  // namespace lb
