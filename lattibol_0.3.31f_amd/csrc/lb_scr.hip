// One-pass Shan-Chen step for fully periodic D3Q19 lattices without solids,
// inlets, Zou-He faces or halo planes (a multiphase box such as spinodal
// decomposition): the sweep of one time step in one launch, with psi in LDS.
//
// A workgroup owns whole x-rows: BY rows of the lattice (plus one frame row on
// each side, periodic in y) over a chunk of zc z-planes, and marches up through
// the chunk.  Each of its NW waves handles up to SPW row segments of 64 nodes
// per plane (segment s = wave + j NW, row s / nseg, x-segment s % nseg), so
// the frame rows are spread over the waves and there is no frame column at
// all: the x-wrap stays inside the row (LDS rows carry one padding element per
// side holding the psi of the opposite end).
//
// Iteration k of the march:
//   A  issue the population loads of plane k (every segment, frame included);
//   B  sweep plane k - 2 (interior segments): pull its populations again (the
//      workgroup read them two iterations before, so they come from L2),
//      moments, the Shan-Chen force from the psi of planes k - 3 .. k - 1 in
//      LDS, BGK, stores;
//   C  wait for the loads of A, psi of plane k into LDS slot k & 3;
//   one workgroup barrier.
// The loads of A are in flight while B computes and stores, so the memory
// pipe does not drain at every plane.  Four LDS slots suffice: the slot
// written at iteration k was last read by the sweep of iteration k - 1,
// which every wave finished before that iteration's barrier.
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

template <class R> LB_DEV R scr_ld(rsrc_t r, uint32_t vo, uint32_t so);
template <> LB_DEV uint16_t scr_ld<uint16_t>(rsrc_t r, uint32_t vo, uint32_t so) {
    return __builtin_amdgcn_raw_buffer_load_b16(r, vo, so, 0);
}
template <> LB_DEV float scr_ld<float>(rsrc_t r, uint32_t vo, uint32_t so) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, vo, so, 0));
}
LB_DEV void scr_st(uint16_t v, rsrc_t r, uint32_t vo, uint32_t so) {
    __builtin_amdgcn_raw_buffer_store_b16(v, r, vo, so, 0);
}
LB_DEV void scr_st(float v, rsrc_t r, uint32_t vo, uint32_t so) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, vo, so, 0);
}

// Wrapped coordinate v in [-1, n] (one step outside at most).
LB_DEV int wrap1(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

// A row segment of the tile: row r of the (BY + 2)-row window (0 and
// rows - 1 are the frame), lattice row y, and the lane's x.
struct Seg {
    int r, y, x;
    bool active;     // the segment exists in this tile (uniform per wave)
    bool interior;   // a row the workgroup sweeps (uniform per wave)
    bool lane;       // the lane's x lies in the lattice
};

// Raw populations of the node at (s.x, s.y, z) after streaming: slot i comes
// from (x - c_x, y - c_y, z - c_z), periodic.  ysb / zsb: element offsets of
// the source rows / planes for c = +1, 0, -1 (index 1 - c); xb: per-lane byte
// offsets of the source columns, same indexing.
template <class S>
LB_DEV void scr_pull(const StepArgs<S> &a, rsrc_t rs, const uint32_t ysb[3], const uint32_t zsb[3],
                     const uint32_t xb[3], typename S::raw *raw) {
    typedef Vel<19> L;
    typedef typename S::raw R;
    constexpr uint32_t ES = sizeof(R);
    static_for<0, 19>([&](auto I) {
        constexpr int i = decltype(I)::value;
        constexpr int cx = L::cx(i), cy = L::cy(i), cz = L::cz(i);
        const uint32_t so = ((uint32_t)i * a.ps + ysb[1 - cy] + zsb[1 - cz]) * ES;
        raw[i] = scr_ld<R>(rs, xb[1 - cx], so);
    });
}

// Element offsets of the planes z - c_z for c_z = +1, 0, -1.
template <class S> LB_DEV void scr_planes(const StepArgs<S> &a, int z, uint32_t zsb[3]) {
    zsb[0] = (uint32_t)wrap1(z - 1, a.nz) * a.sz;
    zsb[1] = (uint32_t)z * a.sz;
    zsb[2] = (uint32_t)wrap1(z + 1, a.nz) * a.sz;
}

template <class S, bool MACROS, int NW, int SPW>
__global__ __launch_bounds__(NW * 64, 4) void k_step_scr(const StepArgs<S> *__restrict__ pa, uint32_t dstep) {
    constexpr int Q = 19;
    typedef typename S::T T;
    typedef typename S::raw R;
    typedef Vel<19> L;
    constexpr uint32_t ES = sizeof(R);
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
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);

    Seg sg[SPW];
    uint32_t ysb[SPW][3], xb[SPW][3];
    static_for<0, SPW>([&](auto J) {
        constexpr int j = decltype(J)::value;
        const int s = w + j * NW;
        const int r = s / nseg, xs = s - r * nseg;
        Seg &g = sg[j];
        g.r = r;
        g.active = r < bye + 2;
        g.interior = r >= 1 && r <= bye;
        g.y = wrap1(y0 - 1 + r, ny);
        g.x = xs * 64 + lane;
        g.lane = g.x < nx;
        for (int c = 0; c < 3; c++) {
            ysb[j][c] = (uint32_t)wrap1(g.y + c - 1, ny) * a.sy;      // source row y - c_y, c_y = 1 - c
            xb[j][c] = (uint32_t)wrap1(g.x + c - 1, nx) * ES;         // source column x - c_x
        }
    });
    const rsrc_t rsa = make_rsrc(a.A, (uint64_t)Q * a.N * ES);
    const rsrc_t rsb = make_rsrc(a.B, (uint64_t)Q * a.N * ES);
    uint32_t step = 0;
    if (a.round_mode) step = ldm(&a.ctr[0]) + dstep;

    for (int k = z0 - 1; k <= z1 + 1; k++) {
        // A: loads of plane k for its psi
        R dr[SPW][19];
        const bool dens = k <= z1;
        const int zk = wrap1(k, nz);
        uint32_t zsk[3];
        scr_planes(a, zk, zsk);
        if (dens) {
            static_for<0, SPW>([&](auto J) {
                constexpr int j = decltype(J)::value;
                if (sg[j].active && sg[j].lane) scr_pull<S>(a, rsa, ysb[j], zsk, xb[j], dr[j]);
            });
        }
        // B: sweep plane q = k - 2 with the psi of planes q - 1, q, q + 1
        const int q = k - 2;
        if (q >= z0) {
            const T *pzm = psl + ((q + 3) & 3) * slot, *pz0 = psl + (q & 3) * slot,
                    *pzp = psl + ((q + 1) & 3) * slot;
            uint32_t zsq[3];
            scr_planes(a, q, zsq);
            static_for<0, SPW>([&](auto J) {
                constexpr int j = decltype(J)::value;
                const Seg &g = sg[j];
                if (!(g.interior && g.lane)) return;
                R rr[19];
                scr_pull<S>(a, rsa, ysb[j], zsq, xb[j], rr);
                T f[19];
                static_for<0, 19>([&](auto I) {
                    constexpr int i = decltype(I)::value;
                    f[i] = S::ld(rr[i], weight<Q, float>(L::wc(i)));
                });
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
                ex = ex + (gp * IF[0]) / (ti * rho);
                ey = ey + (gp * IF[1]) / (ti * rho);
                ez = ez + (gp * IF[2]) / (ti * rho);
                if constexpr (S::DEV) collide_dev<Q, T>(f, drho, ex, ey, ez, ti, a.rest_closure != 0);
                else collide<Q, T>(f, rho, ex, ey, ez, ti, a.rest_closure != 0);
                NodePos p;
                p.x = g.x; p.y = g.y; p.z = q;
                p.id = (uint32_t)g.x + (uint32_t)nx * ((uint32_t)g.y + (uint32_t)ny * (uint32_t)q);
                p.step = step;
                p.seed = a.round_mode ? sr_seed(step, p.id) : 0u;
                if (MACROS) store_macros(a, p.id, (float)rho, (float)ux, (float)uy, (float)uz);
                const uint32_t vo = (uint32_t)g.x * ES;
                const uint32_t rowe = (uint32_t)g.y * a.sy + (uint32_t)q * a.sz;
                auto put = [&](auto sr) {
                    static_for<0, Q>([&](auto I) {
                        constexpr int i = decltype(I)::value;
                        auto v = store_val_m<S, decltype(sr)::value>(p, f[i], weight<Q, float>(L::wc(i)), i);
                        scr_st(v, rsb, vo, ((uint32_t)i * a.ps + rowe) * ES);
                    });
                };
                if constexpr (kRoundable<S>) {
                    if (a.round_mode) put(std::true_type{});
                    else put(std::false_type{});
                } else {
                    put(std::false_type{});
                }
            });
        }
        // C: psi of plane k into slot k & 3 (and the row's padding elements)
        if (dens) {
            T *ps = psl + ((k + 4) & 3) * slot;
            static_for<0, SPW>([&](auto J) {
                constexpr int j = decltype(J)::value;
                const Seg &g = sg[j];
                if (!(g.active && g.lane)) return;
                T f[19];
                static_for<0, 19>([&](auto I) {
                    constexpr int i = decltype(I)::value;
                    f[i] = S::ld(dr[j][i], weight<Q, float>(L::wc(i)));
                });
                T m0, ux, uy, uz, rho;
                if constexpr (S::DEV) {
                    moments_dev<Q, T>(f, m0, ux, uy, uz);
                    rho = T(1) + m0;
                } else {
                    moments<Q, T>(f, m0, ux, uy, uz);
                    rho = m0;
                }
                const T v = sc_psi<T>(rho, a.scrho0);
                const int row = g.r * W;
                ps[row + g.x + 1] = v;
                if (g.x == nx - 1) ps[row] = v;
                if (g.x == 0) ps[row + nx + 1] = v;
            });
        }
        __syncthreads();
    }
}

template <class S, int NW, int SPW>
int launch_v(bool macros, unsigned nb, size_t lds, hipStream_t st, const StepArgs<S> *a, uint32_t dstep) {
    if (macros) hipLaunchKernelGGL((k_step_scr<S, true, NW, SPW>), dim3(nb), dim3(NW * 64), lds, st, a, dstep);
    else hipLaunchKernelGGL((k_step_scr<S, false, NW, SPW>), dim3(nb), dim3(NW * 64), lds, st, a, dstep);
    return (int)hipGetLastError();
}

template <class S>
int launch_s(bool macros, int variant, unsigned nb, size_t lds, hipStream_t st, const void *pa, uint32_t dstep) {
    const StepArgs<S> *a = (const StepArgs<S> *)pa;
    switch (variant) {
    case 0: return launch_v<S, 8, 3>(macros, nb, lds, st, a, dstep);
    case 1: return launch_v<S, 12, 2>(macros, nb, lds, st, a, dstep);
    case 2: return launch_v<S, 16, 2>(macros, nb, lds, st, a, dstep);
    default: return LB_ERR_UNSUPPORTED;
    }
}

}  // namespace

int scr_variant_waves(int variant) { return variant == 0 ? 8 : (variant == 1 ? 12 : (variant == 2 ? 16 : 0)); }
int scr_variant_spw(int variant) { return variant == 0 ? 3 : (variant == 1 ? 2 : (variant == 2 ? 2 : 0)); }

size_t scr_lds_bytes(int nx, int by) { return (size_t)4 * (by + 2) * (nx + 2) * sizeof(float); }

int launch_step_scr(int storage, bool macros, int variant, unsigned nblocks, size_t lds, hipStream_t stream,
                    const void *args, uint32_t dstep) {
    switch (storage) {
    case LB_STORE_F32: return launch_s<StF32>(macros, variant, nblocks, lds, stream, args, dstep);
    case LB_STORE_F16: return launch_s<StF16>(macros, variant, nblocks, lds, stream, args, dstep);
    case LB_STORE_BF16: return launch_s<StBF16>(macros, variant, nblocks, lds, stream, args, dstep);
    default: return LB_ERR_UNSUPPORTED;
    }
}

}  // namespace lb
