// lb_pull.hpp -- building blocks of the fused pull-scheme sweep, shared by
// the single-lattice solver (lb_fused.hip) and the batched multi-lattice
// solver (lb_batch.hip), so that both run the identical per-node code.
//
// Physics per node (the reference's step, collision.cu:168-341 / 357-552,
// restated for a pull scheme):
//   f_i(x) <- f*_i(x - c_i)          pull from the neighbour (coalesced loads;
//                                    only +-1 shifts in x are unaligned)
//   links leaving the lattice: periodic wrap, clamp (zero-gradient outflow),
//   half-way bounce-back at walls (+6 w_i (c_i.u_w) for a moving wall),
//   halo planes in z (pack fused into the sweep, copy, unpack fused into the
//   next sweep); links into solid nodes: half-way bounce-back, or the
//   reference's full-way scheme (solid nodes reverse their populations).
//   Zou-He inlet/outlet (reference), equilibrium inlet planes, body force and
//   Shan-Chen velocity shifts, BGK with the reference expression order.
#pragma once

#include <algorithm>
#include <type_traits>
#include <vector>

#include "lb_common.hpp"

namespace lb {

enum : uint8_t { F_SOLID = 1, F_INLET = 2, F_SLOW = 4 };
constexpr int kThreads = 256;

template <class S> struct StepArgs {
    typedef typename S::raw raw;
    typedef typename S::T T;
    const raw *A;
    raw *B;
    const uint8_t *flags;
    const uint32_t *lmask;  // link masks of F_SLOW nodes (see pull_node)
    const uint32_t *nodew;  // packed node words: flags << 24 | link mask, every node
    int packed;             // 1: read nodew (one load per node), 0: flags (+ lmask)
    float *macros;          // 4*N: rho, ux, uy, uz
    T *psi;                 // Shan-Chen pseudo-potential
    raw *send;              // [send_lo (5 planes, -z pops of z=0) | send_hi (+z pops of z=nz-1)]
    const raw *recv;        // [recv_hi (ghost z=nz) | recv_lo (ghost z=-1)]
    uint32_t N, nxy, P;     // nodes, nodes per z-plane (= halo plane size)
    uint32_t ps, sy, sz;    // population layout: element of (i, x, y, z) = i*ps + x + y*sy + z*sz
    int nx, ny, nz;
    float ti;
    int bc[6];
    float bcu[6][3];
    float bcrho[6];
    int corr_lo[3], corr_hi[3];   // fast-path index corrections at faces
    // Buffer-descriptor accesses of the 16-bit storages (pull_node,
    // store_node): usable when every offset fits 32 bits.  Byte offset of
    // population slot i from the tile's load descriptor A + (e0 - offmax -
    // bias) * esz: lso[i] = (i*ps - off_i + offmax) * esz
    int fast;
    int bias;                     // -(corr_hi[0] + corr_hi[1]) >= 0
    int offmax;                   // largest link offset off_i = cx + cy*sy + cz*sz
    uint32_t lso[19];
    uint32_t sso[19];             // stores: i*ps * esz from the descriptor base B + e0 * esz
    float force[3];
    float scG, scrho0, scpsi_solid;
    int bb_full, zou_he, halo;
    int rest_closure;             // feq_0 = rho - sum_{i>0} feq_i (exact mass balance)
    int round_mode;               // 16-bit stores: 0 nearest-even, 1 stochastic (resolved on the host)
    uint32_t *ctr;                // step counter of the stochastic-rounding stream
    float zh_ux, zh_rho;
    int bx_shift, by, ntx, nty, nblocks, xcd;
    int zc;                       // one-pass Shan-Chen sweep: z-planes per workgroup
    int frame_rev;                // row-slab sweep: the frame rows on the last waves (see k_step_scr)
};

// Workgroups are dispatched round-robin over the 8 XCDs; this bijection
// gives each XCD a contiguous range of tiles so z/y neighbours share an L2.
LB_DEV int xcd_remap(int b, int nb) {
    int q = nb >> 3, r = nb & 7;
    int xcd = b & 7, local = b >> 3;
    return xcd * q + min(xcd, r) + local;
}

// id: node index x + nx (y + ny z) (flags, macros, psi); e: element offset of
// the node's population 0 in the population buffers; e0: that of the tile's
// first node (the same for the whole workgroup).
// step / seed: the stochastic-rounding stream's step and node word (sr_seed).
struct NodePos { int x, y, z; uint32_t id, e, e0; bool valid; uint32_t step, seed; };

// Population layouts.  LB_LAYOUT_SOA: [i][z][y][x], element i*N + id.
// LB_LAYOUT_ROWS: [z][y][i][x], the Q populations of one x-row stored back to
// back (Q*nx contiguous elements per row), so a 256-node row tile writes one
// contiguous run and reads Q short runs from at most 9 rows: far fewer DRAM
// pages open at once than 2Q independent streams.
struct PopLayout { uint32_t ps, sy, sz; };
__host__ __device__ inline PopLayout pop_layout(int layout, int Q, int nx, int ny, int nz) {
    const uint32_t N = (uint32_t)nx * ny * nz;
    if (layout == LB_LAYOUT_ROWS) return {(uint32_t)nx, (uint32_t)Q * nx, (uint32_t)Q * nx * ny};
    return {N, (uint32_t)nx, (uint32_t)nx * ny};
}
// Element offset of node id (1-D kernels: init, reduce, relayout).
LB_DEV uint32_t elem_of(uint32_t id, uint32_t nx, uint32_t ny, uint32_t sy, uint32_t sz) {
    uint32_t r = id / nx, x = id - r * nx;
    uint32_t z = r / ny, y = r - z * ny;
    return x + y * sy + z * sz;
}

template <class S> LB_DEV int tile_of_block(const StepArgs<S> &a) {
    int b = blockIdx.x;
    if (a.xcd) b = xcd_remap(b, a.nblocks);
    return b;
}

// Node of this thread in tile (tx, ty) of plane z (tiles: BX x BY nodes).
template <class S> LB_DEV NodePos node_in_tile(const StepArgs<S> &a, int tx, int ty, int z) {
    int lx = threadIdx.x & ((1 << a.bx_shift) - 1);
    int ly = threadIdx.x >> a.bx_shift;
    NodePos p;
    p.x = (tx << a.bx_shift) + lx;
    p.y = ty * a.by + ly;
    p.z = z;
    p.valid = p.x < a.nx && p.y < a.ny;
    p.id = (uint32_t)p.x + (uint32_t)a.nx * ((uint32_t)p.y + (uint32_t)a.ny * (uint32_t)z);
    p.e0 = ((uint32_t)tx << a.bx_shift) + a.sy * (uint32_t)(ty * a.by) + a.sz * (uint32_t)z;
    p.e = p.e0 + (uint32_t)lx + a.sy * (uint32_t)ly;
    p.step = 0;
    p.seed = 0;
    return p;
}

// Node of this thread inside tile b (tiles: BX x BY nodes of one z-plane).
template <class S> LB_DEV NodePos node_of(const StepArgs<S> &a, int b) {
    int tx = b % a.ntx;
    int r = b / a.ntx;
    int ty = r % a.nty;
    int z = r / a.nty;
    return node_in_tile(a, tx, ty, z);
}

// Global-memory accesses.  The pointers come from the argument block in
// device memory, so the compiler cannot infer their address space and would
// emit FLAT instructions, which also count against the scalar-load counter:
// every s_waitcnt on an argument load would then wait for the population
// traffic in flight.  The casts make them global loads / stores.
// MEM: 0 plain, 1 non-temporal stores, 2 non-temporal loads and stores
// (streaming hints; chosen by the autotuner).
#define LB_GLOBAL __attribute__((address_space(1)))
template <int MEM = 0, class R> LB_DEV R ldm(const R *p) {
    const LB_GLOBAL R *g = (const LB_GLOBAL R *)p;
    if constexpr (MEM >= 2) return __builtin_nontemporal_load(g);
    else return *g;
}
template <int MEM = 0, class R> LB_DEV void stm(R *p, R v) {
    LB_GLOBAL R *g = (LB_GLOBAL R *)p;
    if constexpr (MEM >= 1) __builtin_nontemporal_store(v, g);
    else *g = v;
}
// Raw buffer accesses: a descriptor built once per tile (base, size in
// bytes; 0x00020000 = 32-bit data format, the raw-buffer value for gfx9
// parts), a scalar byte offset per population slot and a per-lane byte
// offset, so a link costs no address arithmetic beyond its face correction,
// and an access past the descriptor's size reads zero (a store is dropped)
// instead of faulting.
typedef __amdgpu_buffer_rsrc_t rsrc_t;
LB_DEV rsrc_t make_rsrc(const void *base, uint64_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0,
                                             (int)(uint32_t)(bytes < 0xffffffffull ? bytes : 0xffffffffull),
                                             0x00020000);
}
// aux 2 = non-temporal
template <int AUX, class R> LB_DEV R buf_ld(rsrc_t r, uint32_t vo, uint32_t so) {
    static_assert(sizeof(R) == 2, "descriptor path: 16-bit storages");
    return __builtin_bit_cast(R, __builtin_amdgcn_raw_buffer_load_b16(r, vo, so, AUX));
}
template <int AUX, class R> LB_DEV void buf_st(R v, rsrc_t r, uint32_t vo, uint32_t so) {
    static_assert(sizeof(R) == 2, "descriptor path: 16-bit storages");
    __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(uint16_t, v), r, vo, so, AUX);
}

// A wave-uniform base plus a 32-bit per-lane byte offset: the compiler issues
// the access with the base in scalar registers and the offset as its only
// per-lane operand, so the address costs no per-lane 64-bit arithmetic.
template <class R> LB_DEV const R *at_bytes(const R *base, uint32_t boff) {
    return (const R *)((const char *)base + boff);
}
template <class R> LB_DEV R *at_bytes(R *base, uint32_t boff) { return (R *)((char *)base + boff); }

// Source coordinate of a link along one axis: periodic wrap, clamp at
// outflow / equilibrium-inlet faces; false where the link leaves through a
// wall face.
LB_DEV bool axis_src(int &s, int n, int bc_lo, int bc_hi) {
    if (s >= 0 && s < n) return true;
    const int k = s < 0 ? bc_lo : bc_hi;
    if (k == LB_BC_PERIODIC) { s = s < 0 ? s + n : s - n; return true; }
    if (k == LB_BC_OUTFLOW || k == LB_BC_INLET_EQ) { s = s < 0 ? 0 : n - 1; return true; }
    return false;
}

// Wall face a link leaves through (axes checked x, y, z), -1 if none.
template <int Q, class S> LB_DEV int wall_face(const StepArgs<S> &a, const NodePos &p, int cx, int cy, int cz) {
    int sx = p.x - cx, sy = p.y - cy, sz = p.z - cz;
    if (!axis_src(sx, a.nx, a.bc[0], a.bc[1])) return sx < 0 ? 0 : 1;
    if (!axis_src(sy, a.ny, a.bc[2], a.bc[3])) return sy < 0 ? 2 : 3;
    if (Q == 19 && !a.halo && !axis_src(sz, a.nz, a.bc[4], a.bc[5])) return sz < 0 ? 4 : 5;
    return -1;
}

// Pull all Q populations of a node into f (compute type).  Every link is
// one independent load: from the upstream neighbour (with periodic / clamp
// index corrections at the faces), from the node itself (opposite slot) for
// a half-way bounce-back link, or from the received halo plane.  Nodes next
// to walls or solids (F_SLOW) carry a link mask: bit i set = link i bounces,
// bit 31 = one of those links bounces off a moving wall.  No per-link
// branches on loaded data, so the loads of a node are all in flight at once.
// Flags and link mask of a node: one packed word, or the flag byte followed by
// the mask of F_SLOW nodes (a second, dependent load for those).  The packed
// form costs 3 more bytes per node and saves a round trip to memory: the
// autotuner picks it for small, latency-bound lattices.
template <class S> LB_DEV void node_word(const StepArgs<S> &a, const NodePos &p, uint8_t &fl, uint32_t &bm) {
    if (a.packed) {
        const uint32_t w = ldm(&a.nodew[p.id]);
        fl = (uint8_t)((w >> 24) & 0x7fu);
        bm = w & 0x80ffffffu;
    } else {
        fl = ldm(&a.flags[p.id]);
        bm = (fl & F_SLOW) ? ldm(&a.lmask[p.id]) : 0u;
    }
}

// GENERIC: always the per-link path (a node outside the tile of its lanes,
// whose element offset need not lie in the tile's descriptor range).
template <int Q, class S, int MEM = 0, bool GENERIC = false>
LB_DEV void pull_node(const StepArgs<S> &a, const NodePos &p, uint32_t bm, typename S::T *f) {
    typedef typename S::T T;
    typedef typename S::raw R;
    typedef Vel<Q> L;
    const int lox = p.x == 0 ? a.corr_lo[0] : 0, hix = p.x == a.nx - 1 ? a.corr_hi[0] : 0;
    const int loy = p.y == 0 ? a.corr_lo[1] : 0, hiy = p.y == a.ny - 1 ? a.corr_hi[1] : 0;
    int loz = 0, hiz = 0;
    if (Q == 19) { loz = p.z == 0 ? a.corr_lo[2] : 0; hiz = p.z == a.nz - 1 ? a.corr_hi[2] : 0; }
    // z-slab halo: links crossing z = -1 / z = nz read the received planes
    // (a tile lies in one z-plane, so these tests are uniform per workgroup)
    const bool glo = Q == 19 && a.halo && p.z == 0;
    const bool ghi = Q == 19 && a.halo && p.z == a.nz - 1;
    // Fast path of the 16-bit storages: no lane of the wave bounces a link,
    // and the tile's links do not cross a halo plane or wrap in z (all
    // uniform per wave).  Every link is then a neighbour load through the
    // tile's descriptor: scalar offset lso[i], per-lane offset (e - e0 + bias)
    // plus the lane's x / y face correction (one add for links with a
    // correction term, none otherwise).  These sweeps are bound by
    // instruction issue, and the per-link address arithmetic below costs
    // them about 10 %.  The 4- and 8-byte sweeps are bound by HBM and run
    // 3-5 % faster with the per-link form (measured), so they keep it.
    if constexpr (sizeof(R) == 2 && !GENERIC) {
        const bool zedge = Q == 19 && ((p.z == 0 && (a.halo || a.corr_lo[2])) ||
                                       (p.z == a.nz - 1 && (a.halo || a.corr_hi[2])));
        if (a.fast && !zedge && __builtin_amdgcn_ballot_w64(bm != 0u) == 0) {
            constexpr int ES = (int)sizeof(R);
            const int64_t b0 = (int64_t)p.e0 - a.offmax - a.bias;   // elements; may be negative
            const rsrc_t rs = make_rsrc((const char *)a.A + b0 * ES, ((int64_t)Q * a.N - b0) * ES);
            const uint32_t vb = (p.e - p.e0 + (uint32_t)a.bias) * (uint32_t)ES;
            const int loxb = lox * ES, hixb = hix * ES, loyb = loy * ES, hiyb = hiy * ES;
            static_for<0, Q>([&](auto I) {
                constexpr int i = decltype(I)::value;
                constexpr int cx = L::cx(i), cy = L::cy(i);
                const int cb = (cx > 0 ? loxb : (cx < 0 ? hixb : 0)) + (cy > 0 ? loyb : (cy < 0 ? hiyb : 0));
                f[i] = S::ld(buf_ld<MEM >= 2 ? 2 : 0, R>(rs, vb + (uint32_t)cb, a.lso[i]), weight<Q, float>(L::wc(i)));
            });
            return;
        }
    }
    static_for<0, Q>([&](auto I) {
        constexpr int i = decltype(I)::value;
        constexpr int cx = L::cx(i), cy = L::cy(i), cz = L::cz(i);
        constexpr int j = L::opp(i);
        const bool bounce = (bm >> i) & 1u;
        if constexpr (Q == 19 && cz != 0) {
            if ((cz > 0 && glo) || (cz < 0 && ghi)) {
                if (bounce) {
                    f[i] = S::ld(ldm(&a.A[(size_t)j * a.ps + p.e]), weight<Q, float>(L::wc(i)));
                } else {
                    int sx = p.x - cx, sy = p.y - cy;
                    axis_src(sx, a.nx, a.bc[0], a.bc[1]);
                    axis_src(sy, a.ny, a.bc[2], a.bc[3]);
                    const uint32_t sxy = (uint32_t)sx + (uint32_t)a.nx * (uint32_t)sy;
                    f[i] = S::ld(ldm(&a.recv[(size_t)((cz > 0 ? 5 : 0) + zslot(i)) * a.P + sxy]),
                                 weight<Q, float>(L::wc(i)));
                }
                return;
            }
        }
        const int corr = (cx > 0 ? lox : (cx < 0 ? hix : 0)) + (cy > 0 ? loy : (cy < 0 ? hiy : 0)) +
                         (cz > 0 ? loz : (cz < 0 ? hiz : 0));
        const int off = cx + cy * (int)a.sy + cz * (int)a.sz;
        const uint32_t src = (uint32_t)((int)p.e - off + corr);
        if constexpr (sizeof(R) == 2) {
            // element indices stay below 2^31 (checked at creation)
            const uint32_t idx = bounce ? (uint32_t)j * a.ps + p.e : (uint32_t)i * a.ps + src;
            f[i] = S::ld(ldm<MEM>(a.A + idx), weight<Q, float>(L::wc(i)));
        } else {
            const R *ptr = bounce ? &a.A[(size_t)j * a.ps + p.e] : &a.A[(size_t)i * a.ps + src];
            f[i] = S::ld(ldm<MEM>(ptr), weight<Q, float>(L::wc(i)));
        }
    });
    if (bm >> 31) {
        // bounced links that left through a moving wall gain 6 w_i (c_i . u_w)
        static_for<1, Q>([&](auto I) {
            constexpr int i = decltype(I)::value;
            if ((bm >> i) & 1u) {
                const int wf = wall_face<Q>(a, p, L::cx(i), L::cy(i), L::cz(i));
                if (wf >= 0 && a.bc[wf] == LB_BC_MOVING_WALL) {
                    T cu = (T)L::cx(i) * (T)a.bcu[wf][0] + (T)L::pcy(i) * (T)a.bcu[wf][1] +
                           (T)L::cz(i) * (T)a.bcu[wf][2];
                    f[i] = f[i] + (T(6) * weight<Q, T>(L::wc(i))) * cu;
                }
            }
        });
    }
}

template <int Q, class T> LB_DEV void moments(const T *f, T &rho, T &ux, T &uy, T &uz) {
    if constexpr (Q == 9) { mom9(f, rho, ux, uy); uz = T(0); }
    else mom19(f, rho, ux, uy, uz);
}

// Stored value of population slot `slot` of node p: nearest-even, or
// stochastic rounding for the 16-bit deviation storages.
template <class S> constexpr bool kRoundable = std::is_same<S, StF16>::value || std::is_same<S, StBF16>::value;

template <class S, bool SR>
LB_DEV typename S::raw store_val_m(const NodePos &p, typename S::T v, float w, int slot) {
    if constexpr (SR && std::is_same<S, StF16>::value) return f32_to_f16_sr(v, sr_mix(p.seed, slot));
    else if constexpr (SR && std::is_same<S, StBF16>::value) return f32_to_bf16_sr(v, sr_mix_hi(p.seed, slot));
    else return S::st(v, w);
}
template <class S>
LB_DEV typename S::raw store_val(const StepArgs<S> &a, const NodePos &p, typename S::T v, float w, int slot) {
    if constexpr (kRoundable<S>) {
        if (a.round_mode) return store_val_m<S, true>(p, v, w, slot);
    }
    return store_val_m<S, false>(p, v, w, slot);
}

// REV (full-way bounce-back at a solid node): slot opp(i) <- f_i.  A
// compile-time flag: a runtime choice between f[i] and f[opp(i)] becomes a
// dynamically indexed private array, i.e. per-lane scratch memory.
// Stores address slot s of the node as the uniform base B + s*ps + e0 plus
// the lane's byte offset (e - e0)*sizeof, the same for every slot.
template <int Q, class S, int MEM = 0, bool REV = false>
LB_DEV void store_node(const StepArgs<S> &a, const NodePos &p, const typename S::T *f) {
    typedef Vel<Q> L;
    typedef typename S::raw R;
    if constexpr (sizeof(R) == 2) {
        // 16-bit storages: the tile's store descriptor (B + e0*2), scalar
        // offset sso[slot] and the per-lane byte offset (e - e0)*2, the same
        // for every slot (or, past 32-bit offsets, a scalar pointer per slot
        // plus that offset).  Uniform branches around all Q stores keep them
        // in one basic block.
        const uint32_t vs = (p.e - p.e0) * (uint32_t)sizeof(R);
        auto put = [&](auto sr, auto buf) {
            constexpr bool BUF = decltype(buf)::value;
            rsrc_t ws;
            if constexpr (BUF) ws = make_rsrc(a.B + p.e0, ((uint64_t)Q * a.N - p.e0) * sizeof(R));
            static_for<0, Q>([&](auto I) {
                constexpr int i = decltype(I)::value;
                constexpr int slot = REV ? L::opp(i) : i;
                auto v = store_val_m<S, decltype(sr)::value>(p, f[i], weight<Q, float>(L::wc(slot)), slot);
                if constexpr (BUF) buf_st<MEM >= 1 ? 2 : 0>(v, ws, vs, a.sso[slot]);
                else stm<MEM>(at_bytes(a.B + ((size_t)slot * a.ps + p.e0), vs), v);
            });
        };
        if (a.fast) {
            if (a.round_mode) put(std::true_type{}, std::true_type{});
            else put(std::false_type{}, std::true_type{});
        } else {
            if (a.round_mode) put(std::true_type{}, std::false_type{});
            else put(std::false_type{}, std::false_type{});
        }
    } else {
        static_for<0, Q>([&](auto I) {
            constexpr int i = decltype(I)::value;
            constexpr int slot = REV ? L::opp(i) : i;
            auto v = store_val<S>(a, p, f[i], weight<Q, float>(L::wc(slot)), slot);
            stm<MEM>(&a.B[(size_t)slot * a.ps + p.e], v);
        });
    }
    if constexpr (Q == 19) {
        if (a.halo && (p.z == a.nz - 1 || p.z == 0)) {
            // pack (get_border): the +z populations of the top plane and the -z
            // populations of the bottom plane, the values just written to B.
            const uint32_t xy = (uint32_t)p.x + (uint32_t)a.nx * (uint32_t)p.y;
            static_for<0, 19>([&](auto I) {
                constexpr int i = decltype(I)::value;
                constexpr int cz = L::cz(i);
                if constexpr (cz != 0) {
                    if ((cz > 0 && p.z == a.nz - 1) || (cz < 0 && p.z == 0)) {
                        auto v = store_val<S>(a, p, f[REV ? L::opp(i) : i], weight<Q, float>(L::wc(i)), i);
                        stm(&a.send[(size_t)((cz > 0 ? 5 : 0) + zslot(i)) * a.P + xy], v);
                    }
                }
            });
        }
    }
}

template <int Q, class T> LB_DEV void feq(T *f, T rho, T ux, T uy, T uz) {
    if constexpr (Q == 9) feq9<T>(f, rho, ux, uy, weight<9, T>(0), weight<9, T>(1), weight<9, T>(2));
    else feq19<T>(f, rho, ux, uy, uz, weight<19, T>(0), weight<19, T>(1), weight<19, T>(2));
}

// BGK relaxation towards feq (the reference expressions, collision.cu:312-330
// and :504-541, evaluated first as equilibria, which rounds identically).
// With rest_closure the rest equilibrium is rho minus the moving ones, so
// the collision conserves mass up to rounding noise (the float lattice
// weights alone sum to 1 + 3e-8 for D3Q19).
template <int Q, class T>
LB_DEV void collide(T *f, T rho, T ux, T uy, T uz, T ti, bool rest_closure) {
    T eq[Q];
    feq<Q, T>(eq, rho, ux, uy, uz);
    if (rest_closure) {
        T s = eq[1];
        static_for<2, Q>([&](auto I) { s = s + eq[decltype(I)::value]; });
        eq[0] = rho - s;
    }
    static_for<0, Q>([&](auto I) {
        constexpr int i = decltype(I)::value;
        f[i] = f[i] - ti * (f[i] - eq[i]);
    });
}

// Deviation form: moments, equilibrium deviations and relaxation.
template <int Q, class T> LB_DEV void moments_dev(const T *s, T &drho, T &ux, T &uy, T &uz) {
    T inv;
    if constexpr (Q == 9) {
        drho = s[0] + s[1] + s[2] + s[3] + s[4] + s[5] + s[6] + s[7] + s[8];
        inv = T(1) / (T(1) + drho);
        ux = (s[1] - s[3] + s[5] - s[6] - s[7] + s[8]) * inv;
        uy = (s[2] - s[4] + s[5] + s[6] - s[7] - s[8]) * inv;
        uz = T(0);
    } else {
        drho = s[0] + s[1] + s[2] + s[3] + s[4] + s[5] + s[6] + s[7] + s[8] + s[9] + s[10] + s[11] +
               s[12] + s[13] + s[14] + s[15] + s[16] + s[17] + s[18];
        inv = T(1) / (T(1) + drho);
        ux = (s[1] + s[7] + s[8] + s[9] + s[10] - s[2] - s[11] - s[12] - s[13] - s[14]) * inv;
        uy = (s[3] + s[7] + s[11] + s[15] + s[16] - s[4] - s[8] - s[12] - s[17] - s[18]) * inv;
        uz = (s[5] + s[9] + s[13] + s[15] + s[17] - s[6] - s[10] - s[14] - s[16] - s[18]) * inv;
    }
}

template <int Q, class T> LB_DEV void feq_dev(T *seq, T drho, T ux, T uy, T uz) {
    T g[Q];
    if constexpr (Q == 9) geq9<T>(g, ux, uy);
    else geq19<T>(g, ux, uy, uz);
    const T rho = T(1) + drho;
    static_for<0, Q>([&](auto I) {
        constexpr int i = decltype(I)::value;
        seq[i] = weight<Q, T>(Vel<Q>::wc(i)) * fma_t(rho, g[i], drho);
    });
}

template <int Q, class T>
LB_DEV void collide_dev(T *s, T drho, T ux, T uy, T uz, T ti, bool rest_closure) {
    T eq[Q];
    feq_dev<Q, T>(eq, drho, ux, uy, uz);
    if (rest_closure) {
        T sum = eq[1];
        static_for<2, Q>([&](auto I) { sum = sum + eq[decltype(I)::value]; });
        eq[0] = drho - sum;
    }
    static_for<0, Q>([&](auto I) {
        constexpr int i = decltype(I)::value;
        s[i] = fma_t(ti, eq[i] - s[i], s[i]);
    });
}

// ---------------------------------------------------------------------------
// The fused step.
// Zou-He inlet (x = 0) / outlet (x = nx-1) in the form matching the storage.
template <int Q, class S> LB_DEV void apply_zou_he(const StepArgs<S> &a, const NodePos &p, typename S::T *f) {
    typedef typename S::T T;
    if (!a.zou_he) return;
    // Two independent ifs, not if / else: with if / else the compiler sinks
    // the two branches' stores (to different f[] slots) into one store with a
    // selected address, which puts f in per-lane scratch memory.
    if (p.x == 0) {
        if constexpr (S::DEV) { if constexpr (Q == 9) zouhe9_in_dev<T>(f, (T)a.zh_ux); else zouhe19_in_dev<T>(f, (T)a.zh_ux); }
        else { if constexpr (Q == 9) zouhe9_in(f, a.zh_ux); else zouhe19_in(f, a.zh_ux); }
    }
    if (p.x == a.nx - 1) {
        if constexpr (S::DEV) { if constexpr (Q == 9) zouhe9_out_dev<T>(f, (T)a.zh_rho); else zouhe19_out_dev<T>(f, (T)a.zh_rho); }
        else { if constexpr (Q == 9) zouhe9_out(f, a.zh_rho); else zouhe19_out(f, a.zh_rho); }
    }
}

template <class S> LB_DEV void store_macros(const StepArgs<S> &a, uint32_t id, float r, float ux, float uy, float uz) {
    stm(&a.macros[id], r);
    stm(&a.macros[a.N + id], ux);
    stm(&a.macros[2 * a.N + id], uy);
    stm(&a.macros[3 * a.N + id], uz);
}

// ---------------------------------------------------------------------------
// The body of the fused step for one node.  With SC set, the Shan-Chen force
// reads the neighbours' psi (written by the density pass) straight from global
// memory: the reads hit L1 / L2, and the sweep needs no LDS tile and no
// workgroup barrier.
template <int Q, class S, bool SC, bool MACROS, int MEM = 0>
LB_DEV void step_body(const StepArgs<S> &a, const NodePos &p) {
    typedef typename S::T T;
    typedef Vel<Q> L;
    if (!p.valid) return;
    uint8_t fl;
    uint32_t bm;
    node_word(a, p, fl, bm);
    T f[Q];
    if (fl & F_SOLID) {
        if (a.bb_full) {
            pull_node<Q, S, MEM>(a, p, bm, f);
            store_node<Q, S, MEM, true>(a, p, f);
        }
        if (MACROS) {
            store_macros(a, p.id, 0.f, 0.f, 0.f, 0.f);
        }
        return;
    }
    // Inlet and fluid nodes end in the same store, so a workgroup that holds
    // both issues each population store once.
    T rho, ux, uy, uz;
    if (fl & F_INLET) {
        const int face = fl >> 4;
        rho = (T)a.bcrho[face]; ux = (T)a.bcu[face][0]; uy = (T)a.bcu[face][1]; uz = (T)a.bcu[face][2];
        if constexpr (S::DEV) feq_dev<Q, T>(f, rho - T(1), ux, uy, uz);
        else feq<Q, T>(f, rho, ux, uy, uz);
    } else {
        pull_node<Q, S, MEM>(a, p, bm, f);
        apply_zou_he<Q, S>(a, p, f);
        T drho = T(0);
        if constexpr (S::DEV) {
            moments_dev<Q, T>(f, drho, ux, uy, uz);
            rho = T(1) + drho;
        } else {
            moments<Q, T>(f, rho, ux, uy, uz);
        }
        const T ti = (T)a.ti;
        T ex = ux, ey = uy, ez = uz;
        if (a.force[0] != 0.0f) ex = ux + (T)a.force[0] / (rho * ti);
        if (a.force[1] != 0.0f) ey = uy + (T)a.force[1] / (rho * ti);
        if (a.force[2] != 0.0f) ez = uz + (T)a.force[2] / (rho * ti);
        if constexpr (SC) {
            // psi index of neighbour x + c_i: the node's index plus c_i's
            // offset, plus a face correction on the lanes at a face the link
            // crosses (uniform per face): across a periodic (or halo) face the
            // neighbour wraps to the opposite plane, across any other face it
            // is clamped to the node's own plane.  Six corrections per lane
            // instead of a wrap test per link.
            const int sx = 1, sy = a.nx, sz = a.nx * a.ny;
            const bool zw_lo = Q == 19 && (a.halo || a.bc[4] == LB_BC_PERIODIC);
            const bool zw_hi = Q == 19 && (a.halo || a.bc[5] == LB_BC_PERIODIC);
            const int xm = p.x == 0 ? (a.bc[0] == LB_BC_PERIODIC ? a.nx * sx : sx) : 0;
            const int xp = p.x == a.nx - 1 ? (a.bc[1] == LB_BC_PERIODIC ? -a.nx * sx : -sx) : 0;
            const int ym = p.y == 0 ? (a.bc[2] == LB_BC_PERIODIC ? a.ny * sy : sy) : 0;
            const int yp = p.y == a.ny - 1 ? (a.bc[3] == LB_BC_PERIODIC ? -a.ny * sy : -sy) : 0;
            int zm = 0, zp = 0;
            if constexpr (Q == 19) {
                zm = p.z == 0 ? (zw_lo ? a.nz * sz : sz) : 0;
                zp = p.z == a.nz - 1 ? (zw_hi ? -a.nz * sz : -sz) : 0;
            }
            T IF[3] = {T(0), T(0), T(0)};
            static_for<1, Q>([&](auto I) {
                constexpr int i = decltype(I)::value;
                constexpr int cx = L::cx(i), cy = L::cy(i), cz = L::cz(i);
                const int corr = (cx > 0 ? xp : (cx < 0 ? xm : 0)) + (cy > 0 ? yp : (cy < 0 ? ym : 0)) +
                                 (cz > 0 ? zp : (cz < 0 ? zm : 0));
                const uint32_t q = p.id + (uint32_t)(cx * sx + cy * sy + cz * sz + corr);
                T wp = weight<Q, T>(L::wc(i)) * (T)ldm(&a.psi[q]);
                if constexpr (cx > 0) IF[0] = IF[0] + wp; else if constexpr (cx < 0) IF[0] = IF[0] - wp;
                if constexpr (L::pcy(i) > 0) IF[1] = IF[1] + wp; else if constexpr (L::pcy(i) < 0) IF[1] = IF[1] - wp;
                if constexpr (cz > 0) IF[2] = IF[2] + wp; else if constexpr (cz < 0) IF[2] = IF[2] - wp;
            });
            T gp = (T)(-a.scG) * (T)ldm(&a.psi[p.id]);
            const T kf = gp / (ti * rho);   // one division for the three components
            ex = fma_t(kf, IF[0], ex);
            ey = fma_t(kf, IF[1], ey);
            if constexpr (Q == 19) ez = fma_t(kf, IF[2], ez);
        }
        if constexpr (S::DEV) collide_dev<Q, T>(f, drho, ex, ey, ez, ti, a.rest_closure != 0);
        else collide<Q, T>(f, rho, ex, ey, ez, ti, a.rest_closure != 0);
    }
    if (MACROS) store_macros(a, p.id, (float)rho, (float)ux, (float)uy, (float)uz);
    store_node<Q, S, MEM>(a, p, f);
}

// Shan-Chen psi(rho) of one node's post-streaming density, with the same
// pull / boundary handling (and summation order) as k_step.
template <int Q, class S, int MEM = 0, bool GENERIC = false>
LB_DEV typename S::T psi_of_node(const StepArgs<S> &a, const NodePos &p) {
    typedef typename S::T T;
    uint8_t fl;
    uint32_t bm;
    node_word(a, p, fl, bm);
    if (fl & F_SOLID) return (T)a.scpsi_solid;
    T rho;
    if (fl & F_INLET) {
        rho = (T)a.bcrho[fl >> 4];
    } else {
        T f[Q];
        pull_node<Q, S, MEM, GENERIC>(a, p, bm, f);
        apply_zou_he<Q, S>(a, p, f);
        T ux, uy, uz;
        if constexpr (S::DEV) {
            T drho;
            moments_dev<Q, T>(f, drho, ux, uy, uz);
            rho = T(1) + drho;
        } else {
            moments<Q, T>(f, rho, ux, uy, uz);
        }
    }
    return sc_psi<T>(rho, a.scrho0);
}

// Shan-Chen density pre-pass: psi of every node into a.psi.
template <int Q, class S, int MEM = 0>
LB_DEV void density_body(const StepArgs<S> &a, const NodePos &p) {
    if (!p.valid) return;
    stm(&a.psi[p.id], psi_of_node<Q, S, MEM>(a, p));
}

// ---------------------------------------------------------------------------
// One-pass Shan-Chen step (D3Q19): a workgroup owns a bx x by column of
// nodes over zc consecutive z-planes and marches up through them.  At plane
// z it pulls the populations of plane z + 1 into registers (the next
// iteration sweeps them), turns their density into psi and writes it to a
// ring of four psi planes in LDS, together with the psi of the one-node frame
// around the tile; after one workgroup barrier it sweeps plane z with the
// psi of planes z - 1, z, z + 1 read from LDS.  Population loads per node:
// one pull of its own plus the frame's share (2 (bx + 2) + 2 by nodes per
// bx by), against two pulls (density pass + sweep) and 19 psi loads in the
// two-pass form.  A frame or out-of-lattice position holds the psi of its
// source node (periodic faces wrap, other faces clamp: the two-pass sweep's
// neighbour index rule), and every psi and every swept node runs the same
// functions as the density pass and the sweep, so the result is bit-identical
// to them.  Four planes in the ring let one barrier per plane suffice: the
// plane written at iteration z + 1 was last read at iteration z - 1, which
// every wave has finished once it passed the barrier of iteration z.
constexpr int kZThreads = 512;  // default workgroup (tiles of 256, 512 or 1,024 threads)
constexpr int kZSlot = 1300;     // largest (bx + 2)(by + 2) of the tile shapes (128 x 8)

// Source coordinate of the psi at position v along an axis of n nodes.
LB_DEV int psi_src(int v, int n, int bc_lo, int bc_hi) {
    if (v < 0) v = bc_lo == LB_BC_PERIODIC ? v + n : 0;
    else if (v >= n) v = bc_hi == LB_BC_PERIODIC ? v - n : n - 1;
    return v < 0 ? 0 : (v >= n ? n - 1 : v);
}

// psi of the source node of lattice position (x, y, z), any position.
template <int Q, class S, int MEM>
LB_DEV typename S::T psi_at(const StepArgs<S> &a, int x, int y, int z) {
    NodePos p;
    p.x = psi_src(x, a.nx, a.bc[0], a.bc[1]);
    p.y = psi_src(y, a.ny, a.bc[2], a.bc[3]);
    p.z = Q == 19 ? psi_src(z, a.nz, a.bc[4], a.bc[5]) : 0;
    p.id = (uint32_t)p.x + (uint32_t)a.nx * ((uint32_t)p.y + (uint32_t)a.ny * (uint32_t)p.z);
    p.e = (uint32_t)p.x + a.sy * (uint32_t)p.y + a.sz * (uint32_t)p.z;
    p.e0 = p.e;
    p.valid = true;
    p.step = 0;
    p.seed = 0;
    return psi_of_node<Q, S, MEM, true>(a, p);
}

// Pulled populations of a node and their moments (m0: rho, or the density
// deviation of the deviation storages), kept from the density phase for the
// sweep of the next iteration.
template <int Q, class S> struct ZState {
    typename S::T f[Q];
    typename S::T m0, ux, uy, uz;
    uint8_t fl;
    uint32_t bm;
};

// Density phase of a node of the lattice: node word, pull and moments as in
// step_body, and the node's psi as psi_of_node returns it.
template <int Q, class S, int MEM>
LB_DEV typename S::T zm_load(const StepArgs<S> &a, const NodePos &p, ZState<Q, S> &st) {
    typedef typename S::T T;
    node_word(a, p, st.fl, st.bm);
    if (st.fl & F_SOLID) {
        if (a.bb_full) pull_node<Q, S, MEM>(a, p, st.bm, st.f);
        return (T)a.scpsi_solid;
    }
    if (st.fl & F_INLET) return sc_psi<T>((T)a.bcrho[st.fl >> 4], a.scrho0);
    pull_node<Q, S, MEM>(a, p, st.bm, st.f);
    apply_zou_he<Q, S>(a, p, st.f);
    T rho;
    if constexpr (S::DEV) {
        moments_dev<Q, T>(st.f, st.m0, st.ux, st.uy, st.uz);
        rho = T(1) + st.m0;
    } else {
        moments<Q, T>(st.f, st.m0, st.ux, st.uy, st.uz);
        rho = st.m0;
    }
    return sc_psi<T>(rho, a.scrho0);
}

// Sweep phase: step_body's arithmetic from the kept state, the psi of the
// node and its neighbours from the LDS planes pz[cz + 1] (node at c, rows W
// apart).
template <int Q, class S, bool MACROS, int MEM>
LB_DEV void zm_sweep(const StepArgs<S> &a, const NodePos &p, ZState<Q, S> &st, const typename S::T *pzm,
                     const typename S::T *pz0, const typename S::T *pzp, int c, int W) {
    typedef typename S::T T;
    typedef Vel<Q> L;
    if (st.fl & F_SOLID) {
        if (a.bb_full) store_node<Q, S, MEM, true>(a, p, st.f);
        if (MACROS) store_macros(a, p.id, 0.f, 0.f, 0.f, 0.f);
        return;
    }
    T rho, ux, uy, uz;
    if (st.fl & F_INLET) {
        const int face = st.fl >> 4;
        rho = (T)a.bcrho[face]; ux = (T)a.bcu[face][0]; uy = (T)a.bcu[face][1]; uz = (T)a.bcu[face][2];
        if constexpr (S::DEV) feq_dev<Q, T>(st.f, rho - T(1), ux, uy, uz);
        else feq<Q, T>(st.f, rho, ux, uy, uz);
    } else {
        T drho = T(0);
        if constexpr (S::DEV) {
            drho = st.m0;
            rho = T(1) + drho;
        } else {
            rho = st.m0;
        }
        ux = st.ux; uy = st.uy; uz = st.uz;
        const T ti = (T)a.ti;
        T ex = ux, ey = uy, ez = uz;
        if (a.force[0] != 0.0f) ex = ux + (T)a.force[0] / (rho * ti);
        if (a.force[1] != 0.0f) ey = uy + (T)a.force[1] / (rho * ti);
        if (a.force[2] != 0.0f) ez = uz + (T)a.force[2] / (rho * ti);
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
        if constexpr (Q == 19) ez = fma_t(kf, IF[2], ez);
        if constexpr (S::DEV) collide_dev<Q, T>(st.f, drho, ex, ey, ez, ti, a.rest_closure != 0);
        else collide<Q, T>(st.f, rho, ex, ey, ez, ti, a.rest_closure != 0);
    }
    if (MACROS) store_macros(a, p.id, (float)rho, (float)ux, (float)uy, (float)uz);
    store_node<Q, S, MEM>(a, p, st.f);
}

// Storages with the one-pass sweep, and its launch (lb_scz.hip): keep = the
// pulled populations stay in registers across the barrier (else pulled again
// for the sweep); threads = bx x by, 256, 512 or 1,024; args = the device
// argument block.
bool scz_storage(int storage);
int launch_step_scz(int storage, bool macros, bool keep, int threads, unsigned nblocks, hipStream_t stream,
                    const void *args, uint32_t dstep);

// The row-slab one-pass sweep of fully periodic lattices (lb_scr.hip): a
// workgroup of nw = 8, 12 or 16 waves over tiles of `by` whole rows; the
// largest tile a storage admits, its LDS bytes, and the launch.
int scr_max_rows(int nw, int nx, int storage);
size_t scr_lds_bytes(int nx, int by);
int launch_step_scr(int storage, bool macros, int nw, int by, int nx, unsigned nblocks, hipStream_t stream,
                    const void *args, uint32_t dstep);

// ---------------------------------------------------------------------------
// Node flags on the host: solid, inlet plane (+ face in the high nibble), and
// F_SLOW with a link mask (lm) for nodes with a link that bounces back: bit i
// when link i leaves through a wall face (axes checked x, y, z; periodic
// faces wrap, outflow / inlet faces clamp, halo z faces pass to the received
// planes) or, without full-way bounce-back, when its upstream node is solid;
// bit 31 when one of those links leaves through a moving wall.
inline void build_flags(const lb_config &g, const std::vector<uint8_t> &solid, std::vector<uint8_t> &fl,
                        std::vector<uint32_t> &lm) {
    const int Q = g.lattice;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const bool halo = g.halo_z && Q == 19;
    fl.assign(solid.size(), 0);
    lm.assign(solid.size(), 0u);
    static const int C9[9][3] = {{0, 0, 0}, {1, 0, 0}, {0, -1, 0}, {-1, 0, 0}, {0, 1, 0},
                                 {1, -1, 0}, {-1, -1, 0}, {-1, 1, 0}, {1, 1, 0}};
    static const int C19[19][3] = {{0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0},
                                   {0, 0, 1}, {0, 0, -1}, {1, 1, 0}, {1, -1, 0}, {1, 0, 1},
                                   {1, 0, -1}, {-1, 1, 0}, {-1, -1, 0}, {-1, 0, 1}, {-1, 0, -1},
                                   {0, 1, 1}, {0, 1, -1}, {0, -1, 1}, {0, -1, -1}};
    const int (*C)[3] = Q == 9 ? C9 : C19;
    const int n[3] = {nx, ny, nz};
    const bool any_solid = std::find(solid.begin(), solid.end(), 1) != solid.end();
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                size_t id = x + (size_t)nx * (y + (size_t)ny * z);
                uint8_t f = 0;
                if (solid[id]) f |= F_SOLID;
                const int p[3] = {x, y, z};
                for (int a = 0; a < 3; a++) {
                    if (a == 2 && Q == 9) break;
                    if (g.bc[2 * a] == LB_BC_INLET_EQ && p[a] == 0) { f |= F_INLET | (uint8_t)((2 * a) << 4); break; }
                    if (g.bc[2 * a + 1] == LB_BC_INLET_EQ && p[a] == n[a] - 1) { f |= F_INLET | (uint8_t)((2 * a + 1) << 4); break; }
                }
                bool on_face = x == 0 || x == nx - 1 || y == 0 || y == ny - 1 || (Q == 19 && (z == 0 || z == nz - 1));
                if (!on_face && !any_solid) { fl[id] = f; continue; }
                uint32_t m = 0;
                for (int i = 1; i < Q; i++) {
                    int s[3] = {x - C[i][0], y - C[i][1], z - C[i][2]};
                    int wf = -1;
                    for (int a = 0; a < 3 && wf < 0; a++) {
                        if (a == 2 && Q == 9) break;
                        if (s[a] >= 0 && s[a] < n[a]) continue;
                        const int face = 2 * a + (s[a] >= n[a] ? 1 : 0);
                        const int k = (a == 2 && halo) ? LB_BC_PERIODIC : g.bc[face];
                        if (k == LB_BC_PERIODIC) s[a] = s[a] < 0 ? s[a] + n[a] : s[a] - n[a];
                        else if (k == LB_BC_OUTFLOW || k == LB_BC_INLET_EQ) s[a] = s[a] < 0 ? 0 : n[a] - 1;
                        else wf = face;
                    }
                    bool bnc = wf >= 0;
                    if (!bnc && !g.bb_full && solid[s[0] + (size_t)nx * (s[1] + (size_t)ny * s[2])]) bnc = true;
                    if (bnc) {
                        m |= 1u << i;
                        if (wf >= 0 && g.bc[wf] == LB_BC_MOVING_WALL) m |= 1u << 31;
                    }
                }
                if (m) { f |= F_SLOW; lm[id] = m; }
                fl[id] = f;
            }
}


// Fast-path index corrections across each face (element units of the
// population layout): periodic wrap or clamp; 0 where the face needs the
// per-link path.
// Scalar offsets and usability of the descriptor path for Q populations and
// tiles of bx x by nodes: scalar + per-lane byte offset must stay below 2^32
// for every load, lso + (e - e0 + bias + correction) * esz, and store,
// sso + (e - e0) * esz.
template <int Q, class S> void fast_path_args(StepArgs<S> &a, int bx, int by) {
    typedef Vel<Q> L;
    const uint64_t es = sizeof(typename S::raw);
    a.bias = -(a.corr_hi[0] + a.corr_hi[1]);
    int64_t off[19], offmax = 0;
    for (int i = 0; i < Q; i++) {
        off[i] = (int64_t)L::cx(i) + (int64_t)L::cy(i) * a.sy + (int64_t)L::cz(i) * a.sz;
        offmax = std::max(offmax, off[i]);
    }
    a.offmax = (int)offmax;
    uint64_t lmax = 0, smax = 0;
    for (int i = 0; i < 19; i++) {
        const uint64_t l = i < Q ? ((uint64_t)i * a.ps + (uint64_t)(offmax - off[i])) * es : 0;
        const uint64_t s = i < Q ? (uint64_t)i * a.ps * es : 0;
        a.lso[i] = (uint32_t)l;
        a.sso[i] = (uint32_t)s;
        lmax = std::max(lmax, l);
        smax = std::max(smax, s);
    }
    const uint64_t lane = (uint64_t)bx + (uint64_t)by * a.sy;   // > e - e0
    const uint64_t lv = (lane + (uint64_t)(a.corr_lo[0] + a.corr_lo[1]) + (uint64_t)a.bias) * es;
    a.fast = lmax + lv < (1ull << 32) && smax + lane * es < (1ull << 32) && offmax < (1ll << 31);
}

inline void face_corrections(const lb_config &g, const PopLayout &L, int lo[3], int hi[3]) {
    const int n[3] = {g.nx, g.ny, g.nz};
    const int stride[3] = {1, (int)L.sy, (int)L.sz};
    for (int a = 0; a < 3; a++) {
        auto corr = [&](int face, int sign) {
            int k = g.bc[face];
            if (k == LB_BC_PERIODIC) return sign * n[a] * stride[a];
            if (k == LB_BC_OUTFLOW || k == LB_BC_INLET_EQ) return sign * stride[a];
            return 0;
        };
        lo[a] = corr(2 * a, +1);
        hi[a] = corr(2 * a + 1, -1);
    }
}

}  // namespace lb
