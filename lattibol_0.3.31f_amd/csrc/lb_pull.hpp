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
    float *macros;          // 4*N: rho, ux, uy, uz
    T *psi;                 // Shan-Chen pseudo-potential
    raw *send;              // [send_lo (5 planes, -z pops of z=0) | send_hi (+z pops of z=nz-1)]
    const raw *recv;        // [recv_hi (ghost z=nz) | recv_lo (ghost z=-1)]
    uint32_t N, nxy, P;     // nodes, nodes per z-plane (= halo plane size)
    int nx, ny, nz;
    float ti;
    int bc[6];
    float bcu[6][3];
    float bcrho[6];
    int corr_lo[3], corr_hi[3];   // fast-path index corrections at faces
    float force[3];
    float scG, scrho0, scpsi_solid;
    int bb_full, zou_he, halo;
    int rest_closure;             // feq_0 = rho - sum_{i>0} feq_i (exact mass balance)
    float zh_ux, zh_rho;
    int bx_shift, by, ntx, nty, nblocks, xcd;
};

// Workgroups are dispatched round-robin over the 8 XCDs; this bijection
// gives each XCD a contiguous range of tiles so z/y neighbours share an L2.
LB_DEV int xcd_remap(int b, int nb) {
    int q = nb >> 3, r = nb & 7;
    int xcd = b & 7, local = b >> 3;
    return xcd * q + min(xcd, r) + local;
}

struct NodePos { int x, y, z; uint32_t id; bool valid; };

template <class S> LB_DEV int tile_of_block(const StepArgs<S> &a) {
    int b = blockIdx.x;
    if (a.xcd) b = xcd_remap(b, a.nblocks);
    return b;
}

// Node of this thread inside tile b (tiles: BX x BY nodes of one z-plane).
template <class S> LB_DEV NodePos node_of(const StepArgs<S> &a, int b) {
    int tx = b % a.ntx;
    int r = b / a.ntx;
    int ty = r % a.nty;
    int z = r / a.nty;
    int lx = threadIdx.x & ((1 << a.bx_shift) - 1);
    int ly = threadIdx.x >> a.bx_shift;
    NodePos p;
    p.x = (tx << a.bx_shift) + lx;
    p.y = ty * a.by + ly;
    p.z = z;
    p.valid = p.x < a.nx && p.y < a.ny;
    p.id = (uint32_t)p.x + (uint32_t)a.nx * ((uint32_t)p.y + (uint32_t)a.ny * (uint32_t)z);
    return p;
}

// Population loads / stores: MEM 0 plain, 1 non-temporal stores, 2 non-
// temporal loads and stores (streaming hints; chosen by the autotuner).
template <int MEM, class R> LB_DEV R ldm(const R *p) {
    if constexpr (MEM >= 2) return __builtin_nontemporal_load(p);
    else return *p;
}
template <int MEM, class R> LB_DEV void stm(R *p, R v) {
    if constexpr (MEM >= 1) __builtin_nontemporal_store(v, p);
    else *p = v;
}

// Pull all Q populations of a fluid node into f (compute type).
template <int Q, class S, int MEM = 0>
LB_DEV void pull_node(const StepArgs<S> &a, const NodePos &p, uint8_t fl, typename S::T *f) {
    typedef typename S::T T;
    typedef Vel<Q> L;
    if (!(fl & F_SLOW)) {
        // Interior or periodic/clamped faces: one index correction per axis.
        int lox = p.x == 0 ? a.corr_lo[0] : 0, hix = p.x == a.nx - 1 ? a.corr_hi[0] : 0;
        int loy = p.y == 0 ? a.corr_lo[1] : 0, hiy = p.y == a.ny - 1 ? a.corr_hi[1] : 0;
        int loz = 0, hiz = 0;
        if (Q == 19) { loz = p.z == 0 ? a.corr_lo[2] : 0; hiz = p.z == a.nz - 1 ? a.corr_hi[2] : 0; }
        static_for<0, Q>([&](auto I) {
            constexpr int i = decltype(I)::value;
            constexpr int cx = L::cx(i), cy = L::cy(i), cz = L::cz(i);
            int corr = (cx > 0 ? lox : (cx < 0 ? hix : 0)) + (cy > 0 ? loy : (cy < 0 ? hiy : 0)) +
                       (cz > 0 ? loz : (cz < 0 ? hiz : 0));
            int off = cx + cy * a.nx + cz * (int)a.nxy;
            uint32_t src = (uint32_t)((int)p.id - off + corr);
            f[i] = S::ld(ldm<MEM>(&a.A[(size_t)i * a.N + src]), weight<Q, float>(L::wc(i)));
        });
        return;
    }
    static_for<0, Q>([&](auto I) {
        constexpr int i = decltype(I)::value;
        constexpr int cx = L::cx(i), cy = L::cy(i), cz = L::cz(i);
        int sx = p.x - cx, sy = p.y - cy, sz = p.z - cz;
        bool bnc = false;
        int wf = -1, ghost = 0;
        if (cx != 0) {
            if (sx < 0 || sx >= a.nx) {
                int face = sx < 0 ? 0 : 1;
                int k = a.bc[face];
                if (k == LB_BC_PERIODIC) sx = sx < 0 ? sx + a.nx : sx - a.nx;
                else if (k == LB_BC_OUTFLOW || k == LB_BC_INLET_EQ) sx = sx < 0 ? 0 : a.nx - 1;
                else { bnc = true; wf = face; }
            }
        }
        if (!bnc && cy != 0) {
            if (sy < 0 || sy >= a.ny) {
                int face = sy < 0 ? 2 : 3;
                int k = a.bc[face];
                if (k == LB_BC_PERIODIC) sy = sy < 0 ? sy + a.ny : sy - a.ny;
                else if (k == LB_BC_OUTFLOW || k == LB_BC_INLET_EQ) sy = sy < 0 ? 0 : a.ny - 1;
                else { bnc = true; wf = face; }
            }
        }
        if (Q == 19 && !bnc && cz != 0) {
            if (sz < 0 || sz >= a.nz) {
                int face = sz < 0 ? 4 : 5;
                if (a.halo) { ghost = sz < 0 ? 1 : 2; sz = sz < 0 ? sz + a.nz : sz - a.nz; }
                else {
                    int k = a.bc[face];
                    if (k == LB_BC_PERIODIC) sz = sz < 0 ? sz + a.nz : sz - a.nz;
                    else if (k == LB_BC_OUTFLOW || k == LB_BC_INLET_EQ) sz = sz < 0 ? 0 : a.nz - 1;
                    else { bnc = true; wf = face; }
                }
            }
        }
        uint32_t sxy = (uint32_t)sx + (uint32_t)a.nx * (uint32_t)sy;
        uint32_t sid = sxy + a.nxy * (uint32_t)sz;
        if (!bnc && !a.bb_full && (a.flags[sid] & F_SOLID)) bnc = true;
        T v;
        if (bnc) {
            constexpr int j = L::opp(i);
            v = S::ld(a.A[(size_t)j * a.N + p.id], weight<Q, float>(L::wc(j)));
            if (wf >= 0 && a.bc[wf] == LB_BC_MOVING_WALL) {
                T cu = (T)L::cx(i) * (T)a.bcu[wf][0] + (T)L::pcy(i) * (T)a.bcu[wf][1] +
                       (T)L::cz(i) * (T)a.bcu[wf][2];
                v = v + (T(6) * weight<Q, T>(L::wc(i))) * cu;
            }
        } else if (Q == 19 && ghost == 1) {
            v = S::ld(a.recv[(size_t)(5 + zslot(i)) * a.P + sxy], weight<Q, float>(L::wc(i)));
        } else if (Q == 19 && ghost == 2) {
            v = S::ld(a.recv[(size_t)zslot(i) * a.P + sxy], weight<Q, float>(L::wc(i)));
        } else {
            v = S::ld(ldm<MEM>(&a.A[(size_t)i * a.N + sid]), weight<Q, float>(L::wc(i)));
        }
        f[i] = v;
    });
}

template <int Q, class T> LB_DEV void moments(const T *f, T &rho, T &ux, T &uy, T &uz) {
    if constexpr (Q == 9) { mom9(f, rho, ux, uy); uz = T(0); }
    else mom19(f, rho, ux, uy, uz);
}

template <int Q, class S, int MEM = 0>
LB_DEV void store_node(const StepArgs<S> &a, const NodePos &p, const typename S::T *f, bool reversed) {
    typedef Vel<Q> L;
    // reversed (full-way bounce-back at a solid node): slot opp(i) <- f_i
    static_for<0, Q>([&](auto I) {
        constexpr int i = decltype(I)::value;
        auto v = S::st(f[i], weight<Q, float>(L::wc(i)));
        stm<MEM>(&a.B[(size_t)(reversed ? L::opp(i) : i) * a.N + p.id], v);
    });
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
                        auto v = S::st(reversed ? f[L::opp(i)] : f[i], weight<Q, float>(L::wc(i)));
                        a.send[(size_t)((cz > 0 ? 5 : 0) + zslot(i)) * a.P + xy] = v;
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
    T rho;
    if constexpr (Q == 9) {
        drho = s[0] + s[1] + s[2] + s[3] + s[4] + s[5] + s[6] + s[7] + s[8];
        rho = T(1) + drho;
        ux = (s[1] - s[3] + s[5] - s[6] - s[7] + s[8]) / rho;
        uy = (s[2] - s[4] + s[5] + s[6] - s[7] - s[8]) / rho;
        uz = T(0);
    } else {
        drho = s[0] + s[1] + s[2] + s[3] + s[4] + s[5] + s[6] + s[7] + s[8] + s[9] + s[10] + s[11] +
               s[12] + s[13] + s[14] + s[15] + s[16] + s[17] + s[18];
        rho = T(1) + drho;
        ux = (s[1] + s[7] + s[8] + s[9] + s[10] - s[2] - s[11] - s[12] - s[13] - s[14]) / rho;
        uy = (s[3] + s[7] + s[11] + s[15] + s[16] - s[4] - s[8] - s[12] - s[17] - s[18]) / rho;
        uz = (s[5] + s[9] + s[13] + s[15] + s[17] - s[6] - s[10] - s[14] - s[16] - s[18]) / rho;
    }
}

template <int Q, class T> LB_DEV void feq_dev(T *seq, T drho, T ux, T uy, T uz) {
    T g[Q];
    if constexpr (Q == 9) geq9<T>(g, ux, uy);
    else geq19<T>(g, ux, uy, uz);
    const T rho = T(1) + drho;
    static_for<0, Q>([&](auto I) {
        constexpr int i = decltype(I)::value;
        seq[i] = weight<Q, T>(Vel<Q>::wc(i)) * (drho + rho * g[i]);
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
        s[i] = s[i] - ti * (s[i] - eq[i]);
    });
}

// Shan-Chen psi tile in LDS: (BX+2) x (BY+2) x 3 planes around the block.
template <int Q, class S>
LB_DEV void load_psi_tile(const StepArgs<S> &a, int b, typename S::T *lds, int bx, int by) {
    int tx = b % a.ntx;
    int r = b / a.ntx;
    int ty = r % a.nty;
    int z = r / a.nty;
    const int W = bx + 2, Hh = by + 2;
    const int planes = Q == 19 ? 3 : 1;
    const int total = W * Hh * planes;
    const int x0 = (tx << a.bx_shift) - 1, y0 = ty * a.by - 1;
    for (int k = threadIdx.x; k < total; k += kThreads) {
        int dx = k % W, rest = k / W;
        int dy = rest % Hh, dz = rest / Hh;
        int gx = x0 + dx, gy = y0 + dy, gz = Q == 19 ? z - 1 + dz : 0;
        // neighbour rule: wrap across periodic (and halo) faces, clamp otherwise
        if (gx < 0) gx = a.bc[0] == LB_BC_PERIODIC ? gx + a.nx : 0;
        if (gx >= a.nx) gx = a.bc[1] == LB_BC_PERIODIC ? gx - a.nx : a.nx - 1;
        if (gx >= a.nx) gx = a.nx - 1;   // tile columns past a ragged edge
        if (gy < 0) gy = a.bc[2] == LB_BC_PERIODIC ? gy + a.ny : 0;
        if (gy >= a.ny) gy = a.bc[3] == LB_BC_PERIODIC ? gy - a.ny : a.ny - 1;
        if (gy >= a.ny) gy = a.ny - 1;
        if (Q == 19) {
            if (gz < 0) gz = (a.halo || a.bc[4] == LB_BC_PERIODIC) ? gz + a.nz : 0;
            if (gz >= a.nz) gz = (a.halo || a.bc[5] == LB_BC_PERIODIC) ? gz - a.nz : a.nz - 1;
        }
        lds[k] = a.psi[(uint32_t)gx + (uint32_t)a.nx * ((uint32_t)gy + (uint32_t)a.ny * (uint32_t)gz)];
    }
}

// ---------------------------------------------------------------------------
// The fused step.
// Zou-He inlet (x = 0) / outlet (x = nx-1) in the form matching the storage.
template <int Q, class S> LB_DEV void apply_zou_he(const StepArgs<S> &a, const NodePos &p, typename S::T *f) {
    typedef typename S::T T;
    if (!a.zou_he) return;
    if (p.x == 0) {
        if constexpr (S::DEV) { if constexpr (Q == 9) zouhe9_in_dev<T>(f, (T)a.zh_ux); else zouhe19_in_dev<T>(f, (T)a.zh_ux); }
        else { if constexpr (Q == 9) zouhe9_in(f, a.zh_ux); else zouhe19_in(f, a.zh_ux); }
    } else if (p.x == a.nx - 1) {
        if constexpr (S::DEV) { if constexpr (Q == 9) zouhe9_out_dev<T>(f, (T)a.zh_rho); else zouhe19_out_dev<T>(f, (T)a.zh_rho); }
        else { if constexpr (Q == 9) zouhe9_out(f, a.zh_rho); else zouhe19_out(f, a.zh_rho); }
    }
}

// ---------------------------------------------------------------------------
// The body of the fused step for one node.  lds_raw holds the psi tile when
// SC is set (loaded by the caller, all threads, before a barrier).
template <int Q, class S, bool SC, bool MACROS, int MEM = 0>
LB_DEV void step_body(const StepArgs<S> &a, const NodePos &p, const unsigned char *lds_raw) {
    typedef typename S::T T;
    typedef Vel<Q> L;
    const int bx = 1 << a.bx_shift;
    if (!p.valid) return;
    const uint8_t fl = a.flags[p.id];
    T f[Q];
    if (fl & F_SOLID) {
        if (a.bb_full) {
            pull_node<Q, S, MEM>(a, p, F_SLOW, f);
            store_node<Q, S, MEM>(a, p, f, true);
        }
        if (MACROS) {
            a.macros[p.id] = 0.f; a.macros[a.N + p.id] = 0.f;
            a.macros[2 * a.N + p.id] = 0.f; a.macros[3 * a.N + p.id] = 0.f;
        }
        return;
    }
    if (fl & F_INLET) {
        int face = fl >> 4;
        T rho = (T)a.bcrho[face], ux = (T)a.bcu[face][0], uy = (T)a.bcu[face][1], uz = (T)a.bcu[face][2];
        if constexpr (S::DEV) feq_dev<Q, T>(f, rho - T(1), ux, uy, uz);
        else feq<Q, T>(f, rho, ux, uy, uz);
        store_node<Q, S, MEM>(a, p, f, false);
        if (MACROS) {
            a.macros[p.id] = (float)rho; a.macros[a.N + p.id] = (float)ux;
            a.macros[2 * a.N + p.id] = (float)uy; a.macros[3 * a.N + p.id] = (float)uz;
        }
        return;
    }
    pull_node<Q, S, MEM>(a, p, fl, f);
    apply_zou_he<Q, S>(a, p, f);
    T rho, ux, uy, uz, drho = T(0);
    if constexpr (S::DEV) {
        moments_dev<Q, T>(f, drho, ux, uy, uz);
        rho = T(1) + drho;
    } else {
        moments<Q, T>(f, rho, ux, uy, uz);
    }
    if (MACROS) {
        a.macros[p.id] = (float)rho; a.macros[a.N + p.id] = (float)ux;
        a.macros[2 * a.N + p.id] = (float)uy; a.macros[3 * a.N + p.id] = (float)uz;
    }
    const T ti = (T)a.ti;
    T ex = ux, ey = uy, ez = uz;
    if (a.force[0] != 0.0f) ex = ux + (T)a.force[0] / (rho * ti);
    if (a.force[1] != 0.0f) ey = uy + (T)a.force[1] / (rho * ti);
    if (a.force[2] != 0.0f) ez = uz + (T)a.force[2] / (rho * ti);
    if constexpr (SC) {
        const T *lds = (const T *)lds_raw;
        const int W = bx + 2, Hh = a.by + 2;
        const int lx = (threadIdx.x & (bx - 1)) + 1, ly = (threadIdx.x >> a.bx_shift) + 1;
        T IF[3] = {T(0), T(0), T(0)};
        static_for<1, Q>([&](auto I) {
            constexpr int i = decltype(I)::value;
            constexpr int cx = L::cx(i), cy = L::cy(i), cz = L::cz(i);
            int k = (lx + cx) + W * ((ly + cy) + Hh * (Q == 19 ? 1 + cz : 0));
            T wp = weight<Q, T>(L::wc(i)) * lds[k];
            if constexpr (cx > 0) IF[0] = IF[0] + wp; else if constexpr (cx < 0) IF[0] = IF[0] - wp;
            if constexpr (L::pcy(i) > 0) IF[1] = IF[1] + wp; else if constexpr (L::pcy(i) < 0) IF[1] = IF[1] - wp;
            if constexpr (cz > 0) IF[2] = IF[2] + wp; else if constexpr (cz < 0) IF[2] = IF[2] - wp;
        });
        const int kc = lx + W * (ly + Hh * (Q == 19 ? 1 : 0));
        T gp = (T)(-a.scG) * lds[kc];
        ex = ex + (gp * IF[0]) / (ti * rho);
        ey = ey + (gp * IF[1]) / (ti * rho);
        if constexpr (Q == 19) ez = ez + (gp * IF[2]) / (ti * rho);
    }
    if constexpr (S::DEV) collide_dev<Q, T>(f, drho, ex, ey, ez, ti, a.rest_closure != 0);
    else collide<Q, T>(f, rho, ex, ey, ez, ti, a.rest_closure != 0);
    store_node<Q, S, MEM>(a, p, f, false);
}

// Shan-Chen density pre-pass: psi(rho) of the post-streaming density, with
// the same pull / boundary handling (and summation order) as k_step.
template <int Q, class S>
LB_DEV void density_body(const StepArgs<S> &a, const NodePos &p) {
    typedef typename S::T T;
    if (!p.valid) return;
    const uint8_t fl = a.flags[p.id];
    if (fl & F_SOLID) { a.psi[p.id] = (T)a.scpsi_solid; return; }
    T rho;
    if (fl & F_INLET) {
        rho = (T)a.bcrho[fl >> 4];
    } else {
        T f[Q];
        pull_node<Q, S>(a, p, fl, f);
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
    a.psi[p.id] = sc_psi<T>(rho, a.scrho0);
}

// ---------------------------------------------------------------------------
// Node flags on the host: solid, inlet plane (+ face in the high nibble),
// and "slow" when a link crosses a wall / halo face or meets a solid node
// (those nodes take the general per-link path of pull_node).
inline void build_flags(const lb_config &g, const std::vector<uint8_t> &solid, std::vector<uint8_t> &fl) {
    const int Q = g.lattice;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const bool halo = g.halo_z && Q == 19;
    fl.assign(solid.size(), 0);
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
                bool slow = false;
                for (int i = 1; i < Q && !slow; i++) {
                    int s[3] = {x - C[i][0], y - C[i][1], z - C[i][2]};
                    for (int a = 0; a < 3 && !slow; a++) {
                        if (a == 2 && Q == 9) break;
                        if (s[a] >= 0 && s[a] < n[a]) continue;
                        int face = 2 * a + (s[a] >= n[a] ? 1 : 0);
                        if (a == 2 && halo) { slow = true; break; }
                        int k = g.bc[face];
                        if (k == LB_BC_PERIODIC) s[a] = s[a] < 0 ? s[a] + n[a] : s[a] - n[a];
                        else if (k == LB_BC_OUTFLOW || k == LB_BC_INLET_EQ) s[a] = s[a] < 0 ? 0 : n[a] - 1;
                        else slow = true;
                    }
                    if (!slow && !g.bb_full && solid[s[0] + (size_t)nx * (s[1] + (size_t)ny * s[2])]) slow = true;
                }
                if ((f & F_SOLID) && g.bb_full) slow = true;
                if (slow) f |= F_SLOW;
                fl[id] = f;
            }
}


// Fast-path index corrections across each face (element units):
// periodic wrap or clamp; 0 where the face needs the per-link path.
inline void face_corrections(const lb_config &g, int lo[3], int hi[3]) {
    const int n[3] = {g.nx, g.ny, g.nz};
    const int stride[3] = {1, g.nx, g.nx * g.ny};
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
