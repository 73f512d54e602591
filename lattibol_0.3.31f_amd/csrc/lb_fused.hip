// lb_fused.hip -- the lb_* solver handle around the fused single-sweep pull
// kernels of lb_pull.hpp (D2Q9 / D3Q19, A -> B double buffering).
//
// Data layout in HBM: one population buffer per time level, either
// [i][z][y][x] (SOA, element i*N + n for node n = x + nx (y + ny z)) or
// [z][y][i][x] (ROWS, the Q populations of an x-row back to back), chosen by
// the autotuner; 1-byte node flags; macro planes rho,ux,uy,uz; Shan-Chen psi
// plane; z halo planes (send/recv).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "lb_pull.hpp"

using namespace lb;

namespace {

// ---------------------------------------------------------------------------
// Kernels: one thread per node, 256-thread workgroups = BX x BY tiles of one
// z-plane, tiles optionally remapped so each XCD sweeps a contiguous slab.
// ---------------------------------------------------------------------------
// The argument block lives in device memory (uniform pointer -> scalar loads);
// passed by value it would be copied to per-lane scratch because the device
// functions index its arrays.
// dstep: position of this sweep in the launch sequence since the step counter
// of the stochastic-rounding stream was last advanced (k_bump).
template <int Q, class S, bool SC, bool MACROS, int MEM = 0>
__global__ __launch_bounds__(kThreads) void k_step(const StepArgs<S> *__restrict__ pa, uint32_t dstep) {
    const StepArgs<S> &a = *pa;
    NodePos p = node_of(a, tile_of_block(a));
    if (a.round_mode) {
        p.step = ldm(&a.ctr[0]) + dstep;
        p.seed = sr_seed(p.step, p.id);
    }
    step_body<Q, S, SC, MACROS, MEM>(a, p);
}

// Advances the step counter of the stochastic-rounding stream after a launch
// sequence of n sweeps (one thread).
__global__ void k_bump(uint32_t *ctr, uint32_t n) {
    if (threadIdx.x == 0) ctr[0] += n;
}

// Temporal blocking for small D2Q9 FP32 lattices, whose steps are bound by
// launch latency rather than HBM: a workgroup loads a 16 x 16 tile plus a
// TB-node halo into LDS, advances it TB steps there (the valid region shrinks
// by one node per step) and stores the tile once.  Each node runs the same
// pull / moving-wall / BGK arithmetic as k_step, so the result is identical
// to TB single steps.  Faces are periodic (the halo wraps) or walls (links
// leaving the lattice bounce, as their link masks say); the host admits no
// other boundary (lb_create: time_block).  One thread per tile cell; each
// step reads the neighbours into registers, waits, then overwrites the
// cell in place (one LDS copy of the tile).
constexpr int kTBI = 16;                        // interior tile edge
constexpr uint32_t kTBSkip = 0x7fffffffu;       // cell not updated (solid / outside)
template <int TB> constexpr int tb_threads() { return (kTBI + 2 * TB) * (kTBI + 2 * TB); }

template <int TB>
__global__ __launch_bounds__(tb_threads<TB>()) void k_step_tb9(const StepArgs<StF32> *__restrict__ pa) {
    typedef Vel<9> L;
    constexpr int E = kTBI + 2 * TB;
    const StepArgs<StF32> &a = *pa;
    __shared__ float fs[9][E * E];
    __shared__ uint32_t bms[E * E];
    const int ntx = (a.nx + kTBI - 1) / kTBI;
    const int x0 = (int)(blockIdx.x % ntx) * kTBI - TB, y0 = (int)(blockIdx.x / ntx) * kTBI - TB;
    const int c = threadIdx.x;
    const int lx = c % E, ly = c / E;
    {
        int gx = x0 + lx, gy = y0 + ly;
        const bool in = axis_src(gx, a.nx, a.bc[0], a.bc[1]) && axis_src(gy, a.ny, a.bc[2], a.bc[3]);
        uint32_t code = kTBSkip;
        if (in) {
            NodePos p{};
            p.x = gx; p.y = gy; p.z = 0;
            p.id = (uint32_t)gx + (uint32_t)a.nx * (uint32_t)gy;
            p.e = (uint32_t)gx + a.sy * (uint32_t)gy;
            uint8_t fl;
            uint32_t bm;
            node_word(a, p, fl, bm);
            if (!(fl & F_SOLID)) code = bm;
            static_for<0, 9>([&](auto I) {
                constexpr int i = decltype(I)::value;
                fs[i][c] = ldm(&a.A[(size_t)i * a.ps + p.e]);
            });
        }
        bms[c] = code;
    }
    __syncthreads();
    const uint32_t bm = bms[c];
    for (int s = 1; s <= TB; s++) {
        const bool act = bm != kTBSkip && lx >= s && lx < E - s && ly >= s && ly < E - s;
        float f[9];
        if (act) {
            static_for<0, 9>([&](auto I) {
                constexpr int i = decltype(I)::value;
                constexpr int j = L::opp(i);
                f[i] = ((bm >> i) & 1u) ? fs[j][c] : fs[i][c - L::cx(i) - L::cy(i) * E];
            });
        }
        __syncthreads();
        if (act) {
            if (bm >> 31) {
                NodePos p{};
                p.x = x0 + lx; p.y = y0 + ly; p.z = 0;   // inside the lattice: the cell is not skipped
                static_for<1, 9>([&](auto I) {
                    constexpr int i = decltype(I)::value;
                    if ((bm >> i) & 1u) {
                        const int wf = wall_face<9>(a, p, L::cx(i), L::cy(i), L::cz(i));
                        if (wf >= 0 && a.bc[wf] == LB_BC_MOVING_WALL) {
                            float cu = (float)L::cx(i) * a.bcu[wf][0] + (float)L::pcy(i) * a.bcu[wf][1] +
                                       (float)L::cz(i) * a.bcu[wf][2];
                            f[i] = f[i] + (6.0f * weight<9, float>(L::wc(i))) * cu;
                        }
                    }
                });
            }
            float rho, ux, uy, uz;
            moments<9, float>(f, rho, ux, uy, uz);
            collide<9, float>(f, rho, ux, uy, uz, a.ti, a.rest_closure != 0);
            static_for<0, 9>([&](auto I) {
                constexpr int i = decltype(I)::value;
                fs[i][c] = f[i];
            });
        }
        __syncthreads();
    }
    const int gx = x0 + lx, gy = y0 + ly;
    if (lx >= TB && lx < TB + kTBI && ly >= TB && ly < TB + kTBI && gx < a.nx && gy < a.ny && bm != kTBSkip) {
        const uint32_t e = (uint32_t)gx + a.sy * (uint32_t)gy;
        static_for<0, 9>([&](auto I) {
            constexpr int i = decltype(I)::value;
            stm(&a.B[(size_t)i * a.ps + e], fs[i][c]);
        });
    }
}

template <int Q, class S, int MEM = 0>
__global__ __launch_bounds__(kThreads) void k_density(const StepArgs<S> *__restrict__ pa) {
    const StepArgs<S> &a = *pa;
    density_body<Q, S, MEM>(a, node_of(a, tile_of_block(a)));
}

// Equilibrium initialisation from device fields.
template <int Q, class S>
__global__ __launch_bounds__(kThreads) void k_init_eq(typename S::raw *A, typename S::raw *B, uint32_t N,
                                                      uint32_t nx, uint32_t ny, PopLayout pl,
                                                      const float *rho, const float *ux, const float *uy,
                                                      const float *uz) {
    typedef typename S::T T;
    uint32_t id = blockIdx.x * kThreads + threadIdx.x;
    if (id >= N) return;
    const uint32_t e = elem_of(id, nx, ny, pl.sy, pl.sz);
    T f[Q];
    if constexpr (S::DEV) feq_dev<Q, T>(f, (T)rho[id] - T(1), (T)ux[id], (T)uy[id], uz ? (T)uz[id] : T(0));
    else feq<Q, T>(f, (T)rho[id], (T)ux[id], (T)uy[id], uz ? (T)uz[id] : T(0));
    static_for<0, Q>([&](auto I) {
        constexpr int i = decltype(I)::value;
        auto v = S::st(f[i], weight<Q, float>(Vel<Q>::wc(i)));
        A[(size_t)i * pl.ps + e] = v;
        B[(size_t)i * pl.ps + e] = v;
    });
}

// FP64 totals of the populations over fluid nodes: per-block partials.
template <int Q, class S>
__global__ __launch_bounds__(kThreads) void k_reduce(const typename S::raw *A, const uint8_t *flags, uint32_t N,
                                                     uint32_t nx, uint32_t ny, PopLayout pl, double *partial) {
    typedef Vel<Q> L;
    double m = 0, px = 0, py = 0, pz = 0;
    for (uint32_t id = blockIdx.x * kThreads + threadIdx.x; id < N; id += gridDim.x * kThreads) {
        if (flags[id] & F_SOLID) continue;
        const uint32_t e = elem_of(id, nx, ny, pl.sy, pl.sz);
        static_for<0, Q>([&](auto I) {
            constexpr int i = decltype(I)::value;
            double v = (double)S::ld(A[(size_t)i * pl.ps + e], weight<Q, float>(L::wc(i)));
            if constexpr (S::DEV) v += (double)weight<Q, float>(L::wc(i));
            m += v;
            if constexpr (L::cx(i) > 0) px += v; else if constexpr (L::cx(i) < 0) px -= v;
            if constexpr (L::pcy(i) > 0) py += v; else if constexpr (L::pcy(i) < 0) py -= v;
            if constexpr (L::cz(i) > 0) pz += v; else if constexpr (L::cz(i) < 0) pz -= v;
        });
    }
    __shared__ double sh[4][kThreads / 64];
    double v4[4] = {wave_sum(m), wave_sum(px), wave_sum(py), wave_sum(pz)};
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) for (int k = 0; k < 4; k++) sh[k][w] = v4[k];
    __syncthreads();
    if (threadIdx.x < 4) {
        double s = 0;
        for (int k = 0; k < kThreads / 64; k++) s += sh[threadIdx.x][k];
        partial[blockIdx.x * 4 + threadIdx.x] = s;
    }
}

// Between the host order [i][z][y][x] and the device layout pl (raw copy).
template <class R>
__global__ __launch_bounds__(kThreads) void k_relayout(const R *src, R *dst, uint32_t N, int Q, uint32_t nx,
                                                       uint32_t ny, PopLayout pl, int to_device) {
    uint32_t id = blockIdx.x * kThreads + threadIdx.x;
    if (id >= N) return;
    const uint32_t e = elem_of(id, nx, ny, pl.sy, pl.sz);
    for (int i = 0; i < Q; i++) {
        if (to_device) dst[(size_t)i * pl.ps + e] = src[(size_t)i * N + id];
        else dst[(size_t)i * N + id] = src[(size_t)i * pl.ps + e];
    }
}

// Fresh z halo from the populations in A: recv slots 0-4 <- the -z
// populations of plane 0, slots 5-9 <- the +z populations of plane nz-1
// (what the pack + copy after a sweep would leave there).
template <class R>
__global__ __launch_bounds__(kThreads) void k_fill_recv(const R *A, R *recv, uint32_t P, uint32_t nx, int nz,
                                                        PopLayout pl) {
    typedef Vel<19> L;
    uint32_t xy = blockIdx.x * kThreads + threadIdx.x;
    if (xy >= P) return;
    const uint32_t y = xy / nx, x = xy - y * nx;
    const uint32_t e0 = x + y * pl.sy, e1 = e0 + (uint32_t)(nz - 1) * pl.sz;
    static_for<0, 19>([&](auto I) {
        constexpr int i = decltype(I)::value;
        if constexpr (L::cz(i) < 0) recv[(size_t)zslot(i) * P + xy] = A[(size_t)i * pl.ps + e0];
        else if constexpr (L::cz(i) > 0) recv[(size_t)(5 + zslot(i)) * P + xy] = A[(size_t)i * pl.ps + e1];
    });
}

// Residual accumulators: sum |u - u_snap|^2 and |u|^2 over fluid nodes,
// then u_snap := u.
__global__ __launch_bounds__(kThreads) void k_residual(const float *macros, float *snap, const uint8_t *flags,
                                                       uint32_t N, int ncomp, double *partial) {
    double num = 0, den = 0;
    for (uint32_t id = blockIdx.x * kThreads + threadIdx.x; id < N; id += gridDim.x * kThreads) {
        if (flags[id] & F_SOLID) continue;
        for (int c = 0; c < ncomp; c++) {
            double u = macros[(size_t)(1 + c) * N + id];
            double s = snap[(size_t)c * N + id];
            num += (u - s) * (u - s);
            den += u * u;
            snap[(size_t)c * N + id] = (float)u;
        }
    }
    __shared__ double sh[2][kThreads / 64];
    double a = wave_sum(num), b = wave_sum(den);
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) { sh[0][w] = a; sh[1][w] = b; }
    __syncthreads();
    if (threadIdx.x < 2) {
        double s = 0;
        for (int k = 0; k < kThreads / 64; k++) s += sh[threadIdx.x][k];
        partial[blockIdx.x * 2 + threadIdx.x] = s;
    }
}

// Final fixed-order sum of per-block partials (one wavefront).
__global__ void k_final_sum(const double *partial, int nblk, int width, double *out) {
    int c = threadIdx.x;
    if (c >= width) return;
    double s = 0;
    for (int b = 0; b < nblk; b++) s += partial[b * width + c];
    out[c] = s;
}

constexpr int kReduceBlocks = 1024;

// Plain streaming copy (16-byte vectors): reads and writes exactly `bytes`;
// the reference point for the HBM traffic counters.
__global__ __launch_bounds__(256) void k_copy16(const uint4 *src, uint4 *dst, size_t n) {
    for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (size_t)gridDim.x * 256) dst[e] = src[e];
}

// Every 32-bit pattern through both BF16 roundings; counts the non-NaN
// inputs where they differ (one atomic per wavefront).
__global__ __launch_bounds__(256) void k_check_bf16(unsigned long long *bad) {
    unsigned long long n = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < (1ull << 32); i += (uint64_t)gridDim.x * 256) {
        const float f = __builtin_bit_cast(float, (uint32_t)i);
        if (f == f && f32_to_bf16(f) != f32_to_bf16_bits(f)) n++;
    }
    for (int off = 32; off > 0; off >>= 1) n += __shfl_xor(n, off, 64);
    if ((threadIdx.x & 63) == 0 && n) atomicAdd(bad, n);
}

// Every FP32 pattern u with two random words (the node word sr_seed(seed, u)
// and its complement): the stochastic-rounding conversions of the stores
// against their bit-pattern forms, BF16 and FP16 mismatches counted apart.
// (2^32 is a multiple of the grid, so every lane runs every iteration and
// the FP16 conversion's wave-wide vote sees full waves.)
__global__ __launch_bounds__(256) void k_check_sr(uint32_t seed, unsigned long long *bad) {
    unsigned long long nb = 0, nh = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < (1ull << 32); i += (uint64_t)gridDim.x * 256) {
        const uint32_t u = (uint32_t)i;
        const float f = __builtin_bit_cast(float, u);
        const uint32_t r0 = sr_seed(seed, u);
        for (int k = 0; k < 2; k++) {
            const uint32_t r = k ? ~r0 : r0;
            if (f32_to_bf16_sr(f, r) != f32_to_bf16_sr_bits(f, r)) nb++;
            if (f32_to_f16_sr<true>(f, r) != f32_to_f16_sr<false>(f, r)) nh++;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        nb += __shfl_xor(nb, off, 64);
        nh += __shfl_xor(nh, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        if (nb) atomicAdd(bad, nb);
        if (nh) atomicAdd(bad + 1, nh);
    }
}

}  // namespace

extern "C" int lb_check_sr_rounding(uint32_t seed, uint64_t *mismatches) {
    if (!mismatches) return LB_ERR_INVALID;
    unsigned long long *d = nullptr;
    int rc = (int)hipMalloc((void **)&d, 2 * sizeof(*d));
    if (rc) return rc;
    rc = (int)hipMemset(d, 0, 2 * sizeof(*d));
    if (!rc) {
        hipLaunchKernelGGL(k_check_sr, dim3(16384), dim3(256), 0, 0, seed, d);
        rc = (int)hipGetLastError();
    }
    unsigned long long h[2] = {0, 0};
    if (!rc) rc = (int)hipMemcpy(h, d, sizeof(h), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    mismatches[0] = (uint64_t)h[0];
    mismatches[1] = (uint64_t)h[1];
    return rc;
}

extern "C" int lb_check_bf16_rounding(uint64_t *mismatches) {
    if (!mismatches) return LB_ERR_INVALID;
    unsigned long long *d = nullptr;
    int rc = (int)hipMalloc((void **)&d, sizeof(*d));
    if (rc) return rc;
    rc = (int)hipMemset(d, 0, sizeof(*d));
    if (!rc) {
        hipLaunchKernelGGL(k_check_bf16, dim3(16384), dim3(256), 0, 0, d);
        rc = (int)hipGetLastError();
    }
    unsigned long long h = 0;
    if (!rc) rc = (int)hipMemcpy(&h, d, sizeof(h), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    *mismatches = (uint64_t)h;
    return rc;
}

extern "C" int lb_copy_device(void *dst, const void *src, size_t bytes, int repeats, float *ms_per_copy) {
    if (!dst || !src || bytes % 16 || repeats < 1) return LB_ERR_INVALID;
    const size_t n = bytes / 16;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    hipLaunchKernelGGL(k_copy16, dim3(4096), dim3(256), 0, 0, (const uint4 *)src, (uint4 *)dst, n);
    (void)hipEventRecord(e0, 0);
    for (int r = 0; r < repeats; r++)
        hipLaunchKernelGGL(k_copy16, dim3(4096), dim3(256), 0, 0, (const uint4 *)src, (uint4 *)dst, n);
    (void)hipEventRecord(e1, 0);
    int rc = (int)hipEventSynchronize(e1);
    float t = 0;
    if (!rc) rc = (int)hipEventElapsedTime(&t, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ms_per_copy) *ms_per_copy = t / repeats;
    return rc ? rc : (int)hipGetLastError();
}

// ===========================================================================
// Handle.
// ===========================================================================
struct lb_ctx {
    lb_config cfg{};
    int Q = 19;
    uint32_t N = 0, P = 0;
    size_t esz = 4;
    hipStream_t stream = nullptr;
    void *buf[2] = {nullptr, nullptr};
    int cur = 0;
    uint8_t *flags = nullptr;
    uint32_t *lmask = nullptr;
    uint32_t *ctr = nullptr;          // step counter of the stochastic-rounding stream (+1 spare word)
    uint32_t *nodew = nullptr;        // packed node words (flags << 24 | link mask)
    int packed = 0;                   // kernels read nodew instead of flags + lmask
    float *macros = nullptr;
    float *snap = nullptr;
    void *psi = nullptr;
    void *halo[2] = {nullptr, nullptr};   // z halo planes; the sweep from buffer s packs into halo[s]
                                          // and unpacks halo[1-s]
    double *partial = nullptr, *dout = nullptr;
    float *tmp_fields = nullptr;
    bool macros_valid = false, snap_valid = false;
    long long steps = 0;
    std::vector<uint8_t> solid;
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    int graph_end[2] = {0, 1};        // buffer holding the result of graph[s]
    // The end of an lb_step call (r plain steps, the macro-writing step and
    // the rounding-stream bump) as a graph per (start buffer, r); the few
    // most recent kept.
    struct Tail { int cur, r, end; hipGraphExec_t g; };
    std::vector<Tail> tails;
    static constexpr size_t kMaxTails = 8;
    int graph_steps = 0;
    int bx_shift = 6, by = 4;
    int xcd = 0;
    int mem = 0;                      // population load/store hints (see ldm/stm)
    int layout = LB_LAYOUT_SOA;       // population layout (LB_LAYOUT_*)
    int tb = 0;                       // steps per temporally blocked launch (0: off)
    // launch geometry and load hints of the Shan-Chen density pass, tuned on
    // their own (its traffic is 19 loads and one store per node)
    int dbx_shift = 6, dby = 4, dxcd = 0, dmem = 0;
    // Shan-Chen D3Q19 step form: 0 density pass + sweep; 1 / 2 the one-pass
    // z-marching sweep (k_step_scz; the sweep's launch fields then hold its
    // tile shape), keeping the pulled populations in registers (1) or pulling
    // them again for the sweep (2), with zc z-planes per workgroup; 3 the
    // row-slab one-pass sweep of fully periodic lattices (k_step_scr; by =
    // rows per tile, scr_nw its waves per workgroup)
    int scz = 0, zc = 24;
    int scr_nw = 8;
    int scr_frame_rev = 0;            // k_step_scr's frame-segment order (tuned with the shape)
    int sc_req = -1;                  // the caller's sc_sweep
    // the tuned two-pass sweep geometry, restored when solids make the
    // row-slab form inapplicable (lb_set_solid)
    int two_bx_shift = 6, two_by = 4, two_xcd = 0, two_mem = 0;
    void *dargs = nullptr;            // argument blocks: sweep from buffer 0 / 1, density pass from 0 / 1
    static constexpr size_t kArgBytes = 640;

    ~lb_ctx() {
        for (auto &g : graph) if (g) (void)hipGraphExecDestroy(g);
        for (auto &t : tails) (void)hipGraphExecDestroy(t.g);
        void *ps[] = {buf[0], buf[1], flags, lmask, ctr, nodew, macros, snap, psi, halo[0], halo[1],
                      partial, dout, tmp_fields, dargs};
        for (void *p : ps) if (p) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }
    PopLayout pl() const { return pop_layout(layout, Q, cfg.nx, cfg.ny, cfg.nz); }
    // (the stream is drained first: a graph launched by an earlier lb_step
    // may still be queued or running)
    void drop_graphs() {
        if (stream && (graph[0] || graph[1] || !tails.empty())) (void)hipStreamSynchronize(stream);
        for (auto &g : graph) if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
        for (auto &t : tails) (void)hipGraphExecDestroy(t.g);
        tails.clear();
    }
};

namespace {

int ilog2(int v) { int s = 0; while ((1 << s) < v) s++; return s; }

// Tile shapes of the one-pass Shan-Chen sweep: 256, 512 or 1,024 threads, the
// frame (2 (bx + 2) + 2 by positions) within the workgroup and the psi plane
// within its LDS slot.
bool scz_shape_ok(int bx, int by) {
    return (bx == 32 || bx == 64 || bx == 128) && (bx * by == 256 || bx * by == 512 || bx * by == 1024) &&
           2 * (bx + 2) + 2 * by <= bx * by && (bx + 2) * (by + 2) <= kZSlot;
}

// The row-slab one-pass sweep (k_step_scr) applies to Shan-Chen D3Q19
// lattices with FP32 / FP16 / BF16 storage and plain accesses whose six
// faces are periodic, with no halo planes, Zou-He faces or solid nodes, rows
// of at most 512 nodes and a population buffer addressable with 32-bit byte
// offsets.
bool scr_lattice_ok(const lb_config &g, uint32_t N) {
    if (g.lattice != 19 || g.sc_G == 0.0f || !scz_storage(g.storage) || g.mem_mode > 0) return false;
    for (int k = 0; k < 6; k++) if (g.bc[k] != LB_BC_PERIODIC) return false;
    if (g.halo_z || g.zou_he || g.nx > 512) return false;
    const uint64_t es = g.storage == LB_STORE_F32 ? 4 : 2;
    return (uint64_t)19 * N * es < (1ull << 32);
}
bool any_solid(const std::vector<uint8_t> &s) { return std::find(s.begin(), s.end(), 1) != s.end(); }
bool scr_shape_ok(int nw, int nx, int by, int storage) {
    return (nw == 6 || nw == 8 || nw == 12 || nw == 16) && by >= 1 && by <= scr_max_rows(nw, nx, storage);
}

// 16-bit stores round stochastically: asked for, or by default on Shan-Chen
// lattices, whose per-step deviation increments fall below half a 16-bit
// spacing and are rounded away with nearest-even.
bool stochastic_rounding(const lb_config &g) {
    if (g.storage != LB_STORE_F16 && g.storage != LB_STORE_BF16) return false;
    return g.round_mode == 1 || (g.round_mode == 0 && g.sc_G != 0.0f);
}

// Argument block of the sweep (dens = false) or of the density pass from
// buffer src; the two differ only in launch geometry.
template <class S> StepArgs<S> make_args(const lb_ctx *c, int src, bool dens = false) {
    StepArgs<S> a{};
    a.A = (const typename S::raw *)c->buf[src];
    a.B = (typename S::raw *)c->buf[1 - src];
    a.flags = c->flags;
    a.lmask = c->lmask;
    a.nodew = c->nodew;
    a.packed = c->packed;
    a.macros = c->macros;
    a.psi = (typename S::T *)c->psi;
    a.send = (typename S::raw *)c->halo[src];
    a.recv = (const typename S::raw *)c->halo[1 - src];
    a.N = c->N;
    const PopLayout pl = c->pl();
    a.ps = pl.ps; a.sy = pl.sy; a.sz = pl.sz;
    a.nxy = (uint32_t)c->cfg.nx * (uint32_t)c->cfg.ny;
    a.P = c->P;
    a.nx = c->cfg.nx; a.ny = c->cfg.ny; a.nz = c->cfg.nz;
    a.ti = c->cfg.tau_inv;
    for (int k = 0; k < 6; k++) {
        a.bc[k] = c->cfg.bc[k];
        a.bcrho[k] = c->cfg.bc_rho[k];
        for (int d = 0; d < 3; d++) a.bcu[k][d] = c->cfg.bc_u[k][d];
    }
    face_corrections(c->cfg, pl, a.corr_lo, a.corr_hi);
    for (int d = 0; d < 3; d++) a.force[d] = c->cfg.force[d];
    a.scG = c->cfg.sc_G; a.scrho0 = c->cfg.sc_rho0; a.scpsi_solid = c->cfg.sc_psi_solid;
    a.bb_full = c->cfg.bb_full; a.zou_he = c->cfg.zou_he; a.rest_closure = c->cfg.rest_closure;
    a.round_mode = stochastic_rounding(c->cfg) ? 1 : 0;
    a.ctr = c->ctr;
    a.halo = c->cfg.halo_z && c->Q == 19;
    a.zh_ux = c->cfg.zh_ux_in; a.zh_rho = c->cfg.zh_rho_out;
    a.bx_shift = dens ? c->dbx_shift : c->bx_shift;
    a.by = dens ? c->dby : c->by;
    a.ntx = (c->cfg.nx + (1 << a.bx_shift) - 1) >> a.bx_shift;
    a.nty = (c->cfg.ny + a.by - 1) / a.by;
    a.nblocks = a.ntx * a.nty * c->cfg.nz;
    a.xcd = dens ? c->dxcd : c->xcd;
    a.zc = c->zc;
    a.frame_rev = c->scz == 3 ? c->scr_frame_rev : 0;
    if (!dens && c->scz == 3) {
        // row-slab tiles: ntx = 64-node segments per row, by rows per tile
        a.bx_shift = 6;
        a.ntx = (c->cfg.nx + 63) / 64;
        a.nty = (c->cfg.ny + a.by - 1) / a.by;
    }
    if (!dens && c->scz) a.nblocks = a.ntx * a.nty * ((c->cfg.nz + c->zc - 1) / c->zc);
    if (!dens && c->scz == 3) a.nblocks = a.nty * ((c->cfg.nz + c->zc - 1) / c->zc);
    if (c->Q == 9) fast_path_args<9>(a, 1 << a.bx_shift, a.by);
    else fast_path_args<19>(a, 1 << a.bx_shift, a.by);
    return a;
}

template <class S> int upload_args_t(lb_ctx *c) {
    static_assert(sizeof(StepArgs<S>) <= lb_ctx::kArgBytes, "argument block size");
    StepArgs<S> h[4] = {make_args<S>(c, 0), make_args<S>(c, 1), make_args<S>(c, 0, true), make_args<S>(c, 1, true)};
    int rc = (int)hipStreamSynchronize(c->stream);
    for (int k = 0; k < 4 && !rc; k++)
        rc = (int)hipMemcpy((char *)c->dargs + k * lb_ctx::kArgBytes, &h[k], sizeof(StepArgs<S>), hipMemcpyHostToDevice);
    return rc;
}

// (Re)write both argument blocks after any change of shape, tuning or flags.
int upload_args(lb_ctx *c) {
    switch (c->cfg.storage) {
    case LB_STORE_F32: return upload_args_t<StF32>(c);
    case LB_STORE_F32S: return upload_args_t<StF32S>(c);
    case LB_STORE_F16: return upload_args_t<StF16>(c);
    case LB_STORE_BF16: return upload_args_t<StBF16>(c);
    default: return upload_args_t<StF64>(c);
    }
}

// The Shan-Chen density pass from buffer src (psi of every node).
template <int Q, class S> void launch_density_t(lb_ctx *c, int src) {
    const StepArgs<S> a = make_args<S>(c, src, true);   // host copy: launch geometry only
    const StepArgs<S> *pd = (const StepArgs<S> *)((const char *)c->dargs + (2 + src) * lb_ctx::kArgBytes);
    if (c->dmem == 2) hipLaunchKernelGGL((k_density<Q, S, 2>), dim3(a.nblocks), dim3(kThreads), 0, c->stream, pd);
    else hipLaunchKernelGGL((k_density<Q, S>), dim3(a.nblocks), dim3(kThreads), 0, c->stream, pd);
}

template <int Q, class S>
int launch_step_t(lb_ctx *c, int src, bool macros, uint32_t dstep) {
    StepArgs<S> a = make_args<S>(c, src);   // host copy: launch geometry only
    const StepArgs<S> *pa = (const StepArgs<S> *)((const char *)c->dargs + src * lb_ctx::kArgBytes);
    const bool sc = c->cfg.sc_G != 0.0f;
    if constexpr (Q == 19) {
        if (sc && c->scz == 3)
            return launch_step_scr(c->cfg.storage, macros, c->scr_nw, c->by, c->cfg.nx, (unsigned)a.nblocks,
                                   c->stream, pa, dstep);
        if (sc && c->scz)
            return launch_step_scz(c->cfg.storage, macros, c->scz == 1, (1 << c->bx_shift) * c->by,
                                   (unsigned)a.nblocks, c->stream, pa, dstep);
    }
    if (sc) {
        launch_density_t<Q, S>(c, src);
        if (macros) hipLaunchKernelGGL((k_step<Q, S, true, true>), dim3(a.nblocks), dim3(kThreads), 0, c->stream, pa, dstep);
        else if (c->mem == 1) hipLaunchKernelGGL((k_step<Q, S, true, false, 1>), dim3(a.nblocks), dim3(kThreads), 0, c->stream, pa, dstep);
        else if (c->mem == 2) hipLaunchKernelGGL((k_step<Q, S, true, false, 2>), dim3(a.nblocks), dim3(kThreads), 0, c->stream, pa, dstep);
        else hipLaunchKernelGGL((k_step<Q, S, true, false>), dim3(a.nblocks), dim3(kThreads), 0, c->stream, pa, dstep);
    } else {
        if (macros) hipLaunchKernelGGL((k_step<Q, S, false, true>), dim3(a.nblocks), dim3(kThreads), 0, c->stream, pa, dstep);
        else if (c->mem == 1) hipLaunchKernelGGL((k_step<Q, S, false, false, 1>), dim3(a.nblocks), dim3(kThreads), 0, c->stream, pa, dstep);
        else if (c->mem == 2) hipLaunchKernelGGL((k_step<Q, S, false, false, 2>), dim3(a.nblocks), dim3(kThreads), 0, c->stream, pa, dstep);
        else hipLaunchKernelGGL((k_step<Q, S, false, false>), dim3(a.nblocks), dim3(kThreads), 0, c->stream, pa, dstep);
    }
    return (int)hipGetLastError();
}

template <int Q> int launch_step_q(lb_ctx *c, int src, bool macros, uint32_t dstep = 0) {
    switch (c->cfg.storage) {
    case LB_STORE_F32: return launch_step_t<Q, StF32>(c, src, macros, dstep);
    case LB_STORE_F32S: return launch_step_t<Q, StF32S>(c, src, macros, dstep);
    case LB_STORE_F16: return launch_step_t<Q, StF16>(c, src, macros, dstep);
    case LB_STORE_BF16: return launch_step_t<Q, StBF16>(c, src, macros, dstep);
    default: return launch_step_t<Q, StF64>(c, src, macros, dstep);
    }
}

// One time step from buffer src.  With the z-slab halo the sweep packs the
// outgoing planes into halo[src] and unpacks halo[1 - src], written by the
// previous sweep: for one subdomain on the ring its -z and +z neighbour is
// the slab itself (recv_hi <- send_lo, recv_lo <- send_hi, slot for slot), so
// the exchange between the two is the identity and alternating the two
// buffers replaces the copy.  A ring of several slabs moves the planes
// between sweeps instead (see lb_halo_buffers).
int launch_step(lb_ctx *c, int src, bool macros, uint32_t dstep = 0) {
    return c->Q == 9 ? launch_step_q<9>(c, src, macros, dstep) : launch_step_q<19>(c, src, macros, dstep);
}

void launch_density(lb_ctx *c, int src) {
    const bool q19 = c->Q == 19;
    switch (c->cfg.storage) {
    case LB_STORE_F32: q19 ? launch_density_t<19, StF32>(c, src) : launch_density_t<9, StF32>(c, src); break;
    case LB_STORE_F32S: q19 ? launch_density_t<19, StF32S>(c, src) : launch_density_t<9, StF32S>(c, src); break;
    case LB_STORE_F16: q19 ? launch_density_t<19, StF16>(c, src) : launch_density_t<9, StF16>(c, src); break;
    case LB_STORE_BF16: q19 ? launch_density_t<19, StBF16>(c, src) : launch_density_t<9, StBF16>(c, src); break;
    default: q19 ? launch_density_t<19, StF64>(c, src) : launch_density_t<9, StF64>(c, src); break;
    }
}

// After n sweeps launched with dstep 0..n-1: advance the rounding stream.
int launch_bump(lb_ctx *c, uint32_t n) {
    if (!stochastic_rounding(c->cfg) || n == 0) return 0;
    hipLaunchKernelGGL(k_bump, dim3(1), dim3(64), 0, c->stream, c->ctr, n);
    return (int)hipGetLastError();
}

// Lattices the temporally blocked kernel reproduces exactly (see k_step_tb9).
bool tb_eligible(const lb_ctx *c, int tb) {
    const lb_config &g = c->cfg;
    if (c->Q != 9 || g.storage != LB_STORE_F32 || g.sc_G != 0.0f || g.zou_he || g.bb_full) return false;
    if (g.force[0] != 0.0f || g.force[1] != 0.0f || g.force[2] != 0.0f) return false;
    // the halo of the last tile row / column reaches tb + 14 nodes past the
    // edge: one periodic wrap must bring every halo node back into the lattice
    if (g.nx < kTBI + tb || g.ny < kTBI + tb) return false;
    for (int k = 0; k < 4; k++)
        if (g.bc[k] != LB_BC_PERIODIC && g.bc[k] != LB_BC_WALL && g.bc[k] != LB_BC_MOVING_WALL) return false;
    return true;
}

// One launch advancing c->tb steps from buffer src into buffer 1 - src.
int launch_tb(lb_ctx *c, int src) {
    const StepArgs<StF32> *pa = (const StepArgs<StF32> *)((const char *)c->dargs + src * lb_ctx::kArgBytes);
    const unsigned nb = (unsigned)(((c->cfg.nx + kTBI - 1) / kTBI) * ((c->cfg.ny + kTBI - 1) / kTBI));
    if (c->tb == 8) hipLaunchKernelGGL(k_step_tb9<8>, dim3(nb), dim3(tb_threads<8>()), 0, c->stream, pa);
    else hipLaunchKernelGGL(k_step_tb9<4>, dim3(nb), dim3(tb_threads<4>()), 0, c->stream, pa);
    return (int)hipGetLastError();
}

// n plain steps from buffer src (temporally blocked launches where enabled);
// returns the buffer holding the result.
int launch_plain_steps(lb_ctx *c, int src, int n, int *rc) {
    int s = 0;
    *rc = 0;
    while (s < n && !*rc) {
        if (c->tb && n - s >= c->tb) { *rc = launch_tb(c, src); s += c->tb; }
        else { *rc = launch_step(c, src, false, (uint32_t)s); s++; }
        src = 1 - src;
    }
    return src;
}

template <int Q, class S> int launch_reduce_t(lb_ctx *c) {
    hipLaunchKernelGGL((k_reduce<Q, S>), dim3(kReduceBlocks), dim3(kThreads), 0, c->stream,
                       (const typename S::raw *)c->buf[c->cur], c->flags, c->N, (uint32_t)c->cfg.nx,
                       (uint32_t)c->cfg.ny, c->pl(), c->partial);
    hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(64), 0, c->stream, c->partial, kReduceBlocks, 4, c->dout);
    return (int)hipGetLastError();
}
template <int Q> int launch_reduce_q(lb_ctx *c) {
    switch (c->cfg.storage) {
    case LB_STORE_F32: return launch_reduce_t<Q, StF32>(c);
    case LB_STORE_F32S: return launch_reduce_t<Q, StF32S>(c);
    case LB_STORE_F16: return launch_reduce_t<Q, StF16>(c);
    case LB_STORE_BF16: return launch_reduce_t<Q, StBF16>(c);
    default: return launch_reduce_t<Q, StF64>(c);
    }
}

template <int Q, class S> int launch_init_t(lb_ctx *c, const float *r, const float *ux, const float *uy, const float *uz) {
    hipLaunchKernelGGL((k_init_eq<Q, S>), dim3((c->N + kThreads - 1) / kThreads), dim3(kThreads), 0, c->stream,
                       (typename S::raw *)c->buf[0], (typename S::raw *)c->buf[1], c->N, (uint32_t)c->cfg.nx,
                       (uint32_t)c->cfg.ny, c->pl(), r, ux, uy, uz);
    return (int)hipGetLastError();
}
template <int Q> int launch_init_q(lb_ctx *c, const float *r, const float *ux, const float *uy, const float *uz) {
    switch (c->cfg.storage) {
    case LB_STORE_F32: return launch_init_t<Q, StF32>(c, r, ux, uy, uz);
    case LB_STORE_F32S: return launch_init_t<Q, StF32S>(c, r, ux, uy, uz);
    case LB_STORE_F16: return launch_init_t<Q, StF16>(c, r, ux, uy, uz);
    case LB_STORE_BF16: return launch_init_t<Q, StBF16>(c, r, ux, uy, uz);
    default: return launch_init_t<Q, StF64>(c, r, ux, uy, uz);
    }
}

// Raw-element dispatch for the layout kernels (2, 4 or 8 byte elements).
template <class F> int by_esz(int esz, F &&f) {
    if (esz == 2) return f(uint16_t{});
    if (esz == 4) return f(uint32_t{});
    return f(uint64_t{});
}

// recv planes from buf[0] (after an initialisation or a population upload).
int refill_recv(lb_ctx *c) {
    if (!c->halo[1]) return 0;
    return by_esz(c->esz, [&](auto t) {
        typedef decltype(t) R;
        hipLaunchKernelGGL((k_fill_recv<R>), dim3((c->P + kThreads - 1) / kThreads), dim3(kThreads), 0, c->stream,
                           (const R *)c->buf[0], (R *)c->halo[1], c->P, (uint32_t)c->cfg.nx, c->cfg.nz, c->pl());
        return (int)hipGetLastError();
    });
}

// Host-order populations <-> device layout through a temporary device buffer.
int relayout(lb_ctx *c, const void *src, void *dst, bool to_device) {
    return by_esz(c->esz, [&](auto t) {
        typedef decltype(t) R;
        hipLaunchKernelGGL((k_relayout<R>), dim3((c->N + kThreads - 1) / kThreads), dim3(kThreads), 0, c->stream,
                           (const R *)src, (R *)dst, c->N, c->Q, (uint32_t)c->cfg.nx, (uint32_t)c->cfg.ny, c->pl(),
                           to_device ? 1 : 0);
        return (int)hipGetLastError();
    });
}

int upload_flags(lb_ctx *c) {
    std::vector<uint8_t> fl;
    std::vector<uint32_t> lm;
    build_flags(c->cfg, c->solid, fl, lm);
    std::vector<uint32_t> nw(c->N);
    for (uint32_t k = 0; k < c->N; k++) nw[k] = ((uint32_t)fl[k] << 24) | lm[k];
    int rc = (int)hipMemcpyAsync(c->flags, fl.data(), c->N, hipMemcpyHostToDevice, c->stream);
    if (!rc) rc = (int)hipMemcpyAsync(c->lmask, lm.data(), (size_t)c->N * 4, hipMemcpyHostToDevice, c->stream);
    if (!rc) rc = (int)hipMemcpyAsync(c->nodew, nw.data(), (size_t)c->N * 4, hipMemcpyHostToDevice, c->stream);
    if (!rc) rc = (int)hipStreamSynchronize(c->stream);
    return rc;
}

// The graph of one chunk of graph_steps plain steps from buffer src, built
// on first use.
int chunk_graph(lb_ctx *c, int src, hipGraphExec_t *out) {
    hipGraphExec_t &ge = c->graph[src];
    if (!ge) {
        hipGraph_t g = nullptr;
        int rc = (int)hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
        if (rc) return rc;
        int lrc = 0;
        c->graph_end[src] = launch_plain_steps(c, src, c->graph_steps, &lrc);
        if (!lrc) lrc = launch_bump(c, (uint32_t)c->graph_steps);
        rc = (int)hipStreamEndCapture(c->stream, &g);
        if (!rc) rc = lrc;
        if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
        rc = (int)hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (rc) { ge = nullptr; return rc; }
    }
    *out = ge;
    return 0;
}

// r plain steps from buffer src, then the macro-writing step, then the bump
// of the rounding stream by r + 1; returns the buffer holding the result.
int launch_tail(lb_ctx *c, int src, int r, int *rc) {
    int end = launch_plain_steps(c, src, r, rc);
    if (!*rc) *rc = launch_step(c, end, true, (uint32_t)r);
    if (!*rc) *rc = launch_bump(c, (uint32_t)r + 1);
    return 1 - end;
}

// The graph of an lb_step call's end from buffer src: r plain steps, the
// macro-writing step and the rounding-stream bump (cached per (src, r)).
int tail_graph(lb_ctx *c, int src, int r, lb_ctx::Tail *out) {
    for (const lb_ctx::Tail &t : c->tails)
        if (t.cur == src && t.r == r) { *out = t; return 0; }
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    int rc = (int)hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
    if (rc) return rc;
    int lrc = 0;
    const int end = launch_tail(c, src, r, &lrc);
    rc = (int)hipStreamEndCapture(c->stream, &g);
    if (!rc) rc = lrc;
    if (!rc) rc = (int)hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    if (g) (void)hipGraphDestroy(g);
    if (rc) return rc;
    if (c->tails.size() >= lb_ctx::kMaxTails) {
        // the evicted graph may still be queued on the stream
        if (int src_rc = (int)hipStreamSynchronize(c->stream)) {
            (void)hipGraphExecDestroy(ge);
            return src_rc;
        }
        (void)hipGraphExecDestroy(c->tails.front().g);
        c->tails.erase(c->tails.begin());
    }
    c->tails.push_back({src, r, end, ge});
    *out = c->tails.back();
    return 0;
}

int run_graph_or_eager(lb_ctx *c, int nsteps) {
    // nsteps plain steps (no macros), chunks of graph_steps through a graph.
    const int chunk = c->graph_steps;
    while (nsteps >= chunk && chunk > 0) {
        hipGraphExec_t ge = nullptr;
        int rc = chunk_graph(c, c->cur, &ge);
        if (!rc) rc = (int)hipGraphLaunch(ge, c->stream);
        if (rc) return rc;
        c->cur = c->graph_end[c->cur];
        nsteps -= chunk;
        c->steps += chunk;
    }
    if (nsteps > 0) {
        int rc = 0;
        c->cur = launch_plain_steps(c, c->cur, nsteps, &rc);
        if (!rc) rc = launch_bump(c, (uint32_t)nsteps);
        if (rc) return rc;
        c->steps += nsteps;
    }
    return 0;
}

// The Shan-Chen forms are timed on a uniform equilibrium at density 0.5 in
// both population buffers, written in the layout they are timed in, rather
// than on the zeroed buffers: a zero deviation (or the state other
// candidates left) would send every FP16 stochastic-rounding store down its
// subnormal / overflow path, which a multiphase run hardly takes (both phases
// sit away from density 1), and a uniform state stays uniform over the
// timed steps.  lb_create zeroes the buffers after tuning.
int fill_tuning_state(lb_ctx *c) {
    float *t = c->tmp_fields;
    int rc = (int)hipMemsetD32Async((hipDeviceptr_t)t, 0x3f000000, c->N, c->stream);   // density 0.5
    if (!rc) rc = (int)hipMemsetAsync(t + c->N, 0, (size_t)3 * c->N * sizeof(float), c->stream);
    if (rc) return rc;
    float *uz = c->Q == 19 ? t + 3 * (size_t)c->N : nullptr;
    return c->Q == 9 ? launch_init_q<9>(c, t, t + c->N, t + 2 * (size_t)c->N, uz)
                     : launch_init_q<19>(c, t, t + c->N, t + 2 * (size_t)c->N, uz);
}

// Launch autotuner: times each candidate (block shape x XCD mapping x
// load/store hints x population layout) over a few launches on the real
// lattice and keeps the fastest.  Runs at creation, before any populations
// are set, so switching the layout moves no data.
int autotune(lb_ctx *c) {
    struct Cand { int bx, by, xcd, mem, layout, packed; };
    std::vector<Cand> cands;
    const int bxs[4] = {32, 64, 128, 256};
    for (int bx : bxs) {
        if (bx > 2 * c->cfg.nx && bx != 32) continue;
        if (c->cfg.block_x > 0 && bx != c->cfg.block_x) continue;
        for (int xcd = 0; xcd < 2; xcd++) {
            if (c->cfg.xcd_remap >= 0 && xcd != c->cfg.xcd_remap) continue;
            for (int mem = 0; mem < 3; mem++) {
                if (c->cfg.mem_mode >= 0 && mem != c->cfg.mem_mode) continue;
                for (int lay = 0; lay < 2; lay++) {
                    if (c->cfg.layout >= 0 && lay != c->cfg.layout) continue;
                    for (int pk = 0; pk < 2; pk++) {
                        // packed node words only pay on small, latency-bound lattices
                        // (on HBM-bound ones their 3 extra bytes per node cost ~2 %)
                        if (c->cfg.node_words >= 0 ? pk != c->cfg.node_words : (pk && c->N > (1u << 20))) continue;
                        cands.push_back({bx, kThreads / bx, xcd, mem, lay, pk});
                    }
                }
            }
        }
    }
    // (none: block_x names a width only the row-slab sweep takes; the two
    // passes then run at a default shape with the fields the caller fixed)
    if (cands.empty())
        cands.push_back({64, 4, std::max(0, c->cfg.xcd_remap), std::max(0, c->cfg.mem_mode), std::max(0, c->cfg.layout),
                         std::max(0, c->cfg.node_words)});
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0)) return LB_ERR_INVALID;
    if (hipEventCreate(&e1)) { (void)hipEventDestroy(e0); return LB_ERR_INVALID; }
    float best = 1e30f;
    Cand bestc = cands[0];
    const int reps = 20;
    for (const Cand &k : cands) {
        c->bx_shift = ilog2(k.bx); c->by = k.by; c->xcd = k.xcd; c->mem = k.mem; c->layout = k.layout;
        c->packed = k.packed;
        c->dbx_shift = c->bx_shift; c->dby = c->by; c->dxcd = c->xcd; c->dmem = c->mem == 2 ? 2 : 0;
        if (int rc = upload_args(c)) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            return rc;
        }
        launch_step(c, 0, false);   // warm
        (void)hipEventRecord(e0, c->stream);
        for (int r = 0; r < reps; r++) launch_step(c, r & 1, false);
        (void)hipEventRecord(e1, c->stream);
        (void)hipEventSynchronize(e1);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (ms < best) { best = ms; bestc = k; }
    }
    c->bx_shift = ilog2(bestc.bx); c->by = bestc.by; c->xcd = bestc.xcd; c->mem = bestc.mem;
    c->layout = bestc.layout;
    c->packed = bestc.packed;
    c->dbx_shift = c->bx_shift; c->dby = c->by; c->dxcd = c->xcd; c->dmem = c->mem == 2 ? 2 : 0;
    int rc = (int)hipGetLastError();
    if (!rc) rc = upload_args(c);
    // Shan-Chen: the density pass's own block shape, XCD remap and load
    // hints (the population layout stays the sweep's)
    if (!rc && c->cfg.sc_G != 0.0f) {
        struct DCand { int bx, xcd, mem; };
        std::vector<DCand> dc;
        for (int bx : bxs) {
            if (bx > 2 * c->cfg.nx && bx != 32) continue;
            if (c->cfg.block_x > 0 && bx != c->cfg.block_x) continue;
            for (int xcd = 0; xcd < 2; xcd++) {
                if (c->cfg.xcd_remap >= 0 && xcd != c->cfg.xcd_remap) continue;
                for (int mem : {0, 2}) {
                    if (c->cfg.mem_mode >= 0 && mem != (c->cfg.mem_mode == 2 ? 2 : 0)) continue;
                    dc.push_back({bx, xcd, mem});
                }
            }
        }
        float dbest = 1e30f;
        DCand dbc{1 << c->dbx_shift, c->dxcd, c->dmem};
        for (const DCand &k : dc) {
            c->dbx_shift = ilog2(k.bx); c->dby = kThreads / k.bx; c->dxcd = k.xcd; c->dmem = k.mem;
            rc = upload_args(c);
            if (rc) break;
            launch_density(c, 0);   // warm
            (void)hipEventRecord(e0, c->stream);
            for (int r = 0; r < reps; r++) launch_density(c, r & 1);
            (void)hipEventRecord(e1, c->stream);
            (void)hipEventSynchronize(e1);
            float ms = 0;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (ms < dbest) { dbest = ms; dbc = k; }
        }
        c->dbx_shift = ilog2(dbc.bx); c->dby = kThreads / dbc.bx; c->dxcd = dbc.xcd; c->dmem = dbc.mem;
        if (!rc) rc = (int)hipGetLastError();
        if (!rc) rc = upload_args(c);
    }
    // Shan-Chen D3Q19: the one-pass z-marching sweep (tile shape x z-chunk x
    // XCD remap x hints x layout) against the tuned two passes
    const bool sc_forms = c->cfg.sc_G != 0.0f && c->Q == 19 && c->cfg.sc_sweep != 0 && scz_storage(c->cfg.storage);
    if (!rc && sc_forms) rc = fill_tuning_state(c);
    if (!rc && sc_forms) {
        auto time_steps = [&]() -> float {
            launch_step(c, 0, false);   // warm
            (void)hipEventRecord(e0, c->stream);
            for (int r = 0; r < reps; r++) launch_step(c, r & 1, false);
            (void)hipEventRecord(e1, c->stream);
            (void)hipEventSynchronize(e1);
            float ms = 0;
            (void)hipEventElapsedTime(&ms, e0, e1);
            return ms;
        };
        struct ZCand { int bx, by, zc, xcd, form, layout, v, rev = 0; };
        const int two_bx = c->bx_shift, two_by = c->by, two_xcd = c->xcd, two_mem = c->mem, two_lay = c->layout;
        c->two_bx_shift = two_bx; c->two_by = two_by; c->two_xcd = two_xcd; c->two_mem = two_mem;
        const float two = c->cfg.sc_sweep > 0 ? 1e30f : time_steps();
        float zbest = 1e30f;
        ZCand zb{64, 8, 24, 0, 1, c->layout, 8};
        const int zcs[5] = {8, 16, 24, 32, 48};
        // tiles of 256 / 512 / 1,024 threads whose frame (2 (bx + 2) + 2 by) fits the workgroup
        const int shapes[8][2] = {{32, 8}, {64, 4}, {32, 16}, {64, 8}, {128, 4}, {32, 32}, {64, 16}, {128, 8}};
        const bool forms12 = c->cfg.sc_sweep < 0 || c->cfg.sc_sweep == 1 || c->cfg.sc_sweep == 2;
        for (const auto &sh : shapes) {
            if (!forms12) break;
            const int bx = sh[0], by = sh[1];
            if (c->cfg.block_x > 0 ? bx != c->cfg.block_x : (bx > 2 * c->cfg.nx && bx != 32)) continue;
            if (c->cfg.block_x > 0 && c->cfg.block_y > 0 && by != c->cfg.block_y) continue;
            for (int zc : zcs) {
                if (c->cfg.sc_zchunk > 0 ? zc != zcs[0] : (zc > c->cfg.nz && zc != zcs[0])) continue;
                const int zcv = c->cfg.sc_zchunk > 0 ? c->cfg.sc_zchunk : zc;
                for (int xcd = 0; xcd < 2; xcd++) {
                    if (c->cfg.xcd_remap >= 0 && xcd != c->cfg.xcd_remap) continue;
                    for (int form : {1, 2}) {
                        if (c->cfg.sc_sweep > 0 && form != c->cfg.sc_sweep) continue;
                        for (int lay = 0; lay < 2; lay++) {
                            if (c->cfg.layout >= 0 && lay != c->cfg.layout) continue;
                            c->scz = form; c->zc = zcv;
                            c->bx_shift = ilog2(bx); c->by = by; c->xcd = xcd; c->mem = 0; c->layout = lay;
                            rc = upload_args(c);
                            if (rc) break;
                            const float ms = time_steps();
                            if (ms < zbest) { zbest = ms; zb = {bx, by, zcv, xcd, form, lay, 0}; }
                        }
                        if (rc) break;
                    }
                    if (rc) break;
                }
                if (rc) break;
            }
            if (rc) break;
        }
        // the row-slab sweep of fully periodic lattices: waves per workgroup x
        // rows per tile x z-chunk with the XCD remap on, then the remap off and
        // the reversed frame-segment order for the best shape; the population
        // layout kept.  The long chunks let
        // short tiles fill the CUs with one workgroup each (192^3 FP32: 3 rows
        // x 48 planes = 256 workgroups, 7 % faster than 5 rows x 32 planes =
        // 234 workgroups; the 16-bit storages keep the 5-row tiles).
        if (!rc && (c->cfg.sc_sweep < 0 || c->cfg.sc_sweep == 3) && scr_lattice_ok(c->cfg, c->N) &&
            !any_solid(c->solid)) {
            const int rzcs[11] = {6, 8, 10, 12, 16, 20, 24, 32, 40, 48, 64};
            c->layout = two_lay;              // the tiled candidates left another layout's state
            rc = fill_tuning_state(c);
            float rbest = 1e30f;
            ZCand rb{64, 0, 0, 1, 3, two_lay, 8};
            auto try_r = [&](int nw, int by, int zcv, int xcd, int rev = 0) {
                c->scz = 3; c->zc = zcv; c->scr_nw = nw; c->scr_frame_rev = rev;
                c->bx_shift = 6; c->by = by; c->xcd = xcd; c->mem = 0; c->layout = two_lay;
                rc = upload_args(c);
                if (rc) return;
                const float ms = time_steps();
                if (ms < rbest) { rbest = ms; rb = {64, by, zcv, xcd, 3, two_lay, nw, rev}; }
            };
            const int xcd0 = c->cfg.xcd_remap >= 0 ? c->cfg.xcd_remap : 1;
            for (int nw : {6, 8, 12, 16}) {
                if (rc) break;
                if (c->cfg.block_x > 0 && c->cfg.block_x != 64 * nw) continue;
                const int bmax = scr_max_rows(nw, c->cfg.nx, c->cfg.storage);
                const int bhi = c->cfg.block_y > 0 ? c->cfg.block_y : bmax;
                const int blo = c->cfg.block_y > 0 ? c->cfg.block_y : std::max(1, bmax - 7);
                for (int by = bhi; by >= blo && !rc; by--) {
                    if (!scr_shape_ok(nw, c->cfg.nx, by, c->cfg.storage)) continue;
                    for (int zc : rzcs) {
                        if (c->cfg.sc_zchunk > 0 ? zc != rzcs[0] : (zc > c->cfg.nz && zc != rzcs[0])) continue;
                        try_r(nw, by, c->cfg.sc_zchunk > 0 ? c->cfg.sc_zchunk : zc, xcd0);
                        if (rc) break;
                    }
                }
            }
            if (!rc && rbest < 1e30f && c->cfg.xcd_remap < 0) try_r(rb.v, rb.by, rb.zc, 1 - xcd0);
            // (which waves take the frame rows moves the step by a few per
            // cent either way, depending on the storage and rounding: 192^3
            // FP16 with stochastic rounding 177 -> 169 us with the frame rows
            // on the last waves, the nearest-even storages slower)
            if (!rc && rbest < 1e30f) try_r(rb.v, rb.by, rb.zc, rb.xcd, 1);
            c->scr_frame_rev = 0;
            if (!rc && rbest < zbest) { zbest = rbest; zb = rb; }
        }
        if (!rc && zbest < two) {
            c->scz = zb.form; c->zc = zb.zc; c->scr_nw = zb.v; c->scr_frame_rev = zb.form == 3 ? zb.rev : 0;
            c->bx_shift = ilog2(zb.bx); c->by = zb.by; c->xcd = zb.xcd; c->mem = 0; c->layout = zb.layout;
        } else {
            c->scz = 0;
            c->bx_shift = two_bx; c->by = two_by; c->xcd = two_xcd; c->mem = two_mem; c->layout = two_lay;
        }
        if (!rc) rc = (int)hipGetLastError();
        if (!rc) rc = upload_args(c);
        if (!rc && c->cfg.sc_sweep > 0 && !c->scz) rc = LB_ERR_UNSUPPORTED;   // no tile shape fits
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

// Temporal blocking on or off: the faster of 64 steps through a graph each
// way (launch-latency-bound lattices gain, large ones lose to the halo
// recomputation).
int tune_tb(lb_ctx *c) {
    float best = 1e30f;
    int best_tb = 0;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0)) return LB_ERR_INVALID;
    if (hipEventCreate(&e1)) { (void)hipEventDestroy(e0); return LB_ERR_INVALID; }
    int rc = 0;
    for (int tb : {0, 4, 8}) {
        if (tb && !tb_eligible(c, tb)) continue;
        c->tb = tb;
        hipGraph_t g = nullptr;
        hipGraphExec_t ge = nullptr;
        rc = (int)hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
        if (rc) break;
        int lrc = 0;
        (void)launch_plain_steps(c, 0, 64, &lrc);
        rc = (int)hipStreamEndCapture(c->stream, &g);
        if (!rc) rc = lrc;
        if (!rc) rc = (int)hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        if (g) (void)hipGraphDestroy(g);
        float ms = 1e30f;
        if (!rc) rc = (int)hipGraphLaunch(ge, c->stream);   // warm
        if (!rc) rc = (int)hipEventRecord(e0, c->stream);
        for (int r = 0; r < 3 && !rc; r++) rc = (int)hipGraphLaunch(ge, c->stream);
        if (!rc) rc = (int)hipEventRecord(e1, c->stream);
        if (!rc) rc = (int)hipEventSynchronize(e1);
        if (!rc) rc = (int)hipEventElapsedTime(&ms, e0, e1);
        if (ge) (void)hipGraphExecDestroy(ge);
        if (rc) break;
        if (ms < best) { best = ms; best_tb = tb; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c->tb = best_tb;
    return rc;
}

}  // namespace

extern "C" int lb_create(const lb_config *cfg, lb_ctx **out) {
    if (!cfg || !out) return LB_ERR_INVALID;
    *out = nullptr;
    if (lb_device_count() <= 0) return LB_ERR_NO_DEVICE;
    if (cfg->lattice != 9 && cfg->lattice != 19) return LB_ERR_INVALID;
    if (cfg->nx < 2 || cfg->ny < 1 || cfg->nz < 1) return LB_ERR_INVALID;
    if (cfg->lattice == 9 && cfg->nz != 1) return LB_ERR_INVALID;
    if (cfg->storage < 0 || cfg->storage > 4) return LB_ERR_INVALID;
    for (int k = 0; k < 6; k++) if (cfg->bc[k] < 0 || cfg->bc[k] > 4) return LB_ERR_INVALID;
    uint64_t N64 = (uint64_t)cfg->nx * cfg->ny * cfg->nz;
    if (N64 * (uint64_t)cfg->lattice >= (1ull << 31)) return LB_ERR_UNSUPPORTED;   // 32-bit node indices
    if (cfg->halo_z && cfg->lattice == 19 && (cfg->bc[4] != LB_BC_PERIODIC || cfg->bc[5] != LB_BC_PERIODIC))
        return LB_ERR_UNSUPPORTED;   // the single-slab ring realises periodic z
    // Shan-Chen reads the psi of z-neighbours, which the halo planes do not
    // carry: a ring of several slabs would need a psi-plane exchange too
    if (cfg->halo_z && cfg->lattice == 19 && cfg->sc_G != 0.0f) return LB_ERR_UNSUPPORTED;
    if (cfg->sc_sweep < -1 || cfg->sc_sweep > 3 || cfg->sc_zchunk < 0) return LB_ERR_INVALID;
    // the row-slab one-pass sweep: block_x = its workgroup (384, 512, 768 or 1,024
    // threads), block_y = rows per tile
    if (cfg->sc_sweep == 3) {
        if (!scr_lattice_ok(*cfg, (uint32_t)N64)) return LB_ERR_UNSUPPORTED;
        const int nw = cfg->block_x / 64;
        if (cfg->block_x > 0 && (cfg->block_x % 64 || (nw != 6 && nw != 8 && nw != 12 && nw != 16)))
            return LB_ERR_UNSUPPORTED;
        if (cfg->block_y > 0 && (cfg->block_x > 0 ? !scr_shape_ok(nw, cfg->nx, cfg->block_y, cfg->storage)
                                                   : !(scr_shape_ok(6, cfg->nx, cfg->block_y, cfg->storage) ||
                                                       scr_shape_ok(8, cfg->nx, cfg->block_y, cfg->storage) ||
                                                       scr_shape_ok(12, cfg->nx, cfg->block_y, cfg->storage) ||
                                                       scr_shape_ok(16, cfg->nx, cfg->block_y, cfg->storage))))
            return LB_ERR_UNSUPPORTED;
    } else if (cfg->block_x && (cfg->block_x & (cfg->block_x - 1) || cfg->block_x > kThreads)) {
        return LB_ERR_INVALID;
    }
    if ((cfg->sc_sweep == 1 || cfg->sc_sweep == 2) && (cfg->lattice != 19 || cfg->sc_G == 0.0f || !scz_storage(cfg->storage) ||
                               cfg->mem_mode > 0 ||
                               (cfg->block_x > 0 && !scz_shape_ok(cfg->block_x, cfg->block_y > 0 ? cfg->block_y
                                                                                       : kZThreads / cfg->block_x))))
        return LB_ERR_UNSUPPORTED;
    if (cfg->mem_mode > 2 || cfg->layout > LB_LAYOUT_ROWS || cfg->round_mode < 0 || cfg->round_mode > 2 ||
        cfg->node_words > 1 ||
        (cfg->time_block != -1 && cfg->time_block != 0 && cfg->time_block != 4 && cfg->time_block != 8))
        return LB_ERR_INVALID;
    lb_ctx *c = new (std::nothrow) lb_ctx();
    if (!c) return LB_ERR_INVALID;
    c->cfg = *cfg;
    c->sc_req = cfg->sc_sweep;
    c->Q = cfg->lattice;
    c->N = (uint32_t)N64;
    c->P = (uint32_t)cfg->nx * cfg->ny;
    c->esz = (cfg->storage == LB_STORE_F32 || cfg->storage == LB_STORE_F32S) ? 4 : (cfg->storage == LB_STORE_F64 ? 8 : 2);
    c->graph_steps = cfg->graph_chunk > 0 ? cfg->graph_chunk : 64;
    c->solid.assign(c->N, 0);
    int rc = (int)hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    const size_t pop_bytes = (size_t)c->Q * c->N * c->esz;
    if (!rc) rc = (int)hipMalloc(&c->buf[0], pop_bytes);
    if (!rc) rc = (int)hipMalloc(&c->buf[1], pop_bytes);
    if (!rc) rc = (int)hipMalloc((void **)&c->flags, c->N);
    if (!rc) rc = (int)hipMalloc((void **)&c->lmask, (size_t)c->N * 4);
    if (!rc) rc = (int)hipMalloc((void **)&c->nodew, (size_t)c->N * 4);
    if (!rc) rc = (int)hipMalloc((void **)&c->ctr, 2 * sizeof(uint32_t));
    if (!rc) rc = (int)hipMemsetAsync(c->ctr, 0, 2 * sizeof(uint32_t), c->stream);
    if (!rc) rc = (int)hipMalloc((void **)&c->macros, (size_t)4 * c->N * sizeof(float));
    if (!rc) rc = (int)hipMalloc((void **)&c->snap, (size_t)3 * c->N * sizeof(float));
    if (!rc) rc = (int)hipMalloc((void **)&c->partial, (size_t)kReduceBlocks * 4 * sizeof(double));
    if (!rc) rc = (int)hipMalloc((void **)&c->dout, 8 * sizeof(double));
    if (!rc) rc = (int)hipMalloc((void **)&c->tmp_fields, (size_t)4 * c->N * sizeof(float));
    if (!rc) rc = (int)hipMalloc(&c->dargs, 4 * lb_ctx::kArgBytes);
    if (!rc && cfg->sc_G != 0.0f) rc = (int)hipMalloc(&c->psi, (size_t)c->N * (cfg->storage == LB_STORE_F64 ? 8 : 4));
    if (!rc && cfg->halo_z && c->Q == 19) {
        for (int h = 0; h < 2 && !rc; h++) {
            rc = (int)hipMalloc(&c->halo[h], (size_t)10 * c->P * c->esz);
            if (!rc) rc = (int)hipMemsetAsync(c->halo[h], 0, (size_t)10 * c->P * c->esz, c->stream);
        }
    }
    if (!rc) rc = (int)hipMemsetAsync(c->buf[0], 0, pop_bytes, c->stream);
    if (!rc) rc = (int)hipMemsetAsync(c->buf[1], 0, pop_bytes, c->stream);
    if (!rc) rc = (int)hipMemsetAsync(c->macros, 0, (size_t)4 * c->N * sizeof(float), c->stream);
    if (!rc && c->psi) rc = (int)hipMemsetAsync(c->psi, 0, (size_t)c->N * (cfg->storage == LB_STORE_F64 ? 8 : 4), c->stream);
    if (!rc) rc = upload_flags(c);
    if (!rc) {
        const bool scq = cfg->sc_G != 0.0f && c->Q == 19;
        if (cfg->block_x > 0 && cfg->xcd_remap >= 0 && cfg->mem_mode >= 0 && cfg->layout >= 0 && cfg->node_words >= 0 &&
            (!scq || cfg->sc_sweep == 0 || (cfg->sc_sweep > 0 && cfg->sc_zchunk > 0))) {
            c->scz = scq ? std::max(0, cfg->sc_sweep) : 0;
            if (c->scz) c->zc = cfg->sc_zchunk;
            if (c->scz == 3) {
                c->scr_nw = cfg->block_x / 64;
                c->bx_shift = 6;
                c->by = cfg->block_y > 0 ? cfg->block_y : scr_max_rows(c->scr_nw, cfg->nx, cfg->storage);
            } else {
                c->bx_shift = ilog2(cfg->block_x);
                c->by = c->scz ? (cfg->block_y > 0 ? cfg->block_y : kZThreads / cfg->block_x)
                               : std::max(1, kThreads / cfg->block_x);
            }
            c->xcd = cfg->xcd_remap;
            c->mem = cfg->mem_mode;
            c->layout = cfg->layout;
            c->packed = cfg->node_words;
            c->dbx_shift = c->bx_shift; c->dby = c->by; c->dxcd = c->xcd; c->dmem = c->mem == 2 ? 2 : 0;
            rc = upload_args(c);
        } else {
            const bool sc_state = cfg->sc_G != 0.0f;
            rc = autotune(c);
            const size_t bytes[3] = {pop_bytes, pop_bytes, (size_t)4 * c->N * sizeof(float)};
            void *bufs[3] = {c->buf[0], c->buf[1], c->macros};
            for (int k = 0; k < 3 && sc_state && !rc; k++) rc = (int)hipMemsetAsync(bufs[k], 0, bytes[k], c->stream);
            if (sc_state && !rc && c->psi)
                rc = (int)hipMemsetAsync(c->psi, 0, (size_t)c->N * (cfg->storage == LB_STORE_F64 ? 8 : 4), c->stream);
        }
    }
    if (!rc && cfg->time_block > 0 && !tb_eligible(c, cfg->time_block)) rc = LB_ERR_UNSUPPORTED;
    if (!rc) {
        if (cfg->time_block > 0) c->tb = cfg->time_block;
        else if (cfg->time_block < 0 && tb_eligible(c, 4)) rc = tune_tb(c);
    }
    if (!rc) rc = (int)hipStreamSynchronize(c->stream);
    if (rc) { delete c; return rc; }
    c->cfg.time_block = c->tb;
    c->cfg.block_x = c->scz == 3 ? 64 * c->scr_nw : 1 << c->bx_shift;
    c->cfg.block_y = c->by;
    c->cfg.xcd_remap = c->xcd;
    c->cfg.mem_mode = c->mem;
    c->cfg.layout = c->layout;
    c->cfg.node_words = c->packed;
    c->cfg.graph_chunk = c->graph_steps;
    c->cfg.sc_sweep = c->scz;
    c->cfg.sc_zchunk = c->scz ? c->zc : 0;
    if (c->cfg.storage == LB_STORE_F16 || c->cfg.storage == LB_STORE_BF16)
        c->cfg.round_mode = stochastic_rounding(c->cfg) ? 1 : 2;
    *out = c;
    return 0;
}

extern "C" int lb_destroy(lb_ctx *c) {
    if (!c) return LB_ERR_INVALID;
    (void)hipStreamSynchronize(c->stream);
    delete c;
    return 0;
}

extern "C" int lb_get_config(lb_ctx *c, lb_config *out) {
    if (!c || !out) return LB_ERR_INVALID;
    *out = c->cfg;
    return 0;
}

extern "C" int lb_set_solid(lb_ctx *c, const uint8_t *solid) {
    if (!c || !solid) return LB_ERR_INVALID;
    for (uint32_t k = 0; k < c->N; k++) c->solid[k] = solid[k] ? 1 : 0;
    c->drop_graphs();
    if (c->scz == 3 && any_solid(c->solid)) {
        // the row-slab sweep has no solid nodes: the two passes take over
        // (same population layout), or the caller's explicit choice fails
        if (c->sc_req == 3) return LB_ERR_UNSUPPORTED;
        c->scz = 0;
        c->bx_shift = c->two_bx_shift; c->by = c->two_by; c->xcd = c->two_xcd; c->mem = c->two_mem;
        c->cfg.sc_sweep = 0;
        c->cfg.sc_zchunk = 0;
        c->cfg.block_x = 1 << c->bx_shift;
        c->cfg.block_y = c->by;
        c->cfg.xcd_remap = c->xcd;
        c->cfg.mem_mode = c->mem;
        if (int rc = upload_args(c)) return rc;
    }
    return upload_flags(c);
}

extern "C" int lb_init_equilibrium_device(lb_ctx *c, const float *r, const float *ux, const float *uy,
                                          const float *uz) {
    if (!c || !r || !ux || !uy) return LB_ERR_INVALID;
    int rc = c->Q == 9 ? launch_init_q<9>(c, r, ux, uy, uz) : launch_init_q<19>(c, r, ux, uy, uz);
    if (rc) return rc;
    c->cur = 0;
    c->steps = 0;
    c->macros_valid = false;
    c->snap_valid = false;
    rc = (int)hipMemsetAsync(c->ctr, 0, 2 * sizeof(uint32_t), c->stream);
    if (rc) return rc;
    if (c->halo[0]) {
        rc = (int)hipMemsetAsync(c->halo[0], 0, (size_t)10 * c->P * c->esz, c->stream);
        if (rc) return rc;
    }
    // A fresh halo: the planes the first sweep pulls across z come from the
    // initial state (as the copy after a sweep would leave them).
    rc = refill_recv(c);
    if (!rc) rc = (int)hipStreamSynchronize(c->stream);
    return rc;
}

extern "C" int lb_init_equilibrium(lb_ctx *c, const float *r, const float *ux, const float *uy, const float *uz) {
    if (!c || !r || !ux || !uy) return LB_ERR_INVALID;
    const size_t nb = (size_t)c->N * sizeof(float);
    float *t = c->tmp_fields;
    int rc = (int)hipMemcpyAsync(t, r, nb, hipMemcpyHostToDevice, c->stream);
    if (!rc) rc = (int)hipMemcpyAsync(t + c->N, ux, nb, hipMemcpyHostToDevice, c->stream);
    if (!rc) rc = (int)hipMemcpyAsync(t + 2 * (size_t)c->N, uy, nb, hipMemcpyHostToDevice, c->stream);
    if (!rc && uz) rc = (int)hipMemcpyAsync(t + 3 * (size_t)c->N, uz, nb, hipMemcpyHostToDevice, c->stream);
    if (rc) return rc;
    return lb_init_equilibrium_device(c, t, t + c->N, t + 2 * (size_t)c->N, uz ? t + 3 * (size_t)c->N : nullptr);
}

extern "C" int lb_init_uniform(lb_ctx *c, float rho, float ux, float uy, float uz) {
    if (!c) return LB_ERR_INVALID;
    std::vector<float> h((size_t)4 * c->N);
    std::fill(h.begin(), h.begin() + c->N, rho);
    std::fill(h.begin() + c->N, h.begin() + 2 * (size_t)c->N, ux);
    std::fill(h.begin() + 2 * (size_t)c->N, h.begin() + 3 * (size_t)c->N, uy);
    std::fill(h.begin() + 3 * (size_t)c->N, h.end(), uz);
    return lb_init_equilibrium(c, h.data(), h.data() + c->N, h.data() + 2 * (size_t)c->N, h.data() + 3 * (size_t)c->N);
}

extern "C" int lb_set_populations(lb_ctx *c, const void *host) {
    if (!c || !host) return LB_ERR_INVALID;
    size_t bytes = (size_t)c->Q * c->N * c->esz;
    int rc = 0;
    if (c->layout == LB_LAYOUT_SOA) {
        rc = (int)hipMemcpyAsync(c->buf[0], host, bytes, hipMemcpyHostToDevice, c->stream);
    } else {
        void *tmp = nullptr;
        rc = (int)hipMalloc(&tmp, bytes);
        if (!rc) rc = (int)hipMemcpyAsync(tmp, host, bytes, hipMemcpyHostToDevice, c->stream);
        if (!rc) rc = relayout(c, tmp, c->buf[0], true);
        if (!rc) rc = (int)hipStreamSynchronize(c->stream);
        if (tmp) (void)hipFree(tmp);
    }
    if (!rc) rc = (int)hipMemcpyAsync(c->buf[1], c->buf[0], bytes, hipMemcpyDeviceToDevice, c->stream);
    if (!rc) rc = (int)hipMemsetAsync(c->ctr, 0, 2 * sizeof(uint32_t), c->stream);
    c->cur = 0;
    c->steps = 0;
    c->macros_valid = false;
    c->snap_valid = false;
    if (!rc) rc = refill_recv(c);
    if (!rc) rc = (int)hipStreamSynchronize(c->stream);
    return rc;
}

extern "C" int lb_get_populations(lb_ctx *c, void *host) {
    if (!c || !host) return LB_ERR_INVALID;
    size_t bytes = (size_t)c->Q * c->N * c->esz;
    if (c->layout == LB_LAYOUT_SOA) {
        int rc = (int)hipMemcpyAsync(host, c->buf[c->cur], bytes, hipMemcpyDeviceToHost, c->stream);
        if (!rc) rc = (int)hipStreamSynchronize(c->stream);
        return rc;
    }
    void *tmp = nullptr;
    int rc = (int)hipMalloc(&tmp, bytes);
    if (!rc) rc = relayout(c, c->buf[c->cur], tmp, false);
    if (!rc) rc = (int)hipMemcpyAsync(host, tmp, bytes, hipMemcpyDeviceToHost, c->stream);
    if (!rc) rc = (int)hipStreamSynchronize(c->stream);
    if (tmp) (void)hipFree(tmp);
    return rc;
}

extern "C" int lb_step(lb_ctx *c, int nsteps) {
    if (!c || nsteps < 0) return LB_ERR_INVALID;
    if (nsteps == 0) return 0;
    const int chunk = c->graph_steps;
    const int full = chunk > 0 ? (nsteps - 1) / chunk * chunk : 0;
    int rc = run_graph_or_eager(c, full);
    if (rc) return rc;
    const int r = nsteps - 1 - full;
    lb_ctx::Tail t{};
    rc = tail_graph(c, c->cur, r, &t);
    if (!rc) rc = (int)hipGraphLaunch(t.g, c->stream);
    if (rc) return rc;
    c->cur = t.end;
    c->steps += r + 1;
    c->macros_valid = true;
    return 0;
}

extern "C" int lb_step_prepare(lb_ctx *c, int nsteps) {
    if (!c || nsteps < 0) return LB_ERR_INVALID;
    if (nsteps == 0) return 0;
    const int chunk = c->graph_steps;
    int cur = c->cur, n = chunk > 0 ? (nsteps - 1) / chunk : 0;
    for (int k = 0; k < n; k++) {
        hipGraphExec_t ge = nullptr;
        if (int rc = chunk_graph(c, cur, &ge)) return rc;
        cur = c->graph_end[cur];
    }
    lb_ctx::Tail t{};
    return tail_graph(c, cur, nsteps - 1 - n * chunk, &t);
}

extern "C" int lb_sync(lb_ctx *c) {
    if (!c) return LB_ERR_INVALID;
    return (int)hipStreamSynchronize(c->stream);
}

extern "C" int lb_get_macros(lb_ctx *c, float *rho, float *ux, float *uy, float *uz) {
    if (!c) return LB_ERR_INVALID;
    if (!c->macros_valid) return LB_ERR_NOT_READY;
    float *dst[4] = {rho, ux, uy, uz};
    int rc = 0;
    for (int k = 0; k < 4 && !rc; k++)
        if (dst[k]) rc = (int)hipMemcpyAsync(dst[k], c->macros + (size_t)k * c->N, (size_t)c->N * 4, hipMemcpyDeviceToHost, c->stream);
    if (!rc) rc = (int)hipStreamSynchronize(c->stream);
    return rc;
}

extern "C" int lb_macros_device(lb_ctx *c, float **out) {
    if (!c || !out) return LB_ERR_INVALID;
    *out = c->macros;
    return 0;
}

extern "C" int lb_reduce(lb_ctx *c, double out[4]) {
    if (!c || !out) return LB_ERR_INVALID;
    int rc = c->Q == 9 ? launch_reduce_q<9>(c) : launch_reduce_q<19>(c);
    if (!rc) rc = (int)hipMemcpyAsync(out, c->dout, 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (!rc) rc = (int)hipStreamSynchronize(c->stream);
    return rc;
}

extern "C" int lb_residual(lb_ctx *c, double *out) {
    if (!c || !out) return LB_ERR_INVALID;
    if (!c->macros_valid) return LB_ERR_NOT_READY;
    const int ncomp = c->Q == 9 ? 2 : 3;
    hipLaunchKernelGGL(k_residual, dim3(kReduceBlocks), dim3(kThreads), 0, c->stream, c->macros, c->snap, c->flags,
                       c->N, ncomp, c->partial);
    hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(64), 0, c->stream, c->partial, kReduceBlocks, 2, c->dout);
    double h[2] = {0, 0};
    int rc = (int)hipGetLastError();
    if (!rc) rc = (int)hipMemcpyAsync(h, c->dout, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (!rc) rc = (int)hipStreamSynchronize(c->stream);
    if (rc) return rc;
    *out = !c->snap_valid ? 1.0 : (h[1] > 0 ? std::sqrt(h[0] / h[1]) : 0.0);
    c->snap_valid = true;
    return 0;
}

extern "C" int lb_read_value(lb_ctx *c, int x, int y, int z, float out[4]) {
    if (!c || !out) return LB_ERR_INVALID;
    if (x < 0 || y < 0 || z < 0 || x >= c->cfg.nx || y >= c->cfg.ny || z >= c->cfg.nz) return LB_ERR_INVALID;
    if (!c->macros_valid) return LB_ERR_NOT_READY;
    size_t id = x + (size_t)c->cfg.nx * (y + (size_t)c->cfg.ny * z);
    const int comp[4] = {1, 2, 3, 0};   // ux, uy, uz, rho (read_value.cu order)
    int rc = 0;
    for (int k = 0; k < 4 && !rc; k++)
        rc = (int)hipMemcpyAsync(out + k, c->macros + (size_t)comp[k] * c->N + id, 4, hipMemcpyDeviceToHost, c->stream);
    if (!rc) rc = (int)hipStreamSynchronize(c->stream);
    return rc;
}

extern "C" int lb_steps_done(lb_ctx *c, long long *out) {
    if (!c || !out) return LB_ERR_INVALID;
    *out = c->steps;
    return 0;
}

extern "C" int lb_time_step_kernel(lb_ctx *c, int n, float *ms) {
    if (!c || !ms || n <= 0) return LB_ERR_INVALID;
    // Only the sweep kernels (the halo copy is excluded), back to back
    // through a captured graph so launch gaps do not count as kernel time.
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    int rc = (int)hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
    if (rc) return rc;
    int src = c->cur;
    for (int s = 0; s < n; s++) {
        if (c->Q == 9) launch_step_q<9>(c, src, false, (uint32_t)s);
        else launch_step_q<19>(c, src, false, (uint32_t)s);
        src = 1 - src;
    }
    launch_bump(c, (uint32_t)n);
    rc = (int)hipStreamEndCapture(c->stream, &g);
    if (rc) return rc;
    rc = (int)hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (rc) return rc;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    rc = (int)hipGraphLaunch(ge, c->stream);   // warm
    if (!rc) rc = (int)hipEventRecord(e0, c->stream);
    if (!rc) rc = (int)hipGraphLaunch(ge, c->stream);
    if (!rc) rc = (int)hipEventRecord(e1, c->stream);
    if (!rc) rc = (int)hipEventSynchronize(e1);
    float t = 0;
    if (!rc) rc = (int)hipEventElapsedTime(&t, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipGraphExecDestroy(ge);
    // 2n sweeps ran: the state advanced; keep the buffer parity consistent.
    if ((2 * n) & 1) c->cur = 1 - c->cur;
    c->steps += 2 * (long long)n;
    c->macros_valid = false;
    *ms = t / (float)n;
    return rc;
}

extern "C" int lb_halo_buffers(lb_ctx *c, void **send_lo, void **send_hi, void **recv_lo, void **recv_hi) {
    if (!c) return LB_ERR_INVALID;
    if (!c->halo[0]) return LB_ERR_UNSUPPORTED;
    const size_t pb5 = (size_t)5 * c->P * c->esz;
    void *send = c->halo[c->cur], *recv = c->halo[1 - c->cur];   // the next sweep's pair
    if (send_lo) *send_lo = send;
    if (send_hi) *send_hi = (char *)send + pb5;
    if (recv_hi) *recv_hi = recv;
    if (recv_lo) *recv_lo = (char *)recv + pb5;
    return 0;
}
