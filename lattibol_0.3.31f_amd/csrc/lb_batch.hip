// lb_batch.hip -- many independent D2Q9 channels of uneven size stepped
// together: one launch per time step covers every tile of every active
// lattice (a block table maps workgroup -> (lattice, tile)).  Run to steady
// state, the whole loop lives on the device: one captured hipGraph holds
// many chunks of k steps, each followed by the residual, the retirement of
// converged (or diverged) lattices and the compaction of the block table,
// all by kernels; the graph is built once and replayed, and the host reads
// two counters per replay.
//
// Each lattice runs the exact per-node code of the single-lattice solver
// (step_body in lb_pull.hpp) with the configuration of lattice_config():
//   Zou-He velocity inlet u = (u_in, 0) at x = 0 and pressure outlet rho = 1
//   at x = nx-1 (the reference's pair, collision.cu:211-262), rows 0 and ny-1
//   half-way bounce-back walls, solids half-way bounce-back, rest closure,
//   FP32 deviation storage (s_i = f_i - w_i, collision in deviation form),
// so a lattice of the batch is bit-identical to lb_create() with that
// configuration.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "lb_pull.hpp"

using namespace lb;

namespace {

constexpr int kBX = 64, kBY = 4;   // 256-thread tiles of 64 x 4 nodes

// Per-lattice state of a run to steady state, kept on the device: still
// stepping, steps taken, snapshot primed, buffer holding the current state,
// last residual (1 until a snapshot exists, +inf when diverged).
struct BatchState { int active, steps, snap_valid, cur; double res; };

struct BatchLat {
    int nx, ny;
    uint32_t N;
    float ti, u_in;
    uint64_t pop_off;     // element offset of population 0 in A / B
    uint64_t node_off;    // node offset in flags / fields
    int ntx, nty;
};

// Configuration of every lattice of the batch (also used to build its node
// flags, so a member is identical to lb_create() with this configuration).
inline lb_config lattice_config(int nx, int ny, float ti, float u_in) {
    lb_config c{};
    c.lattice = 9; c.nx = nx; c.ny = ny; c.nz = 1;
    c.storage = LB_STORE_F32S;
    c.tau_inv = ti;
    c.bc[0] = LB_BC_OUTFLOW; c.bc[1] = LB_BC_OUTFLOW; c.bc[2] = LB_BC_WALL; c.bc[3] = LB_BC_WALL;
    c.zou_he = 1; c.zh_ux_in = u_in; c.zh_rho_out = 1.0f;
    c.rest_closure = 1;
    return c;
}

// Argument block of lattice L reading buffer `src` (A = 0 / B = 1).
StepArgs<StF32S> lattice_args(const BatchLat &L, float *A, float *B, int src, const uint8_t *flags,
                              const uint32_t *lmask, float *macros, int layout) {
    const lb_config c = lattice_config(L.nx, L.ny, L.ti, L.u_in);
    StepArgs<StF32S> a{};
    float *S0 = src == 0 ? A : B, *S1 = src == 0 ? B : A;
    a.A = S0 + L.pop_off;
    a.B = S1 + L.pop_off;
    a.flags = flags + L.node_off;
    a.lmask = lmask + L.node_off;
    a.macros = macros + 4 * L.node_off;
    a.N = L.N;
    a.nxy = L.N;
    a.P = (uint32_t)L.nx;
    a.nx = L.nx; a.ny = L.ny; a.nz = 1;
    a.ti = c.tau_inv;
    for (int k = 0; k < 6; k++) {
        a.bc[k] = c.bc[k];
        a.bcrho[k] = c.bc_rho[k];
    }
    const PopLayout pl = pop_layout(layout, 9, L.nx, L.ny, 1);
    a.ps = pl.ps; a.sy = pl.sy; a.sz = pl.sz;
    face_corrections(c, pl, a.corr_lo, a.corr_hi);
    a.zou_he = 1; a.zh_ux = c.zh_ux_in; a.zh_rho = c.zh_rho_out;
    a.rest_closure = 1;
    a.bx_shift = 6; a.by = kBY;
    a.ntx = L.ntx; a.nty = L.nty;
    a.nblocks = L.ntx * L.nty;
    fast_path_args<9>(a, 1 << a.bx_shift, a.by);
    return a;
}

// counts (device, may be null): {active lattices, active tiles} of the
// compacted table; workgroups past the active tiles exit at once.
template <bool MACROS, int MEM>
__global__ __launch_bounds__(kThreads) void k_batch_step(const StepArgs<StF32S> *__restrict__ args,
                                                         const int2 *__restrict__ table, int src,
                                                         const int *__restrict__ counts) {
    if (counts && (int)blockIdx.x >= counts[1]) return;
    const int2 e = table[blockIdx.x];          // (lattice, local tile)
    const StepArgs<StF32S> &a = args[2 * e.x + src];
    step_body<9, StF32S, false, MACROS, MEM>(a, node_of(a, e.y));
}

// Per-tile residual partials of the velocity change since the snapshot, plus
// the largest |u| (a diverging lattice can hold huge, slowly changing values
// that the relative residual alone would call converged).
__global__ __launch_bounds__(kThreads) void k_batch_residual(const BatchLat *lats, const int2 *table,
                                                             const float *macros, float *snap,
                                                             const uint8_t *flags, double *partial,
                                                             const int *counts) {
    if ((int)blockIdx.x >= counts[1]) return;
    const int2 e = table[blockIdx.x];
    const BatchLat L = lats[e.x];
    int tx = e.y % L.ntx, ty = e.y / L.ntx;
    int x = tx * kBX + (threadIdx.x & (kBX - 1)), y = ty * kBY + (threadIdx.x >> 6);
    double num = 0, den = 0, umax = 0;
    if (x < L.nx && y < L.ny) {
        uint32_t id = (uint32_t)x + (uint32_t)L.nx * (uint32_t)y;
        if (!(flags[L.node_off + id] & F_SOLID)) {
            const float *m = macros + 4 * L.node_off;
            float *sn = snap + 2 * L.node_off;
            for (int c = 0; c < 2; c++) {
                double u = m[(size_t)(1 + c) * L.N + id];
                double s = sn[(size_t)c * L.N + id];
                num += (u - s) * (u - s);
                den += u * u;
                double au = fabs(u);
                umax = (au <= 1e30) ? fmax(umax, au) : INFINITY;   // NaN / inf -> inf
                sn[(size_t)c * L.N + id] = (float)u;
            }
        }
    }
    __shared__ double sh[3][kThreads / 64];
    double s0 = wave_sum(num), s1 = wave_sum(den);
    double s2 = umax;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s2 = fmax(s2, __shfl_xor(s2, off, 64));
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) { sh[0][w] = s0; sh[1][w] = s1; sh[2][w] = s2; }
    __syncthreads();
    if (threadIdx.x < 3) {
        double t = 0;
        for (int k = 0; k < kThreads / 64; k++) t = threadIdx.x < 2 ? t + sh[threadIdx.x][k] : fmax(t, sh[2][k]);
        partial[(size_t)blockIdx.x * 3 + threadIdx.x] = t;
    }
}

// One thread per active lattice: fixed-order sum over its tiles, then the
// lattice's state: steps += k, the buffer now holding it, the residual, and
// retirement when converged or diverged.  A lattice whose |u| exceeds
// kUDiverged (Mach ~0.9) or turns non-finite reports +inf.  The first check
// after an initialisation only primes the snapshot (residual 1).
constexpr double kUDiverged = 0.5;
__global__ void k_batch_final(const int *act, const int2 *ranges, const int *counts, const double *partial,
                              BatchState *st, double tol, int k, int parity_after) {
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= counts[0]) return;
    int2 r = ranges[a];   // (first block, count)
    double num = 0, den = 0, umax = 0;
    for (int b = r.x; b < r.x + r.y; b++) {
        num += partial[3 * b];
        den += partial[3 * b + 1];
        umax = fmax(umax, partial[3 * b + 2]);
    }
    // divergence is judged on the velocities alone (umax is +inf when any
    // velocity is NaN or inf), never on the residual value itself
    double v = den > 0 ? sqrt(num / den) : 0.0;
    if (umax > kUDiverged) v = INFINITY;
    BatchState &s = st[act[a]];
    s.steps += k;
    s.cur = parity_after;
    const bool diverged = isinf(v);
    s.res = (s.snap_valid || diverged) ? v : 1.0;
    if (diverged || (s.snap_valid && v < tol)) s.active = 0;
    s.snap_valid = 1;
}

// Compacted work list of the active lattices, in lattice order (one
// workgroup): act[a] = lattice, ranges[a] = (first tile, tile count),
// table[first + t] = (lattice, t), counts = {lattices, tiles}.  The same
// table the host builds, so residual sums run in the same order.
constexpr int kCompactThreads = 1024;
__global__ __launch_bounds__(kCompactThreads) void k_batch_compact(const BatchLat *lats, const BatchState *st, int n,
                                                                   int *act, int2 *ranges, int2 *table, int *counts) {
    __shared__ int s_l[kCompactThreads], s_t[kCompactThreads];
    __shared__ int base_l, base_t;
    const int t = threadIdx.x;
    if (t == 0) { base_l = 0; base_t = 0; }
    __syncthreads();
    for (int c0 = 0; c0 < n; c0 += kCompactThreads) {
        const int i = c0 + t;
        const int on = (i < n && st[i].active) ? 1 : 0;
        const int nt = on ? lats[i].ntx * lats[i].nty : 0;
        s_l[t] = on;
        s_t[t] = nt;
        __syncthreads();
        for (int off = 1; off < kCompactThreads; off <<= 1) {   // inclusive scans
            const int al = t >= off ? s_l[t - off] : 0, at = t >= off ? s_t[t - off] : 0;
            __syncthreads();
            s_l[t] += al;
            s_t[t] += at;
            __syncthreads();
        }
        if (on) {
            const int a = base_l + s_l[t] - 1, first = base_t + s_t[t] - nt;
            act[a] = i;
            ranges[a] = int2{first, nt};
            for (int q = 0; q < nt; q++) table[first + q] = int2{i, q};
        }
        __syncthreads();
        if (t == 0) { base_l += s_l[kCompactThreads - 1]; base_t += s_t[kCompactThreads - 1]; }
        __syncthreads();
    }
    if (t == 0) { counts[0] = base_l; counts[1] = base_t; }
}

// Initial populations from (rho, ux, uy) fields (NULL: rest).  neq = 0:
// equilibrium.  neq = 1: equilibrium plus the first-order Chapman-Enskog
// non-equilibrium part implied by the velocity field.  Before collision it
// is f_i^neq = -(w_i rho tau / cs^2) Q_i:S (Q_i = c_i c_i - cs^2 I,
// S = (grad u + grad u^T) / 2); the buffers hold post-collision values,
// (1 - 1/tau) times that, i.e. -(w_i rho (tau - 1) / cs^2) Q_i:S.  Central
// differences, one-sided next to solids and lattice edges.  Physical axes:
// y = -row.
__device__ float grad_1d(const float *u, uint32_t id, int stride, bool has_m, bool has_p) {
    if (has_m && has_p) return 0.5f * (u[id + stride] - u[id - stride]);
    if (has_p) return u[id + stride] - u[id];
    if (has_m) return u[id] - u[id - stride];
    return 0.0f;
}

__global__ __launch_bounds__(kThreads) void k_batch_init(const BatchLat *lats, const int2 *table, float *A, float *B,
                                                         const float *fields, const uint8_t *flags, int neq,
                                                         int layout) {
    const int2 e = table[blockIdx.x];
    const BatchLat L = lats[e.x];
    int tx = e.y % L.ntx, ty = e.y / L.ntx;
    int x = tx * kBX + (threadIdx.x & (kBX - 1)), y = ty * kBY + (threadIdx.x >> 6);
    if (x >= L.nx || y >= L.ny) return;
    uint32_t id = (uint32_t)x + (uint32_t)L.nx * (uint32_t)y;
    float rho = 1.0f, ux = 0.0f, uy = 0.0f;
    float f[9];
    if (fields) {
        const float *F = fields + 3 * L.node_off;
        rho = F[id]; ux = F[L.N + id]; uy = F[2 * (size_t)L.N + id];
    }
    feq_dev<9, float>(f, rho - 1.0f, ux, uy, 0.0f);
    const uint8_t *fl = flags + L.node_off;
    if (fields && neq && !(fl[id] & F_SOLID)) {
        const float *Ux = fields + 3 * L.node_off + L.N, *Uy = Ux + L.N;
        const bool xm = x > 0 && !(fl[id - 1] & F_SOLID), xp = x < L.nx - 1 && !(fl[id + 1] & F_SOLID);
        const bool rm = y > 0 && !(fl[id - L.nx] & F_SOLID), rp = y < L.ny - 1 && !(fl[id + L.nx] & F_SOLID);
        const int nx = L.nx;
        const float dxux = grad_1d(Ux, id, 1, xm, xp), dxuy = grad_1d(Uy, id, 1, xm, xp);
        const float dyux = -grad_1d(Ux, id, nx, rm, rp), dyuy = -grad_1d(Uy, id, nx, rm, rp);
        const float sxx = dxux, syy = dyuy, sxy = 0.5f * (dxuy + dyux);
        const float tau_m1 = 1.0f / L.ti - 1.0f;
        static_for<0, 9>([&](auto I) {
            constexpr int i = decltype(I)::value;
            constexpr float cx = (float)Vel<9>::cx(i), cy = (float)Vel<9>::pcy(i);
            const float qs = cx * cx * sxx + 2.0f * cx * cy * sxy + cy * cy * syy - (sxx + syy) * (1.0f / 3.0f);
            f[i] = f[i] - 3.0f * weight<9, float>(Vel<9>::wc(i)) * rho * tau_m1 * qs;
        });
    }
    const PopLayout pl = pop_layout(layout, 9, L.nx, L.ny, 1);
    const uint32_t e0 = (uint32_t)x + pl.sy * (uint32_t)y;
    for (int i = 0; i < 9; i++) {
        A[L.pop_off + (size_t)i * pl.ps + e0] = f[i];
        B[L.pop_off + (size_t)i * pl.ps + e0] = f[i];
    }
}

}  // namespace

struct lb_batch {
    int n = 0;
    std::vector<BatchLat> lats;
    std::vector<int> cur;          // buffer holding each lattice's current state
    std::vector<char> active;
    std::vector<char> snap_valid;
    int parity = 0;                // buffer the active lattices read next
    uint64_t pop_total = 0, node_total = 0;
    hipStream_t stream = nullptr;
    BatchLat *d_lats = nullptr;
    StepArgs<StF32S> *d_args = nullptr;   // 2 per lattice (source buffer A / B)
    int2 *d_table = nullptr, *d_ranges = nullptr;
    float *A = nullptr, *B = nullptr, *macros = nullptr, *snap = nullptr;
    uint8_t *flags = nullptr;
    uint32_t *lmask = nullptr;
    double *partial = nullptr;
    BatchState *d_state = nullptr;
    int *d_act = nullptr, *d_counts = nullptr;   // compacted active lattices; {lattices, tiles}
    std::vector<int2> table, ranges;
    std::vector<int> act_list;
    int nblocks = 0, total_tiles = 0;
    // graphs of the device-side loop: `chunks` chunks of k steps from buffer parity
    struct LoopGraph { int chunks, k, parity; double tol; hipGraphExec_t g; };
    std::vector<LoopGraph> loops;
    std::vector<char> macros_valid;
    int layout = LB_LAYOUT_SOA;    // population layout of every lattice (autotuned)
    int mem = 0;                   // load / store hints (autotuned, see ldm / stm)

    ~lb_batch() {
        for (auto &l : loops) (void)hipGraphExecDestroy(l.g);
        void *ps[] = {d_lats, d_args, d_table, d_ranges, A, B, macros, snap, flags, lmask, partial, d_state, d_act,
                      d_counts};
        for (void *p : ps) if (p) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

// Block table over the active lattices (all lattices when all_lattices).
int rebuild_table(lb_batch *b, bool all_lattices = false) {
    b->table.clear();
    b->ranges.clear();
    b->act_list.clear();
    for (int i = 0; i < b->n; i++) {
        if (!all_lattices && !b->active[i]) continue;
        const BatchLat &L = b->lats[i];
        int nt = L.ntx * L.nty;
        b->ranges.push_back(int2{(int)b->table.size(), nt});
        b->act_list.push_back(i);
        for (int t = 0; t < nt; t++) b->table.push_back(int2{i, t});
    }
    b->nblocks = (int)b->table.size();
    if (b->nblocks == 0) return 0;
    int rc = (int)hipMemcpyAsync(b->d_table, b->table.data(), b->table.size() * sizeof(int2), hipMemcpyHostToDevice, b->stream);
    if (!rc) rc = (int)hipMemcpyAsync(b->d_ranges, b->ranges.data(), b->ranges.size() * sizeof(int2), hipMemcpyHostToDevice, b->stream);
    if (!rc) rc = (int)hipStreamSynchronize(b->stream);
    return rc;
}

// One step of the lattices in the block table: the host-built table
// (counts null, nblocks workgroups) or the device-compacted one (a grid over
// every tile of every lattice, workgroups past counts[1] idle).
void launch_batch_step(lb_batch *b, int parity, bool macros, const int *counts = nullptr) {
    const dim3 g(counts ? b->total_tiles : b->nblocks), t(kThreads);
    const StepArgs<StF32S> *a = b->d_args;
    if (macros) hipLaunchKernelGGL((k_batch_step<true, 0>), g, t, 0, b->stream, a, b->d_table, parity, counts);
    else if (b->mem == 1) hipLaunchKernelGGL((k_batch_step<false, 1>), g, t, 0, b->stream, a, b->d_table, parity, counts);
    else if (b->mem == 2) hipLaunchKernelGGL((k_batch_step<false, 2>), g, t, 0, b->stream, a, b->d_table, parity, counts);
    else hipLaunchKernelGGL((k_batch_step<false, 0>), g, t, 0, b->stream, a, b->d_table, parity, counts);
}

// Both argument blocks of every lattice for the current layout.
int upload_batch_args(lb_batch *b) {
    std::vector<StepArgs<StF32S>> hargs(2 * (size_t)b->n);
    for (int i = 0; i < b->n; i++)
        for (int src = 0; src < 2; src++)
            hargs[2 * i + src] = lattice_args(b->lats[i], b->A, b->B, src, b->flags, b->lmask, b->macros, b->layout);
    int rc = (int)hipMemcpyAsync(b->d_args, hargs.data(), hargs.size() * sizeof(StepArgs<StF32S>),
                                 hipMemcpyHostToDevice, b->stream);
    if (!rc) rc = (int)hipStreamSynchronize(b->stream);
    return rc;
}

// Layout x load/store-hint autotuner over a few steps of the whole batch
// (at creation, before any populations are set: nothing to relayout).
int autotune_batch(lb_batch *b) {
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0)) return LB_ERR_INVALID;
    if (hipEventCreate(&e1)) { (void)hipEventDestroy(e0); return LB_ERR_INVALID; }
    float best = 1e30f;
    int best_layout = 0, best_mem = 0;
    for (int lay = 0; lay < 2; lay++)
        for (int mem = 0; mem < 3; mem++) {
            b->layout = lay;
            b->mem = mem;
            if (int rc = upload_batch_args(b)) {
                (void)hipEventDestroy(e0);
                (void)hipEventDestroy(e1);
                return rc;
            }
            launch_batch_step(b, 0, false);
            (void)hipEventRecord(e0, b->stream);
            for (int r = 0; r < 6; r++) launch_batch_step(b, r & 1, false);
            (void)hipEventRecord(e1, b->stream);
            (void)hipEventSynchronize(e1);
            float ms = 0;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (ms < best) { best = ms; best_layout = lay; best_mem = mem; }
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->layout = best_layout;
    b->mem = best_mem;
    int rc = (int)hipGetLastError();
    if (!rc) rc = upload_batch_args(b);
    return rc;
}

// One chunk of the device-side loop: k steps from buffer parity (the last
// writes the macros), residual partials, per-lattice state, compaction.
void launch_loop_chunk(lb_batch *b, int parity, int k, double tol) {
    int par = parity;
    for (int s = 0; s < k; s++) { launch_batch_step(b, par, s == k - 1, b->d_counts); par = 1 - par; }
    hipLaunchKernelGGL(k_batch_residual, dim3(b->total_tiles), dim3(kThreads), 0, b->stream, b->d_lats, b->d_table,
                       b->macros, b->snap, b->flags, b->partial, b->d_counts);
    hipLaunchKernelGGL(k_batch_final, dim3((b->n + 63) / 64), dim3(64), 0, b->stream, b->d_act, b->d_ranges, b->d_counts,
                       b->partial, b->d_state, tol, k, par);
    hipLaunchKernelGGL(k_batch_compact, dim3(1), dim3(kCompactThreads), 0, b->stream, b->d_lats, b->d_state, b->n, b->d_act,
                       b->d_ranges, b->d_table, b->d_counts);
}

// The captured loop of `chunks` chunks from buffer parity (built once).
int loop_graph(lb_batch *b, int chunks, int k, int parity, double tol, hipGraphExec_t *out) {
    for (const auto &l : b->loops)
        if (l.chunks == chunks && l.k == k && l.parity == parity && l.tol == tol) { *out = l.g; return 0; }
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    int rc = (int)hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
    if (rc) return rc;
    int par = parity;
    for (int c = 0; c < chunks; c++) {
        launch_loop_chunk(b, par, k, tol);
        if (k & 1) par = 1 - par;
    }
    const int lrc = (int)hipGetLastError();
    rc = (int)hipStreamEndCapture(b->stream, &g);
    if (!rc) rc = lrc;
    if (!rc) rc = (int)hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    if (g) (void)hipGraphDestroy(g);
    if (rc) return rc;
    if (b->loops.size() >= 4) { (void)hipGraphExecDestroy(b->loops.front().g); b->loops.erase(b->loops.begin()); }
    b->loops.push_back({chunks, k, parity, tol, ge});
    *out = ge;
    return 0;
}

}  // namespace

extern "C" int lb_batch_create(int n, const lb_batch_lattice *lat, const uint8_t *solids, lb_batch **out) {
    if (n <= 0 || !lat || !solids || !out) return LB_ERR_INVALID;
    *out = nullptr;
    if (lb_device_count() <= 0) return LB_ERR_NO_DEVICE;
    lb_batch *b = new (std::nothrow) lb_batch();
    if (!b) return LB_ERR_INVALID;
    b->n = n;
    b->lats.resize(n);
    uint64_t pop = 0, node = 0;
    size_t total_tiles = 0;
    std::vector<uint8_t> flags_all;
    std::vector<uint32_t> lmask_all;
    for (int i = 0; i < n; i++) {
        const lb_batch_lattice &s = lat[i];
        if (s.nx < 2 || s.ny < 3 || (uint64_t)s.nx * s.ny * 9 >= (1ull << 31)) { delete b; return LB_ERR_INVALID; }
        BatchLat &L = b->lats[i];
        L.nx = s.nx; L.ny = s.ny; L.N = (uint32_t)s.nx * s.ny;
        L.ti = s.tau_inv; L.u_in = s.u_in;
        L.pop_off = pop; L.node_off = node;
        L.ntx = (s.nx + kBX - 1) / kBX; L.nty = (s.ny + kBY - 1) / kBY;
        total_tiles += (size_t)L.ntx * L.nty;
        // node flags with the same rules as the single-lattice solver
        const lb_config cfg = lattice_config(s.nx, s.ny, s.tau_inv, s.u_in);
        std::vector<uint8_t> sol(solids + s.solid_offset, solids + s.solid_offset + L.N), fl;
        for (auto &v : sol) v = v ? 1 : 0;
        std::vector<uint32_t> lm;
        build_flags(cfg, sol, fl, lm);
        flags_all.insert(flags_all.end(), fl.begin(), fl.end());
        lmask_all.insert(lmask_all.end(), lm.begin(), lm.end());
        pop += 9ull * L.N;
        node += L.N;
    }
    if (total_tiles >= (size_t)INT32_MAX) { delete b; return LB_ERR_UNSUPPORTED; }
    b->total_tiles = (int)total_tiles;
    b->pop_total = pop;
    b->node_total = node;
    b->cur.assign(n, 0);
    b->active.assign(n, 1);
    b->snap_valid.assign(n, 0);
    b->macros_valid.assign(n, 0);
    int rc = (int)hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (!rc) rc = (int)hipMalloc((void **)&b->d_lats, n * sizeof(BatchLat));
    if (!rc) rc = (int)hipMalloc((void **)&b->d_args, 2 * n * sizeof(StepArgs<StF32S>));
    if (!rc) rc = (int)hipMalloc((void **)&b->d_table, total_tiles * sizeof(int2));
    if (!rc) rc = (int)hipMalloc((void **)&b->d_ranges, n * sizeof(int2));
    if (!rc) rc = (int)hipMalloc((void **)&b->A, pop * sizeof(float));
    if (!rc) rc = (int)hipMalloc((void **)&b->B, pop * sizeof(float));
    if (!rc) rc = (int)hipMalloc((void **)&b->macros, 4 * node * sizeof(float));
    if (!rc) rc = (int)hipMalloc((void **)&b->snap, 2 * node * sizeof(float));
    if (!rc) rc = (int)hipMalloc((void **)&b->flags, node);
    if (!rc) rc = (int)hipMalloc((void **)&b->lmask, node * 4);
    if (!rc) rc = (int)hipMalloc((void **)&b->partial, total_tiles * 3 * sizeof(double));
    if (!rc) rc = (int)hipMalloc((void **)&b->d_state, n * sizeof(BatchState));
    if (!rc) rc = (int)hipMalloc((void **)&b->d_act, n * sizeof(int));
    if (!rc) rc = (int)hipMalloc((void **)&b->d_counts, 2 * sizeof(int));
    if (!rc) rc = (int)hipMemcpyAsync(b->d_lats, b->lats.data(), n * sizeof(BatchLat), hipMemcpyHostToDevice, b->stream);
    if (!rc) rc = (int)hipMemcpyAsync(b->flags, flags_all.data(), node, hipMemcpyHostToDevice, b->stream);
    if (!rc) rc = (int)hipMemcpyAsync(b->lmask, lmask_all.data(), node * 4, hipMemcpyHostToDevice, b->stream);
    if (!rc) rc = (int)hipMemsetAsync(b->macros, 0, 4 * node * sizeof(float), b->stream);
    if (!rc) rc = (int)hipStreamSynchronize(b->stream);
    if (!rc) rc = (int)hipMemsetAsync(b->A, 0, pop * sizeof(float), b->stream);
    if (!rc) rc = (int)hipMemsetAsync(b->B, 0, pop * sizeof(float), b->stream);
    if (!rc) rc = rebuild_table(b);
    if (!rc) rc = autotune_batch(b);
    if (!rc) rc = lb_batch_init_rest(b);
    if (rc) { delete b; return rc; }
    *out = b;
    return 0;
}

extern "C" int lb_batch_destroy(lb_batch *b) {
    if (!b) return LB_ERR_INVALID;
    (void)hipStreamSynchronize(b->stream);
    delete b;
    return 0;
}

static int batch_init(lb_batch *b, const float *fields, int neq) {
    std::fill(b->active.begin(), b->active.end(), 1);
    std::fill(b->snap_valid.begin(), b->snap_valid.end(), 0);
    std::fill(b->macros_valid.begin(), b->macros_valid.end(), 0);
    std::fill(b->cur.begin(), b->cur.end(), 0);
    b->parity = 0;
    int rc = rebuild_table(b);
    if (rc) return rc;
    hipLaunchKernelGGL(k_batch_init, dim3(b->nblocks), dim3(kThreads), 0, b->stream, b->d_lats, b->d_table, b->A, b->B, fields,
                       b->flags, neq, b->layout);
    rc = (int)hipGetLastError();
    // the first residual check only primes the snapshot, but its value must
    // still be finite: start the snapshot from zero velocities
    if (!rc) rc = (int)hipMemsetAsync(b->snap, 0, 2 * b->node_total * sizeof(float), b->stream);
    if (!rc) rc = (int)hipStreamSynchronize(b->stream);
    return rc;
}

extern "C" int lb_batch_init_rest(lb_batch *b) {
    if (!b) return LB_ERR_INVALID;
    return batch_init(b, nullptr, 0);
}

extern "C" int lb_batch_init_fields_device(lb_batch *b, const float *d_fields) {
    if (!b || !d_fields) return LB_ERR_INVALID;
    return batch_init(b, d_fields, 0);
}

extern "C" int lb_batch_init_fields_device_ex(lb_batch *b, const float *d_fields, int mode) {
    if (!b || !d_fields || mode < 0 || mode > 1) return LB_ERR_INVALID;
    return batch_init(b, d_fields, mode);
}

extern "C" int lb_batch_step(lb_batch *b, int nsteps) {
    if (!b || nsteps < 0) return LB_ERR_INVALID;
    if (b->nblocks == 0 || nsteps == 0) return 0;
    for (int s = 0; s < nsteps; s++) {
        launch_batch_step(b, b->parity, s == nsteps - 1);
        b->parity = 1 - b->parity;
    }
    for (int i : b->act_list) { b->cur[i] = b->parity; b->macros_valid[i] = 1; }
    int rc = (int)hipGetLastError();
    if (!rc) rc = (int)hipStreamSynchronize(b->stream);
    return rc;
}

// Chunks of k steps per replay of the loop graph (about 0.2 s of work for
// the 512-lattice ensemble at k = 100): the host checks once per replay.
constexpr int kLoopChunks = 20;

extern "C" int lb_batch_run_to_steady(lb_batch *b, double tol, int k, int max_steps, int *steps_out, double *res_out) {
    if (!b || k < 2 || max_steps < 0) return LB_ERR_INVALID;
    // device state from the host's view (lattices retired earlier stay so)
    std::vector<BatchState> st(b->n);
    for (int i = 0; i < b->n; i++) st[i] = BatchState{b->active[i], 0, b->snap_valid[i], b->cur[i], 1.0};
    int rc = (int)hipMemcpyAsync(b->d_state, st.data(), st.size() * sizeof(BatchState), hipMemcpyHostToDevice, b->stream);
    if (!rc) {
        hipLaunchKernelGGL(k_batch_compact, dim3(1), dim3(kCompactThreads), 0, b->stream, b->d_lats, b->d_state, b->n,
                           b->d_act, b->d_ranges, b->d_table, b->d_counts);
        rc = (int)hipGetLastError();
    }
    int counts[2] = {0, 0};
    if (!rc) rc = (int)hipMemcpyAsync(counts, b->d_counts, sizeof(counts), hipMemcpyDeviceToHost, b->stream);
    if (!rc) rc = (int)hipStreamSynchronize(b->stream);
    int done = 0;
    while (!rc && counts[0] > 0 && done + k <= max_steps) {
        const int chunks = std::min(kLoopChunks, (max_steps - done) / k);
        hipGraphExec_t g = nullptr;
        rc = loop_graph(b, chunks, k, b->parity, tol, &g);
        if (!rc) rc = (int)hipGraphLaunch(g, b->stream);
        if (!rc) rc = (int)hipMemcpyAsync(counts, b->d_counts, sizeof(counts), hipMemcpyDeviceToHost, b->stream);
        if (!rc) rc = (int)hipStreamSynchronize(b->stream);
        if (((long long)chunks * k) & 1) b->parity = 1 - b->parity;
        done += chunks * k;
    }
    if (!rc) rc = (int)hipMemcpyAsync(st.data(), b->d_state, st.size() * sizeof(BatchState), hipMemcpyDeviceToHost, b->stream);
    if (!rc) rc = (int)hipStreamSynchronize(b->stream);
    if (rc) return rc;
    for (int i = 0; i < b->n; i++) {
        b->active[i] = (char)st[i].active;
        b->snap_valid[i] = (char)st[i].snap_valid;
        b->cur[i] = st[i].cur;
        if (st[i].steps > 0) b->macros_valid[i] = 1;
        if (steps_out) steps_out[i] = st[i].steps;
        if (res_out) res_out[i] = st[i].res;
    }
    return rebuild_table(b);   // the host's table for lb_batch_step
}

extern "C" int lb_batch_get_macros(lb_batch *b, int i, float *rho, float *ux, float *uy) {
    if (!b || i < 0 || i >= b->n) return LB_ERR_INVALID;
    if (!b->macros_valid[i]) return LB_ERR_NOT_READY;
    const BatchLat &L = b->lats[i];
    float *dst[3] = {rho, ux, uy};
    int rc = 0;
    for (int c = 0; c < 3 && !rc; c++)
        if (dst[c]) rc = (int)hipMemcpyAsync(dst[c], b->macros + 4 * L.node_off + (size_t)c * L.N, L.N * sizeof(float),
                                             hipMemcpyDeviceToHost, b->stream);
    if (!rc) rc = (int)hipStreamSynchronize(b->stream);
    return rc;
}

extern "C" int lb_batch_get_populations(lb_batch *b, int i, float *host) {
    if (!b || !host || i < 0 || i >= b->n) return LB_ERR_INVALID;
    const BatchLat &L = b->lats[i];
    const float *src = (b->cur[i] == 0 ? b->A : b->B) + L.pop_off;
    if (b->layout == LB_LAYOUT_SOA) {
        int rc = (int)hipMemcpyAsync(host, src, 9ull * L.N * sizeof(float), hipMemcpyDeviceToHost, b->stream);
        if (!rc) rc = (int)hipStreamSynchronize(b->stream);
        return rc;
    }
    // rows layout: [y][i][x] on the device, [i][y][x] for the caller
    std::vector<float> tmp(9ull * L.N);
    int rc = (int)hipMemcpyAsync(tmp.data(), src, tmp.size() * sizeof(float), hipMemcpyDeviceToHost, b->stream);
    if (!rc) rc = (int)hipStreamSynchronize(b->stream);
    if (!rc)
        for (int y = 0; y < L.ny; y++)
            for (int q = 0; q < 9; q++)
                std::memcpy(host + (size_t)q * L.N + (size_t)y * L.nx, &tmp[((size_t)y * 9 + q) * L.nx],
                            L.nx * sizeof(float));
    return rc;
}

extern "C" int64_t lb_batch_total_nodes(lb_batch *b) { return b ? (int64_t)b->node_total : -1; }
