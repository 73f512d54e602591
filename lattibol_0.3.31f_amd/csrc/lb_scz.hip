// One-pass Shan-Chen step for D3Q19: the z-marching sweep with the psi
// planes in LDS (ZState, zm_load and zm_sweep in lb_pull.hpp).  Its own
// translation unit: the kernel is large and compiles in parallel with the
// rest of the library.
#include <hip/hip_runtime.h>

#include "lb_pull.hpp"

namespace lb {
namespace {

// Block = bx x by = NT threads (256, 512 or 1,024), grid = ntx x nty tiles x
// ceil(nz / zc) z-chunks (XCD remap optional).  Iteration k of the march
// (k = z0 - 1 .. z1) computes the psi of plane k into LDS slot k & 3, waits at
// the barrier, then sweeps plane k - 1 with the psi of planes k - 2, k - 1, k.
// KEEP: the density phase keeps the pulled populations of plane k in
// registers for the next iteration's sweep; otherwise the sweep pulls them
// again (from L2: the workgroup read them one plane earlier), which halves
// the registers the populations hold across the barrier.  Threads 0 ..
// 2 (bx + 2) + 2 by - 1 also compute the psi of one frame position each per
// plane.
template <class S, bool MACROS, int NT, bool KEEP>
__global__ __launch_bounds__(NT) void k_step_scz(const StepArgs<S> *__restrict__ pa, uint32_t dstep) {
    constexpr int Q = 19;
    constexpr int MEM = 0;
    typedef typename S::T T;
    const StepArgs<S> &a = *pa;
    __shared__ T psl[4 * kZSlot];
    const int bx = 1 << a.bx_shift, by = a.by, W = bx + 2;
    const int plane = W * (by + 2);
    int b = tile_of_block(a);
    const int tx = b % a.ntx;
    b /= a.ntx;
    const int ty = b % a.nty;
    const int z0 = (b / a.nty) * a.zc, z1 = min(a.nz, z0 + a.zc);
    const int lx = threadIdx.x & (bx - 1), ly = threadIdx.x >> a.bx_shift;
    const int gx = tx * bx + lx, gy = ty * by + ly;
    const bool inside = gx < a.nx && gy < a.ny;
    const int c = (lx + 1) + (ly + 1) * W;
    // frame position of this thread: rows y = -1 and y = by, then columns
    // x = -1 and x = bx
    const bool framer = (int)threadIdx.x < 2 * W + 2 * by;
    int fx = 0, fy = 0;
    {
        int r = threadIdx.x;
        if (r < W) { fx = r - 1; fy = -1; }
        else if (r < 2 * W) { fx = r - W - 1; fy = by; }
        else { r -= 2 * W; if (r < by) { fx = -1; fy = r; } else { fx = bx; fy = r - by; } }
    }
    const int cf = (fx + 1) + (fy + 1) * W;
    uint32_t step = 0;
    if (a.round_mode) step = ldm(&a.ctr[0]) + dstep;
    ZState<Q, S> cur, nxt;
    for (int k = z0 - 1; k <= z1; k++) {
        T *slot = psl + ((k + 4) & 3) * plane;
        T v;
        if (inside && k >= 0 && k < a.nz) {
            if constexpr (KEEP) v = zm_load<Q, S, MEM>(a, node_in_tile(a, tx, ty, k), nxt);
            else v = psi_of_node<Q, S, MEM>(a, node_in_tile(a, tx, ty, k));
        } else {
            v = psi_at<Q, S, MEM>(a, gx, gy, k);
        }
        slot[c] = v;
        if (framer) slot[cf] = psi_at<Q, S, MEM>(a, tx * bx + fx, ty * by + fy, k);
        __syncthreads();
        if (inside && k > z0) {
            const int z = k - 1;
            NodePos p = node_in_tile(a, tx, ty, z);
            if (a.round_mode) {
                p.step = step;
                p.seed = sr_seed(step, p.id);
            }
            if constexpr (!KEEP) (void)zm_load<Q, S, MEM>(a, p, cur);
            zm_sweep<Q, S, MACROS, MEM>(a, p, cur, psl + ((z + 3) & 3) * plane, psl + (z & 3) * plane,
                                        psl + ((z + 1) & 3) * plane, c, W);
        }
        if constexpr (KEEP) cur = nxt;
    }
}

template <class S, int NT, bool KEEP>
int launch_n(bool macros, unsigned nb, hipStream_t st, const StepArgs<S> *a, uint32_t dstep) {
    if (macros) hipLaunchKernelGGL((k_step_scz<S, true, NT, KEEP>), dim3(nb), dim3(NT), 0, st, a, dstep);
    else hipLaunchKernelGGL((k_step_scz<S, false, NT, KEEP>), dim3(nb), dim3(NT), 0, st, a, dstep);
    return (int)hipGetLastError();
}
template <class S>
int launch_s(bool macros, bool form, int nt, unsigned nb, hipStream_t st, const void *pa, uint32_t dstep) {
    const StepArgs<S> *a = (const StepArgs<S> *)pa;
    if (nt == 256) return form ? launch_n<S, 256, true>(macros, nb, st, a, dstep) : launch_n<S, 256, false>(macros, nb, st, a, dstep);
    if (nt == 512) return form ? launch_n<S, 512, true>(macros, nb, st, a, dstep) : launch_n<S, 512, false>(macros, nb, st, a, dstep);
    if (nt == 1024) return form ? launch_n<S, 1024, true>(macros, nb, st, a, dstep) : launch_n<S, 1024, false>(macros, nb, st, a, dstep);
    return LB_ERR_UNSUPPORTED;
}

}  // namespace

bool scz_storage(int storage) {
    return storage == LB_STORE_F32 || storage == LB_STORE_F16 || storage == LB_STORE_BF16;
}

int launch_step_scz(int storage, bool macros, bool form, int threads, unsigned nblocks, hipStream_t stream,
                    const void *args, uint32_t dstep) {
    switch (storage) {
    case LB_STORE_F32: return launch_s<StF32>(macros, form, threads, nblocks, stream, args, dstep);
    case LB_STORE_F16: return launch_s<StF16>(macros, form, threads, nblocks, stream, args, dstep);
    case LB_STORE_BF16: return launch_s<StBF16>(macros, form, threads, nblocks, stream, args, dstep);
    default: return LB_ERR_UNSUPPORTED;
    }
}

}  // namespace lb
