// lb_ref.hip -- the reference's extern "C" launchers with reference semantics,
// and the reference host driver (Simulation{D2Q9,D3Q19}) as lb_refsim_*.
//
// Semantics follow src/streaming.cu, src/collision.cu, src/get_border.cu,
// src/copy_border.cu, src/apply_border.cu, src/read_value.cu and
// src/simulation_d2q9.cu / simulation_d3q19.cu (cited per function).  The
// launch shapes are this library's own: one thread per node on a 1-D grid of
// 256-thread workgroups (four wave64s) for the sweeps, one thread per edge
// node for the halo pack/unpack instead of the reference's full-lattice grid.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "lb_common.hpp"

using namespace lb;

namespace {

constexpr int kBlock = 256;

inline unsigned nblocks(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// Population traffic of the sweeps is streamed once per kernel: non-temporal
// loads and stores keep it from evicting the flags, borders and plot data.
__device__ __forceinline__ float ldn(const float *p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ void stn(float *p, float v) { __builtin_nontemporal_store(v, p); }

inline bool shape_ok(lb_dim3 g, lb_dim3 b, int w, int h, int d) {
    return (int64_t)g.x * b.x == w && (int64_t)g.y * b.y == h && (int64_t)g.z * b.z == d && w > 0 &&
           h > 0 && d > 0;
}

// ---------------------------------------------------------------------------
// Streaming: in-place swap of the positive-half slot pairs with periodic
// neighbours inside the subdomain (streaming.cu:24-74, 84-163).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_ref_stream9(DataD2Q9 d) {
    const int W = d.width, H = d.height;
    int64_t o0 = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (o0 >= (int64_t)W * H) return;
    if (d.solid_bound[o0]) return;
    int x = (int)(o0 % W), y = (int)(o0 / W);
    int xn = x > 0 ? x - 1 : W - 1;
    int xp = x < W - 1 ? x + 1 : 0;
    int yp = y > 0 ? y - 1 : H - 1;
    int64_t o1 = xp + (int64_t)y * W, o2 = x + (int64_t)yp * W;
    int64_t o3 = xp + (int64_t)yp * W, o4 = xn + (int64_t)yp * W;
    float t;
    t = ldn(&d.f1[o0]); stn(&d.f1[o0], ldn(&d.f3[o1])); stn(&d.f3[o1], t);
    t = ldn(&d.f2[o0]); stn(&d.f2[o0], ldn(&d.f4[o2])); stn(&d.f4[o2], t);
    t = ldn(&d.f5[o0]); stn(&d.f5[o0], ldn(&d.f7[o3])); stn(&d.f7[o3], t);
    t = ldn(&d.f6[o0]); stn(&d.f6[o0], ldn(&d.f8[o4])); stn(&d.f8[o4], t);
}

__global__ __launch_bounds__(kBlock) void k_ref_stream19(DataD3Q19 d) {
    const int W = d.width, H = d.height, D = d.depth;
    int64_t o0 = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (o0 >= (int64_t)W * H * D) return;
    if (d.solid_bound[o0]) return;
    int x = (int)(o0 % W);
    int y = (int)((o0 / W) % H);
    int z = (int)(o0 / ((int64_t)W * H));
    int xp = x < W - 1 ? x + 1 : 0;
    int yn = y > 0 ? y - 1 : H - 1, yp = y < H - 1 ? y + 1 : 0;
    int zn = z > 0 ? z - 1 : D - 1, zp = z < D - 1 ? z + 1 : 0;
    auto id = [&](int a, int b, int c) { return a + (int64_t)W * (b + (int64_t)H * c); };
    float t;
#define LB_SWAP(A, B, OFF) { int64_t o = OFF; t = ldn(&d.A[o0]); stn(&d.A[o0], ldn(&d.B[o])); stn(&d.B[o], t); }
    LB_SWAP(f1, f2, id(xp, y, z));
    LB_SWAP(f3, f4, id(x, yp, z));
    LB_SWAP(f5, f6, id(x, y, zp));
    LB_SWAP(f7, f12, id(xp, yp, z));
    LB_SWAP(f8, f11, id(xp, yn, z));
    LB_SWAP(f9, f14, id(xp, y, zp));
    LB_SWAP(f10, f13, id(xp, y, zn));
    LB_SWAP(f15, f18, id(x, yp, zp));
    LB_SWAP(f16, f17, id(x, yp, zn));
#undef LB_SWAP
}

// ---------------------------------------------------------------------------
// Collision D2Q9 (collision.cu:168-341): swap back to natural slots, Zou-He
// velocity inlet on x=0 of subdomain 0 / pressure outlet on x=W-1 of the last
// subdomain, moments, body-force shift, then BGK (G == 0) or psi/ueq store.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_ref_coll9(float *plot, DataD2Q9 d, float w0, float w1,
                                                      float w5, float ti, float ux_in,
                                                      float rho_out, float Fx, float Fy, float G, int zou_he) {
    const int W = d.width, H = d.height;
    const int64_t N = (int64_t)W * H;
    int64_t o = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (o >= N) return;
    int x = (int)(o % W);
    if (!d.solid[o]) {
        float v[9] = {ldn(&d.f0[o]), ldn(&d.f3[o]), ldn(&d.f4[o]), ldn(&d.f1[o]), ldn(&d.f2[o]),
                      ldn(&d.f7[o]), ldn(&d.f8[o]), ldn(&d.f5[o]), ldn(&d.f6[o])};
        if (zou_he) {
            if (x == 0 && d.dev == 0) zouhe9_in(v, ux_in);
            else if (x == W - 1 && d.dev == d.deviceCount - 1) zouhe9_out(v, rho_out);
        }
        float rho, ux, uy;
        mom9(v, rho, ux, uy);
        float ex = ux, ey = uy;
        if (Fx != 0) ex = ux + Fx / (rho * ti);
        if (Fy != 0) ey = uy + Fy / (rho * ti);
        d.ux[o] = ux; d.uy[o] = uy; d.rho[o] = rho;
        if (G != 0) {
            d.psi[o] = (float)(4. * exp_cw(-200. / (double)rho));
            d.ueq_x[o] = ex;
            d.ueq_y[o] = ey;
        } else {
            *(float4 *)&plot[o * 4] = make_float4(ux, uy, 0.f, rho);
            bgk9<float>(v, rho, ex, ey, ti, w0, w1, w5);
        }
        stn(&d.f0[o], v[0]); stn(&d.f1[o], v[1]); stn(&d.f2[o], v[2]); stn(&d.f3[o], v[3]); stn(&d.f4[o], v[4]);
        stn(&d.f5[o], v[5]); stn(&d.f6[o], v[6]); stn(&d.f7[o], v[7]); stn(&d.f8[o], v[8]);
    } else {
        *(float4 *)&plot[o * 4] = make_float4(0.f, 0.f, 0.f, 0.f);
        plot[(o + N) * 4] = 0.f;
        d.ux[o] = 0.f; d.uy[o] = 0.f; d.rho[o] = 0.f;
    }
}

// Shan-Chen second pass (collision.cu:35-166): psi stencil force, adhesion,
// velocity shift, BGK.  Neighbours wrap periodically over the subdomain.
__global__ __launch_bounds__(kBlock) void k_ref_sc9(float *plot, DataD2Q9 d, float w0, float w1,
                                                    float w5, float ti, float G, float Gads) {
    const int W = d.width, H = d.height;
    int64_t o = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (o >= (int64_t)W * H) return;
    if (d.solid[o]) return;
    int x = (int)(o % W), y = (int)(o / W);
    int xn = x > 0 ? x - 1 : W - 1, xp = x < W - 1 ? x + 1 : 0;
    int yn = y > 0 ? y - 1 : H - 1, yp = y < H - 1 ? y + 1 : 0;
    int64_t m1 = xp + (int64_t)y * W, m2 = x + (int64_t)yp * W, m3 = xn + (int64_t)y * W;
    int64_t m4 = x + (int64_t)yn * W, m5 = xp + (int64_t)yp * W, m6 = xn + (int64_t)yp * W;
    int64_t m7 = xn + (int64_t)yn * W, m8 = xp + (int64_t)yn * W;
    const float *psi = d.psi;
    float IFx = 0, IFy = 0;
    IFx = IFx + w1 * psi[m1];
    IFx = IFx - w1 * psi[m3];
    IFx = IFx + w5 * psi[m5];
    IFx = IFx - w5 * psi[m6];
    IFx = IFx - w5 * psi[m7];
    IFx = IFx + w5 * psi[m8];
    IFy = IFy - w1 * psi[m2];
    IFy = IFy + w1 * psi[m4];
    IFy = IFy - w5 * psi[m5];
    IFy = IFy - w5 * psi[m6];
    IFy = IFy + w5 * psi[m7];
    IFy = IFy + w5 * psi[m8];
    float po = psi[o], rho = d.rho[o];
    IFx = -G * po * IFx;
    IFy = -G * po * IFy;
    float ex = d.ueq_x[o] + (IFx / (ti * rho));
    float ey = d.ueq_y[o] + (IFy / (ti * rho));
    if (Gads != 0) {
        const bool *s = d.solid;
        float Fsx = 0, Fsy = 0;
        if (s[m1]) Fsx = Fsx + w1;
        if (s[m2]) Fsy = Fsy - w1;
        if (s[m3]) Fsx = Fsx - w1;
        if (s[m4]) Fsy = Fsy + w1;
        if (s[m5]) { Fsx = Fsx + w5; Fsy = Fsy - w5; }
        if (s[m6]) { Fsx = Fsx - w5; Fsy = Fsy - w5; }
        if (s[m7]) { Fsx = Fsx - w5; Fsy = Fsy + w5; }
        if (s[m8]) { Fsx = Fsx + w5; Fsy = Fsy + w5; }
        Fsx = -Gads * po * Fsx;
        Fsy = -Gads * po * Fsy;
        ex = ex + (Fsx / (ti * rho));
        ey = ey + (Fsy / (ti * rho));
    }
    *(float4 *)&plot[o * 4] = make_float4(d.ux[o], d.uy[o], 0.f, rho);
    float v[9] = {ldn(&d.f0[o]), ldn(&d.f1[o]), ldn(&d.f2[o]), ldn(&d.f3[o]), ldn(&d.f4[o]),
                  ldn(&d.f5[o]), ldn(&d.f6[o]), ldn(&d.f7[o]), ldn(&d.f8[o])};
    bgk9<float>(v, rho, ex, ey, ti, w0, w1, w5);
    stn(&d.f0[o], v[0]); stn(&d.f1[o], v[1]); stn(&d.f2[o], v[2]); stn(&d.f3[o], v[3]); stn(&d.f4[o], v[4]);
    stn(&d.f5[o], v[5]); stn(&d.f6[o], v[6]); stn(&d.f7[o], v[7]); stn(&d.f8[o], v[8]);
}

// Collision D3Q19 (collision.cu:357-552).
__global__ __launch_bounds__(kBlock) void k_ref_coll19(float *plot, DataD3Q19 d, float w0,
                                                       float w1, float w7, float ti,
                                                       float ux_in, float rho_out, int zou_he) {
    const int W = d.width;
    const int64_t N = (int64_t)W * d.height * d.depth;
    int64_t o = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (o >= N) return;
    int x = (int)(o % W);
    if (!d.solid[o]) {
        // natural slots after the local swap (collision.cu:375-410)
        float v[19] = {ldn(&d.f0[o]),  ldn(&d.f2[o]),  ldn(&d.f1[o]),  ldn(&d.f4[o]),  ldn(&d.f3[o]),
                       ldn(&d.f6[o]),  ldn(&d.f5[o]),  ldn(&d.f12[o]), ldn(&d.f11[o]), ldn(&d.f14[o]),
                       ldn(&d.f13[o]), ldn(&d.f8[o]),  ldn(&d.f7[o]),  ldn(&d.f10[o]), ldn(&d.f9[o]),
                       ldn(&d.f18[o]), ldn(&d.f17[o]), ldn(&d.f16[o]), ldn(&d.f15[o])};
        if (zou_he) {
            if (x == 0 && d.dev == 0) zouhe19_in(v, ux_in);
            else if (x == W - 1 && d.dev == d.deviceCount - 1) zouhe19_out(v, rho_out);
        }
        float rho, ux, uy, uz;
        mom19(v, rho, ux, uy, uz);
        d.ux[o] = ux; d.uy[o] = uy; d.uz[o] = uz; d.rho[o] = rho;
        *(float4 *)&plot[o * 4] = make_float4(ux, uy, uz, rho);
        bgk19<float>(v, rho, ux, uy, uz, ti, w0, w1, w7);
        float *dst[19] = {d.f0,  d.f1,  d.f2,  d.f3,  d.f4,  d.f5,  d.f6,  d.f7,  d.f8, d.f9,
                          d.f10, d.f11, d.f12, d.f13, d.f14, d.f15, d.f16, d.f17, d.f18};
#pragma unroll
        for (int i = 0; i < 19; i++) stn(&dst[i][o], v[i]);
    } else {
        *(float4 *)&plot[o * 4] = make_float4(0.f, 0.f, 0.f, 0.f);
        d.ux[o] = 0.f; d.uy[o] = 0.f; d.uz[o] = 0.f; d.rho[o] = 0.f;
    }
}

// Halo pack/unpack over the edge columns only (get_border.cu:26-98,
// apply_border.cu:25-97).
__global__ __launch_bounds__(kBlock) void k_ref_get9(DataD2Q9 d) {
    int y = blockIdx.x * kBlock + threadIdx.x;
    if (y >= d.height) return;
    const int W = d.width;
    int64_t oR = (W - 1) + (int64_t)y * W, oL = (int64_t)y * W;
    d.border_f1[y] = d.f1[oR]; d.border_f5[y] = d.f5[oR]; d.border_f8[y] = d.f8[oR];
    d.border_f3[y] = d.f3[oL]; d.border_f6[y] = d.f6[oL]; d.border_f7[y] = d.f7[oL];
}
__global__ __launch_bounds__(kBlock) void k_ref_apply9(DataD2Q9 d) {
    int y = blockIdx.x * kBlock + threadIdx.x;
    if (y >= d.height) return;
    const int W = d.width;
    int64_t oR = (W - 1) + (int64_t)y * W, oL = (int64_t)y * W;
    d.f3[oL] = d.nb_border_f3[y]; d.f7[oL] = d.nb_border_f7[y]; d.f6[oL] = d.nb_border_f6[y];
    d.f1[oR] = d.nb_border_f1[y]; d.f8[oR] = d.nb_border_f8[y]; d.f5[oR] = d.nb_border_f5[y];
}
__global__ __launch_bounds__(kBlock) void k_ref_get19(DataD3Q19 d) {
    int64_t bd = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int W = d.width, H = d.height;
    if (bd >= (int64_t)H * d.depth) return;
    int y = (int)(bd % H);
    int64_t z = bd / H;
    int64_t oL = (int64_t)W * (y + H * z), oR = oL + W - 1;
    d.border_f1[bd] = d.f1[oR]; d.border_f7[bd] = d.f7[oR]; d.border_f8[bd] = d.f8[oR];
    d.border_f9[bd] = d.f9[oR]; d.border_f10[bd] = d.f10[oR];
    d.border_f2[bd] = d.f2[oL]; d.border_f11[bd] = d.f11[oL]; d.border_f12[bd] = d.f12[oL];
    d.border_f13[bd] = d.f13[oL]; d.border_f14[bd] = d.f14[oL];
}
__global__ __launch_bounds__(kBlock) void k_ref_apply19(DataD3Q19 d) {
    int64_t bd = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int W = d.width, H = d.height;
    if (bd >= (int64_t)H * d.depth) return;
    int y = (int)(bd % H);
    int64_t z = bd / H;
    int64_t oL = (int64_t)W * (y + H * z), oR = oL + W - 1;
    d.f2[oL] = d.nb_border_f2[bd]; d.f11[oL] = d.nb_border_f11[bd]; d.f12[oL] = d.nb_border_f12[bd];
    d.f13[oL] = d.nb_border_f13[bd]; d.f14[oL] = d.nb_border_f14[bd];
    d.f1[oR] = d.nb_border_f1[bd]; d.f7[oR] = d.nb_border_f7[bd]; d.f8[oR] = d.nb_border_f8[bd];
    d.f9[oR] = d.nb_border_f9[bd]; d.f10[oR] = d.nb_border_f10[bd];
}

// read_value.cu:27-51
__global__ void k_ref_read_value(const float *plot, float *value, int W, int H, int x, int y, int z) {
    int64_t o = x + (int64_t)y * W + (int64_t)z * W * H;
    value[0] = plot[o * 4];
    value[1] = plot[o * 4 + 1];
    value[2] = plot[o * 4 + 2];
    value[3] = plot[o * 4 + 3];
}

void check_launch() { lb_hip_err(hipGetLastError()); }

}  // namespace

// ===========================================================================
// Reference entry points.
// ===========================================================================
extern "C" void streaming_d2q9(DataD2Q9 *data, lb_dim3 grid, lb_dim3 block) {
    if (!data) { lb_set_error(LB_ERR_INVALID); return; }
    if (!shape_ok(grid, block, data->width, data->height, 1)) { lb_set_error(LB_ERR_SHAPE); return; }
    hipLaunchKernelGGL(k_ref_stream9, dim3(nblocks((int64_t)data->width * data->height)), dim3(kBlock), 0, 0, *data);
    check_launch();
}

extern "C" void streaming_d3q19(DataD3Q19 *data, lb_dim3 grid, lb_dim3 block) {
    if (!data) { lb_set_error(LB_ERR_INVALID); return; }
    if (!shape_ok(grid, block, data->width, data->height, data->depth)) { lb_set_error(LB_ERR_SHAPE); return; }
    hipLaunchKernelGGL(k_ref_stream19, dim3(nblocks((int64_t)data->width * data->height * data->depth)),
                       dim3(kBlock), 0, 0, *data);
    check_launch();
}

// The collision launches with the Zou-He inlet / outlet on (the reference's
// launchers, collision.cu:211-262, always apply them) or off (x periodic over
// the ring of subdomains: the host driver's switch).
namespace {
void launch_collision_d2q9(float *plot_data, DataD2Q9 *data, float w0, float w1, float w5, float tau_inv,
                           float ux_in, float rho_out, lb_dim3 grid, lb_dim3 block, float Fx, float Fy, float G,
                           float Gads, int zou_he) {
    if (!data || !plot_data) { lb_set_error(LB_ERR_INVALID); return; }
    if (!shape_ok(grid, block, data->width, data->height, 1)) { lb_set_error(LB_ERR_SHAPE); return; }
    unsigned nb = nblocks((int64_t)data->width * data->height);
    hipLaunchKernelGGL(k_ref_coll9, dim3(nb), dim3(kBlock), 0, 0, plot_data, *data, w0, w1, w5, tau_inv,
                       ux_in, rho_out, Fx, Fy, G, zou_he);
    check_launch();
    if (G < 0) {
        hipLaunchKernelGGL(k_ref_sc9, dim3(nb), dim3(kBlock), 0, 0, plot_data, *data, w0, w1, w5, tau_inv, G, Gads);
        check_launch();
    }
}

void launch_collision_d3q19(float *plot_data, DataD3Q19 *data, float w0, float w1, float w7, float tau_inv,
                            float ux_in, float rho_out, lb_dim3 grid, lb_dim3 block, int zou_he) {
    if (!data || !plot_data) { lb_set_error(LB_ERR_INVALID); return; }
    if (!shape_ok(grid, block, data->width, data->height, data->depth)) { lb_set_error(LB_ERR_SHAPE); return; }
    hipLaunchKernelGGL(k_ref_coll19, dim3(nblocks((int64_t)data->width * data->height * data->depth)),
                       dim3(kBlock), 0, 0, plot_data, *data, w0, w1, w7, tau_inv, ux_in, rho_out, zou_he);
    check_launch();
}
}  // namespace

extern "C" void collision_d2q9(float *plot_data, DataD2Q9 *data, float w0, float w1, float w5,
                               float tau_inv, float ux_in, float uy_in, float rho_out, float u_min,
                               float u_max, float rho_min, float rho_max, lb_dim3 grid,
                               lb_dim3 block, float Fx, float Fy, float G, float Gads) {
    (void)uy_in; (void)u_min; (void)u_max; (void)rho_min; (void)rho_max;
    launch_collision_d2q9(plot_data, data, w0, w1, w5, tau_inv, ux_in, rho_out, grid, block, Fx, Fy, G, Gads, 1);
}

extern "C" void collision_d3q19(float *plot_data, DataD3Q19 *data, float w0, float w1, float w7,
                                float tau_inv, float ux_in, float uy_in, float uz_in, float rho_out,
                                float u_min, float u_max, float rho_min, float rho_max,
                                lb_dim3 grid, lb_dim3 block) {
    (void)uy_in; (void)uz_in; (void)u_min; (void)u_max; (void)rho_min; (void)rho_max;
    launch_collision_d3q19(plot_data, data, w0, w1, w7, tau_inv, ux_in, rho_out, grid, block, 1);
}

extern "C" void get_border_d2q9(DataD2Q9 *data, lb_dim3 grid, lb_dim3 block) {
    if (!data) { lb_set_error(LB_ERR_INVALID); return; }
    if (!shape_ok(grid, block, data->width, data->height, 1)) { lb_set_error(LB_ERR_SHAPE); return; }
    hipLaunchKernelGGL(k_ref_get9, dim3(nblocks(data->height)), dim3(kBlock), 0, 0, *data);
    check_launch();
}

extern "C" void get_border_d3q19(DataD3Q19 *data, lb_dim3 grid, lb_dim3 block) {
    if (!data) { lb_set_error(LB_ERR_INVALID); return; }
    if (!shape_ok(grid, block, data->width, data->height, data->depth)) { lb_set_error(LB_ERR_SHAPE); return; }
    hipLaunchKernelGGL(k_ref_get19, dim3(nblocks((int64_t)data->height * data->depth)), dim3(kBlock), 0, 0, *data);
    check_launch();
}

extern "C" void apply_border_d2q9(DataD2Q9 *data, lb_dim3 grid, lb_dim3 block) {
    if (!data) { lb_set_error(LB_ERR_INVALID); return; }
    if (!shape_ok(grid, block, data->width, data->height, 1)) { lb_set_error(LB_ERR_SHAPE); return; }
    hipLaunchKernelGGL(k_ref_apply9, dim3(nblocks(data->height)), dim3(kBlock), 0, 0, *data);
    check_launch();
}

extern "C" void apply_border_d3q19(DataD3Q19 *data, lb_dim3 grid, lb_dim3 block) {
    if (!data) { lb_set_error(LB_ERR_INVALID); return; }
    if (!shape_ok(grid, block, data->width, data->height, data->depth)) { lb_set_error(LB_ERR_SHAPE); return; }
    hipLaunchKernelGGL(k_ref_apply19, dim3(nblocks((int64_t)data->height * data->depth)), dim3(kBlock), 0, 0, *data);
    check_launch();
}

// copy_border.cu:26-81: nb_border_* <- border_* with stream-ordered copies.
// Border planes of copy_border_devices_*: one launch for all of them when
// every pointer is device memory of the current device (the subdomains of
// one GPU), else one hipMemcpyAsync per plane (peer or host memory).
namespace {
struct PlaneCopy { const float *src[10]; float *dst[10]; };
__global__ __launch_bounds__(kBlock) void k_copy_planes(PlaneCopy c, int64_t n) {
    const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (e < n) c.dst[blockIdx.y][e] = c.src[blockIdx.y][e];
}

bool on_current_device(const void *p) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) return false;
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) { (void)hipGetLastError(); return false; }
    return a.type == hipMemoryTypeDevice && a.device == dev;
}

void copy_planes(float *const *dst, float *const *src, int k, int64_t n, hipStream_t st) {
    bool local = true;
    for (int i = 0; i < k && local; i++) local = on_current_device(dst[i]) && on_current_device(src[i]);
    if (local) {
        PlaneCopy c{};
        for (int i = 0; i < k; i++) { c.src[i] = src[i]; c.dst[i] = dst[i]; }
        hipLaunchKernelGGL(k_copy_planes, dim3(nblocks(n), k), dim3(kBlock), 0, st, c, n);
        check_launch();
        return;
    }
    for (int i = 0; i < k; i++)
        if (lb_hip_err(hipMemcpyAsync(dst[i], src[i], sizeof(float) * (size_t)n, hipMemcpyDefault, st))) return;
}
}  // namespace

extern "C" void copy_border_devices_d2q9(float *nb1, float *nb5, float *nb8, float *nb3, float *nb6,
                                         float *nb7, float *b1, float *b5, float *b8, float *b3,
                                         float *b6, float *b7, int height, lb_stream_t *s) {
    if (height <= 0) { lb_set_error(LB_ERR_INVALID); return; }
    hipStream_t st = s ? (hipStream_t)*s : nullptr;
    float *dst[6] = {nb1, nb5, nb8, nb3, nb6, nb7};
    float *src[6] = {b1, b5, b8, b3, b6, b7};
    for (int k = 0; k < 6; k++)
        if (!dst[k] || !src[k]) { lb_set_error(LB_ERR_INVALID); return; }
    copy_planes(dst, src, 6, height, st);
}

extern "C" void copy_border_devices_d3q19(float *nb1, float *nb7, float *nb8, float *nb9, float *nb10,
                                          float *nb2, float *nb11, float *nb12, float *nb13,
                                          float *nb14, float *b1, float *b7, float *b8, float *b9,
                                          float *b10, float *b2, float *b11, float *b12, float *b13,
                                          float *b14, int height, int depth, lb_stream_t *s) {
    if (height <= 0 || depth <= 0) { lb_set_error(LB_ERR_INVALID); return; }
    hipStream_t st = s ? (hipStream_t)*s : nullptr;
    float *dst[10] = {nb1, nb7, nb8, nb9, nb10, nb2, nb11, nb12, nb13, nb14};
    float *src[10] = {b1, b7, b8, b9, b10, b2, b11, b12, b13, b14};
    for (int k = 0; k < 10; k++)
        if (!dst[k] || !src[k]) { lb_set_error(LB_ERR_INVALID); return; }
    copy_planes(dst, src, 10, (int64_t)height * depth, st);
}

extern "C" void read_value_2d(float *plot_data, DataD2Q9 *data, int x, int y, int z) {
    if (!data || !plot_data || !data->valueCur) { lb_set_error(LB_ERR_INVALID); return; }
    hipLaunchKernelGGL(k_ref_read_value, dim3(1), dim3(1), 0, 0, plot_data, data->valueCur, data->width,
                       data->height, x, y, z);
    check_launch();
    lb_hip_err(hipDeviceSynchronize());
}

extern "C" void read_value_3d(float *plot_data, DataD3Q19 *data, int x, int y, int z) {
    if (!data || !plot_data || !data->valueCur) { lb_set_error(LB_ERR_INVALID); return; }
    hipLaunchKernelGGL(k_ref_read_value, dim3(1), dim3(1), 0, 0, plot_data, data->valueCur, data->width,
                       data->height, x, y, z);
    check_launch();
    lb_hip_err(hipDeviceSynchronize());
}

// ===========================================================================
// Reference host driver.
// ===========================================================================
namespace {

const float kW9[3] = {(float)(16. / 36.), (float)(4. / 36.), (float)(1. / 36.)};
const float kW19[3] = {(float)(12. / 36.), (float)(2. / 36.), (float)(1. / 36.)};

// Host exp with the device's operation sequence (initialisation psi).
double host_exp_cw(double x) {
    const double INV_LN2 = 1.44269504088896338700e+00;
    const double LN2_HI = 6.93147180369123816490e-01;
    const double LN2_LO = 1.90821492927058770002e-10;
    if (x != x) return x;
    if (x > 709.0) return HUGE_VAL;
    if (x < -708.0) return 0.0;
    double kd = std::nearbyint(x * INV_LN2);
    double r = (x - kd * LN2_HI) - kd * LN2_LO;
    const double c[14] = {1.0 / 6227020800.0, 1.0 / 479001600.0, 1.0 / 39916800.0, 1.0 / 3628800.0,
                          1.0 / 362880.0,     1.0 / 40320.0,     1.0 / 5040.0,      1.0 / 720.0,
                          1.0 / 120.0,        1.0 / 24.0,        1.0 / 6.0,         0.5,
                          1.0,                1.0};
    double p = c[0];
    for (int k = 1; k < 14; k++) p = p * r + c[k];
    uint64_t bits = (uint64_t)(int64_t)((int)kd + 1023) << 52;
    double scale;
    std::memcpy(&scale, &bits, 8);
    return p * scale;
}

struct Buf {
    void *p = nullptr;
    ~Buf() { if (p) (void)hipFree(p); }
};

}  // namespace

struct lb_refsim {
    lb_ref_params prm{};
    bool is3d = false;
    int nsub = 1, W = 0, H = 0, D = 1;   // W = subdomain width
    int64_t N = 0;
    float tau_inv = 0.f;
    std::vector<DataD2Q9> d2;
    std::vector<DataD3Q19> d3;
    std::vector<float *> plot;
    std::vector<void *> allocs;
    float *value_dev = nullptr;

    ~lb_refsim() {
        for (void *p : allocs) (void)hipFree(p);
    }
    template <class T> T *alloc(size_t n) {
        void *p = nullptr;
        if (hipMalloc(&p, n * sizeof(T) + 16) != hipSuccess) return nullptr;
        allocs.push_back(p);
        return (T *)p;
    }
};

namespace {

// Host initialisation, simulation_d2q9.cu:52-303: equilibrium with the
// reference's float products promoted to double by its double literals.
void feq_ref9(float *F, int64_t N, int64_t k, float r, float ex, float ey) {
    const float w0 = kW9[0], w1 = kW9[1], w2 = kW9[2];
    float sq = ex * ex + ey * ey;
    auto e = [&](float wr, float cu) {
        return (float)((double)wr * (1. + 3. * (double)cu + 4.5 * (double)cu * (double)cu - 1.5 * (double)sq));
    };
    F[0 * N + k] = (float)((double)(w0 * r) * (1. - 1.5 * (double)sq));
    F[1 * N + k] = e(w1 * r, +ex);
    F[2 * N + k] = e(w1 * r, +ey);
    F[3 * N + k] = e(w1 * r, -ex);
    F[4 * N + k] = e(w1 * r, -ey);
    F[5 * N + k] = e(w2 * r, ex + ey);
    F[6 * N + k] = e(w2 * r, -ex + ey);
    F[7 * N + k] = e(w2 * r, -ex - ey);
    F[8 * N + k] = e(w2 * r, ex - ey);
}

int init_d2q9(lb_refsim *s, const uint8_t *all) {
    const lb_ref_params &p = s->prm;
    const int W = s->W, H = s->H, nd = s->nsub;
    const int64_t N = s->N;
    const float w1 = kW9[1], w2 = kW9[2];
    const float ti = s->tau_inv;
    std::vector<float> F(9 * N), UX(N, 0.f), UY(N, 0.f), RHO(N, 0.f), PSI(N, 0.f), EX(N, 0.f), EY(N, 0.f);
    UX[0] = p.ux; UY[0] = p.uy; RHO[0] = p.rho;
    EX[0] = p.ux; EY[0] = p.uy;
    if (p.Fx != 0) EX[0] = UX[0] + p.Fx / (ti * RHO[0]);
    if (p.Fy != 0) EY[0] = UY[0] + p.Fy / (ti * RHO[0]);
    if (p.G != 0) {
        if (nd != 1) return LB_ERR_UNSUPPORTED;
        // Shan-Chen start (simulation_d2q9.cu:126-274): uniform rho, psi in
        // double, forcing-shifted ueq with the initialiser's IFy sign and
        // its sequential reuse of ueq[0].
        for (int64_t o = 0; o < N; o++)
            if (!all[o]) {
                RHO[o] = p.rho;
                PSI[o] = (float)(4. * host_exp_cw(-200. / (double)RHO[o]));
                UX[o] = UX[0]; UY[o] = UY[0];
            }
        for (int y = 0; y < H; y++) {
            int yp = y < H - 1 ? y + 1 : 0, yn = y > 0 ? y - 1 : H - 1;
            for (int x = 0; x < W; x++) {
                int xp = x < W - 1 ? x + 1 : 0, xn = x > 0 ? x - 1 : W - 1;
                int64_t c = (int64_t)y * W + x;
                if (all[c]) continue;
                int64_t o1 = (int64_t)y * W + xp, o2 = (int64_t)yp * W + x, o3 = (int64_t)y * W + xn;
                int64_t o4 = (int64_t)yn * W + x, o5 = (int64_t)yp * W + xp, o6 = (int64_t)yp * W + xn;
                int64_t o7 = (int64_t)yn * W + xn, o8 = (int64_t)yn * W + xp;
                float IFx = 0, IFy = 0;
                IFx = IFx + w1 * PSI[o1]; IFx = IFx - w1 * PSI[o3];
                IFx = IFx + w2 * PSI[o5]; IFx = IFx - w2 * PSI[o6];
                IFx = IFx - w2 * PSI[o7]; IFx = IFx + w2 * PSI[o8];
                IFy = IFy + w1 * PSI[o2]; IFy = IFy - w1 * PSI[o4];
                IFy = IFy + w2 * PSI[o5]; IFy = IFy + w2 * PSI[o6];
                IFy = IFy - w2 * PSI[o7]; IFy = IFy - w2 * PSI[o8];
                IFx = -p.G * PSI[c] * IFx;
                IFy = -p.G * PSI[c] * IFy;
                EX[c] = EX[0] + (IFx / (ti * RHO[c]));
                EY[c] = EY[0] + (IFy / (ti * RHO[c]));
                float Fsx = 0, Fsy = 0;
                if (all[o1]) Fsx = Fsx + w1;
                if (all[o2]) Fsy = Fsy - w1;
                if (all[o3]) Fsx = Fsx - w1;
                if (all[o4]) Fsy = Fsy + w1;
                if (all[o5]) { Fsx = Fsx + w2; Fsy = Fsy - w2; }
                if (all[o6]) { Fsx = Fsx - w2; Fsy = Fsy - w2; }
                if (all[o7]) { Fsx = Fsx - w2; Fsy = Fsy + w2; }
                if (all[o8]) { Fsx = Fsx + w2; Fsy = Fsy + w2; }
                Fsx = -p.Gads * PSI[c] * Fsx;
                Fsy = -p.Gads * PSI[c] * Fsy;
                EX[c] = EX[c] + (Fsx / (ti * RHO[c]));
                EY[c] = EY[c] + (Fsy / (ti * RHO[c]));
                feq_ref9(F.data(), N, c, RHO[c], EX[c], EY[c]);
            }
        }
        for (int64_t o = 0; o < N; o++)
            if (all[o]) {
                RHO[o] = 0; UX[o] = 0; UY[o] = 0; PSI[o] = 0;
                for (int i = 0; i < 9; i++) F[i * N + o] = kW9[i == 0 ? 0 : (i < 5 ? 1 : 2)] * p.rho;
            }
    } else {
        feq_ref9(F.data(), N, 0, RHO[0], EX[0], EY[0]);
        for (int64_t i = 0; i < N; i++) {
            UX[i] = UX[0]; UY[i] = UY[0]; RHO[i] = RHO[0];
            for (int q = 0; q < 9; q++) F[q * N + i] = F[q * N];
        }
    }
    std::vector<uint8_t> sol(N), sb(N);
    for (int d = 0; d < nd; d++) {
        DataD2Q9 &g = s->d2[d];
        for (int y = 0; y < H; y++)
            for (int x = 0; x < W; x++) sol[x + (int64_t)y * W] = all[(x + d * W) + (int64_t)y * W * nd];
        std::fill(sb.begin(), sb.end(), 0);
        // simulation_d2q9.cu:318-334
        for (int y = 0; y < H; y++)
            for (int x = 1; x < W - 1; x++) {
                int yn = y > 0 ? y - 1 : H - 1, yp = y < H - 1 ? y + 1 : 0;
                auto S = [&](int a, int b) { return sol[a + (int64_t)b * W] != 0; };
                sb[x + (int64_t)y * W] = S(x, y) && S(x + 1, y) && S(x, yp) && S(x - 1, y) && S(x, yn) &&
                                         S(x + 1, yp) && S(x - 1, yp) && S(x - 1, yn) && S(x + 1, yn);
            }
        float *fp[9] = {g.f0, g.f1, g.f2, g.f3, g.f4, g.f5, g.f6, g.f7, g.f8};
        for (int i = 0; i < 9; i++)
            if (hipMemcpy(fp[i], F.data() + i * N, N * 4, hipMemcpyHostToDevice) != hipSuccess) return LB_ERR_INVALID;
        (void)hipMemcpy(g.ux, UX.data(), N * 4, hipMemcpyHostToDevice);
        (void)hipMemcpy(g.uy, UY.data(), N * 4, hipMemcpyHostToDevice);
        (void)hipMemcpy(g.rho, RHO.data(), N * 4, hipMemcpyHostToDevice);
        (void)hipMemcpy(g.psi, PSI.data(), N * 4, hipMemcpyHostToDevice);
        (void)hipMemcpy(g.ueq_x, EX.data(), N * 4, hipMemcpyHostToDevice);
        (void)hipMemcpy(g.ueq_y, EY.data(), N * 4, hipMemcpyHostToDevice);
        (void)hipMemcpy(g.solid, sol.data(), N, hipMemcpyHostToDevice);
        if (lb_hip_err(hipMemcpy(g.solid_bound, sb.data(), N, hipMemcpyHostToDevice))) return LB_ERR_INVALID;
    }
    return 0;
}

// simulation_d3q19.cu:50-319.
int init_d3q19(lb_refsim *s, const uint8_t *all) {
    const lb_ref_params &p = s->prm;
    const int W = s->W, H = s->H, D = s->D, nd = s->nsub;
    const int64_t N = s->N;
    const float w0 = kW19[0], w1 = kW19[1], w2 = kW19[2];
    const float a = p.ux, b = p.uy, c = p.uz, r = p.rho;
    float sq = a * a + b * b + c * c;
    auto e = [&](float wr, float cu) {
        return (float)((double)wr * (1. + 3. * (double)cu + 4.5 * (double)cu * (double)cu - 1.5 * (double)sq));
    };
    float q[19] = {(float)((double)(w0 * r) * (1. - 1.5 * (double)sq)),
                   e(w1 * r, +a), e(w1 * r, -a), e(w1 * r, +b), e(w1 * r, -b), e(w1 * r, +c), e(w1 * r, -c),
                   e(w2 * r, +a + b), e(w2 * r, +a - b), e(w2 * r, +a + c), e(w2 * r, +a - c),
                   e(w2 * r, -a + b), e(w2 * r, -a - b), e(w2 * r, -a + c), e(w2 * r, -a - c),
                   e(w2 * r, +b + c), e(w2 * r, +b - c), e(w2 * r, -b + c), e(w2 * r, -b - c)};
    std::vector<float> tmp(N);
    std::vector<uint8_t> sol(N), sb(N);
    for (int d = 0; d < nd; d++) {
        DataD3Q19 &g = s->d3[d];
        float *fp[19] = {g.f0, g.f1, g.f2, g.f3, g.f4, g.f5, g.f6, g.f7, g.f8, g.f9,
                         g.f10, g.f11, g.f12, g.f13, g.f14, g.f15, g.f16, g.f17, g.f18};
        for (int i = 0; i < 19; i++) {
            std::fill(tmp.begin(), tmp.end(), q[i]);
            if (lb_hip_err(hipMemcpy(fp[i], tmp.data(), N * 4, hipMemcpyHostToDevice))) return LB_ERR_INVALID;
        }
        float *mp[4] = {g.ux, g.uy, g.uz, g.rho};
        float mv[4] = {a, b, c, r};
        for (int k = 0; k < 4; k++) {
            std::fill(tmp.begin(), tmp.end(), mv[k]);
            (void)hipMemcpy(mp[k], tmp.data(), N * 4, hipMemcpyHostToDevice);
        }
        for (int z = 0; z < D; z++)
            for (int y = 0; y < H; y++)
                for (int x = 0; x < W; x++)
                    sol[x + (int64_t)W * (y + (int64_t)H * z)] =
                        all[(x + d * W) + (int64_t)y * W * nd + (int64_t)z * W * nd * H];
        std::fill(sb.begin(), sb.end(), 0);
        // simulation_d3q19.cu:166-204
        for (int z = 0; z < D; z++)
            for (int y = 0; y < H; y++)
                for (int x = 1; x < W - 1; x++) {
                    int yn = y > 0 ? y - 1 : H - 1, yp = y < H - 1 ? y + 1 : 0;
                    int zn = z > 0 ? z - 1 : D - 1, zp = z < D - 1 ? z + 1 : 0;
                    bool all19 = true;
                    const int xs[3] = {x - 1, x, x + 1}, ys[3] = {yn, y, yp}, zs[3] = {zn, z, zp};
                    for (int i = 0; i < 3 && all19; i++)
                        for (int j = 0; j < 3 && all19; j++)
                            for (int k = 0; k < 3 && all19; k++) {
                                int nonzero = (i != 1) + (j != 1) + (k != 1);
                                if (nonzero == 3) continue;  // D3Q19 has no corner links
                                all19 = sol[xs[i] + (int64_t)W * (ys[j] + (int64_t)H * zs[k])] != 0;
                            }
                    sb[x + (int64_t)W * (y + (int64_t)H * z)] = all19;
                }
        (void)hipMemcpy(g.solid, sol.data(), N, hipMemcpyHostToDevice);
        if (lb_hip_err(hipMemcpy(g.solid_bound, sb.data(), N, hipMemcpyHostToDevice))) return LB_ERR_INVALID;
    }
    return 0;
}

}  // namespace

extern "C" int lb_refsim_create(const lb_ref_params *p, const uint8_t *solids, lb_refsim **out) {
    if (!p || !solids || !out) return LB_ERR_INVALID;
    *out = nullptr;
    if (lb_device_count() <= 0) return LB_ERR_NO_DEVICE;
    if (p->deviceCount < 1 || p->width < 2 || p->height < 1 || p->depth < 1) return LB_ERR_INVALID;
    if (p->width % p->deviceCount) return LB_ERR_SHAPE;
    lb_refsim *s = new (std::nothrow) lb_refsim();
    if (!s) return LB_ERR_INVALID;
    s->prm = *p;
    s->is3d = p->depth > 1;
    s->nsub = p->deviceCount;
    s->W = p->width / p->deviceCount;
    s->H = p->height;
    s->D = p->depth;
    s->N = (int64_t)s->W * s->H * s->D;
    s->tau_inv = (float)(1. / (3. * (double)p->viscosity + 0.5));   // input.cpp:168-171
    const int64_t N = s->N;
    s->value_dev = s->alloc<float>(4);
    if (!s->value_dev) { delete s; return LB_ERR_INVALID; }
    int rc = 0;
    if (!s->is3d) {
        s->d2.resize(s->nsub);
        for (int d = 0; d < s->nsub; d++) {
            DataD2Q9 g{};
            g.width = s->W; g.height = s->H; g.deviceCount = s->nsub; g.dev = d;
            g.ux = s->alloc<float>(N); g.uy = s->alloc<float>(N); g.rho = s->alloc<float>(N);
            g.solid = s->alloc<bool>(N); g.solid_bound = s->alloc<bool>(N);
            float **fp[9] = {&g.f0, &g.f1, &g.f2, &g.f3, &g.f4, &g.f5, &g.f6, &g.f7, &g.f8};
            for (int i = 0; i < 9; i++) *fp[i] = s->alloc<float>(N);
            float **bp[12] = {&g.border_f1, &g.border_f5, &g.border_f8, &g.border_f3, &g.border_f6,
                              &g.border_f7, &g.nb_border_f1, &g.nb_border_f5, &g.nb_border_f8,
                              &g.nb_border_f3, &g.nb_border_f6, &g.nb_border_f7};
            for (int k = 0; k < 12; k++) *bp[k] = s->alloc<float>(s->H);
            g.psi = s->alloc<float>(N); g.ueq_x = s->alloc<float>(N); g.ueq_y = s->alloc<float>(N);
            g.plotBuffer = s->alloc<float>(8 * N);
            g.valueCur = s->value_dev;
            g.stream = nullptr;
            bool ok = g.ux && g.uy && g.rho && g.solid && g.solid_bound && g.psi && g.ueq_x && g.ueq_y &&
                      g.plotBuffer && g.f8 && g.nb_border_f7;
            if (!ok) { delete s; return LB_ERR_INVALID; }
            (void)hipMemset(g.plotBuffer, 0, 8 * N * sizeof(float));
            s->d2[d] = g;
            s->plot.push_back(g.plotBuffer);
        }
        rc = init_d2q9(s, solids);
    } else {
        s->d3.resize(s->nsub);
        const int64_t P = (int64_t)s->H * s->D;
        for (int d = 0; d < s->nsub; d++) {
            DataD3Q19 g{};
            g.width = s->W; g.height = s->H; g.depth = s->D; g.deviceCount = s->nsub; g.dev = d;
            g.ux = s->alloc<float>(N); g.uy = s->alloc<float>(N); g.uz = s->alloc<float>(N);
            g.rho = s->alloc<float>(N);
            g.solid = s->alloc<bool>(N); g.solid_bound = s->alloc<bool>(N);
            float **fp[19] = {&g.f0, &g.f1, &g.f2, &g.f3, &g.f4, &g.f5, &g.f6, &g.f7, &g.f8, &g.f9,
                              &g.f10, &g.f11, &g.f12, &g.f13, &g.f14, &g.f15, &g.f16, &g.f17, &g.f18};
            for (int i = 0; i < 19; i++) *fp[i] = s->alloc<float>(N);
            float **bp[20] = {&g.border_f1, &g.border_f7, &g.border_f8, &g.border_f9, &g.border_f10,
                              &g.border_f2, &g.border_f11, &g.border_f12, &g.border_f13, &g.border_f14,
                              &g.nb_border_f1, &g.nb_border_f7, &g.nb_border_f8, &g.nb_border_f9,
                              &g.nb_border_f10, &g.nb_border_f2, &g.nb_border_f11, &g.nb_border_f12,
                              &g.nb_border_f13, &g.nb_border_f14};
            for (int k = 0; k < 20; k++) *bp[k] = s->alloc<float>(P);
            g.plotBuffer = s->alloc<float>(4 * N);
            g.valueCur = s->value_dev;
            g.stream = nullptr;
            bool ok = g.ux && g.uy && g.uz && g.rho && g.solid && g.solid_bound && g.plotBuffer && g.f18 &&
                      g.nb_border_f14;
            if (!ok) { delete s; return LB_ERR_INVALID; }
            (void)hipMemset(g.plotBuffer, 0, 4 * N * sizeof(float));
            s->d3[d] = g;
            s->plot.push_back(g.plotBuffer);
        }
        rc = init_d3q19(s, solids);
    }
    if (rc == 0) rc = lb_hip_err(hipDeviceSynchronize());
    if (rc) { delete s; return rc; }
    *out = s;
    return 0;
}

extern "C" int lb_refsim_destroy(lb_refsim *s) {
    if (!s) return LB_ERR_INVALID;
    (void)hipDeviceSynchronize();
    delete s;
    return 0;
}

// simulation_d2q9.cu:462-554 / simulation_d3q19.cu:352-430, without the
// per-step device synchronisation and 16-byte read-back: every launch is
// ordered on the legacy default stream and valueCur is read once at the end.
extern "C" int lb_refsim_run(lb_refsim *s, int steps) {
    if (!s || steps < 0) return LB_ERR_INVALID;
    (void)lb_last_error();
    const int nd = s->nsub;
    const lb_ref_params &p = s->prm;
    lb_dim3 block{32, 4, 1};
    lb_dim3 grid{(unsigned)s->W / 32, (unsigned)s->H / 4, (unsigned)s->D};
    // The reference's fixed (32,4,1) blocks need W%32 == 0 and H%4 == 0
    // (simulation.cu:43-44); other shapes are described as one block.
    if (s->W % 32 || s->H % 4) { block = lb_dim3{(unsigned)s->W, (unsigned)s->H, 1}; grid = lb_dim3{1, 1, (unsigned)s->D}; }
    for (int st = 0; st < steps; st++) {
        if (!s->is3d) {
            for (int d = 0; d < nd; d++) streaming_d2q9(&s->d2[d], grid, block);
            if (nd > 1) {
                for (int d = 0; d < nd; d++) get_border_d2q9(&s->d2[d], grid, block);
                for (int d = 0; d < nd; d++) {
                    int left = d == 0 ? nd - 1 : d - 1, right = d == nd - 1 ? 0 : d + 1;
                    DataD2Q9 &g = s->d2[d], &R = s->d2[right], &L = s->d2[left];
                    lb_stream_t st0 = nullptr;
                    copy_border_devices_d2q9(g.nb_border_f1, g.nb_border_f5, g.nb_border_f8, g.nb_border_f3,
                                             g.nb_border_f6, g.nb_border_f7, R.border_f1, R.border_f5,
                                             R.border_f8, L.border_f3, L.border_f6, L.border_f7, s->H, &st0);
                }
                for (int d = 0; d < nd; d++) apply_border_d2q9(&s->d2[d], grid, block);
            }
            for (int d = 0; d < nd; d++)
                launch_collision_d2q9(s->plot[d], &s->d2[d], kW9[0], kW9[1], kW9[2], s->tau_inv, p.ux, p.rho, grid,
                                      block, p.Fx, p.Fy, p.G, p.Gads, p.zou_he);
        } else {
            for (int d = 0; d < nd; d++) streaming_d3q19(&s->d3[d], grid, block);
            if (nd > 1) {
                for (int d = 0; d < nd; d++) get_border_d3q19(&s->d3[d], grid, block);
                for (int d = 0; d < nd; d++) {
                    int left = d == 0 ? nd - 1 : d - 1, right = d == nd - 1 ? 0 : d + 1;
                    DataD3Q19 &g = s->d3[d], &R = s->d3[right], &L = s->d3[left];
                    lb_stream_t st0 = nullptr;
                    copy_border_devices_d3q19(g.nb_border_f1, g.nb_border_f7, g.nb_border_f8, g.nb_border_f9,
                                              g.nb_border_f10, g.nb_border_f2, g.nb_border_f11, g.nb_border_f12,
                                              g.nb_border_f13, g.nb_border_f14, R.border_f1, R.border_f7,
                                              R.border_f8, R.border_f9, R.border_f10, L.border_f2, L.border_f11,
                                              L.border_f12, L.border_f13, L.border_f14, s->H, s->D, &st0);
                }
                for (int d = 0; d < nd; d++) apply_border_d3q19(&s->d3[d], grid, block);
            }
            for (int d = 0; d < nd; d++)
                launch_collision_d3q19(s->plot[d], &s->d3[d], kW19[0], kW19[1], kW19[2], s->tau_inv, p.ux, p.rho,
                                       grid, block, p.zou_he);
        }
        hipLaunchKernelGGL(k_ref_read_value, dim3(1), dim3(1), 0, 0, s->plot[0], s->value_dev, s->W, s->H,
                           p.cursor_x, p.cursor_y, p.cursor_z);
        int e = lb_last_error();
        if (e) return e;
    }
    return lb_hip_err(hipDeviceSynchronize());
}

extern "C" int lb_refsim_value_cur(lb_refsim *s, float out[4]) {
    if (!s || !out) return LB_ERR_INVALID;
    return lb_hip_err(hipMemcpy(out, s->value_dev, 4 * sizeof(float), hipMemcpyDeviceToHost));
}

extern "C" int lb_refsim_download(lb_refsim *s, int dev, float *f, float *ux, float *uy, float *uz,
                                  float *rho, float *plot) {
    if (!s || dev < 0 || dev >= s->nsub) return LB_ERR_INVALID;
    const int64_t N = s->N;
    int rc = lb_hip_err(hipDeviceSynchronize());
    if (rc) return rc;
    auto get = [&](float *dst, const float *src, int64_t n) {
        if (dst && src && !rc) rc = lb_hip_err(hipMemcpy(dst, src, n * sizeof(float), hipMemcpyDeviceToHost));
    };
    if (!s->is3d) {
        DataD2Q9 &g = s->d2[dev];
        const float *fp[9] = {g.f0, g.f1, g.f2, g.f3, g.f4, g.f5, g.f6, g.f7, g.f8};
        if (f) for (int i = 0; i < 9; i++) get(f + i * N, fp[i], N);
        get(ux, g.ux, N); get(uy, g.uy, N); get(rho, g.rho, N);
        if (uz) std::memset(uz, 0, N * sizeof(float));
        get(plot, s->plot[dev], 8 * N);
    } else {
        DataD3Q19 &g = s->d3[dev];
        const float *fp[19] = {g.f0, g.f1, g.f2, g.f3, g.f4, g.f5, g.f6, g.f7, g.f8, g.f9,
                               g.f10, g.f11, g.f12, g.f13, g.f14, g.f15, g.f16, g.f17, g.f18};
        if (f) for (int i = 0; i < 19; i++) get(f + i * N, fp[i], N);
        get(ux, g.ux, N); get(uy, g.uy, N); get(uz, g.uz, N); get(rho, g.rho, N);
        get(plot, s->plot[dev], 4 * N);
    }
    return rc;
}

extern "C" int lb_refsim_data_d2q9(lb_refsim *s, int dev, DataD2Q9 *out) {
    if (!s || !out || s->is3d || dev < 0 || dev >= s->nsub) return LB_ERR_INVALID;
    *out = s->d2[dev];
    return 0;
}

extern "C" int lb_refsim_data_d3q19(lb_refsim *s, int dev, DataD3Q19 *out) {
    if (!s || !out || !s->is3d || dev < 0 || dev >= s->nsub) return LB_ERR_INVALID;
    *out = s->d3[dev];
    return 0;
}
