// lb_unet.hip -- convolutional U-Net surrogate for steady channel flow:
// forward, backward and Adam on the GPU with hand-written kernels.
//
// Convolutions are implicit GEMMs on gfx950 BF16 matrix cores
// (v_mfma_f32_32x32x16_bf16: 32x32 output tile, K = 16 per instruction,
// FP32 accumulation).  Activations are NHWC BF16, weights BF16 copies of FP32
// master weights, activation gradients FP32 (rounded to BF16 only as MFMA
// operands).
//
//   forward   out[m, co]  = sum_{tap, ci} in[m + tap, ci] W[co, tap, ci]      M = pixels, N = Co, K = 9 Ci
//   bwd data  din[m, ci]  = sum_{tap, co} dout[m - tap, co] W[co, tap, ci]    = forward with W'[ci, 8-tap, co]
//   bwd wgt   dW[co, k]   = sum_m dout[m, co] in[m + tap(k), ci(k)]           split over pixel chunks
//
// Network (levels L = 4, base C = 32, 3x3 convs, ReLU, 2x2 average pooling,
// nearest upsampling with skip concatenation, 1x1 output conv):
//   enc l: conv(C_in -> C 2^l), conv(C 2^l -> C 2^l), pool      (l = 0..L-2)
//   bottom: conv(C 2^(L-2) -> C 2^(L-1)), conv(C 2^(L-1) -> C 2^(L-1))
//   dec l: upcat(prev, enc l), conv(-> C 2^l), conv(C 2^l -> C 2^l)   (l = L-2..0)
//   head: conv1x1(C -> out_ch)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "lb_common.hpp"

namespace {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ bf16 to_bf16(float f) {
    return __builtin_bit_cast(bf16, lb::f32_to_bf16(f));
}
__device__ __forceinline__ float from_bf16(bf16 h) {
    return lb::bf16_to_f32(__builtin_bit_cast(uint16_t, h));
}

// 8 consecutive channels starting at p (16-byte aligned) as an MFMA operand.
__device__ __forceinline__ bf16x8 load8(const bf16 *p) {
    return __builtin_bit_cast(bf16x8, *(const u16x8 *)p);
}
__device__ __forceinline__ bf16x8 load8(const float *p) {
    const float4 a = *(const float4 *)p, b = *(const float4 *)(p + 4);
    u16x8 r;
    r[0] = lb::f32_to_bf16(a.x); r[1] = lb::f32_to_bf16(a.y); r[2] = lb::f32_to_bf16(a.z); r[3] = lb::f32_to_bf16(a.w);
    r[4] = lb::f32_to_bf16(b.x); r[5] = lb::f32_to_bf16(b.y); r[6] = lb::f32_to_bf16(b.z); r[7] = lb::f32_to_bf16(b.w);
    return __builtin_bit_cast(bf16x8, r);
}
__device__ __forceinline__ bf16x8 zero8() { return __builtin_bit_cast(bf16x8, u16x8{0, 0, 0, 0, 0, 0, 0, 0}); }

// BF16 bits of an activation (FP32 activations round to nearest even).
__device__ __forceinline__ uint16_t bf16_bits(bf16 v) { return __builtin_bit_cast(uint16_t, v); }
__device__ __forceinline__ uint16_t bf16_bits(float v) { return lb::f32_to_bf16(v); }

template <class T> __device__ __forceinline__ float to_f(T v);
template <> __device__ __forceinline__ float to_f<float>(float v) { return v; }
template <> __device__ __forceinline__ float to_f<bf16>(bf16 v) { return from_bf16(v); }
template <class T> __device__ __forceinline__ T from_f(float v);
template <> __device__ __forceinline__ float from_f<float>(float v) { return v; }
template <> __device__ __forceinline__ bf16 from_f<bf16>(float v) { return to_bf16(v); }

// ---------------------------------------------------------------------------
// 3x3 convolution (pad 1, stride 1), implicit GEMM.  One wavefront computes
// a 32-pixel x (32 NB)-channel tile; a workgroup holds 4 wavefronts along the
// pixel dimension.  Ci % 16 == 0 and Co % 32 == 0.
// MFMA operand layouts (32x32x16): A row i = lane % 32 holds k = 8 (lane/32)
// .. +7; B column j = lane % 32 holds the same k; accumulator r of lane l is
// element (8 (r/4) + 4 (l/32) + r%4, l % 32).
// ---------------------------------------------------------------------------
// KS = 4: the 4 wavefronts share one 32-pixel tile and each sums a quarter
// of K (partial tiles added in LDS): 4x the workgroups for the small pixel
// counts of the deep levels.
template <class TIn, class TOut, int NB, bool BIAS_RELU, int KS = 1>
__global__ __launch_bounds__(256) void k_conv3x3(const TIn *__restrict__ in, const bf16 *__restrict__ W,
                                                 const float *__restrict__ bias, TOut *__restrict__ out,
                                                 int npix, int H, int Wd, int Ci, int Co) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m0 = KS == 1 ? (blockIdx.x * 4 + wave) * 32 : blockIdx.x * 32;
    const int n0 = blockIdx.y * 32 * NB;
    if (m0 >= npix) return;
    const int row = lane & 31, kh = (lane >> 5) * 8;
    const int m = m0 + row;
    const int HW = H * Wd;
    const bool mvalid = m < npix;
    const int nimg = mvalid ? m / HW : 0;
    const int rem = mvalid ? m - nimg * HW : 0;
    const int h = rem / Wd, w = rem - (rem / Wd) * Wd;
    f32x16 acc[NB];
    for (int b = 0; b < NB; b++)
        for (int r = 0; r < 16; r++) acc[b][r] = 0.f;
    const int K = 9 * Ci;
    int kbeg = 0, kend = K;
    if (KS > 1) {
        const int steps = K / 16;
        kbeg = (wave * steps / KS) * 16;
        kend = ((wave + 1) * steps / KS) * 16;
    }
    // Batches of U K-steps: all operand loads of a batch are issued before
    // its MFMAs, so a wavefront waits for memory once per batch, not once per
    // step.  (tap, channel group) advance incrementally (Ci % 16 == 0).
    constexpr int U = 4;
    int tap = kbeg / Ci, cc = kbeg - tap * Ci;
    const bf16 *wrow[NB];
    for (int b = 0; b < NB; b++) wrow[b] = W + (size_t)(n0 + 32 * b + row) * K + kh;
    for (int k0 = kbeg; k0 < kend; k0 += 16 * U) {
        bf16x8 a[U], bw[U][NB];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int k = k0 + 16 * u;
            a[u] = zero8();
            for (int b = 0; b < NB; b++) bw[u][b] = zero8();
            if (k < kend) {
                const int hh = h + tap / 3 - 1, ww = w + tap % 3 - 1;
                if (mvalid && hh >= 0 && hh < H && ww >= 0 && ww < Wd)
                    a[u] = load8(in + ((size_t)(nimg * H + hh) * Wd + ww) * Ci + cc + kh);
                for (int b = 0; b < NB; b++) bw[u][b] = load8(wrow[b] + k);
                cc += 16;
                if (cc == Ci) { cc = 0; tap++; }
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++)
            for (int b = 0; b < NB; b++) acc[b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[u], bw[u][b], acc[b], 0, 0, 0);
    }
    if constexpr (KS == 1) {
        for (int b = 0; b < NB; b++) {
            const int co = n0 + 32 * b + row;
            const float bv = BIAS_RELU ? bias[co] : 0.f;
            for (int r = 0; r < 16; r++) {
                const int i = 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                const int mm = m0 + i;
                if (mm >= npix) continue;
                float v = acc[b][r];
                if (BIAS_RELU) { v += bv; v = v > 0.f ? v : 0.f; }
                out[(size_t)mm * Co + co] = from_f<TOut>(v);
            }
        }
    } else {
        __shared__ float sh[KS][NB * 16][64];
        for (int b = 0; b < NB; b++)
            for (int r = 0; r < 16; r++) sh[wave][b * 16 + r][lane] = acc[b][r];
        __syncthreads();
        for (int e = threadIdx.x; e < NB * 16 * 64; e += 256) {
            const int l = e & 63, br = e >> 6, b = br >> 4, r = br & 15;
            float v = 0.f;
            for (int w = 0; w < KS; w++) v += sh[w][br][l];   // fixed order
            const int i = 8 * (r >> 2) + 4 * (l >> 5) + (r & 3);
            const int mm = m0 + i, co = n0 + 32 * b + (l & 31);
            if (mm >= npix) continue;
            if (BIAS_RELU) { v += bias[co]; v = v > 0.f ? v : 0.f; }
            out[(size_t)mm * Co + co] = from_f<TOut>(v);
        }
    }
}

// ---------------------------------------------------------------------------
// 3x3 convolution through LDS tiles (W % SW == 0, Ci % 32 == 0,
// Co % (32 NB) == 0, H % (4 PM 32 / SW) == 0).  A workgroup computes 4 PM
// MFMA pixel tiles x 32 NB output channels; an MFMA pixel tile is 32 / SW
// image rows x SW pixels (SW = 32: one row segment; SW = 16: two rows), and
// wavefront w owns tiles w PM .. w PM + PM - 1.  SPLIT: blockIdx.z takes a
// share of the input-channel stages and writes its FP32 partial tile to
// part[z][pixel][co] (k_split_sum adds the shares in order, then bias and
// ReLU), for layers with too few pixels to fill the GPU.  Input channels
// go in stages of 32: the stage's input window (4 PM + 2 rows x 34 pixels,
// zero outside the image) and weight block (32 NB channels x 9 taps x 32)
// are written to LDS once, converted to BF16 on the way (FP32 input of the
// input-gradient form), and every tap reads its A operand from the shifted
// window, so an input element leaves HBM/L2 about (4 PM + 2) 34 / (4 PM 32)
// times instead of 9 times.  The next stage's global loads are issued into
// registers before the current stage's MFMAs.  K order: stage, tap, 16-channel
// half (FP32 accumulation).
// ---------------------------------------------------------------------------
constexpr int kTileCK = 32;                  // input channels per stage
constexpr int kTilePS = kTileCK + 8;         // LDS stride of one window pixel (bf16)
constexpr int kTileWS = 9 * kTileCK + 8;     // LDS stride of one output channel's weights (bf16)

// Eight consecutive channels held in registers between their global load and
// their LDS store: BF16 as loaded, FP32 converted (nearest even) only at the
// store, so the loads of the next stage stay in flight during the MFMAs.
template <class T> struct Piece8;
template <> struct Piece8<bf16> {
    u16x8 v;
    __device__ __forceinline__ void zero() { v = u16x8{0, 0, 0, 0, 0, 0, 0, 0}; }
    __device__ __forceinline__ void load(const bf16 *p) { v = *(const u16x8 *)p; }
    __device__ __forceinline__ u16x8 bits() const { return v; }
};
template <> struct Piece8<float> {
    float4 a, b;
    __device__ __forceinline__ void zero() { a = b = float4{0.f, 0.f, 0.f, 0.f}; }
    __device__ __forceinline__ void load(const float *p) { a = *(const float4 *)p; b = *(const float4 *)(p + 4); }
    __device__ __forceinline__ u16x8 bits() const {
        u16x8 r;
        r[0] = lb::f32_to_bf16(a.x); r[1] = lb::f32_to_bf16(a.y); r[2] = lb::f32_to_bf16(a.z); r[3] = lb::f32_to_bf16(a.w);
        r[4] = lb::f32_to_bf16(b.x); r[5] = lb::f32_to_bf16(b.y); r[6] = lb::f32_to_bf16(b.z); r[7] = lb::f32_to_bf16(b.w);
        return r;
    }
};

// DPRE (input gradient of a conv whose input is the previous conv's ReLU
// output act, TOut = bf16, no K split): the epilogue applies that conv's ReLU
// backward, dpre = v * (act > 0), stores dpre as BF16 (nearest even) and
// writes the bias-gradient partials of the workgroup's pixels,
// bpart[blockIdx.x][co] = sum of the unrounded dpre (lane order, then the
// lane pair holding the same channel, then the four wavefronts in order).
template <class TIn, class TOut, int PM, int NB, bool BIAS_RELU, int SW = 32, bool SPLIT = false, bool DPRE = false>
__global__ __launch_bounds__(256) void k_conv3x3_tile(const TIn *__restrict__ in, const TIn *__restrict__ low, int cl,
                                                      const bf16 *__restrict__ W, const float *__restrict__ bias,
                                                      TOut *__restrict__ out, float *__restrict__ part, int npix, int H,
                                                      int Wd, int Ci, int Co, const bf16 *__restrict__ act = nullptr,
                                                      float *__restrict__ bpart = nullptr, TOut *__restrict__ out2 = nullptr,
                                                      int csplit = 0) {
    static_assert(!DPRE || (!SPLIT && !BIAS_RELU), "ReLU-backward epilogue: no K split, no bias");
    constexpr int RPT = 32 / SW;                 // image rows per MFMA pixel tile
    constexpr int ROWS = 4 * PM * RPT, IR = ROWS + 2, IC = SW + 2;
    constexpr int NCO = 32 * NB;
    constexpr int NIN = IR * IC * 4;             // 16-byte pieces of the input window
    constexpr int NW = NCO * 9 * 4;              // 16-byte pieces of the weight block
    constexpr int RIN = (NIN + 255) / 256, RW = (NW + 255) / 256;
    __shared__ __attribute__((aligned(16))) uint16_t s_in[IR * IC * kTilePS];
    __shared__ __attribute__((aligned(16))) uint16_t s_w[NCO * kTileWS];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row = lane & 31, kh = (lane >> 5) * 8;
    const int segs = Wd / SW, rblocks = H / ROWS;
    int t = blockIdx.x;
    const int seg = t % segs;
    t /= segs;
    const int rb = t % rblocks, img = t / rblocks;
    const int x0 = seg * SW, y0 = rb * ROWS;
    const int n0 = blockIdx.y * NCO;
    const int nst_all = Ci / kTileCK;
    const int sbeg = SPLIT ? (int)(blockIdx.z * nst_all / gridDim.z) : 0;
    const int send = SPLIT ? (int)((blockIdx.z + 1) * nst_all / gridDim.z) : nst_all;
    const int ry = row / SW, rx = row % SW;      // this lane's pixel in its MFMA tile

    // register prefetch of one stage
    Piece8<TIn> pin[RIN];
    Piece8<bf16> pw[RW];
    auto fetch = [&](int c0) {
#pragma unroll
        for (int j = 0; j < RIN; j++) {
            const int e = threadIdx.x + 256 * j;
            pin[j].zero();
            if (e < NIN) {
                const int q = e & 3, p = e >> 2;
                const int iy = p / IC, ix = p - iy * IC;
                const int yy = y0 + iy - 1, xx = x0 + ix - 1;
                if (yy >= 0 && yy < H && xx >= 0 && xx < Wd) {
                    const int c = c0 + q * 8;
                    if (c < cl)
                        pin[j].load(low + (((size_t)img * (H >> 1) + (yy >> 1)) * (Wd >> 1) + (xx >> 1)) * cl + c);
                    else
                        pin[j].load(in + (((size_t)img * H + yy) * Wd + xx) * (Ci - cl) + (c - cl));
                }
            }
        }
#pragma unroll
        for (int j = 0; j < RW; j++) {
            const int e = threadIdx.x + 256 * j;
            pw[j].zero();
            if (e < NW) {
                const int q = e & 3, r = e >> 2;
                const int co = r / 9, tap = r - co * 9;
                pw[j].load(W + ((size_t)(n0 + co) * 9 + tap) * Ci + c0 + q * 8);
            }
        }
    };
    auto store = [&]() {
#pragma unroll
        for (int j = 0; j < RIN; j++) {
            const int e = threadIdx.x + 256 * j;
            if (e < NIN) *(u16x8 *)&s_in[(e >> 2) * kTilePS + (e & 3) * 8] = pin[j].bits();
        }
#pragma unroll
        for (int j = 0; j < RW; j++) {
            const int e = threadIdx.x + 256 * j;
            if (e < NW) {
                const int q = e & 3, r = e >> 2;
                const int co = r / 9, tap = r - co * 9;
                *(u16x8 *)&s_w[co * kTileWS + tap * kTileCK + q * 8] = pw[j].bits();
            }
        }
    };

    f32x16 acc[PM][NB];
#pragma unroll
    for (int j = 0; j < PM; j++)
#pragma unroll
        for (int b = 0; b < NB; b++)
            for (int r = 0; r < 16; r++) acc[j][b][r] = 0.f;

    if (sbeg < send) fetch(sbeg * kTileCK);
    for (int s = sbeg; s < send; s++) {
        __syncthreads();                 // every wavefront is done with the previous stage
        store();
        __syncthreads();
        if (s + 1 < send) fetch((s + 1) * kTileCK);
#pragma unroll
        for (int tap = 0; tap < 9; tap++) {
            const int dy = tap / 3, dx = tap % 3;
#pragma unroll
            for (int kk = 0; kk < 2; kk++) {
                bf16x8 a[PM], bw[NB];
#pragma unroll
                for (int j = 0; j < PM; j++)
                    a[j] = __builtin_bit_cast(
                        bf16x8, *(const u16x8 *)&s_in[(((wave * PM + j) * RPT + ry + dy) * IC + rx + dx) * kTilePS +
                                                      kk * 16 + kh]);
#pragma unroll
                for (int b = 0; b < NB; b++)
                    bw[b] = __builtin_bit_cast(
                        bf16x8, *(const u16x8 *)&s_w[(b * 32 + row) * kTileWS + tap * kTileCK + kk * 16 + kh]);
#pragma unroll
                for (int j = 0; j < PM; j++)
#pragma unroll
                    for (int b = 0; b < NB; b++)
                        acc[j][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[j], bw[b], acc[j][b], 0, 0, 0);
            }
        }
    }
    if constexpr (DPRE) {
        __shared__ float s_b[4][32 * NB];
        float bs[NB];
#pragma unroll
        for (int b = 0; b < NB; b++) bs[b] = 0.f;
#pragma unroll
        for (int j = 0; j < PM; j++) {
            const int ty0 = y0 + (wave * PM + j) * RPT;
#pragma unroll
            for (int b = 0; b < NB; b++) {
                const int co = n0 + 32 * b + row;
                for (int r = 0; r < 16; r++) {
                    const int i = 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                    const size_t pix = ((size_t)img * H + ty0 + i / SW) * Wd + x0 + i % SW;
                    const float av = lb::bf16_to_f32(__builtin_bit_cast(uint16_t, act[pix * Co + co]));
                    const float v = av > 0.f ? acc[j][b][r] : 0.f;
                    out[pix * Co + co] = from_f<TOut>(v);
                    bs[b] += v;
                }
            }
        }
#pragma unroll
        for (int b = 0; b < NB; b++) {
            bs[b] += __shfl_xor(bs[b], 32);
            if (lane < 32) s_b[wave][32 * b + lane] = bs[b];
        }
        __syncthreads();
        if (threadIdx.x < 32 * NB) {
            const int t = threadIdx.x;
            bpart[(size_t)blockIdx.x * Co + n0 + t] = ((s_b[0][t] + s_b[1][t]) + s_b[2][t]) + s_b[3][t];
        }
    } else {
#pragma unroll
        for (int j = 0; j < PM; j++) {
            const int ty0 = y0 + (wave * PM + j) * RPT;
#pragma unroll
            for (int b = 0; b < NB; b++) {
                const int co = n0 + 32 * b + row;
                const float bv = (BIAS_RELU && !SPLIT) ? bias[co] : 0.f;
                // out2: channels [0, csplit) go to out ([pixel][csplit]),
                // the rest to out2 ([pixel][Co - csplit]); csplit % 32 == 0,
                // so a 32-channel group has one destination
                TOut *dst = out;
                int dc = Co, dco = co;
                if (out2) {
                    const bool hi = n0 + 32 * b >= csplit;
                    dst = hi ? out2 : out;
                    dc = hi ? Co - csplit : csplit;
                    dco = hi ? co - csplit : co;
                }
                for (int r = 0; r < 16; r++) {
                    const int i = 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                    const size_t pix = ((size_t)img * H + ty0 + i / SW) * Wd + x0 + i % SW;
                    float v = acc[j][b][r];
                    if (SPLIT) {
                        part[((size_t)blockIdx.z * npix + pix) * Co + co] = v;
                    } else {
                        if (BIAS_RELU) { v += bv; v = v > 0.f ? v : 0.f; }
                        dst[pix * dc + dco] = from_f<TOut>(v);
                    }
                }
            }
        }
    }
}

// Sum of the K-split shares of k_conv3x3_tile in order, then bias + ReLU
// (BIAS_RELU) and the output type; four channels per thread.
template <class TOut, bool BIAS_RELU>
__global__ void k_split_sum(const float *__restrict__ part, int ks, size_t n, int Co, const float *__restrict__ bias,
                            TOut *__restrict__ out) {
    const size_t e = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
    if (e >= n) return;
    float4 v = *(const float4 *)(part + e);
    for (int z = 1; z < ks; z++) {
        const float4 w = *(const float4 *)(part + (size_t)z * n + e);
        v.x += w.x; v.y += w.y; v.z += w.z; v.w += w.w;
    }
    float r[4] = {v.x, v.y, v.z, v.w};
    const int co = (int)(e % Co);
    for (int k = 0; k < 4; k++) {
        float x = r[k];
        if (BIAS_RELU) { x += bias[co + k]; x = x > 0.f ? x : 0.f; }
        out[e + k] = from_f<TOut>(x);
    }
}

// Weight gradient partials: block (co tile, k tile, pixel chunk); each of the
// 4 wavefronts reduces a quarter of the chunk, the block sums them in LDS.
template <class TIn, class TD = float>
__global__ __launch_bounds__(256) void k_conv3x3_wgrad(const TD *__restrict__ dout, const TIn *__restrict__ in,
                                                       float *__restrict__ partial, int npix, int H, int Wd,
                                                       int Ci, int Co, int chunk) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int co0 = blockIdx.x * 32, kt0 = blockIdx.y * 32;
    const int K = 9 * Ci;
    const int row = lane & 31, kh = (lane >> 5) * 8;
    // this lane's column k of the tile
    const int k = kt0 + row;
    const bool kvalid = k < K;
    const int tap = kvalid ? k / Ci : 0, ci = kvalid ? k - (k / Ci) * Ci : 0;
    const int dy = tap / 3 - 1, dx = tap % 3 - 1;
    const int HW = H * Wd;
    const int pbeg = blockIdx.z * chunk + wave * (chunk / 4);
    const int pend = min(npix, pbeg + chunk / 4);
    f32x16 acc;
    for (int r = 0; r < 16; r++) acc[r] = 0.f;
    // unrolled so the loads of several steps are in flight before their MFMAs
#pragma unroll 8
    for (int p0 = pbeg; p0 < pend; p0 += 16) {
        u16x8 av, bv;
        // pixel coordinates of this lane's first element, then stepped along w
        int p = p0 + kh;
        int n = p / HW, h = (p - n * HW) / Wd, w = p - n * HW - h * Wd;
        for (int j = 0; j < 8; j++, p++) {
            uint16_t gb = 0;
            uint16_t xb = 0;
            if (p < pend) {
                gb = bf16_bits(dout[(size_t)p * Co + co0 + row]);
                const int hh = h + dy, ww = w + dx;
                if (kvalid && hh >= 0 && hh < H && ww >= 0 && ww < Wd)
                    xb = bf16_bits(in[((size_t)(n * H + hh) * Wd + ww) * Ci + ci]);
            }
            av[j] = gb;
            bv[j] = xb;
            if (++w == Wd) { w = 0; if (++h == H) { h = 0; n++; } }
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, av), __builtin_bit_cast(bf16x8, bv), acc, 0, 0, 0);
    }
    __shared__ float sh[4][32][33];
    for (int r = 0; r < 16; r++) {
        const int i = 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        sh[wave][i][row] = acc[r];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < 32 * 32; e += 256) {
        const int i = e >> 5, j = e & 31;
        const float s = sh[0][i][j] + sh[1][i][j] + sh[2][i][j] + sh[3][i][j];
        if (kt0 + j < K) partial[((size_t)blockIdx.z * Co + co0 + i) * K + kt0 + j] = s;
    }
}

// NHWC -> channel-major [C][npix] BF16 (64-pixel x 64-channel tiles through
// LDS: coalesced reads along channels, coalesced writes along pixels).
template <class T>
__global__ __launch_bounds__(256) void k_to_cmajor(const T *__restrict__ src, bf16 *__restrict__ dst, int npix, int C) {
    __shared__ uint16_t tile[64][65];
    const int m0 = blockIdx.x * 64, c0 = blockIdx.y * 64;
    const int t = threadIdx.x & 63, r0 = threadIdx.x >> 6;
    for (int r = r0; r < 64; r += 4) {
        const int m = m0 + r, c = c0 + t;
        tile[t][r] = (m < npix && c < C) ? bf16_bits(src[(size_t)m * C + c]) : (uint16_t)0;
    }
    __syncthreads();
    for (int r = r0; r < 64; r += 4) {
        const int c = c0 + r, m = m0 + t;
        if (c < C && m < npix) dst[(size_t)c * npix + m] = __builtin_bit_cast(bf16, tile[r][t]);
    }
}

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ u32x4 ld8w(const bf16 *p) { return *(const u32x4 *)p; }

// Weight gradient from channel-major BF16 operands (dout_t [Co][npix],
// in_t [Ci][npix]): a lane's 8 consecutive pixels of one channel are one
// 16-byte load (A: output channel co, B: input channel ci of tap (dy, dx),
// pixels shifted by dy rows and dx columns).  A 16-pixel step stays in one
// image row (W % 16 == 0), so the row shift dy only masks whole groups and
// the column shift dx = +-1 is a one-element funnel shift across two aligned
// loads, with the zero padding at the row ends.  Same blocks, pixel split
// and partial layout as k_conv3x3_wgrad.
__global__ __launch_bounds__(256) void k_conv3x3_wgrad_cm(const bf16 *__restrict__ dout_t, const bf16 *__restrict__ in_t,
                                                          float *__restrict__ partial, int npix, int H, int Wd, int Ci,
                                                          int Co, int chunk) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int co0 = blockIdx.x * 32, kt0 = blockIdx.y * 32;
    const int K = 9 * Ci;
    const int row = lane & 31, kh = (lane >> 5) * 8;
    const int k = kt0 + row;
    const bool kvalid = k < K;
    const int tap = kvalid ? k / Ci : 4, ci = kvalid ? k - (k / Ci) * Ci : 0;
    const int dy = tap / 3 - 1, dx = tap % 3 - 1;
    const int HW = H * Wd;
    const int pbeg = blockIdx.z * chunk + wave * (chunk / 4);
    const int pend = min(npix, pbeg + chunk / 4);
    const bf16 *arow = dout_t + (size_t)(co0 + row) * npix;
    const bf16 *brow = in_t + (size_t)ci * npix;
    f32x16 acc;
    for (int r = 0; r < 16; r++) acc[r] = 0.f;
    int p = pbeg + kh;
    int n = p / HW, h = (p - n * HW) / Wd, w = p - n * HW - h * Wd;
    // batches of U pixel steps: loads first, then the MFMAs (one memory wait
    // per batch); steps past pend carry zero operands
    constexpr int U = 4;
    for (int p0 = pbeg; p0 < pend; p0 += 16 * U) {
        u32x4 av[U], cur[U];
        unsigned edge[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            av[u] = u32x4{0u, 0u, 0u, 0u};
            cur[u] = u32x4{0u, 0u, 0u, 0u};
            edge[u] = 0u;
            if (p0 + 16 * u < pend) {
                av[u] = ld8w(arow + p0 + 16 * u + kh);
                const int hh = h + dy;
                if (kvalid && hh >= 0 && hh < H) {
                    const bf16 *q = brow + (size_t)(n * H + hh) * Wd + w;
                    cur[u] = ld8w(q);
                    if (dx < 0 && w > 0) edge[u] = ld8w(q - 8)[3];
                    if (dx > 0 && w + 8 < Wd) edge[u] = ld8w(q + 8)[0];
                }
                w += 16;
                if (w >= Wd) { w -= Wd; if (++h == H) { h = 0; n++; } }
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            u32x4 bv = cur[u];
            if (dx < 0) {          // element j <- x[w + j - 1]
                bv[0] = __builtin_amdgcn_alignbit(cur[u][0], edge[u], 16);
                bv[1] = __builtin_amdgcn_alignbit(cur[u][1], cur[u][0], 16);
                bv[2] = __builtin_amdgcn_alignbit(cur[u][2], cur[u][1], 16);
                bv[3] = __builtin_amdgcn_alignbit(cur[u][3], cur[u][2], 16);
            } else if (dx > 0) {   // element j <- x[w + j + 1]
                bv[0] = __builtin_amdgcn_alignbit(cur[u][1], cur[u][0], 16);
                bv[1] = __builtin_amdgcn_alignbit(cur[u][2], cur[u][1], 16);
                bv[2] = __builtin_amdgcn_alignbit(cur[u][3], cur[u][2], 16);
                bv[3] = __builtin_amdgcn_alignbit(edge[u], cur[u][3], 16);
            }
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, av[u]), __builtin_bit_cast(bf16x8, bv),
                                                          acc, 0, 0, 0);
        }
    }
    __shared__ float sh[4][32][33];
    for (int r = 0; r < 16; r++) {
        const int i = 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        sh[wave][i][row] = acc[r];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < 32 * 32; e += 256) {
        const int i = e >> 5, j = e & 31;
        const float s = sh[0][i][j] + sh[1][i][j] + sh[2][i][j] + sh[3][i][j];
        if (kt0 + j < K) partial[((size_t)blockIdx.z * Co + co0 + i) * K + kt0 + j] = s;
    }
}

// ---------------------------------------------------------------------------
// Weight gradient through LDS tiles (Co % 32 == 0, Ci % 8 == 0; TW = 32:
// W % 32 == 0, H % 8 == 0; TW = 16: W % 16 == 0, H % 16 == 0): block (32
// output channels, 32 input channels, chunk of pixel tiles).  A pixel tile
// is 256 / TW image rows x TW pixels.  Per tile the gradient block (256
// pixels x 32 channels, FP32 -> BF16) and the input window (tile rows + 2 x
// TW + 4 pixels x 32 channels, zero outside the image) are written to LDS
// channel-major, four pixels per 8-byte store, so that an MFMA operand (8
// consecutive pixels of one channel) is one 16-byte LDS read.  Wavefront w
// takes a quarter of the tile's rows (four 16-pixel K-steps) and keeps all
// nine taps' 32 x 32 accumulators; the column shift of a tap is a funnel
// shift across an aligned 16-byte and 4-byte read of the window row.  Two
// workgroups share a CU (at most 256 registers per lane), so one stages its
// next tile while the other runs its MFMAs.  The four wavefronts' tiles are
// added in LDS in a fixed
// order and written as the chunk's partial (partial[chunk][co][tap][ci]).
// ---------------------------------------------------------------------------
constexpr int kWgXS = 520;      // LDS stride of one input channel (<= 18 window rows; 260 dwords = 4 mod 64 banks)
constexpr int kWgGS = 264;      // LDS stride of one output channel (256 pixels; 132 dwords = 4 mod 64 banks)

__device__ __forceinline__ uint2 pack4(uint16_t a, uint16_t b, uint16_t c, uint16_t d) {
    return uint2{(uint32_t)a | ((uint32_t)b << 16), (uint32_t)c | ((uint32_t)d << 16)};
}

template <int TW, class TD = float>
__global__ __launch_bounds__(256, 2) void k_conv3x3_wgrad_tile(const TD *__restrict__ dout, const bf16 *__restrict__ in,
                                                            const bf16 *__restrict__ low, int cl,
                                                            float *__restrict__ partial, int H, int Wd, int Ci, int Co,
                                                            int ntiles, int tpw) {
    __shared__ __attribute__((aligned(16))) uint16_t sX[32 * kWgXS];
    __shared__ __attribute__((aligned(16))) uint16_t sG[32 * kWgGS];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int col = lane & 31, kh = (lane >> 5) * 8;
    const int co0 = blockIdx.x * 32, ci0 = blockIdx.y * 32;
    const int K = 9 * Ci;
    constexpr int TR = 256 / TW;                 // tile rows
    constexpr int WIN = TW == 32 ? 40 : 24;      // LDS stride of a window row (TW + 4 pixels used)
    constexpr int NQX = (TW + 4) / 4;            // 4-pixel groups of a window row
    constexpr int QPR = TW / 4;                  // 4-pixel groups of a tile row
    constexpr int KPR = TW / 16;                 // 16-pixel K-steps of a tile row
    const int segs = Wd / TW, rblocks = H / TR;
    const int t0 = blockIdx.z * tpw, t1 = min(ntiles, t0 + tpw);
    constexpr int NXQ = (TR + 2) * NQX * 4;      // window: rows x 4-pixel groups x four 8-channel groups
    constexpr int RX = (NXQ + 255) / 256;
    u16x8 xq[RX][4];
    Piece8<TD> gq[4];
    auto fetch = [&](int tile) {
        const int seg = tile % segs, rest = tile / segs;
        const int y0 = (rest % rblocks) * TR, img = rest / rblocks, x0 = seg * TW;
#pragma unroll
        for (int j = 0; j < RX; j++) {
            const int e = tid + 256 * j;
#pragma unroll
            for (int p = 0; p < 4; p++) xq[j][p] = u16x8{0, 0, 0, 0, 0, 0, 0, 0};
            if (e < NXQ) {
                const int q = e & 3, r = e >> 2;
                const int iy = r / NQX, qx = r - iy * NQX;
                const int yy = y0 + iy - 1, c = ci0 + 8 * q;
                if (yy >= 0 && yy < H && c < Ci) {
#pragma unroll
                    for (int p = 0; p < 4; p++) {
                        const int xx = x0 + 4 * qx + p - 1;
                        if (xx >= 0 && xx < Wd)
                            xq[j][p] = c < cl ? *(const u16x8 *)(low + (((size_t)img * (H >> 1) + (yy >> 1)) * (Wd >> 1) +
                                                                        (xx >> 1)) * cl + c)
                                              : *(const u16x8 *)(in + (((size_t)img * H + yy) * Wd + xx) * (Ci - cl) + (c - cl));
                    }
                }
            }
        }
        {   // gradient: 64 4-pixel groups x four 8-channel groups = one piece per thread
            const int q = tid & 3, r = tid >> 2;
            const int iy = r / QPR, qx = r % QPR;
            const TD *g = dout + (((size_t)img * H + y0 + iy) * Wd + x0 + 4 * qx) * Co + co0 + 8 * q;
#pragma unroll
            for (int p = 0; p < 4; p++) gq[p].load(g + (size_t)p * Co);
        }
    };
    auto store = [&]() {
#pragma unroll
        for (int j = 0; j < RX; j++) {
            const int e = tid + 256 * j;
            if (e < NXQ) {
                const int q = e & 3, r = e >> 2;
                const int iy = r / NQX, qx = r - iy * NQX;
                const int base = iy * WIN + 4 * qx;
#pragma unroll
                for (int k = 0; k < 8; k++)
                    *(uint2 *)&sX[(8 * q + k) * kWgXS + base] = pack4(xq[j][0][k], xq[j][1][k], xq[j][2][k], xq[j][3][k]);
            }
        }
        {
            const int q = tid & 3, r = tid >> 2;
            const int base = 4 * r;              // tile pixel (r / QPR) TW + 4 (r % QPR)
            const u16x8 b0 = gq[0].bits(), b1 = gq[1].bits(), b2 = gq[2].bits(), b3 = gq[3].bits();
#pragma unroll
            for (int k = 0; k < 8; k++) *(uint2 *)&sG[(8 * q + k) * kWgGS + base] = pack4(b0[k], b1[k], b2[k], b3[k]);
        }
    };

    f32x16 acc[9];
#pragma unroll
    for (int t = 0; t < 9; t++)
        for (int r = 0; r < 16; r++) acc[t][r] = 0.f;
    for (int t = t0; t < t1; t++) {
        __syncthreads();                 // every wavefront is done with the previous tile
        fetch(t);
        store();
        __syncthreads();
#pragma unroll
        for (int s = 0; s < 4; s++) {
            const int trow = wave * (4 / KPR) + s / KPR, pxb = (s % KPR) * 16 + kh;
            const bf16x8 a = __builtin_bit_cast(bf16x8, *(const u16x8 *)&sG[col * kWgGS + trow * TW + pxb]);
#pragma unroll
            for (int dy = 0; dy < 3; dy++) {
                const uint16_t *xp = &sX[col * kWgXS + (trow + dy) * WIN + pxb];
                const u32x4 d = *(const u32x4 *)xp;
                const uint32_t d4 = *(const uint32_t *)(xp + 8);
                const u32x4 b1 = u32x4{__builtin_amdgcn_alignbit(d[1], d[0], 16), __builtin_amdgcn_alignbit(d[2], d[1], 16),
                                       __builtin_amdgcn_alignbit(d[3], d[2], 16), __builtin_amdgcn_alignbit(d4, d[3], 16)};
                const u32x4 b2 = u32x4{d[1], d[2], d[3], d4};
                acc[3 * dy] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, __builtin_bit_cast(bf16x8, d), acc[3 * dy], 0, 0, 0);
                acc[3 * dy + 1] =
                    __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, __builtin_bit_cast(bf16x8, b1), acc[3 * dy + 1], 0, 0, 0);
                acc[3 * dy + 2] =
                    __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, __builtin_bit_cast(bf16x8, b2), acc[3 * dy + 2], 0, 0, 0);
            }
        }
    }
    __syncthreads();
    float *red = (float *)sX;            // 4 x 32 x 33 floats
#pragma unroll
    for (int tap = 0; tap < 9; tap++) {
        for (int r = 0; r < 16; r++) {
            const int i = 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            red[(wave * 32 + i) * 33 + col] = acc[tap][r];
        }
        __syncthreads();
        for (int e = tid; e < 32 * 32; e += 256) {
            const int i = e >> 5, j = e & 31;
            const float sum = ((red[i * 33 + j] + red[(32 + i) * 33 + j]) + red[(64 + i) * 33 + j]) + red[(96 + i) * 33 + j];
            if (ci0 + j < Ci) partial[((size_t)blockIdx.z * Co + co0 + i) * K + tap * Ci + ci0 + j] = sum;
        }
        __syncthreads();
    }
}


// Fixed-order sum of many chunks in one pass, for one or two independent
// arrays (segment a: blocks [0, nblk_a), segment b: the rest): thread
// (element el of a 16-element group, phase ph of 16) adds chunks ph, ph + 16,
// ph + 32, ... in order; the 16 phase sums are then added in LDS in phase
// order (for nchunk <= 16 this is the plain in-order sum).
struct SumSeg {
    const float *partial;
    int nchunk;
    size_t n;
    float *out;
};
__global__ __launch_bounds__(256) void k_sum_chunks16(SumSeg a, SumSeg b, unsigned nblk_a) {
    const bool first = blockIdx.x < nblk_a;
    const float *partial = first ? a.partial : b.partial;
    const int nchunk = first ? a.nchunk : b.nchunk;
    const size_t n = first ? a.n : b.n;
    float *out = first ? a.out : b.out;
    const unsigned blk = first ? blockIdx.x : blockIdx.x - nblk_a;
    const int el = threadIdx.x & 15, ph = threadIdx.x >> 4;
    const size_t e = (size_t)blk * 16 + el;
    float s = 0.f;
    if (e < n) {
#pragma unroll 8
        for (int c = ph; c < nchunk; c += 16) s += partial[(size_t)c * n + e];
    }
    __shared__ float sh[16][17];
    sh[ph][el] = s;
    __syncthreads();
    if (threadIdx.x < 16 && e < n) {
        float v = 0.f;
        for (int p = 0; p < 16; p++) v += sh[p][el];
        out[e] = v;
    }
}

// The same sums for segments whose n is a multiple of 4 (16-byte aligned
// arrays), four adjacent elements per thread, with the same per-element
// order: a segment of more than 16 chunks takes blocks of 16 element quads x
// 16 phases as above; one of at most 16 chunks (where the phase sums above
// are the plain in-order sum) takes 256 quads per block, each thread adding
// the chunks in order with no LDS step.
inline unsigned sum4_blocks(const SumSeg &g) {
    const size_t q = g.n / 4;
    return (unsigned)(g.nchunk > 16 ? (q + 15) / 16 : (q + 255) / 256);
}
__global__ __launch_bounds__(256) void k_sum_chunks16_x4(SumSeg a, SumSeg b, unsigned nblk_a) {
    const bool first = blockIdx.x < nblk_a;
    const SumSeg g = first ? a : b;
    const unsigned blk = first ? blockIdx.x : blockIdx.x - nblk_a;
    const size_t nq = g.n / 4;
    const float4 *part = (const float4 *)g.partial;
    if (g.nchunk <= 16) {
        const size_t e = (size_t)blk * 256 + threadIdx.x;
        if (e >= nq) return;
        float4 s = float4{0.f, 0.f, 0.f, 0.f};
        for (int c = 0; c < g.nchunk; c++) {
            const float4 v = part[(size_t)c * nq + e];
            s.x += v.x; s.y += v.y; s.z += v.z; s.w += v.w;
        }
        ((float4 *)g.out)[e] = s;
        return;
    }
    const int el = threadIdx.x & 15, ph = threadIdx.x >> 4;
    const size_t e = (size_t)blk * 16 + el;
    float4 s = float4{0.f, 0.f, 0.f, 0.f};
    if (e < nq) {
#pragma unroll 8
        for (int c = ph; c < g.nchunk; c += 16) {
            const float4 v = part[(size_t)c * nq + e];
            s.x += v.x; s.y += v.y; s.z += v.z; s.w += v.w;
        }
    }
    __shared__ float4 sh[16][17];
    sh[ph][el] = s;
    __syncthreads();
    if (threadIdx.x < 16 && e < nq) {
        float4 v = float4{0.f, 0.f, 0.f, 0.f};
        for (int p = 0; p < 16; p++) {
            const float4 w = sh[p][el];
            v.x += w.x; v.y += w.y; v.z += w.z; v.w += w.w;
        }
        ((float4 *)g.out)[e] = v;
    }
}

// Bias-gradient chunking of k_relu_bwd_bias: at most 1,024 chunks of >= 64
// pixels (more chunks keep more loads in flight here but cost more in the
// chunk sum, whose threads add up to nchunk / 16 partials each).
struct BiasChunks { int chunk, nchunk; };
inline BiasChunks bias_chunks(int npix) {
    const int chunk = npix > 64 * 1024 ? (npix + 1023) / 1024 : 64;
    return {chunk, (npix + chunk - 1) / chunk};
}

// ReLU backward on the stored post-activation, dpre = d * (act > 0), and the
// bias-gradient partials of the same pass, partial[chunk][co] =
// sum_{m in chunk} dpre[m, co] (FP32).  dpre is stored as BF16 (nearest
// even): its consumers, the weight gradient and the input gradient, round it
// so before their MFMAs, so they see the same operands.  Block = one pixel chunk x all Co channels
// (Co % 4 == 0, Co <= 1024): thread (4-channel group, pixel phase) walks the
// chunk with stride 256 / (Co / 4); the phases are added in LDS in a fixed
// order.
// dpool != null: the incoming gradient is d[m] + dpool[parent of m] / 4 (the
// skip gradient plus the 2x2 average-pool backward of an H x W level), summed
// here instead of in a pass of its own.
// dup != null (d unused): the incoming gradient is the nearest-upsampling
// backward of dup, a 2H x 2W gradient with Co channels: the sum of m's four
// children, (c00 + c01) + (c10 + c11).
// hdy != null (d unused; the last conv, whose output feeds the 1x1 head):
// the incoming gradient is the head's input gradient, sum_o hdy[m, o]
// hw[o, c] in o order (hw staged in LDS as floats, ho <= 8, Co <= 256),
// formed here rather than written and read back.
__global__ __launch_bounds__(256) void k_relu_bwd_bias(const float *__restrict__ d, const bf16 *__restrict__ act,
                                                       bf16 *__restrict__ dpre, int npix, int Co, int chunk,
                                                       float *__restrict__ partial, const float *__restrict__ dpool,
                                                       int H, int Wd, const float *__restrict__ dup = nullptr,
                                                       const float *__restrict__ hdy = nullptr,
                                                       const bf16 *__restrict__ hw = nullptr, int ho = 0) {
    const int tpp = Co >> 2, nph = 256 / tpp;
    const int t = threadIdx.x, cg = t % tpp, ph = t / tpp;
    __shared__ float hws[8 * 256];
    if (hdy) {                           // uniform: the whole workgroup takes it
        for (int e = t; e < ho * Co; e += 256) hws[e] = from_bf16(hw[e]);
        __syncthreads();
    }
    const int pbeg = blockIdx.x * chunk, pend = min(npix, pbeg + chunk);
    float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
    if (ph < nph) {
#pragma unroll 4
        for (int m = pbeg + ph; m < pend; m += nph) {
            const size_t o = (size_t)m * Co + 4 * cg;
            float4 g;
            if (dup) {
                const int x = m % Wd, r = m / Wd, y = r % H, b = r / H;
                const float *q = dup + (((size_t)b * (2 * H) + 2 * y) * (2 * Wd) + 2 * x) * Co + 4 * cg;
                const float4 a = *(const float4 *)q, bb = *(const float4 *)(q + Co);
                const float4 cc = *(const float4 *)(q + (size_t)2 * Wd * Co), dd = *(const float4 *)(q + (size_t)2 * Wd * Co + Co);
                g = float4{(a.x + bb.x) + (cc.x + dd.x), (a.y + bb.y) + (cc.y + dd.y), (a.z + bb.z) + (cc.z + dd.z),
                           (a.w + bb.w) + (cc.w + dd.w)};
            } else if (hdy) {
                float gy[8];
#pragma unroll
                for (int k = 0; k < 8; k++) gy[k] = k < ho ? hdy[(size_t)m * ho + k] : 0.f;
                float rr[4];
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    float sum = 0.f;
#pragma unroll
                    for (int q = 0; q < 8; q++)
                        if (q < ho) sum += gy[q] * hws[q * Co + 4 * cg + k];
                    rr[k] = sum;
                }
                g = float4{rr[0], rr[1], rr[2], rr[3]};
            } else {
                g = *(const float4 *)(d + o);
            }
            if (dpool) {
                const int x = m % Wd, r = m / Wd, y = r % H, b = r / H;
                const float4 sp = *(const float4 *)(dpool + (((size_t)b * (H / 2) + y / 2) * (Wd / 2) + x / 2) * Co + 4 * cg);
                g = float4{g.x + 0.25f * sp.x, g.y + 0.25f * sp.y, g.z + 0.25f * sp.z, g.w + 0.25f * sp.w};
            }
            const uint2 a = *(const uint2 *)(act + o);
            float4 r;
            r.x = lb::bf16_to_f32((uint16_t)(a.x & 0xffffu)) > 0.f ? g.x : 0.f;
            r.y = lb::bf16_to_f32((uint16_t)(a.x >> 16)) > 0.f ? g.y : 0.f;
            r.z = lb::bf16_to_f32((uint16_t)(a.y & 0xffffu)) > 0.f ? g.z : 0.f;
            r.w = lb::bf16_to_f32((uint16_t)(a.y >> 16)) > 0.f ? g.w : 0.f;
            *(uint2 *)(dpre + o) = uint2{(uint32_t)lb::f32_to_bf16(r.x) | ((uint32_t)lb::f32_to_bf16(r.y) << 16),
                                          (uint32_t)lb::f32_to_bf16(r.z) | ((uint32_t)lb::f32_to_bf16(r.w) << 16)};
            s0 += r.x; s1 += r.y; s2 += r.z; s3 += r.w;
        }
    }
    __shared__ float sh[1024];
    if (ph < nph) {
        float *q = sh + ph * Co + 4 * cg;
        q[0] = s0; q[1] = s1; q[2] = s2; q[3] = s3;
    }
    __syncthreads();
    if (t < Co) {
        float v = 0.f;
        for (int p = 0; p < nph; p++) v += sh[p * Co + t];
        partial[(size_t)blockIdx.x * Co + t] = v;
    }
}

// 2x2 average pooling, eight channels per thread (C % 8 == 0).
__global__ void k_avgpool(const bf16 *in, bf16 *out, int B, int H, int Wd, int C) {
    const unsigned C8 = C >> 3, H2 = H / 2, W2 = Wd / 2;
    const unsigned e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (unsigned)B * H2 * W2 * C8) return;
    const unsigned c = (e % C8) * 8, p = e / C8;
    const unsigned x = p % W2, r = p / W2, y = r % H2, b = r / H2;
    const bf16 *q = in + (((size_t)b * H + 2 * y) * Wd + 2 * x) * C + c;
    const u16x8 v00 = *(const u16x8 *)q, v01 = *(const u16x8 *)(q + C);
    const u16x8 v10 = *(const u16x8 *)(q + (size_t)Wd * C), v11 = *(const u16x8 *)(q + (size_t)Wd * C + C);
    u16x8 o;
    for (int k = 0; k < 8; k++) {
        const float sum = (lb::bf16_to_f32(v00[k]) + lb::bf16_to_f32(v01[k])) +
                          (lb::bf16_to_f32(v10[k]) + lb::bf16_to_f32(v11[k]));
        o[k] = lb::f32_to_bf16(0.25f * sum);
    }
    *(u16x8 *)(out + (size_t)p * C + c) = o;
}

// out[b, y, x, :] = [low[b, y/2, x/2, 0:Cl] | skip[b, y, x, 0:Cs]], eight
// channels per thread (Cl % 8 == 0, Cs % 8 == 0).
__global__ void k_upcat(const bf16 *low, const bf16 *skip, bf16 *out, int B, int H, int Wd, int Cl, int Cs) {
    const unsigned C8 = (Cl + Cs) >> 3;
    const unsigned e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (unsigned)B * H * Wd * C8) return;
    const unsigned c = (e % C8) * 8, p = e / C8;
    const unsigned x = p % Wd, r = p / Wd, y = r % H, b = r / H;
    const u16x8 v = c < (unsigned)Cl ? *(const u16x8 *)(low + (((size_t)b * (H / 2) + y / 2) * (Wd / 2) + x / 2) * Cl + c)
                                     : *(const u16x8 *)(skip + (size_t)p * Cs + (c - Cl));
    *(u16x8 *)(out + (size_t)p * (Cl + Cs) + c) = v;
}

// Backward of upcat: dlow = sum of the 4 children; dskip = slice (four
// channels per thread).
__global__ void k_upcat_bwd_low(const float *dcat, float *dlow, int B, int H, int Wd, int Cl, int Cs) {
    const unsigned C4 = Cl >> 2, H2 = H / 2, W2 = Wd / 2, C = Cl + Cs;
    const unsigned e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (unsigned)B * H2 * W2 * C4) return;
    const unsigned c = (e % C4) * 4, p = e / C4;
    const unsigned x = p % W2, r = p / W2, y = r % H2, b = r / H2;
    const float *q = dcat + (((size_t)b * H + 2 * y) * Wd + 2 * x) * C + c;
    const float4 a = *(const float4 *)q, bb = *(const float4 *)(q + C);
    const float4 cc = *(const float4 *)(q + (size_t)Wd * C), dd = *(const float4 *)(q + (size_t)Wd * C + C);
    *(float4 *)(dlow + (size_t)p * Cl + c) = float4{(a.x + bb.x) + (cc.x + dd.x), (a.y + bb.y) + (cc.y + dd.y),
                                                    (a.z + bb.z) + (cc.z + dd.z), (a.w + bb.w) + (cc.w + dd.w)};
}
__global__ void k_upcat_bwd_skip(const float *dcat, float *dskip, int B, int H, int Wd, int Cl, int Cs) {
    const unsigned C4 = Cs >> 2;
    const unsigned e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (unsigned)B * H * Wd * C4) return;
    const unsigned c = (e % C4) * 4, p = e / C4;
    *(float4 *)(dskip + (size_t)p * Cs + c) = *(const float4 *)(dcat + (size_t)p * (Cl + Cs) + Cl + c);
}

// 1x1 output head (FP32), one pixel per thread (C % 8 == 0, C <= 256,
// O <= 8): y[m, o] = sum_c x[m, c] Wh[o, c] + bh[o], c ascending.  The
// weights are staged in LDS as floats once per workgroup.
__global__ __launch_bounds__(256) void k_head_fwd(const bf16 *x, const bf16 *Wh, const float *bh, float *y, int npix,
                                                  int C, int O) {
    __shared__ float wsh[8 * 256];
    for (int e = threadIdx.x; e < O * C; e += 256) wsh[e] = from_bf16(Wh[e]);
    __syncthreads();
    const unsigned m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= (unsigned)npix) return;
    float s[8];
    for (int o = 0; o < 8; o++) s[o] = 0.f;
    for (int c = 0; c < C; c += 8) {
        const u16x8 v = *(const u16x8 *)(x + (size_t)m * C + c);
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const float xv = lb::bf16_to_f32(v[k]);
#pragma unroll
            for (int o = 0; o < 8; o++)
                if (o < O) s[o] += xv * wsh[o * C + c + k];
        }
    }
#pragma unroll
    for (int o = 0; o < 8; o++)
        if (o < O) y[(size_t)m * O + o] = s[o] + bh[o];
}
// dx[m, c] = sum_o dy[m, o] Wh[o, c], four channels per thread (weights
// staged in LDS as floats, O <= 8, C <= 256).
__global__ __launch_bounds__(256) void k_head_bwd_data(const float *dy, const bf16 *Wh, float *dx, int npix, int C,
                                                       int O) {
    __shared__ float wsh[8 * 256];
    for (int e = threadIdx.x; e < O * C; e += 256) wsh[e] = from_bf16(Wh[e]);
    __syncthreads();
    const unsigned C4 = C >> 2;
    const unsigned e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (unsigned)npix * C4) return;
    const unsigned c = (e % C4) * 4, m = e / C4;
    float g[8];
#pragma unroll
    for (int o = 0; o < 8; o++) g[o] = o < O ? dy[(size_t)m * O + o] : 0.f;
    float r[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        float sum = 0.f;
#pragma unroll
        for (int o = 0; o < 8; o++)
            if (o < O) sum += g[o] * wsh[o * C + c + k];
        r[k] = sum;
    }
    *(float4 *)(dx + (size_t)m * C + c) = float4{r[0], r[1], r[2], r[3]};
}
// dWh[o, c] and dbh[o] partials of one pixel chunk (C % 8 == 0, C <= 256,
// O <= 8): thread (8-channel group, pixel phase) loads 16 bytes of the input
// per pixel and accumulates all O outputs for its channels (and the bias
// sums); the phases are added in LDS in a fixed order; partial[chunk][o][c],
// with the bias at c = C; k_sum_chunks16 adds the chunks.
__global__ __launch_bounds__(256) void k_head_bwd_w(const float *dy, const bf16 *x, int npix, int C, int O, int chunk,
                                                    float *partial) {
    const int tpp = C >> 3, nph = 256 / tpp;
    const int t = threadIdx.x, cg = t % tpp, ph = t / tpp;
    const int pbeg = blockIdx.x * chunk, pend = min(npix, pbeg + chunk);
    float acc[8][8], accb[8];
    for (int o = 0; o < 8; o++) {
        accb[o] = 0.f;
        for (int k = 0; k < 8; k++) acc[o][k] = 0.f;
    }
    if (ph < nph)
#pragma unroll 2
        for (int m = pbeg + ph; m < pend; m += nph) {
            const u16x8 v = *(const u16x8 *)(x + (size_t)m * C + 8 * cg);
            float xv[8];
            for (int k = 0; k < 8; k++) xv[k] = lb::bf16_to_f32(v[k]);
            for (int o = 0; o < 8; o++)
                if (o < O) {
                    const float g = dy[(size_t)m * O + o];
                    for (int k = 0; k < 8; k++) acc[o][k] += g * xv[k];
                    accb[o] += g;
                }
        }
    __shared__ float sh[64 * 8 * 33];            // nph x O x (C + 1) <= 2048 / C x 8 x (C + 1)
    const int np = O * (C + 1);
    if (ph < nph)
        for (int o = 0; o < 8; o++)
            if (o < O) {
                for (int k = 0; k < 8; k++) sh[ph * np + o * (C + 1) + 8 * cg + k] = acc[o][k];
                if (cg == 0) sh[ph * np + o * (C + 1) + C] = accb[o];
            }
    __syncthreads();
    for (int e = t; e < np; e += 256) {
        float v = 0.f;
        for (int p = 0; p < nph; p++) v += sh[p * np + e];
        partial[(size_t)blockIdx.x * np + e] = v;
    }
}
// partial (o, c) index -> dW [o][c], bias (o, C) -> db[o]
__global__ void k_head_scatter(const float *sum, int C, int O, float *dW, float *db) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= O * (C + 1)) return;
    const int o = e / (C + 1), c = e % (C + 1);
    if (c < C) dW[o * C + c] = sum[e];
    else db[o] = sum[e];
}

// Per-block sums (double) of the mask weights: partial[block].
__global__ __launch_bounds__(256) void k_mask_sum(const float *mask, int npix, double *partial) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    __shared__ double sh[256];
    sh[threadIdx.x] = e < (size_t)npix ? (double)mask[e] : 0.0;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

// Masked mean-squared error: dy = 2 (y - t) mask / count with count = O x
// (sum of the mask weights, *wsum; zero gradients for an all-zero mask), loss
// partials.
__global__ __launch_bounds__(256) void k_mse(const float *y, const float *t, const float *mask, float *dy, int npix,
                                             int O, const double *wsum, double *partial) {
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const double cnt = *wsum * O;
    const float inv_count = cnt > 0 ? (float)(1.0 / cnt) : 0.f;
    double l = 0;
    if (e < (size_t)npix * O) {
        float w = mask[e / O];
        float d = (y[e] - t[e]) * w;
        dy[e] = 2.f * d * inv_count;
        l = (double)d * d;
    }
    __shared__ double sh[256];
    sh[threadIdx.x] = l;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}
// Sum of n doubles by one 256-thread block: strided per-thread sums, then a
// fixed tree (deterministic).
__global__ __launch_bounds__(256) void k_sum_d(const double *p, int n, double *out) {
    double s = 0;
    for (int i = threadIdx.x; i < n; i += 256) s += p[i];
    __shared__ double sh[256];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

// Adam on FP32 masters, then refresh the BF16 copy.
__global__ void k_adam(float *w, const float *g, float *mo, float *vo, bf16 *wb, size_t n, float lr, float b1,
                       float b2, float eps, float c1, float c2) {
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    float gi = g[e];
    float m = b1 * mo[e] + (1.f - b1) * gi;
    float v = b2 * vo[e] + (1.f - b2) * gi * gi;
    mo[e] = m;
    vo[e] = v;
    float upd = lr * (m / c1) / (sqrtf(v / c2) + eps);
    float nw = w[e] - upd;
    w[e] = nw;
    if (wb) wb[e] = to_bf16(nw);
}

__global__ void k_to_bf16(const float *w, bf16 *wb, size_t n) {
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) wb[e] = to_bf16(w[e]);
}

// W'[ci, tap', co] = W[co, 8 - tap', ci] (flipped, transposed: backward
// data) for every conv in one launch.  tab (device memory) holds per conv
// {wt offset, wb offset, Ci, Co}; element e of the packed flipped copies
// belongs to the last conv whose wt offset is <= e.
__global__ void k_flip_transpose_all(const bf16 *wb, bf16 *wt, const long long *tab, int nconv, size_t total) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    int j = 0;
    while (j + 1 < nconv && (size_t)tab[4 * (j + 1)] <= e) j++;
    const size_t local = e - (size_t)tab[4 * j];
    const size_t woff = (size_t)tab[4 * j + 1];
    const int Ci = (int)tab[4 * j + 2], Co = (int)tab[4 * j + 3];
    const int co = (int)(local % Co);
    const size_t r = local / Co;
    const int tap = (int)(r % 9), ci = (int)(r / 9);
    wt[e] = wb[woff + ((size_t)co * 9 + (8 - tap)) * Ci + ci];
}

// Input tensor: fp32 (B, H, W, cin) -> bf16 (B, H, W, cpad), zero padded.
// One pixel per thread, written as 16-byte pieces (cpad % 8 == 0, cin <= cpad).
__global__ void k_pack_input(const float *x, bf16 *out, size_t npix, int cin, int cpad) {
    const size_t m = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= npix) return;
    for (int c0 = 0; c0 < cpad; c0 += 8) {
        u16x8 v;
        for (int k = 0; k < 8; k++) v[k] = c0 + k < cin ? lb::f32_to_bf16(x[m * cin + c0 + k]) : (uint16_t)0;
        *(u16x8 *)(out + m * cpad + c0) = v;
    }
}

// Standardised prediction -> lattice fields of sample b, packed
// [rho | ux | uy] (hw nodes each):  ux = u_in (y0 s0 + m0),
// uy = u_in (y1 s1 + m1), rho = 1 + u_in (y2 s2 + m2).
__global__ void k_unet_fields(const float *y, int B, int hw, const float *uin, float m0, float s0, float m1,
                              float s1, float m2, float s2, float *fields) {
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)B * hw) return;
    const int b = (int)(e / hw), p = (int)(e - (size_t)b * hw);
    const float u = uin[b];
    const float *yy = y + e * 3;
    float *f = fields + (size_t)b * 3 * hw;
    f[p] = 1.f + u * (yy[2] * s2 + m2);
    f[hw + p] = u * (yy[0] * s0 + m0);
    f[2 * hw + p] = u * (yy[1] * s1 + m1);
}

inline unsigned nb256(size_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

// ===========================================================================
// Handle
// ===========================================================================
struct UConv {
    int ci, co;              // ci padded to a multiple of 16
    size_t woff, boff;       // offsets in the parameter arrays
};

struct lb_unet {
    lb_unet_desc d{};
    int cin_pad = 16;
    std::vector<UConv> convs;           // encoder, bottom, decoder convs in forward order
    size_t n_params = 0, head_woff = 0, head_boff = 0;
    float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;   // FP32 master / grad / Adam
    bf16 *wb = nullptr, *wt = nullptr;                                // BF16 weights, flipped copies
    std::vector<size_t> wt_off;
    size_t wt_total = 0;
    long long *flip_tab = nullptr;   // device: per conv {wt offset, wb offset, Ci, Co}
    // activations (bf16) and gradients (fp32)
    std::vector<bf16 *> act;       // per conv output
    std::vector<bf16 *> pool;      // per encoder level
    std::vector<bf16 *> cat;       // per decoder level
    // per decoder level, set by the last forward: the first conv read its
    // input [upsampled low | skip] in place (cat[l] not written), low's pointer
    std::vector<char> cat_direct;
    std::vector<const bf16 *> cat_low;
    bf16 *xin = nullptr;
    float *y = nullptr, *dy = nullptr;
    float *gA = nullptr, *gB = nullptr, *gC = nullptr;   // gradient scratch (largest tensor)
    size_t big = 0;                                      // their size (elements)
    bf16 *tA = nullptr, *tB = nullptr;                   // channel-major copies for the weight gradient
    std::vector<float *> gskip;    // gradient of each encoder output (skip)
    float *partial = nullptr;      // chunk partials of the weight and head gradients, K-split conv shares
    float *bpart = nullptr;        // chunk partials of a bias gradient
    size_t bpart_n = 0;
    size_t partial_n = 0;
    double *lpart = nullptr, *lsum = nullptr;
    std::vector<void *> allocs;
    hipStream_t stream = nullptr;
    int steps = 0;

    ~lb_unet() {
        for (void *p : allocs) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }
    template <class T> T *alloc(size_t n) {
        void *p = nullptr;
        if (hipMalloc(&p, n * sizeof(T) + 64) != hipSuccess) return nullptr;
        allocs.push_back(p);
        return (T *)p;
    }
};

namespace {

struct Level { int H, W, C; };

// Conv index layout: enc l (2 convs), bottom (2), dec l (2), l descending.
inline int enc_conv(int l) { return 2 * l; }

// Pixels per chunk of the head's weight-gradient partials.
constexpr int kHeadChunk = 512;

// Fixed-order sum of nchunk partial arrays of n floats (and, in the same
// launch, of a second set when b.n > 0).
inline bool quad_ok(const SumSeg &g) {
    return g.n % 4 == 0 && (uintptr_t)g.partial % 16 == 0 && (uintptr_t)g.out % 16 == 0;
}
void sum_chunks2(lb_unet *u, SumSeg a, SumSeg b) {
    if (quad_ok(a) && quad_ok(b)) {
        const unsigned na = sum4_blocks(a), nb = sum4_blocks(b);
        if (na + nb == 0) return;
        hipLaunchKernelGGL(k_sum_chunks16_x4, dim3(na + nb), dim3(256), 0, u->stream, a, b, na);
        return;
    }
    const unsigned na = (unsigned)((a.n + 15) / 16), nb = (unsigned)((b.n + 15) / 16);
    if (na + nb == 0) return;
    hipLaunchKernelGGL(k_sum_chunks16, dim3(na + nb), dim3(256), 0, u->stream, a, b, na);
}
void sum_chunks(lb_unet *u, const float *partial, int nchunk, size_t n, float *out) {
    sum_chunks2(u, SumSeg{partial, nchunk, n, out}, SumSeg{nullptr, 0, 0, nullptr});
}

// Launch plan of the LDS-tile conv: segment width (32, 16; 0 = shape not
// supported), MFMA tiles per wavefront, 32-channel groups per workgroup and
// K-split shares.  32-pixel segments take the largest tile that still gives
// at least two workgroups per CU; 16-pixel segments (small images) split the
// input-channel stages over up to enough workgroups to fill the GPU.
struct ConvPlan { int sw, pm, nb, ks; };
inline ConvPlan conv_plan(int npix, int H, int Wd, int Ci, int Co) {
    if (Ci % kTileCK || Co % 32 || H < 1 || Wd < 1) return {0, 0, 0, 0};
    const int B = npix / (H * Wd);
    const int nb = Co % 64 == 0 ? 2 : 1;
    if (Wd % 32 == 0 && H % 4 == 0) {
        const int pm = (H % 8 == 0 && (size_t)B * (H / 8) * (Wd / 32) * (Co / (32 * nb)) >= 512) ? 2 : 1;
        return {32, pm, nb, 1};
    }
    if (Wd % 16 == 0 && H % 8 == 0) {
        const long wgs = (long)B * (H / 8) * (Wd / 16) * (Co / (32 * nb));
        const int nst = Ci / kTileCK;
        int ks = 1;
        while (ks * 2 <= nst && wgs * ks < 512) ks *= 2;
        return {16, 1, nb, ks};
    }
    return {0, 0, 0, 0};
}
// FP32 scratch of a K-split conv (elements).
inline size_t conv_split_scratch(int npix, int H, int Wd, int Ci, int Co) {
    const ConvPlan pl = conv_plan(npix, H, Wd, Ci, Co);
    return pl.ks > 1 ? (size_t)pl.ks * npix * Co : 0;
}

// LDS-tile launch for eligible shapes; returns false (nothing launched) when
// the shape does not fit the tile kernel or a K-split finds no scratch
// (u->partial, u->partial_n elements).  low != null: the input is the
// concatenation [nearest-upsampled low (cl channels) | in (Ci - cl channels)]
// read in place (cl % 32 == 0).
template <class TIn, class TOut, bool BR>
bool conv_tile_launch(lb_unet *u, const TIn *in, const bf16 *W, const float *bias, TOut *out, int npix, int H, int Wd,
                      int Ci, int Co, const TIn *low = nullptr, int cl = 0) {
    const ConvPlan pl = conv_plan(npix, H, Wd, Ci, Co);
    if (!pl.sw) return false;
    float *part = nullptr;
    if (pl.ks > 1) {
        if (!u->partial || u->partial_n < (size_t)pl.ks * npix * Co) return false;
        part = u->partial;
    }
    const int B = npix / (H * Wd);
    const int rows = 4 * pl.pm * (32 / pl.sw);
    const dim3 grd((unsigned)(B * (H / rows) * (Wd / pl.sw)), (unsigned)(Co / (32 * pl.nb)), (unsigned)pl.ks), blk(256);
    if (pl.sw == 32) {
        if (pl.nb == 2 && pl.pm == 2)
            hipLaunchKernelGGL((k_conv3x3_tile<TIn, TOut, 2, 2, BR>), grd, blk, 0, u->stream, in, low, cl, W, bias, out, part, npix, H,
                               Wd, Ci, Co);
        else if (pl.nb == 2)
            hipLaunchKernelGGL((k_conv3x3_tile<TIn, TOut, 1, 2, BR>), grd, blk, 0, u->stream, in, low, cl, W, bias, out, part, npix, H,
                               Wd, Ci, Co);
        else if (pl.pm == 2)
            hipLaunchKernelGGL((k_conv3x3_tile<TIn, TOut, 2, 1, BR>), grd, blk, 0, u->stream, in, low, cl, W, bias, out, part, npix, H,
                               Wd, Ci, Co);
        else
            hipLaunchKernelGGL((k_conv3x3_tile<TIn, TOut, 1, 1, BR>), grd, blk, 0, u->stream, in, low, cl, W, bias, out, part, npix, H,
                               Wd, Ci, Co);
    } else if (pl.ks > 1) {
        if (pl.nb == 2)
            hipLaunchKernelGGL((k_conv3x3_tile<TIn, TOut, 1, 2, BR, 16, true>), grd, blk, 0, u->stream, in, low, cl, W, bias, out, part,
                               npix, H, Wd, Ci, Co);
        else
            hipLaunchKernelGGL((k_conv3x3_tile<TIn, TOut, 1, 1, BR, 16, true>), grd, blk, 0, u->stream, in, low, cl, W, bias, out, part,
                               npix, H, Wd, Ci, Co);
        const size_t n = (size_t)npix * Co;
        hipLaunchKernelGGL((k_split_sum<TOut, BR>), dim3(nb256(n / 4)), blk, 0, u->stream, part, pl.ks, n, Co, bias, out);
    } else {
        if (pl.nb == 2)
            hipLaunchKernelGGL((k_conv3x3_tile<TIn, TOut, 1, 2, BR, 16>), grd, blk, 0, u->stream, in, low, cl, W, bias, out, part, npix,
                               H, Wd, Ci, Co);
        else
            hipLaunchKernelGGL((k_conv3x3_tile<TIn, TOut, 1, 1, BR, 16>), grd, blk, 0, u->stream, in, low, cl, W, bias, out, part, npix,
                               H, Wd, Ci, Co);
    }
    return true;
}

// Input gradient of a conv with the previous conv's ReLU backward and bias
// partials fused into the epilogue (k_conv3x3_tile DPRE): dout = this conv's
// pre-activation gradient (Ci = its output channels), Wt its flipped,
// transposed weights, act = its input (the previous conv's output, Co
// channels).  Writes dpre (BF16) and bpart[chunk][Co]; returns the chunk
// count, 0 (nothing launched) when the shape takes no unsplit LDS-tile plan.
int dgrad_dpre_chunks(int npix, int H, int Wd, int Ci, int Co) {
    const ConvPlan pl = conv_plan(npix, H, Wd, Ci, Co);
    if (!pl.sw || pl.ks > 1) return 0;
    const int B = npix / (H * Wd);
    return B * (H / (4 * pl.pm * (32 / pl.sw))) * (Wd / pl.sw);
}
int dgrad_dpre_launch(lb_unet *u, const bf16 *dout, const bf16 *Wt, const bf16 *act, bf16 *dpre, float *bpart, int npix,
                      int H, int Wd, int Ci, int Co) {
    const int nchunk = dgrad_dpre_chunks(npix, H, Wd, Ci, Co);
    if (!nchunk) return 0;
    const ConvPlan pl = conv_plan(npix, H, Wd, Ci, Co);
    const dim3 grd((unsigned)nchunk, (unsigned)(Co / (32 * pl.nb)), 1u), blk(256);
    const bf16 *nol = nullptr;
    if (pl.sw == 32) {
        if (pl.nb == 2 && pl.pm == 2)
            hipLaunchKernelGGL((k_conv3x3_tile<bf16, bf16, 2, 2, false, 32, false, true>), grd, blk, 0, u->stream, dout, nol, 0,
                               Wt, nullptr, dpre, nullptr, npix, H, Wd, Ci, Co, act, bpart);
        else if (pl.nb == 2)
            hipLaunchKernelGGL((k_conv3x3_tile<bf16, bf16, 1, 2, false, 32, false, true>), grd, blk, 0, u->stream, dout, nol, 0,
                               Wt, nullptr, dpre, nullptr, npix, H, Wd, Ci, Co, act, bpart);
        else if (pl.pm == 2)
            hipLaunchKernelGGL((k_conv3x3_tile<bf16, bf16, 2, 1, false, 32, false, true>), grd, blk, 0, u->stream, dout, nol, 0,
                               Wt, nullptr, dpre, nullptr, npix, H, Wd, Ci, Co, act, bpart);
        else
            hipLaunchKernelGGL((k_conv3x3_tile<bf16, bf16, 1, 1, false, 32, false, true>), grd, blk, 0, u->stream, dout, nol, 0,
                               Wt, nullptr, dpre, nullptr, npix, H, Wd, Ci, Co, act, bpart);
    } else {
        if (pl.nb == 2)
            hipLaunchKernelGGL((k_conv3x3_tile<bf16, bf16, 1, 2, false, 16, false, true>), grd, blk, 0, u->stream, dout, nol, 0,
                               Wt, nullptr, dpre, nullptr, npix, H, Wd, Ci, Co, act, bpart);
        else
            hipLaunchKernelGGL((k_conv3x3_tile<bf16, bf16, 1, 1, false, 16, false, true>), grd, blk, 0, u->stream, dout, nol, 0,
                               Wt, nullptr, dpre, nullptr, npix, H, Wd, Ci, Co, act, bpart);
    }
    return nchunk;
}

// Input gradient of a decoder level's first conv (Co = cl upsampled-low
// channels + the skip channels) written as two tensors: channels [0, cl) to
// out_low ([pixel][cl]), the rest to out_skip ([pixel][Co - cl]), so neither
// half is copied out of a concatenated gradient.  False (nothing launched)
// unless the shape takes an unsplit LDS-tile plan and cl % 32 == 0.
bool dgrad_split_launch(lb_unet *u, const bf16 *in, const bf16 *W, float *out_low, float *out_skip, int npix, int H,
                        int Wd, int Ci, int Co, int cl) {
    const ConvPlan pl = conv_plan(npix, H, Wd, Ci, Co);
    if (!pl.sw || pl.ks > 1 || cl <= 0 || cl >= Co || cl % 32) return false;
    const int B = npix / (H * Wd);
    const int rows = 4 * pl.pm * (32 / pl.sw);
    const dim3 grd((unsigned)(B * (H / rows) * (Wd / pl.sw)), (unsigned)(Co / (32 * pl.nb)), 1u), blk(256);
    const bf16 *nl = nullptr;
    if (pl.sw == 32) {
        if (pl.nb == 2 && pl.pm == 2)
            hipLaunchKernelGGL((k_conv3x3_tile<bf16, float, 2, 2, false>), grd, blk, 0, u->stream, in, nl, 0, W, nullptr,
                               out_low, nullptr, npix, H, Wd, Ci, Co, nullptr, nullptr, out_skip, cl);
        else if (pl.nb == 2)
            hipLaunchKernelGGL((k_conv3x3_tile<bf16, float, 1, 2, false>), grd, blk, 0, u->stream, in, nl, 0, W, nullptr,
                               out_low, nullptr, npix, H, Wd, Ci, Co, nullptr, nullptr, out_skip, cl);
        else if (pl.pm == 2)
            hipLaunchKernelGGL((k_conv3x3_tile<bf16, float, 2, 1, false>), grd, blk, 0, u->stream, in, nl, 0, W, nullptr,
                               out_low, nullptr, npix, H, Wd, Ci, Co, nullptr, nullptr, out_skip, cl);
        else
            hipLaunchKernelGGL((k_conv3x3_tile<bf16, float, 1, 1, false>), grd, blk, 0, u->stream, in, nl, 0, W, nullptr,
                               out_low, nullptr, npix, H, Wd, Ci, Co, nullptr, nullptr, out_skip, cl);
    } else if (pl.nb == 2) {
        hipLaunchKernelGGL((k_conv3x3_tile<bf16, float, 1, 2, false, 16>), grd, blk, 0, u->stream, in, nl, 0, W, nullptr,
                           out_low, nullptr, npix, H, Wd, Ci, Co, nullptr, nullptr, out_skip, cl);
    } else {
        hipLaunchKernelGGL((k_conv3x3_tile<bf16, float, 1, 1, false, 16>), grd, blk, 0, u->stream, in, nl, 0, W, nullptr,
                           out_low, nullptr, npix, H, Wd, Ci, Co, nullptr, nullptr, out_skip, cl);
    }
    return true;
}

template <class TIn, class TOut, bool BR>
void conv_launch(lb_unet *u, const TIn *in, const bf16 *W, const float *bias, TOut *out, int npix, int H, int Wd,
                 int Ci, int Co, bool gather_only = false) {
    if (!gather_only && conv_tile_launch<TIn, TOut, BR>(u, in, W, bias, out, npix, H, Wd, Ci, Co, (const TIn *)nullptr, 0))
        return;
    dim3 blk(256);
    const int nb = Co % 64 == 0 ? 2 : 1;
    // fewer than ~4 workgroups per CU with one 32-pixel tile per wavefront:
    // split K over the wavefronts instead
    const bool ksplit = (size_t)((npix + 127) / 128) * (Co / (32 * nb)) < 1024 && (9 * Ci) / 16 >= 8;
    if (nb == 2) {
        if (ksplit)
            hipLaunchKernelGGL((k_conv3x3<TIn, TOut, 2, BR, 4>), dim3((npix + 31) / 32, Co / 64), blk, 0, u->stream, in, W,
                               bias, out, npix, H, Wd, Ci, Co);
        else
            hipLaunchKernelGGL((k_conv3x3<TIn, TOut, 2, BR>), dim3((npix + 127) / 128, Co / 64), blk, 0, u->stream, in, W,
                               bias, out, npix, H, Wd, Ci, Co);
    } else {
        if (ksplit)
            hipLaunchKernelGGL((k_conv3x3<TIn, TOut, 1, BR, 4>), dim3((npix + 31) / 32, Co / 32), blk, 0, u->stream, in, W,
                               bias, out, npix, H, Wd, Ci, Co);
        else
            hipLaunchKernelGGL((k_conv3x3<TIn, TOut, 1, BR>), dim3((npix + 127) / 128, Co / 32), blk, 0, u->stream, in, W,
                               bias, out, npix, H, Wd, Ci, Co);
    }
}

// Pixel-tile chunks of the LDS-tile weight gradient: tiles per workgroup
// chosen for about 512 workgroups (two per CU); 0 when the shape does not fit
// the kernel.
struct WgradTiles { int ntiles, tpw, nchunk, tw; };
inline WgradTiles wgrad_tiles(int npix, int H, int Wd, int Ci, int Co) {
    if (Co % 32 || Ci % 8) return {0, 0, 0, 0};
    const int tw = (Wd % 32 == 0 && H % 8 == 0) ? 32 : (Wd % 16 == 0 && H % 16 == 0) ? 16 : 0;
    if (!tw) return {0, 0, 0, 0};
    const int ntiles = npix / 256;
    const long base = (long)(Co / 32) * ((Ci + 31) / 32);
    const int tpw = (int)std::max(1L, (ntiles * base + 511) / 512);
    return {ntiles, tpw, (ntiles + tpw - 1) / tpw, tw};
}
// The first conv of a decoder level can read [upsampled low | skip] in place
// in both its forward and its weight gradient.
inline bool cat_in_place(const lb_unet *u, int npix, int H, int Wd, int cl, int Ci, int Co) {
    if (cl % 32) return false;
    if (!conv_plan(npix, H, Wd, Ci, Co).sw || !wgrad_tiles(npix, H, Wd, Ci, Co).ntiles) return false;
    return conv_split_scratch(npix, H, Wd, Ci, Co) <= u->partial_n;
}

// Chunk count of the per-wavefront weight-gradient kernels (1,024 pixels each).
inline int wgrad_chunks_legacy(int npix) { return (npix + 1023) / 1024; }

// Weight gradient dW[co][tap][ci] from dout (NHWC, FP32 or BF16; every kernel
// rounds it to BF16 to nearest even as an MFMA operand) and the conv input
// (BF16 NHWC).  Shapes the LDS-tile kernel supports take it (method < 0 or
// 2); otherwise, with W % 16 == 0 and channel-major scratch (tA >= npix Co,
// tB >= npix Ci elements), both operands are transposed first and the
// vector-load kernel runs (method 1); otherwise the gathering kernel
// (method 0).
// extra: a second chunk sum (the bias gradient) done in the same launch as
// the weight gradient's.
// low != null: the input is the concatenation [upsampled low | in] read in
// place (LDS-tile kernel only; the caller checks wgrad_tiles and cl % 32).
template <class TIn, class TD>
void wgrad_launch(lb_unet *u, const TD *dout, const TIn *in, float *dW, int npix, int H, int Wd, int Ci, int Co,
                  bf16 *tA, bf16 *tB, int method = -1, SumSeg extra = SumSeg{nullptr, 0, 0, nullptr},
                  const bf16 *low = nullptr, int cl = 0) {
    const int K = 9 * Ci;
    const WgradTiles wt = wgrad_tiles(npix, H, Wd, Ci, Co);
    if ((method < 0 || method == 2) && wt.ntiles) {
        const dim3 grd(Co / 32, (Ci + 31) / 32, wt.nchunk);
        if (wt.tw == 32)
            hipLaunchKernelGGL((k_conv3x3_wgrad_tile<32, TD>), grd, dim3(256), 0, u->stream, dout, (const bf16 *)in, low, cl, u->partial, H,
                               Wd, Ci, Co, wt.ntiles, wt.tpw);
        else
            hipLaunchKernelGGL((k_conv3x3_wgrad_tile<16, TD>), grd, dim3(256), 0, u->stream, dout, (const bf16 *)in, low, cl, u->partial, H,
                               Wd, Ci, Co, wt.ntiles, wt.tpw);
        sum_chunks2(u, SumSeg{u->partial, wt.nchunk, (size_t)Co * K, dW}, extra);
        return;
    }
    const int chunk = 1024;
    const int nchunk = (npix + chunk - 1) / chunk;
    dim3 grd(Co / 32, (K + 31) / 32, nchunk);
    if (method != 0 && tA && tB && Wd % 16 == 0) {
        hipLaunchKernelGGL(k_to_cmajor<TD>, dim3((npix + 63) / 64, (Co + 63) / 64), dim3(256), 0, u->stream, dout, tA,
                           npix, Co);
        hipLaunchKernelGGL(k_to_cmajor<TIn>, dim3((npix + 63) / 64, (Ci + 63) / 64), dim3(256), 0, u->stream, in, tB,
                           npix, Ci);
        hipLaunchKernelGGL(k_conv3x3_wgrad_cm, grd, dim3(256), 0, u->stream, tA, tB, u->partial, npix, H, Wd, Ci, Co,
                           chunk);
    } else {
        hipLaunchKernelGGL((k_conv3x3_wgrad<TIn, TD>), grd, dim3(256), 0, u->stream, dout, in, u->partial, npix, H, Wd, Ci, Co,
                           chunk);
    }
    sum_chunks2(u, SumSeg{u->partial, nchunk, (size_t)Co * K, dW}, extra);
}

std::vector<Level> levels(const lb_unet *u) {
    std::vector<Level> L;
    for (int l = 0; l < u->d.levels; l++) L.push_back({u->d.height >> l, u->d.width >> l, u->d.base << l});
    return L;
}

}  // namespace

int lb_unet_refresh(lb_unet *u);

extern "C" int lb_unet_create(const lb_unet_desc *desc, uint64_t seed, lb_unet **out) {
    if (!desc || !out) return LB_ERR_INVALID;
    *out = nullptr;
    if (lb_device_count() <= 0) return LB_ERR_NO_DEVICE;
    const int Lv = desc->levels;
    if (Lv < 2 || Lv > 6 || desc->base % 32 || desc->base > 256 || (desc->base << (Lv - 1)) > 1024 ||
        desc->in_ch < 1 || desc->in_ch > 16 || desc->out_ch < 1 || desc->out_ch > 8 ||
        desc->max_batch < 1 || desc->height % (1 << (Lv - 1)) || desc->width % (1 << (Lv - 1)))
        return LB_ERR_INVALID;
    lb_unet *u = new (std::nothrow) lb_unet();
    if (!u) return LB_ERR_INVALID;
    u->d = *desc;
    std::vector<Level> L = levels(u);
    // conv list
    auto add = [&](int ci, int co) {
        UConv c{ci, co, u->n_params, 0};
        u->n_params += (size_t)co * 9 * ci;
        c.boff = u->n_params;
        u->n_params += co;
        u->convs.push_back(c);
    };
    for (int l = 0; l < Lv - 1; l++) {
        add(l == 0 ? u->cin_pad : L[l - 1].C, L[l].C);
        add(L[l].C, L[l].C);
    }
    add(L[Lv - 2].C, L[Lv - 1].C);
    add(L[Lv - 1].C, L[Lv - 1].C);
    for (int l = Lv - 2; l >= 0; l--) {
        add(L[l + 1].C + L[l].C, L[l].C);
        add(L[l].C, L[l].C);
    }
    u->head_woff = u->n_params;
    u->n_params += (size_t)desc->out_ch * L[0].C;
    u->head_boff = u->n_params;
    u->n_params += desc->out_ch;
    const size_t P = u->n_params;
    int rc = (int)hipStreamCreateWithFlags(&u->stream, hipStreamNonBlocking);
    u->w = u->alloc<float>(P); u->g = u->alloc<float>(P); u->m = u->alloc<float>(P); u->v = u->alloc<float>(P);
    u->wb = u->alloc<bf16>(P);
    size_t wt_total = 0;
    std::vector<long long> htab;
    for (auto &c : u->convs) {
        u->wt_off.push_back(wt_total);
        htab.insert(htab.end(), {(long long)wt_total, (long long)c.woff, (long long)c.ci, (long long)c.co});
        wt_total += (size_t)c.co * 9 * c.ci;
    }
    u->wt_total = wt_total;
    u->wt = u->alloc<bf16>(wt_total);
    u->flip_tab = u->alloc<long long>(htab.size());
    if (!rc && u->flip_tab)
        rc = (int)hipMemcpyAsync(u->flip_tab, htab.data(), htab.size() * sizeof(long long), hipMemcpyHostToDevice, u->stream);
    // He-uniform initialisation from SplitMix64 (host), zero biases
    std::vector<float> hw(P, 0.f);
    uint64_t st = seed;
    auto uni = [&]() {
        uint64_t z = (st += 0x9e3779b97f4a7c15ULL);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    };
    for (size_t ci = 0; ci < u->convs.size(); ci++) {
        const UConv &c = u->convs[ci];
        int fan_in = 9 * (ci == 0 ? desc->in_ch : c.ci);
        double a = std::sqrt(6.0 / fan_in);
        for (int co = 0; co < c.co; co++)
            for (int t = 0; t < 9; t++)
                for (int k = 0; k < c.ci; k++) {
                    double r = (2 * uni() - 1) * a;
                    if (ci == 0 && k >= desc->in_ch) r = 0;
                    hw[c.woff + ((size_t)co * 9 + t) * c.ci + k] = (float)r;
                }
    }
    {
        double a = std::sqrt(6.0 / L[0].C) * 0.1;
        for (size_t e = 0; e < (size_t)desc->out_ch * L[0].C; e++) hw[u->head_woff + e] = (float)((2 * uni() - 1) * a);
    }
    // activations
    const int B = desc->max_batch;
    for (size_t ci = 0; ci < u->convs.size(); ci++) {
        int l;
        if (ci < (size_t)2 * (Lv - 1)) l = (int)ci / 2;
        else if (ci < (size_t)2 * Lv) l = Lv - 1;
        else l = Lv - 2 - (int)(ci - 2 * Lv) / 2;
        u->act.push_back(u->alloc<bf16>((size_t)B * L[l].H * L[l].W * u->convs[ci].co));
    }
    for (int l = 0; l < Lv - 1; l++) u->pool.push_back(u->alloc<bf16>((size_t)B * L[l + 1].H * L[l + 1].W * L[l].C));
    for (int l = 0; l < Lv - 1; l++) u->cat.push_back(u->alloc<bf16>((size_t)B * L[l].H * L[l].W * (L[l + 1].C + L[l].C)));
    for (int l = 0; l < Lv - 1; l++) u->gskip.push_back(u->alloc<float>((size_t)B * L[l].H * L[l].W * L[l].C));
    size_t big = 0;
    for (int l = 0; l < Lv; l++) {
        size_t cmax = l < Lv - 1 ? (size_t)(L[l + 1].C + L[l].C) : (size_t)L[l].C;
        big = std::max(big, (size_t)B * L[l].H * L[l].W * std::max(cmax, (size_t)u->cin_pad));
    }
    u->big = big;
    u->gA = u->alloc<float>(big); u->gB = u->alloc<float>(big); u->gC = u->alloc<float>(big);
    u->tA = u->alloc<bf16>(big); u->tB = u->alloc<bf16>(big);
    u->xin = u->alloc<bf16>((size_t)B * L[0].H * L[0].W * u->cin_pad);
    u->y = u->alloc<float>((size_t)B * L[0].H * L[0].W * desc->out_ch);
    u->dy = u->alloc<float>((size_t)B * L[0].H * L[0].W * desc->out_ch);
    // weight-gradient partials: max over convs of nchunk * Co * K
    for (size_t ci = 0; ci < u->convs.size(); ci++) {
        int l;
        if (ci < (size_t)2 * (Lv - 1)) l = (int)ci / 2;
        else if (ci < (size_t)2 * Lv) l = Lv - 1;
        else l = Lv - 2 - (int)(ci - 2 * Lv) / 2;
        size_t npix = (size_t)B * L[l].H * L[l].W;
        const WgradTiles wt = wgrad_tiles((int)npix, L[l].H, L[l].W, u->convs[ci].ci, u->convs[ci].co);
        size_t nchunk = std::max((size_t)wgrad_chunks_legacy((int)npix), (size_t)wt.nchunk);
        const size_t nw = (size_t)u->convs[ci].co * 9 * u->convs[ci].ci;
        u->partial_n = std::max(u->partial_n, nchunk * nw);
        u->bpart_n = std::max(u->bpart_n, (size_t)bias_chunks((int)npix).nchunk * u->convs[ci].co);
        // bias partials of the previous conv from this conv's fused input gradient
        u->bpart_n = std::max(u->bpart_n, (size_t)dgrad_dpre_chunks((int)npix, L[l].H, L[l].W, u->convs[ci].co,
                                                                    u->convs[ci].ci) * u->convs[ci].ci);
        // K-split scratch of the forward and input-gradient convs
        u->partial_n = std::max(u->partial_n, conv_split_scratch((int)npix, L[l].H, L[l].W, u->convs[ci].ci, u->convs[ci].co));
        u->partial_n = std::max(u->partial_n, conv_split_scratch((int)npix, L[l].H, L[l].W, u->convs[ci].co, u->convs[ci].ci));
    }
    {
        // head: chunk partials followed by their sum
        const size_t npix0 = (size_t)B * L[0].H * L[0].W;
        const size_t nchunk = (npix0 + kHeadChunk - 1) / kHeadChunk;
        u->partial_n = std::max(u->partial_n, (nchunk + 1) * desc->out_ch * (L[0].C + 1));
    }
    u->partial = u->alloc<float>(u->partial_n);
    u->bpart = u->alloc<float>(u->bpart_n);
    u->lpart = u->alloc<double>(nb256((size_t)B * L[0].H * L[0].W * desc->out_ch));
    u->lsum = u->alloc<double>(2);
    bool ok = !rc && u->w && u->g && u->m && u->v && u->wb && u->wt && u->flip_tab && u->gA && u->gB && u->gC && u->tA && u->tB &&
              u->xin && u->y &&
              u->dy && u->partial && u->bpart && u->lpart && u->lsum;
    for (auto p : u->act) ok = ok && p;
    for (auto p : u->pool) ok = ok && p;
    for (auto p : u->cat) ok = ok && p;
    for (auto p : u->gskip) ok = ok && p;
    if (!ok) { delete u; return rc ? rc : LB_ERR_INVALID; }
    rc = (int)hipMemcpyAsync(u->w, hw.data(), P * 4, hipMemcpyHostToDevice, u->stream);
    if (!rc) rc = (int)hipMemsetAsync(u->m, 0, P * 4, u->stream);
    if (!rc) rc = (int)hipMemsetAsync(u->v, 0, P * 4, u->stream);
    if (!rc) rc = (int)hipMemsetAsync(u->g, 0, P * 4, u->stream);
    if (rc) { delete u; return rc; }
    rc = lb_unet_refresh(u);
    if (!rc) rc = (int)hipStreamSynchronize(u->stream);
    if (rc) { delete u; return rc; }
    *out = u;
    return 0;
}

// BF16 copies and flipped/transposed copies from the FP32 masters.
int lb_unet_refresh(lb_unet *u) {
    hipLaunchKernelGGL(k_to_bf16, dim3(nb256(u->n_params)), dim3(256), 0, u->stream, u->w, u->wb, u->n_params);
    hipLaunchKernelGGL(k_flip_transpose_all, dim3(nb256(u->wt_total)), dim3(256), 0, u->stream, u->wb, u->wt, u->flip_tab,
                       (int)u->convs.size(), u->wt_total);
    return (int)hipGetLastError();
}

extern "C" int lb_unet_destroy(lb_unet *u) {
    if (!u) return LB_ERR_INVALID;
    (void)hipStreamSynchronize(u->stream);
    delete u;
    return 0;
}

extern "C" int64_t lb_unet_param_count(lb_unet *u) { return u ? (int64_t)u->n_params : -1; }

extern "C" int lb_unet_get_params(lb_unet *u, float *host, float *grad_host) {
    if (!u) return LB_ERR_INVALID;
    int rc = 0;
    if (host) rc = (int)hipMemcpyAsync(host, u->w, u->n_params * 4, hipMemcpyDeviceToHost, u->stream);
    if (!rc && grad_host) rc = (int)hipMemcpyAsync(grad_host, u->g, u->n_params * 4, hipMemcpyDeviceToHost, u->stream);
    if (!rc) rc = (int)hipStreamSynchronize(u->stream);
    return rc;
}

extern "C" int lb_unet_set_params(lb_unet *u, const float *host) {
    if (!u || !host) return LB_ERR_INVALID;
    int rc = (int)hipMemcpyAsync(u->w, host, u->n_params * 4, hipMemcpyHostToDevice, u->stream);
    if (!rc) rc = lb_unet_refresh(u);
    if (!rc) rc = (int)hipStreamSynchronize(u->stream);
    return rc;
}

// Forward on a device batch x (B, H, W, in_ch) FP32 -> y (B, H, W, out_ch).
extern "C" int lb_unet_forward(lb_unet *u, const float *d_x, int B, float *d_y) {
    if (!u || !d_x || B < 1 || B > u->d.max_batch) return LB_ERR_INVALID;
    std::vector<Level> L = levels(u);
    const int Lv = u->d.levels;
    size_t npix0 = (size_t)B * L[0].H * L[0].W;
    hipLaunchKernelGGL(k_pack_input, dim3(nb256(npix0)), dim3(256), 0, u->stream, d_x, u->xin, npix0,
                       u->d.in_ch, u->cin_pad);
    const bf16 *cur = u->xin;
    int ci = 0;
    for (int l = 0; l < Lv; l++) {
        const int npix = B * L[l].H * L[l].W;
        for (int k = 0; k < 2; k++, ci++) {
            const UConv &c = u->convs[ci];
            conv_launch<bf16, bf16, true>(u, cur, u->wb + c.woff, u->w + c.boff, u->act[ci], npix, L[l].H, L[l].W, c.ci, c.co);
            cur = u->act[ci];
        }
        if (l < Lv - 1) {
            size_t n = (size_t)B * L[l + 1].H * L[l + 1].W * L[l].C / 8;
            hipLaunchKernelGGL(k_avgpool, dim3(nb256(n)), dim3(256), 0, u->stream, cur, u->pool[l], B, L[l].H, L[l].W, L[l].C);
            cur = u->pool[l];
        }
    }
    u->cat_direct.assign(Lv, 0);
    u->cat_low.assign(Lv, nullptr);
    for (int l = Lv - 2; l >= 0; l--) {
        const int npix = B * L[l].H * L[l].W;
        const UConv &c0 = u->convs[ci];
        const bf16 *skip = u->act[enc_conv(l) + 1];
        u->cat_low[l] = cur;
        if (cat_in_place(u, npix, L[l].H, L[l].W, L[l + 1].C, c0.ci, c0.co) &&
            conv_tile_launch<bf16, bf16, true>(u, skip, u->wb + c0.woff, u->w + c0.boff, u->act[ci], npix, L[l].H, L[l].W,
                                               c0.ci, c0.co, cur, L[l + 1].C)) {
            u->cat_direct[l] = 1;
        } else {
            size_t n = (size_t)npix * (L[l + 1].C + L[l].C) / 8;
            hipLaunchKernelGGL(k_upcat, dim3(nb256(n)), dim3(256), 0, u->stream, cur, skip, u->cat[l], B, L[l].H, L[l].W,
                               L[l + 1].C, L[l].C);
            conv_launch<bf16, bf16, true>(u, u->cat[l], u->wb + c0.woff, u->w + c0.boff, u->act[ci], npix, L[l].H, L[l].W,
                                          c0.ci, c0.co);
        }
        cur = u->act[ci];
        ci++;
        const UConv &c1 = u->convs[ci];
        conv_launch<bf16, bf16, true>(u, cur, u->wb + c1.woff, u->w + c1.boff, u->act[ci], npix, L[l].H, L[l].W, c1.ci, c1.co);
        cur = u->act[ci];
        ci++;
    }
    float *y = d_y ? d_y : u->y;
    hipLaunchKernelGGL(k_head_fwd, dim3(nb256(npix0)), dim3(256), 0, u->stream, cur, u->wb + u->head_woff,
                       u->w + u->head_boff, y, (int)npix0, L[0].C, u->d.out_ch);
    if (y != u->y)
        (void)hipMemcpyAsync(u->y, y, npix0 * u->d.out_ch * 4, hipMemcpyDeviceToDevice, u->stream);
    int rc = (int)hipGetLastError();
    // an output for the caller is complete on return (any stream or the host
    // may read it); the training path (d_y NULL) stays asynchronous
    if (!rc && d_y) rc = (int)hipStreamSynchronize(u->stream);
    return rc;
}

// Masked MSE of the last forward against d_t (B, H, W, out_ch) with d_mask
// (B, H, W) weights; fills all gradients.  *loss (host) receives the mean.
extern "C" int lb_unet_backward(lb_unet *u, const float *d_t, const float *d_mask, int B, double *loss) {
    if (!u || !d_t || !d_mask || B < 1 || B > u->d.max_batch) return LB_ERR_INVALID;
    std::vector<Level> L = levels(u);
    const int Lv = u->d.levels;
    const int O = u->d.out_ch;
    const int npix0 = B * L[0].H * L[0].W;
    // lsum[1] = sum of the mask weights, lsum[0] = sum of squared weighted
    // errors, both reduced on the device (no host round trip before the
    // backward pass)
    const unsigned nbm = nb256((size_t)npix0), nbl = nb256((size_t)npix0 * O);
    hipLaunchKernelGGL(k_mask_sum, dim3(nbm), dim3(256), 0, u->stream, d_mask, npix0, u->lpart);
    hipLaunchKernelGGL(k_sum_d, dim3(1), dim3(256), 0, u->stream, u->lpart, (int)nbm, u->lsum + 1);
    hipLaunchKernelGGL(k_mse, dim3(nbl), dim3(256), 0, u->stream, u->y, d_t, d_mask, u->dy, npix0, O, u->lsum + 1, u->lpart);
    hipLaunchKernelGGL(k_sum_d, dim3(1), dim3(256), 0, u->stream, u->lpart, (int)nbl, u->lsum);
    int rc = 0;
    // head
    const int C0 = L[0].C;
    const int last = (int)u->convs.size() - 1;
    {
        // partials in the weight-gradient buffer (free until the first wgrad_launch)
        const int chunk = kHeadChunk, nchunk = (npix0 + chunk - 1) / chunk, np = O * (C0 + 1);
        hipLaunchKernelGGL(k_head_bwd_w, dim3(nchunk), dim3(256), 0, u->stream, u->dy, u->act[last], npix0, C0, O, chunk,
                           u->partial);
        sum_chunks(u, u->partial, nchunk, (size_t)np, u->partial + (size_t)nchunk * np);
        hipLaunchKernelGGL(k_head_scatter, dim3(nb256(np)), dim3(256), 0, u->stream, u->partial + (size_t)nchunk * np, C0, O,
                           u->g + u->head_woff, u->g + u->head_boff);
    }
    // the head's input gradient is formed by the first ReLU backward
    // (k_relu_bwd_bias hdy mode), not written to gA
    bool head_pending = C0 <= 256 && O <= 8;
    if (!head_pending)
        hipLaunchKernelGGL(k_head_bwd_data, dim3(nb256((size_t)npix0 * C0 / 4)), dim3(256), 0, u->stream, u->dy,
                           u->wb + u->head_woff, u->gA, npix0, C0, O);
    // gA = d(out of current conv), post-activation gradient
    float *gcur = u->gA, *gtmp = u->gB;
    // pre-activation gradients (BF16): the two halves of gC, so a fused input
    // gradient can write the next conv's while it reads this conv's
    bf16 *gpre_buf[2] = {(bf16 *)u->gC, (bf16 *)u->gC + u->big};
    int gsel = 0;
    // set by a fused input gradient: the next conv's dpre (gpre_buf[gsel]) and
    // its bias partials in bpart (pre_nchunk chunks) are ready
    bool pre_ready = false;
    int pre_nchunk = 0;
    int ci = last;
    // set after a decoder level's first conv wrote its input gradient as
    // [low | skip] tensors: gcur holds the full-resolution gradient of the
    // upsampled low part, whose upsampling backward the next ReLU backward
    // sums itself
    bool up_pending = false;
    // dpool: the incoming gradient of this conv's output is gcur (its skip
    // gradient) plus the average-pool backward of dpool (see k_relu_bwd_bias).
    // fuse: this conv's input is the ReLU output of the conv processed next
    // (same level): its input gradient also applies that conv's ReLU backward
    // and bias partials (dgrad_dpre_launch), where the plan allows.
    // skip_out: a decoder level's first conv; its input gradient goes to
    // gcur (low part, skip_cl channels) and skip_out (skip part) where the
    // plan allows (split_done), else to gcur as the whole [low | skip].
    bool split_done = false;
    auto conv_bwd = [&](int cidx, const bf16 *input, int l, bool need_din, const bf16 *low = nullptr, int cl = 0,
                        const float *dpool = nullptr, bool fuse = false, float *skip_out = nullptr, int skip_cl = 0) {
        const UConv &c = u->convs[cidx];
        const int npix = B * L[l].H * L[l].W;
        bf16 *gpre = gpre_buf[gsel];
        SumSeg bseg;
        if (pre_ready) {
            bseg = SumSeg{u->bpart, pre_nchunk, (size_t)c.co, u->g + c.boff};
            pre_ready = false;
        } else {
            // ReLU backward + bias-gradient partials, summed in the weight
            // gradient's chunk-sum launch
            const BiasChunks bc = bias_chunks(npix);
            hipLaunchKernelGGL(k_relu_bwd_bias, dim3(bc.nchunk), dim3(256), 0, u->stream, dpool ? u->gskip[l] : gcur,
                               u->act[cidx], gpre, npix, c.co, bc.chunk, u->bpart, dpool, L[l].H, L[l].W,
                               up_pending ? (const float *)gcur : nullptr, head_pending ? (const float *)u->dy : nullptr,
                               u->wb + u->head_woff, O);
            up_pending = false;
            head_pending = false;
            bseg = SumSeg{u->bpart, bc.nchunk, (size_t)c.co, u->g + c.boff};
        }
        wgrad_launch<bf16, bf16>(u, gpre, input, u->g + c.woff, npix, L[l].H, L[l].W, c.ci, c.co, u->tA, u->tB, -1, bseg,
                                 low, cl);
        if (!need_din) return;
        if (fuse) {
            const int nch = dgrad_dpre_launch(u, gpre, u->wt + u->wt_off[cidx], input, gpre_buf[1 - gsel], u->bpart, npix,
                                              L[l].H, L[l].W, c.co, c.ci);
            if (nch) {
                gsel = 1 - gsel;
                pre_ready = true;
                pre_nchunk = nch;
                return;
            }
        }
        split_done = skip_out && dgrad_split_launch(u, gpre, u->wt + u->wt_off[cidx], gtmp, skip_out, npix, L[l].H,
                                                    L[l].W, c.co, c.ci, skip_cl);
        if (!split_done)
            conv_launch<bf16, float, false>(u, gpre, u->wt + u->wt_off[cidx], nullptr, gtmp, npix, L[l].H, L[l].W, c.co,
                                            c.ci);
        std::swap(gcur, gtmp);
    };
    for (int l = 0; l <= Lv - 2; l++) {
        const int lev = l;   // decoder levels in backward order: 0, 1, ..., Lv-2
        conv_bwd(ci, u->act[ci - 1], lev, true, nullptr, 0, nullptr, true);
        ci--;
        if (u->cat_direct.size() > (size_t)lev && u->cat_direct[lev])
            conv_bwd(ci, u->act[enc_conv(lev) + 1], lev, true, u->cat_low[lev], L[lev + 1].C, nullptr, false,
                     u->gskip[lev], L[lev + 1].C);
        else
            conv_bwd(ci, u->cat[lev], lev, true, nullptr, 0, nullptr, false, u->gskip[lev], L[lev + 1].C);
        ci--;
        if (split_done) {               // gcur = d(low) at this level's resolution, gskip[lev] written
            up_pending = true;
            continue;
        }
        // gcur = d(cat[lev]); split into d(low) and d(skip)
        const int npix = B * L[lev].H * L[lev].W;
        size_t nsk = (size_t)npix * L[lev].C / 4;
        hipLaunchKernelGGL(k_upcat_bwd_skip, dim3(nb256(nsk)), dim3(256), 0, u->stream, gcur, u->gskip[lev], B, L[lev].H,
                           L[lev].W, L[lev + 1].C, L[lev].C);
        size_t nlow = (size_t)B * L[lev + 1].H * L[lev + 1].W * L[lev + 1].C / 4;
        hipLaunchKernelGGL(k_upcat_bwd_low, dim3(nb256(nlow)), dim3(256), 0, u->stream, gcur, gtmp, B, L[lev].H, L[lev].W,
                           L[lev + 1].C, L[lev].C);
        std::swap(gcur, gtmp);
    }
    // bottom (level Lv-1) and encoder, deepest first
    const float *dpool = nullptr;
    for (int l = Lv - 1; l >= 0; l--) {
        const int c2 = 2 * l + 1, c1 = 2 * l;
        // the gradient of the encoder output act[c2] is its skip gradient
        // plus the pool backward of d(pool[l]), summed in the ReLU backward
        conv_bwd(c2, u->act[c1], l, true, nullptr, 0, dpool, true);
        const bf16 *in1 = l == 0 ? u->xin : u->pool[l - 1];
        conv_bwd(c1, in1, l, l > 0);
        dpool = l > 0 ? gcur : nullptr;    // d(pool[l - 1])
    }
    rc = (int)hipGetLastError();
    if (!rc && loss) {
        double s[2] = {0, 0};
        rc = (int)hipMemcpyAsync(s, u->lsum, 16, hipMemcpyDeviceToHost, u->stream);
        if (!rc) rc = (int)hipStreamSynchronize(u->stream);
        const double cnt = s[1] * O;
        if (!rc && cnt <= 0) rc = LB_ERR_INVALID;
        *loss = cnt > 0 ? s[0] / cnt : 0.0;
    }
    return rc;
}

extern "C" int lb_unet_adam(lb_unet *u, float lr, float beta1, float beta2, float eps) {
    if (!u) return LB_ERR_INVALID;
    u->steps++;
    float c1 = 1.f - std::pow(beta1, (float)u->steps), c2 = 1.f - std::pow(beta2, (float)u->steps);
    hipLaunchKernelGGL(k_adam, dim3(nb256(u->n_params)), dim3(256), 0, u->stream, u->w, u->g, u->m, u->v, u->wb, u->n_params,
                       lr, beta1, beta2, eps, c1, c2);
    hipLaunchKernelGGL(k_flip_transpose_all, dim3(nb256(u->wt_total)), dim3(256), 0, u->stream, u->wb, u->wt, u->flip_tab,
                       (int)u->convs.size(), u->wt_total);
    return (int)hipGetLastError();
}

extern "C" int lb_unet_sync(lb_unet *u) {
    if (!u) return LB_ERR_INVALID;
    return (int)hipStreamSynchronize(u->stream);
}

// Building block exposed for testing and reuse: one 3x3 convolution on
// device tensors (NHWC).  mode 0: bf16 in -> bf16 out with bias + ReLU;
// mode 1: fp32 in -> fp32 out, no bias / activation; modes 2 / 3: the same
// forms on the gathering kernel only (the LDS-tile kernel is otherwise taken
// for every shape it supports).
extern "C" int lb_conv3x3_bf16(int mode, const void *d_in, const void *d_w, const float *d_bias, void *d_out, int B,
                               int H, int W, int Ci, int Co, lb_stream_t stream) {
    if (Ci % 16 || Co % 32 || B < 1 || H < 1 || W < 1 || mode < 0 || mode > 3) return LB_ERR_INVALID;
    lb_unet tmp;
    tmp.stream = (hipStream_t)stream;
    const int npix = B * H * W;
    const bool gather = mode >= 2;
    // a K-split plan needs scratch: allocated here (and the call then waits)
    const size_t scratch = gather ? 0 : conv_split_scratch(npix, H, W, Ci, Co);
    if (scratch) {
        tmp.partial = tmp.alloc<float>(scratch);
        tmp.partial_n = scratch;
        if (!tmp.partial) { tmp.stream = nullptr; return LB_ERR_INVALID; }
    }
    if (mode % 2 == 0)
        conv_launch<bf16, bf16, true>(&tmp, (const bf16 *)d_in, (const bf16 *)d_w, d_bias, (bf16 *)d_out, npix, H, W, Ci,
                                      Co, gather);
    else
        conv_launch<float, float, false>(&tmp, (const float *)d_in, (const bf16 *)d_w, nullptr, (float *)d_out, npix, H, W,
                                         Ci, Co, gather);
    int rc = (int)hipGetLastError();
    if (!rc && scratch) rc = (int)hipStreamSynchronize(tmp.stream);
    tmp.stream = nullptr;
    return rc;
}

// ---------------------------------------------------------------------------
// lb_conv3x3_bf16_noalloc: lb_conv3x3_bf16 for callers that cannot allocate
// (code under hipGraph capture, serving loops).  The K-split shares of the
// 16-pixel tile plan go to a fixed per-device array instead of a hipMalloc'd
// buffer; every share writes its whole region before k_split_sum reads it,
// so nothing needs clearing between calls.  Calls that overlap on one GPU
// (two streams, or two host threads) share the array and race.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocScratchFloats = size_t(4) << 20;   // 16 MiB per GPU
static __device__ __align__(256) float scratch[kNoAllocScratchFloats];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float *g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device (each GPU holds its own copy of the array).
int noalloc_scratch(float **out) {
    int dev = 0;
    int rc = (int)hipGetDevice(&dev);
    if (rc) return rc;
    if (dev < 0 || dev >= kMaxDevices) return LB_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        rc = (int)hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (rc) return rc;
        if (!p) return LB_ERR_INVALID;
        g_scratch_ptr[dev] = (float *)p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}
}  // namespace

extern "C" int lb_conv3x3_bf16_noalloc(int mode, const void *d_in, const void *d_w, const float *d_bias, void *d_out,
                                       int B, int H, int W, int Ci, int Co, lb_stream_t stream) {
    if (Ci % 16 || Co % 32 || B < 1 || H < 1 || W < 1 || mode < 0 || mode > 3) return LB_ERR_INVALID;
    if (!d_in || !d_w || !d_out || (mode % 2 == 0 && !d_bias)) return LB_ERR_INVALID;
    if ((int64_t)B * H * W > INT32_MAX) return LB_ERR_INVALID;
    const int npix = B * H * W;
    const bool gather = mode >= 2;
    const size_t need = gather ? 0 : conv_split_scratch(npix, H, W, Ci, Co);
    if (need > kNoAllocScratchFloats) return LB_ERR_SCRATCH;
    lb_unet tmp;
    tmp.stream = (hipStream_t)stream;
    if (need) {
        int rc = noalloc_scratch(&tmp.partial);
        if (rc) { tmp.stream = nullptr; return rc; }
        tmp.partial_n = kNoAllocScratchFloats;
    }
    if (mode % 2 == 0)
        conv_launch<bf16, bf16, true>(&tmp, (const bf16 *)d_in, (const bf16 *)d_w, d_bias, (bf16 *)d_out, npix, H, W, Ci,
                                      Co, gather);
    else
        conv_launch<float, float, false>(&tmp, (const float *)d_in, (const bf16 *)d_w, nullptr, (float *)d_out, npix, H, W,
                                         Ci, Co, gather);
    tmp.partial = nullptr;   // not the handle's to free
    tmp.stream = nullptr;
    return (int)hipGetLastError();
}

// Weight gradient of one 3x3 convolution (test and diagnostic entry point;
// allocates its scratch).  method 0: gathering kernel, 1: channel-major
// operands (needs W % 16 == 0), 2: LDS tiles (needs W % 32 == 0 and H % 8 == 0,
// or W % 16 == 0 and H % 16 == 0).
extern "C" int lb_conv3x3_wgrad_bf16(int method, const float *d_dout, const void *d_in, float *d_dw, int B, int H, int W,
                                     int Ci, int Co, lb_stream_t stream) {
    if (Ci % 16 || Co % 32 || B < 1 || H < 1 || W < 1 || method < 0 || method > 2) return LB_ERR_INVALID;
    if (method == 1 && W % 16) return LB_ERR_UNSUPPORTED;
    const int npix = B * H * W;
    const WgradTiles wt = wgrad_tiles(npix, H, W, Ci, Co);
    if (method == 2 && !wt.ntiles) return LB_ERR_UNSUPPORTED;
    const size_t nchunk = method == 2 ? (size_t)wt.nchunk : (size_t)wgrad_chunks_legacy(npix);
    lb_unet tmp;
    tmp.stream = (hipStream_t)stream;
    tmp.partial = tmp.alloc<float>(nchunk * Co * 9 * Ci);
    bf16 *tA = method ? tmp.alloc<bf16>((size_t)npix * Co) : nullptr;
    bf16 *tB = method ? tmp.alloc<bf16>((size_t)npix * Ci) : nullptr;
    int rc = 0;
    if (!tmp.partial || (method && (!tA || !tB))) rc = LB_ERR_INVALID;
    if (!rc) {
        wgrad_launch<bf16, float>(&tmp, d_dout, (const bf16 *)d_in, d_dw, npix, H, W, Ci, Co, tA, tB, method);
        rc = (int)hipGetLastError();
    }
    if (!rc) rc = (int)hipStreamSynchronize(tmp.stream);
    tmp.stream = nullptr;
    return rc;
}

// Predicted fields for the batched solver's warm start (see k_unet_fields).
extern "C" int lb_unet_to_fields(lb_unet *u, const float *d_y, int B, const float *d_uin, const float *stats6,
                                 float *d_fields) {
    if (!u || !d_y || !d_uin || !stats6 || !d_fields || B < 1 || u->d.out_ch != 3) return LB_ERR_INVALID;
    const int hw = u->d.height * u->d.width;
    hipLaunchKernelGGL(k_unet_fields, dim3(nb256((size_t)B * hw)), dim3(256), 0, u->stream, d_y, B, hw, d_uin, stats6[0],
                       stats6[1], stats6[2], stats6[3], stats6[4], stats6[5], d_fields);
    int rc = (int)hipGetLastError();
    if (!rc) rc = (int)hipStreamSynchronize(u->stream);
    return rc;
}
