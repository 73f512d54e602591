// ref_caller.cpp -- a compiled C++ caller of liblattibol_amd.so that keeps the
// reference's host code shape: the Data structs are filled field by field as
// the reference constructors do (simulation_d2q9.cu:304-399,
// simulation_d3q19.cu:151-261), and the step loop is the reference's runLBM
// (simulation_d2q9.cu:462-554, simulation_d3q19.cu:352-430) with its call
// sites unchanged: HIP's dim3 grid / block, data[dev].stream passed by
// reference to copy_border_devices_*, the probe after every step.
//
// Usage: ref_caller <input.bin> <output.bin>
//   input:  int32 {Q, ndev, W (per subdomain), H, D, steps, cx, cy, cz}
//           float32 {ux_in, uy_in, uz_in, rho_out, tau_inv, Fx, Fy, G, Gads}
//           per subdomain: f (Q*N float32, f0 plane first), solid (N bytes),
//           solid_bound (N bytes), rho, ux, uy (N float32 each) and for Q = 9
//           psi, ueq_x, ueq_y (N float32 each)
//   output: per subdomain f, rho, ux, uy, (uz for Q = 19), plot (8N / 4N
//           float32), then valueCur (4 float32) and lb_last_error() (int32).
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "lattibol_amd.h"

// Field offsets of the reference structs on LP64 (lattibol.h:29-63 and
// :65-97: pointers 8 bytes, int 4 bytes, cudaStream_t a pointer): a caller
// compiled against the reference header reads and writes the same bytes.
#define AT(T, f, off) static_assert(offsetof(T, f) == (off), #T "." #f);
AT(DataD2Q9, ux, 0) AT(DataD2Q9, uy, 8) AT(DataD2Q9, rho, 16) AT(DataD2Q9, solid, 24)
AT(DataD2Q9, solid_bound, 32) AT(DataD2Q9, f0, 40) AT(DataD2Q9, f1, 48) AT(DataD2Q9, f2, 56)
AT(DataD2Q9, f3, 64) AT(DataD2Q9, f4, 72) AT(DataD2Q9, f5, 80) AT(DataD2Q9, f6, 88) AT(DataD2Q9, f7, 96)
AT(DataD2Q9, f8, 104) AT(DataD2Q9, border_f1, 112) AT(DataD2Q9, border_f5, 120) AT(DataD2Q9, border_f8, 128)
AT(DataD2Q9, border_f3, 136) AT(DataD2Q9, border_f6, 144) AT(DataD2Q9, border_f7, 152)
AT(DataD2Q9, nb_border_f1, 160) AT(DataD2Q9, nb_border_f5, 168) AT(DataD2Q9, nb_border_f8, 176)
AT(DataD2Q9, nb_border_f3, 184) AT(DataD2Q9, nb_border_f6, 192) AT(DataD2Q9, nb_border_f7, 200)
AT(DataD2Q9, width, 208) AT(DataD2Q9, height, 212) AT(DataD2Q9, deviceCount, 216) AT(DataD2Q9, dev, 220)
AT(DataD2Q9, plotBuffer, 224) AT(DataD2Q9, valueCur, 232) AT(DataD2Q9, stream, 240) AT(DataD2Q9, psi, 248)
AT(DataD2Q9, ueq_x, 256) AT(DataD2Q9, ueq_y, 264)
static_assert(sizeof(DataD2Q9) == 272, "DataD2Q9 size");
AT(DataD3Q19, ux, 0) AT(DataD3Q19, uy, 8) AT(DataD3Q19, uz, 16) AT(DataD3Q19, rho, 24) AT(DataD3Q19, solid, 32)
AT(DataD3Q19, solid_bound, 40) AT(DataD3Q19, f0, 48) AT(DataD3Q19, f1, 56) AT(DataD3Q19, f2, 64)
AT(DataD3Q19, f3, 72) AT(DataD3Q19, f4, 80) AT(DataD3Q19, f5, 88) AT(DataD3Q19, f6, 96) AT(DataD3Q19, f7, 104)
AT(DataD3Q19, f8, 112) AT(DataD3Q19, f9, 120) AT(DataD3Q19, f10, 128) AT(DataD3Q19, f11, 136)
AT(DataD3Q19, f12, 144) AT(DataD3Q19, f13, 152) AT(DataD3Q19, f14, 160) AT(DataD3Q19, f15, 168)
AT(DataD3Q19, f16, 176) AT(DataD3Q19, f17, 184) AT(DataD3Q19, f18, 192) AT(DataD3Q19, border_f1, 200)
AT(DataD3Q19, border_f7, 208) AT(DataD3Q19, border_f8, 216) AT(DataD3Q19, border_f9, 224)
AT(DataD3Q19, border_f10, 232) AT(DataD3Q19, border_f2, 240) AT(DataD3Q19, border_f11, 248)
AT(DataD3Q19, border_f12, 256) AT(DataD3Q19, border_f13, 264) AT(DataD3Q19, border_f14, 272)
AT(DataD3Q19, nb_border_f1, 280) AT(DataD3Q19, nb_border_f7, 288) AT(DataD3Q19, nb_border_f8, 296)
AT(DataD3Q19, nb_border_f9, 304) AT(DataD3Q19, nb_border_f10, 312) AT(DataD3Q19, nb_border_f2, 320)
AT(DataD3Q19, nb_border_f11, 328) AT(DataD3Q19, nb_border_f12, 336) AT(DataD3Q19, nb_border_f13, 344)
AT(DataD3Q19, nb_border_f14, 352) AT(DataD3Q19, width, 360) AT(DataD3Q19, height, 364)
AT(DataD3Q19, depth, 368) AT(DataD3Q19, deviceCount, 372) AT(DataD3Q19, dev, 376)
AT(DataD3Q19, plotBuffer, 384) AT(DataD3Q19, valueCur, 392) AT(DataD3Q19, stream, 400)
static_assert(sizeof(DataD3Q19) == 408, "DataD3Q19 size");
#undef AT

#define CHECK(x)                                                                          \
    do {                                                                                  \
        hipError_t e_ = (x);                                                              \
        if (e_ != hipSuccess) {                                                           \
            std::fprintf(stderr, "ref_caller: %s at line %d: %s\n", #x, __LINE__,         \
                         hipGetErrorString(e_));                                          \
            std::exit(2);                                                                 \
        }                                                                                 \
    } while (0)

namespace {

FILE *in_f, *out_f;

template <class T> void rd(T *p, size_t n) {
    if (std::fread(p, sizeof(T), n, in_f) != n) { std::fprintf(stderr, "ref_caller: short input\n"); std::exit(2); }
}
template <class T> void wr(const T *p, size_t n) { std::fwrite(p, sizeof(T), n, out_f); }

// Allocates n elements on the device and fills them from the input file.
template <class T> T *upload(size_t n) {
    std::vector<unsigned char> h(n * sizeof(T));
    rd(h.data(), h.size());
    T *d = nullptr;
    CHECK(hipMalloc((void **)&d, h.size()));
    CHECK(hipMemcpy(d, h.data(), h.size(), hipMemcpyHostToDevice));
    return d;
}
float *dalloc(size_t n) {
    float *d = nullptr;
    CHECK(hipMalloc((void **)&d, n * sizeof(float)));
    CHECK(hipMemset(d, 0, n * sizeof(float)));
    return d;
}
void download(const float *d, size_t n) {
    std::vector<float> h(n);
    CHECK(hipMemcpy(h.data(), d, n * sizeof(float), hipMemcpyDeviceToHost));
    wr(h.data(), n);
}

struct Params {
    int Q, deviceCount, width, height, depth, updateSteps, xCur, yCur, zCur;
    float ux_in, uy_in, uz_in, rho_out, tauInv, Fx, Fy, G, Gads;
};

// The lattice weights as the reference's constructors assign them
// (simulation_d2q9.cu:56-58, simulation_d3q19.cu:54-56).
int run_d2q9(const Params &p) {
    const int deviceCount = p.deviceCount, width = p.width, height = p.height;
    const size_t size = (size_t)width * height;
    const float w0 = 16. / 36., w1 = 4. / 36., w2 = 1. / 36.;
    const float minU = 0, maxU = 0, minRho = 0, maxRho = 0;
    std::vector<DataD2Q9> data(deviceCount);
    std::vector<float *> plotData(deviceCount);
    for (int i = 0; i < deviceCount; i++) {
        DataD2Q9 &d = data[i];
        d.width = width;
        d.height = height;
        d.deviceCount = deviceCount;
        d.dev = i;
        float **f[9] = {&d.f0, &d.f1, &d.f2, &d.f3, &d.f4, &d.f5, &d.f6, &d.f7, &d.f8};
        for (float **q : f) *q = upload<float>(size);
        d.solid = upload<bool>(size);
        d.solid_bound = upload<bool>(size);
        d.rho = upload<float>(size);
        d.ux = upload<float>(size);
        d.uy = upload<float>(size);
        d.psi = upload<float>(size);
        d.ueq_x = upload<float>(size);
        d.ueq_y = upload<float>(size);
        float **b[12] = {&d.border_f1, &d.border_f5, &d.border_f8, &d.border_f3, &d.border_f6, &d.border_f7,
                         &d.nb_border_f1, &d.nb_border_f5, &d.nb_border_f8, &d.nb_border_f3, &d.nb_border_f6,
                         &d.nb_border_f7};
        for (float **q : b) *q = dalloc(height);
        d.valueCur = dalloc(4);
        hipStream_t s;
        CHECK(hipStreamCreate(&s));
        d.stream = s;
        plotData[i] = dalloc(8 * size);   // the display buffer (a GL PBO in the reference)
        d.plotBuffer = i ? dalloc(8 * size) : nullptr;
    }
    dim3 grid(width / 32, height / 4, 1), block(32, 4, 1);
    const float ux_in = p.ux_in, uy_in = p.uy_in, rho_out = p.rho_out, tauInv = p.tauInv;
    const float Fx = p.Fx, Fy = p.Fy, G = p.G, Gads = p.Gads;
    int left = 0, right = 0;
    float valueCur[4] = {0, 0, 0, 0};
    for (int subStep = 0; subStep < p.updateSteps; subStep++) {
        for (int dev = 0; dev < deviceCount; dev++) streaming_d2q9(&data[dev], grid, block);
        if (deviceCount > 1) {
            for (int dev = 0; dev < deviceCount; dev++) get_border_d2q9(&data[dev], grid, block);
            for (int dev = 0; dev < deviceCount; dev++) {
                if (dev > 0 && dev < (deviceCount - 1)) { left = dev - 1; right = dev + 1; }
                if (dev == 0) { left = deviceCount - 1; right = dev + 1; }
                if (dev == deviceCount - 1) { left = dev - 1; right = 0; }
                copy_border_devices_d2q9(data[dev].nb_border_f1, data[dev].nb_border_f5, data[dev].nb_border_f8,
                                         data[dev].nb_border_f3, data[dev].nb_border_f6, data[dev].nb_border_f7,
                                         data[right].border_f1, data[right].border_f5, data[right].border_f8,
                                         data[left].border_f3, data[left].border_f6, data[left].border_f7,
                                         height, data[dev].stream);
            }
            CHECK(hipDeviceSynchronize());
            for (int dev = 0; dev < deviceCount; dev++) apply_border_d2q9(&data[dev], grid, block);
        }
        CHECK(hipDeviceSynchronize());
        for (int dev = 0; dev < deviceCount; dev++) {
            if (dev == 0)
                collision_d2q9(plotData[dev], &data[dev], w0, w1, w2, tauInv, ux_in, uy_in, rho_out, minU, maxU,
                               minRho, maxRho, grid, block, Fx, Fy, G, Gads);
            else {
                collision_d2q9(data[dev].plotBuffer, &data[dev], w0, w1, w2, tauInv, ux_in, uy_in, rho_out, minU,
                               maxU, minRho, maxRho, grid, block, Fx, Fy, G, Gads);
                CHECK(hipMemcpy(plotData[dev], data[dev].plotBuffer, sizeof(float) * size * 8, hipMemcpyDefault));
            }
        }
        CHECK(hipDeviceSynchronize());
        read_value_2d(plotData[0], &data[0], p.xCur, p.yCur, p.zCur);
        CHECK(hipMemcpy(&valueCur, data[0].valueCur, sizeof(float) * 4, hipMemcpyDefault));
    }
    CHECK(hipDeviceSynchronize());
    for (int i = 0; i < deviceCount; i++) {
        const DataD2Q9 &d = data[i];
        const float *f[9] = {d.f0, d.f1, d.f2, d.f3, d.f4, d.f5, d.f6, d.f7, d.f8};
        for (const float *q : f) download(q, size);
        download(d.rho, size);
        download(d.ux, size);
        download(d.uy, size);
        download(plotData[i], 8 * size);
    }
    wr(valueCur, 4);
    return 0;
}

int run_d3q19(const Params &p) {
    const int deviceCount = p.deviceCount, width = p.width, height = p.height, depth = p.depth;
    const size_t size = (size_t)width * height * depth;
    const float w0 = 12. / 36., w1 = 2. / 36., w2 = 1. / 36.;
    const float minU = 0, maxU = 0, minRho = 0, maxRho = 0;
    std::vector<DataD3Q19> data(deviceCount);
    std::vector<float *> plotData(deviceCount);
    for (int i = 0; i < deviceCount; i++) {
        DataD3Q19 &d = data[i];
        d.width = width;
        d.height = height;
        d.depth = depth;
        d.deviceCount = deviceCount;
        d.dev = i;
        float **f[19] = {&d.f0, &d.f1, &d.f2, &d.f3, &d.f4, &d.f5, &d.f6, &d.f7, &d.f8, &d.f9, &d.f10,
                         &d.f11, &d.f12, &d.f13, &d.f14, &d.f15, &d.f16, &d.f17, &d.f18};
        for (float **q : f) *q = upload<float>(size);
        d.solid = upload<bool>(size);
        d.solid_bound = upload<bool>(size);
        d.rho = upload<float>(size);
        d.ux = upload<float>(size);
        d.uy = upload<float>(size);
        d.uz = dalloc(size);
        float **b[20] = {&d.border_f1, &d.border_f7, &d.border_f8, &d.border_f9, &d.border_f10,
                         &d.border_f2, &d.border_f11, &d.border_f12, &d.border_f13, &d.border_f14,
                         &d.nb_border_f1, &d.nb_border_f7, &d.nb_border_f8, &d.nb_border_f9, &d.nb_border_f10,
                         &d.nb_border_f2, &d.nb_border_f11, &d.nb_border_f12, &d.nb_border_f13, &d.nb_border_f14};
        for (float **q : b) *q = dalloc((size_t)height * depth);
        d.valueCur = dalloc(4);
        hipStream_t s;
        CHECK(hipStreamCreate(&s));
        d.stream = s;
        plotData[i] = dalloc(4 * size);
        d.plotBuffer = i ? dalloc(4 * size) : nullptr;
    }
    dim3 grid(width / 32, height / 4, depth), block(32, 4, 1);
    const float ux_in = p.ux_in, uy_in = p.uy_in, uz_in = p.uz_in, rho_out = p.rho_out, tauInv = p.tauInv;
    int left = 0, right = 0;
    float valueCur[4] = {0, 0, 0, 0};
    for (int subStep = 0; subStep < p.updateSteps; subStep++) {
        for (int dev = 0; dev < deviceCount; dev++) streaming_d3q19(&data[dev], grid, block);
        if (deviceCount > 1) {
            for (int dev = 0; dev < deviceCount; dev++) get_border_d3q19(&data[dev], grid, block);
            CHECK(hipDeviceSynchronize());
            for (int dev = 0; dev < deviceCount; dev++) {
                if (dev > 0 && dev < (deviceCount - 1)) { left = dev - 1; right = dev + 1; }
                if (dev == 0) { left = deviceCount - 1; right = dev + 1; }
                if (dev == deviceCount - 1) { left = dev - 1; right = 0; }
                copy_border_devices_d3q19(
                    data[dev].nb_border_f1, data[dev].nb_border_f7, data[dev].nb_border_f8, data[dev].nb_border_f9,
                    data[dev].nb_border_f10, data[dev].nb_border_f2, data[dev].nb_border_f11,
                    data[dev].nb_border_f12, data[dev].nb_border_f13, data[dev].nb_border_f14,
                    data[right].border_f1, data[right].border_f7, data[right].border_f8, data[right].border_f9,
                    data[right].border_f10, data[left].border_f2, data[left].border_f11, data[left].border_f12,
                    data[left].border_f13, data[left].border_f14, height, depth, data[dev].stream);
            }
            CHECK(hipDeviceSynchronize());
            for (int dev = 0; dev < deviceCount; dev++) apply_border_d3q19(&data[dev], grid, block);
        }
        CHECK(hipDeviceSynchronize());
        for (int dev = 0; dev < deviceCount; dev++) {
            if (dev == 0)
                collision_d3q19(plotData[dev], &data[dev], w0, w1, w2, tauInv, ux_in, uy_in, uz_in, rho_out, minU,
                                maxU, minRho, maxRho, grid, block);
            else {
                collision_d3q19(data[dev].plotBuffer, &data[dev], w0, w1, w2, tauInv, ux_in, uy_in, uz_in, rho_out,
                                minU, maxU, minRho, maxRho, grid, block);
                CHECK(hipMemcpy(plotData[dev], data[dev].plotBuffer, sizeof(float) * size * 4, hipMemcpyDefault));
            }
        }
        CHECK(hipDeviceSynchronize());
        read_value_3d(plotData[0], &data[0], p.xCur, p.yCur, p.zCur);
        CHECK(hipMemcpy(&valueCur, data[0].valueCur, sizeof(float) * 4, hipMemcpyDefault));
    }
    CHECK(hipDeviceSynchronize());
    for (int i = 0; i < deviceCount; i++) {
        const DataD3Q19 &d = data[i];
        const float *f[19] = {d.f0, d.f1, d.f2, d.f3, d.f4, d.f5, d.f6, d.f7, d.f8, d.f9,
                              d.f10, d.f11, d.f12, d.f13, d.f14, d.f15, d.f16, d.f17, d.f18};
        for (const float *q : f) download(q, size);
        download(d.rho, size);
        download(d.ux, size);
        download(d.uy, size);
        download(d.uz, size);
        download(plotData[i], 4 * size);
    }
    wr(valueCur, 4);
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: ref_caller <input.bin> <output.bin>\n");
        return 2;
    }
    in_f = std::fopen(argv[1], "rb");
    out_f = std::fopen(argv[2], "wb");
    if (!in_f || !out_f) {
        std::fprintf(stderr, "ref_caller: cannot open files\n");
        return 2;
    }
    Params p;
    int32_t iv[9];
    float fv[9];
    rd(iv, 9);
    rd(fv, 9);
    p.Q = iv[0]; p.deviceCount = iv[1]; p.width = iv[2]; p.height = iv[3]; p.depth = iv[4];
    p.updateSteps = iv[5]; p.xCur = iv[6]; p.yCur = iv[7]; p.zCur = iv[8];
    p.ux_in = fv[0]; p.uy_in = fv[1]; p.uz_in = fv[2]; p.rho_out = fv[3]; p.tauInv = fv[4];
    p.Fx = fv[5]; p.Fy = fv[6]; p.G = fv[7]; p.Gads = fv[8];
    // the reference's launch geometry needs these (simulation.cu:43-44)
    if (p.width % 32 || p.height % 4 || p.deviceCount < 1 || (p.Q != 9 && p.Q != 19)) {
        std::fprintf(stderr, "ref_caller: lattice shape does not fit the (32, 4, 1) blocks\n");
        return 2;
    }
    std::printf("ref_caller: D%dQ%d %d x %d x %d, %d subdomain(s), %d steps\n", p.Q == 9 ? 2 : 3, p.Q, p.width,
                p.height, p.depth, p.deviceCount, p.updateSteps);
    std::fflush(stdout);
    int rc = p.Q == 9 ? run_d2q9(p) : run_d3q19(p);
    int32_t err = lb_last_error();
    wr(&err, 1);
    std::fclose(out_f);
    std::fclose(in_f);
    std::printf("ref_caller: done, lb_last_error %d\n", err);
    return rc;
}
