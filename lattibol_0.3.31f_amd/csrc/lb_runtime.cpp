// lb_runtime.cpp -- error state and small host utilities of the C ABI.
#include <hip/hip_runtime.h>

#include <atomic>

#include "../../include/lattibol_amd.h"

static std::atomic<int> g_last_error{0};

extern "C" void lb_set_error(int code) {
    int expected = 0;
    if (code != 0) g_last_error.compare_exchange_strong(expected, code);
}

int lb_hip_err(hipError_t e) {
    if (e != hipSuccess) {
        lb_set_error((int)e);
        return (int)e;
    }
    return 0;
}

extern "C" int lb_last_error(void) { return g_last_error.exchange(0); }

extern "C" const char *lb_error_string(int code) {
    switch (code) {
    case LB_OK: return "success";
    case LB_ERR_INVALID: return "invalid argument";
    case LB_ERR_SHAPE: return "grid*block does not match the lattice shape";
    case LB_ERR_UNSUPPORTED: return "unsupported configuration";
    case LB_ERR_NO_DEVICE: return "no usable GPU device";
    case LB_ERR_NOT_READY: return "requested state has not been computed";
    case LB_ERR_SCRATCH: return "work exceeds the fixed device scratch of the non-allocating entry point";
    default: break;
    }
    if (code > 0) return hipGetErrorString((hipError_t)code);
    return "unknown error";
}

extern "C" int lb_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

// Device-memory plumbing for hosts without a GPU framework (Python via ctypes).
extern "C" int lb_device_alloc(size_t bytes, void **out) {
    if (!out) return LB_ERR_INVALID;
    *out = nullptr;
    return (int)hipMalloc(out, bytes ? bytes : 1);
}

extern "C" int lb_device_free(void *p) { return p ? (int)hipFree(p) : 0; }

// The solver, batch and U-Net handles run on non-blocking streams, which do
// not order against the null stream: wait for the whole device first, so an
// upload never overwrites a buffer that queued kernels still read and a
// download never reads one they still write.
extern "C" int lb_memcpy_h2d(void *dst, const void *src, size_t bytes) {
    int rc = (int)hipDeviceSynchronize();
    return rc ? rc : (int)hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
}

extern "C" int lb_memcpy_d2h(void *dst, const void *src, size_t bytes) {
    int rc = (int)hipDeviceSynchronize();
    return rc ? rc : (int)hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
}

extern "C" int lb_device_synchronize(void) { return (int)hipDeviceSynchronize(); }
