// Device runtime: device checks, zero-fill kernel, device-event timers.
#include "runtime.hpp"

namespace pti {
namespace dev {

int require_device(const char *module) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        pti_ComplainError(module, hip_code(e), __FILE__, __LINE__, hipGetErrorString(e));
        return hip_code(e);
    }
    if (n <= 0) {
        int code = hip_code(hipErrorNoDevice);
        pti_ComplainError(module, code, __FILE__, __LINE__, "no HIP device");
        return code;
    }
    return 0;
}

// Clears n4 float4s at p4 and then ntail (< 4) floats at tail.
__global__ void __launch_bounds__(256) zero_kernel(float4 *__restrict__ p4, uint64_t n4, float *__restrict__ tail,
                                                   uint32_t ntail) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (; i < n4; i += step) p4[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (blockIdx.x == 0 && threadIdx.x < ntail) tail[threadIdx.x] = 0.f;
}

__global__ void __launch_bounds__(256) zero_scalar_kernel(float *__restrict__ p, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (; i < n; i += step) p[i] = 0.f;
}

hipError_t zero_floats(float *p, uint64_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    if ((uintptr_t)p % 16 == 0) {
        uint64_t n4 = n / 4;
        uint32_t ntail = (uint32_t)(n - n4 * 4);
        uint64_t blocks = (n4 + 255) / 256;
        if (blocks > 8192) blocks = 8192;
        if (blocks == 0) blocks = 1;
        hipLaunchKernelGGL(zero_kernel, dim3((uint32_t)blocks), dim3(256), 0, s, reinterpret_cast<float4 *>(p), n4,
                           p + n4 * 4, ntail);
    } else {
        uint64_t blocks = (n + 255) / 256;
        if (blocks > 8192) blocks = 8192;
        hipLaunchKernelGGL(zero_scalar_kernel, dim3((uint32_t)blocks), dim3(256), 0, s, p, n);
    }
    return hipGetLastError();
}

}  // namespace dev

// Device-event timer hooks used by ptiNewTimer(..., use_cuda != 0).
struct DevTimer {
    hipEvent_t a, b;
};

int dev_timer_create(void **state) {
    int rc = dev::require_device("Timer New");
    if (rc) return rc;
    DevTimer *t = new DevTimer();
    PTI_HIP(hipEventCreate(&t->a), "Timer New");
    PTI_HIP(hipEventCreate(&t->b), "Timer New");
    *state = t;
    return 0;
}
int dev_timer_start(void *state) {
    PTI_HIP(hipEventRecord(static_cast<DevTimer *>(state)->a, nullptr), "Timer Start");
    return 0;
}
int dev_timer_stop(void *state) {
    DevTimer *t = static_cast<DevTimer *>(state);
    PTI_HIP(hipEventRecord(t->b, nullptr), "Timer Stop");
    PTI_HIP(hipEventSynchronize(t->b), "Timer Stop");
    return 0;
}
double dev_timer_elapsed(void *state) {
    DevTimer *t = static_cast<DevTimer *>(state);
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, t->a, t->b) != hipSuccess) return -1.0;
    return ms * 1e-3;
}
void dev_timer_destroy(void *state) {
    DevTimer *t = static_cast<DevTimer *>(state);
    if (!t) return;
    (void)hipEventDestroy(t->a);
    (void)hipEventDestroy(t->b);
    delete t;
}

}  // namespace pti

extern "C" int ptiDevCount(void) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) return -pti::dev::hip_code(e);
    return n;
}
