// Device runtime shared by the HIP sources of libHiParTI.so: error mapping,
// owning device buffers, and launch helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../host/internal.h"

namespace pti {
namespace dev {

// Map a HIP status to the reference's error space (PTIERR_CUDA_ERROR + code).
inline int hip_code(hipError_t e) { return e == hipSuccess ? 0 : (int)PTIERR_CUDA_ERROR + (int)e; }

#define PTI_HIP(expr, module)                                                      \
    do {                                                                           \
        hipError_t _e = (expr);                                                    \
        if (_e != hipSuccess) {                                                    \
            pti_ComplainError((module), ::pti::dev::hip_code(_e), __FILE__, __LINE__, \
                              hipGetErrorString(_e));                              \
            return ::pti::dev::hip_code(_e);                                       \
        }                                                                          \
    } while (0)

// Checks that a HIP device is usable; records and returns an error otherwise.
int require_device(const char *module);

// Owning device allocation.
struct Buffer {
    void *ptr = nullptr;
    size_t bytes = 0;
    Buffer() = default;
    Buffer(const Buffer &) = delete;
    Buffer &operator=(const Buffer &) = delete;
    Buffer(Buffer &&o) noexcept : ptr(o.ptr), bytes(o.bytes) { o.ptr = nullptr; o.bytes = 0; }
    Buffer &operator=(Buffer &&o) noexcept {
        release();
        ptr = o.ptr;
        bytes = o.bytes;
        o.ptr = nullptr;
        o.bytes = 0;
        return *this;
    }
    ~Buffer() { release(); }
    hipError_t alloc(size_t n) {
        release();
        bytes = n;
        return hipMalloc(&ptr, n ? n : 256);
    }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    template <typename T>
    T *as() const { return static_cast<T *>(ptr); }
};

// Zero `n` floats starting at p (a kernel, not hipMemsetAsync, so that the
// call stays valid inside captured HIP graphs).
hipError_t zero_floats(float *p, uint64_t n, hipStream_t s);

}  // namespace dev
}  // namespace pti
