// Internal helpers shared by the host-side sources of libHiParTI.so.
#pragma once

#include "../../../include/HiParTI.h"

#include <cerrno>
#include <cstring>
#include <cstdint>
#include <thread>
#include <vector>
#include <algorithm>
#include <functional>

// Record an error (thread-local, as ptiGetLastError reports it) and return it.
#define PTI_FAIL(code, module, reason)                                   \
    do {                                                                 \
        int _pti_code = (code);                                          \
        pti_ComplainError((module), _pti_code, __FILE__, __LINE__, (reason)); \
        return _pti_code;                                                \
    } while (0)

#define PTI_CHECK(expr, module)                                          \
    do {                                                                 \
        int _pti_rc = (expr);                                            \
        if (_pti_rc != 0) {                                              \
            pti_ComplainError((module), _pti_rc, __FILE__, __LINE__, #expr); \
            return _pti_rc;                                              \
        }                                                                \
    } while (0)

#define PTI_CHECK_ALLOC(ptr, module)                                     \
    do {                                                                 \
        if (!(ptr)) PTI_FAIL(PTIERR_OS_ERROR + ENOMEM, (module), "out of memory"); \
    } while (0)

namespace pti {

// Number of host threads to use when the caller passes nthreads <= 0.
int default_threads();

// Run body(begin, end, tid) over [0, n) split into `nthreads` contiguous ranges.
void parallel_for(uint64_t n, int nthreads,
                  const std::function<void(uint64_t, uint64_t, int)> &body);

// Quiet flag for conversion progress output (HIPARTI_VERBOSE=1 enables it).
bool verbose();

}  // namespace pti
