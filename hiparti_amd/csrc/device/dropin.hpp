// Helpers for the drop-in entry points that take host ptiMatrix arrays.
#pragma once

#include "runtime.hpp"

namespace pti {
namespace dev {

// mats_order[1..N-1] must name every mode other than `mode` exactly once.
// The reference reads whatever modes it names (src/sptensor/mttkrp.c:84-97),
// so a repeated entry there squares one factor and drops another; the
// kernels here form the product over all non-target modes and size-check the
// factors they gather, so such an order is rejected instead.
inline bool order_is_permutation(ptiIndex nmodes, const ptiIndex mats_order[], ptiIndex mode) {
    uint32_t seen = 1u << mode;
    for (ptiIndex i = 1; i < nmodes; ++i) {
        const ptiIndex m = mats_order[i];
        if (m >= nmodes || (seen >> m & 1u)) return false;
        seen |= 1u << m;
    }
    return true;
}

// Reference shape rules for MTTKRP arguments (src/sptensor/mttkrp.c:58-65):
// every factor has the output's column count and ndims[m] rows.  Also checks
// that mats_order[1..N-1] is a permutation of the other modes, that all
// matrices share one row stride (the reference silently assumes it) and that
// the output can hold the rows.
inline int check_mttkrp_args(ptiIndex nmodes, const ptiIndex *ndims, ptiMatrix *const mats[],
                             const ptiIndex mats_order[], ptiIndex mode, const char *module) {
    if (nmodes < 2 || nmodes > 8) PTI_FAIL(PTIERR_VALUE_ERROR, module, "nmodes must be in [2, 8]");
    if (mode >= nmodes) PTI_FAIL(PTIERR_VALUE_ERROR, module, "mode out of range");
    for (ptiIndex i = 0; i < nmodes; ++i) {
        if (mats[i]->ncols != mats[nmodes]->ncols)
            PTI_FAIL(PTIERR_SHAPE_MISMATCH, module, "mats[i]->cols != mats[nmodes]->ncols");
        if (mats[i]->nrows != ndims[i]) PTI_FAIL(PTIERR_SHAPE_MISMATCH, module, "mats[i]->nrows != ndims[i]");
        if (mats[i]->stride != mats[0]->stride) PTI_FAIL(PTIERR_SHAPE_MISMATCH, module, "factor strides differ");
    }
    if (mats[nmodes]->stride != mats[0]->stride) PTI_FAIL(PTIERR_SHAPE_MISMATCH, module, "output stride differs");
    if (mats[nmodes]->cap < mats[mode]->nrows) PTI_FAIL(PTIERR_SHAPE_MISMATCH, module, "output has too few rows");
    if (!order_is_permutation(nmodes, mats_order, mode))
        PTI_FAIL(PTIERR_VALUE_ERROR, module, "mats_order[1..N-1] must be a permutation of the other modes");
    return 0;
}

// Device copies of the factor matrices a drop-in MTTKRP call reads, plus its
// output buffer.
struct StagedFactors {
    Buffer fac[8];
    Buffer out;
    float *ptrs[9] = {nullptr};
};

inline int stage_factors(ptiIndex nmodes, ptiMatrix *const mats[], const ptiIndex mats_order[], ptiIndex mode,
                         StagedFactors &sf, hipStream_t s, const char *module) {
    const ptiIndex stride = mats[0]->stride;
    for (ptiIndex i = 1; i < nmodes; ++i) {
        const ptiIndex m = mats_order[i];
        if (sf.fac[m].ptr) continue;
        size_t bytes = (size_t)mats[m]->nrows * stride * sizeof(float);
        PTI_HIP(sf.fac[m].alloc(bytes), module);
        if (bytes) PTI_HIP(hipMemcpyAsync(sf.fac[m].ptr, mats[m]->values, bytes, hipMemcpyHostToDevice, s), module);
        sf.ptrs[m] = sf.fac[m].as<float>();
    }
    size_t obytes = (size_t)mats[mode]->nrows * stride * sizeof(float);
    PTI_HIP(sf.out.alloc(obytes), module);
    sf.ptrs[nmodes] = sf.out.as<float>();
    // Modes that are not read (the target mode) still need a valid pointer.
    for (ptiIndex m = 0; m < nmodes; ++m)
        if (!sf.ptrs[m]) sf.ptrs[m] = sf.out.as<float>();
    return 0;
}

inline int fetch_output(ptiIndex nmodes, ptiMatrix *const mats[], ptiIndex mode, StagedFactors &sf, hipStream_t s,
                        const char *module) {
    size_t obytes = (size_t)mats[mode]->nrows * mats[0]->stride * sizeof(float);
    if (obytes) PTI_HIP(hipMemcpyAsync(mats[nmodes]->values, sf.out.ptr, obytes, hipMemcpyDeviceToHost, s), module);
    PTI_HIP(hipStreamSynchronize(s), module);
    return 0;
}

}  // namespace dev
}  // namespace pti
