// The reference's OpenMP and matrix-tiling MTTKRP entry points
// (include/includes/sptensors.h:184-203,258-300; src/sptensor/mttkrp_omp.c,
// src/sptensor/hicoo/mttkrp.c:313-542, src/sptensor/hicoo/mttkrp_omp.c).
// They compute the same product as ptiMTTKRP / ptiMTTKRPHiCOO with different
// CPU parallelisations; here every one of them runs the GPU kernels.  Thread
// counts, block-parallelism flags, scheduling choices, reduction buffers and
// lock pools are accepted and not used.  ptiRankMatrix arguments are viewed
// as ptiMatrix (same memory layout, 8-bit ncols and stride).
#include "internal.h"

namespace {

int hicoo_rank(const ptiSparseTensorHiCOO *hitsr, ptiRankMatrix *mats[], const ptiIndex mats_order[], ptiIndex mode) {
    const ptiIndex N = hitsr->nmodes;
    if (N < 1 || N > 8) PTI_FAIL(PTIERR_VALUE_ERROR, "HiCOO SpTns MTTKRP", "nmodes must be in [1, 8]");
    ptiMatrix views[9];
    ptiMatrix *ptrs[9];
    for (ptiIndex m = 0; m <= N; ++m) {
        if (mats[m]->stride == 0)
            PTI_FAIL(PTIERR_VALUE_ERROR, "HiCOO SpTns MTTKRP", "ptiRankMatrix stride is 0 (ncols > 248)");
        views[m].nrows = mats[m]->nrows;
        views[m].ncols = mats[m]->ncols;
        views[m].cap = mats[m]->cap;
        views[m].stride = mats[m]->stride;
        views[m].values = mats[m]->values;
        ptrs[m] = &views[m];
    }
    return ptiMTTKRPHiCOO(hitsr, ptrs, mats_order, mode);
}

}  // namespace

extern "C" {

int ptiOmpMTTKRP(ptiSparseTensor const *const X, ptiMatrix *mats[], ptiIndex const mats_order[], ptiIndex const mode,
                 const int tk) {
    (void)tk;
    return ptiMTTKRP(X, mats, mats_order, mode);
}

int ptiOmpMTTKRP_Reduce(ptiSparseTensor const *const X, ptiMatrix *mats[], ptiMatrix *copy_mats[],
                        ptiIndex const mats_order[], ptiIndex const mode, const int tk) {
    (void)copy_mats;
    (void)tk;
    return ptiMTTKRP(X, mats, mats_order, mode);
}

int ptiOmpMTTKRP_Lock(ptiSparseTensor const *const X, ptiMatrix *mats[], ptiIndex const mats_order[],
                      ptiIndex const mode, const int tk, ptiMutexPool *lock_pool) {
    (void)tk;
    (void)lock_pool;
    return ptiMTTKRP(X, mats, mats_order, mode);
}

int ptiOmpMTTKRPHiCOO(ptiSparseTensorHiCOO const *const hitsr, ptiMatrix *mats[], ptiIndex const mats_order[],
                      ptiIndex const mode, const int tk, const int tb) {
    (void)tk;
    (void)tb;
    return ptiMTTKRPHiCOO(hitsr, mats, mats_order, mode);
}

int ptiMTTKRPHiCOO_MatrixTiling(ptiSparseTensorHiCOO const *const hitsr, ptiRankMatrix *mats[],
                                ptiIndex const mats_order[], ptiIndex const mode) {
    return hicoo_rank(hitsr, mats, mats_order, mode);
}

int ptiOmpMTTKRPHiCOO_MatrixTiling(ptiSparseTensorHiCOO const *const hitsr, ptiRankMatrix *mats[],
                                   ptiIndex const mats_order[], ptiIndex const mode, const int tk, const int tb) {
    (void)tk;
    (void)tb;
    return hicoo_rank(hitsr, mats, mats_order, mode);
}

int ptiOmpMTTKRPHiCOO_MatrixTiling_Scheduled(ptiSparseTensorHiCOO const *const hitsr, ptiRankMatrix *mats[],
                                             ptiIndex const mats_order[], ptiIndex const mode, const int tk,
                                             const int tb, int balanced) {
    (void)tk;
    (void)tb;
    (void)balanced;
    return hicoo_rank(hitsr, mats, mats_order, mode);
}

int ptiOmpMTTKRPHiCOO_MatrixTiling_Scheduled_Reduce(ptiSparseTensorHiCOO const *const hitsr, ptiRankMatrix *mats[],
                                                    ptiRankMatrix *copy_mats[], ptiIndex const mats_order[],
                                                    ptiIndex const mode, const int tk, const int tb, int balanced) {
    (void)copy_mats;
    (void)tk;
    (void)tb;
    (void)balanced;
    return hicoo_rank(hitsr, mats, mats_order, mode);
}

int ptiOmpMTTKRPHiCOO_MatrixTiling_Scheduled_Reduce_Two(ptiSparseTensorHiCOO const *const hitsr,
                                                        ptiRankMatrix *mats[], ptiRankMatrix *copy_mats[],
                                                        ptiIndex const mats_order[], ptiIndex const mode,
                                                        const int tk, const int tb) {
    (void)copy_mats;
    (void)tk;
    (void)tb;
    return hicoo_rank(hitsr, mats, mats_order, mode);
}

}  // extern "C"
