// Dense row-major matrices in the ptiMatrix layout: stride = ncols rounded up
// to a multiple of 8, 32-byte aligned storage of cap*stride values
// (reference: src/matrix/matrix.c:36-55).  The column norms follow
// src/matrix/matrix.c:389-553 (2-norm; max-norm clamped to >= 1), computed
// here with FP64 column accumulators.
#include "internal.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>

extern "C" {

int ptiNewMatrix(ptiMatrix *mtx, ptiIndex const nrows, ptiIndex const ncols) {
    mtx->nrows = nrows;
    mtx->ncols = ncols;
    mtx->cap = nrows != 0 ? nrows : 1;
    mtx->stride = ncols == 0 ? 8 : ((ncols - 1) / 8 + 1) * 8;
    void *p = nullptr;
    size_t bytes = (size_t)mtx->cap * mtx->stride * sizeof(ptiValue);
    if (posix_memalign(&p, 8 * sizeof(ptiValue), bytes ? bytes : 32) != 0) p = nullptr;
    mtx->values = static_cast<ptiValue *>(p);
    if (!mtx->values) PTI_FAIL(PTIERR_OS_ERROR + ENOMEM, "Mtx New", "out of memory");
    return 0;
}

int ptiRandomizeMatrix(ptiMatrix *mtx) {
    for (ptiIndex i = 0; i < mtx->nrows; ++i)
        for (ptiIndex j = 0; j < mtx->ncols; ++j)
            mtx->values[(size_t)i * mtx->stride + j] = (ptiValue)(i + j + 1);
    return 0;
}

int ptiConstantMatrix(ptiMatrix *const mtx, ptiValue const val) {
    for (ptiIndex i = 0; i < mtx->nrows; ++i)
        for (ptiIndex j = 0; j < mtx->ncols; ++j) mtx->values[(size_t)i * mtx->stride + j] = val;
    return 0;
}

int ptiCopyMatrix(ptiMatrix *dest, const ptiMatrix *src) {
    PTI_CHECK(ptiNewMatrix(dest, src->nrows, src->ncols), "Mtx Copy");
    std::memcpy(dest->values, src->values, (size_t)src->nrows * src->stride * sizeof(ptiValue));
    return 0;
}

void ptiFreeMatrix(ptiMatrix *mtx) {
    std::free(mtx->values);
    mtx->values = nullptr;
    mtx->nrows = 0;
    mtx->ncols = 0;
    mtx->cap = 0;
    mtx->stride = 0;
}

int ptiDumpMatrix(ptiMatrix *mtx, FILE *fp) {
    if (std::fprintf(fp, "%" PRIu32 " x %" PRIu32 " matrix\n", mtx->nrows, mtx->ncols) < 0)
        PTI_FAIL(PTIERR_OS_ERROR + errno, "Mtx Dump", std::strerror(errno));
    for (ptiIndex i = 0; i < mtx->nrows; ++i) {
        for (ptiIndex j = 0; j < mtx->ncols; ++j)
            std::fprintf(fp, "%.2f\t", (double)mtx->values[(size_t)i * mtx->stride + j]);
        std::fprintf(fp, "\n");
    }
    std::fprintf(fp, "\n");
    return 0;
}

int ptiMatrix2Norm(ptiMatrix *const A, ptiValue *const lambda) {
    const ptiIndex R = A->ncols, S = A->stride;
    std::vector<double> acc(R, 0.0);
    for (ptiIndex i = 0; i < A->nrows; ++i)
        for (ptiIndex j = 0; j < R; ++j) {
            double v = A->values[(size_t)i * S + j];
            acc[j] += v * v;
        }
    for (ptiIndex j = 0; j < R; ++j) lambda[j] = (ptiValue)std::sqrt(acc[j]);
    for (ptiIndex i = 0; i < A->nrows; ++i)
        for (ptiIndex j = 0; j < R; ++j) A->values[(size_t)i * S + j] /= lambda[j];
    return 0;
}

int ptiMatrixMaxNorm(ptiMatrix *const A, ptiValue *const lambda) {
    const ptiIndex R = A->ncols, S = A->stride;
    for (ptiIndex j = 0; j < R; ++j) lambda[j] = 0;
    for (ptiIndex i = 0; i < A->nrows; ++i)
        for (ptiIndex j = 0; j < R; ++j)
            if (A->values[(size_t)i * S + j] > lambda[j]) lambda[j] = A->values[(size_t)i * S + j];
    for (ptiIndex j = 0; j < R; ++j)
        if (lambda[j] < 1) lambda[j] = 1;
    for (ptiIndex i = 0; i < A->nrows; ++i)
        for (ptiIndex j = 0; j < R; ++j) A->values[(size_t)i * S + j] /= lambda[j];
    return 0;
}

}  // extern "C"
