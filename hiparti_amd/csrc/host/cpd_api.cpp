// Host-side helpers of the CP-ALS API: ptiRankMatrix / ptiRankKruskalTensor
// containers, the Gram Hadamard product, final lambda and the Kruskal fit
// (reference: src/matrix/rankmatrix.c, src/matrix/matrix.c:340-385,556-572,
// src/ktensor/ktensor.c:77-240, src/ktensor/rankktensor.c).  These are the
// reference's host utilities for callers that drive their own ALS loop;
// ptiCpdAls / ptiCpdAlsHiCOO do the same work on the device.
//
// Gram matrices ("ata") are read only in their row-major upper triangle
// (entry [i][j], j >= i), the triangle a column-major ssyrk('L') of the
// row-major factor writes.  The reference's KruskalTensorFrobeniusNormSquared
// reads the other triangle (ktensor.c:128-143), which nothing writes; here
// the written one is used, so the fit depends only on the inputs.
#include "internal.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>

namespace {

// A ptiRankMatrix has ptiMatrix's memory layout with 8-bit ncols/stride.
ptiMatrix view_of(const ptiRankMatrix *r) {
    ptiMatrix m;
    m.nrows = r->nrows;
    m.ncols = r->ncols;
    m.cap = r->cap;
    m.stride = r->stride;
    m.values = r->values;
    return m;
}

// aTa[nmodes] = Hadamard product of aTa[k] (upper triangles), k != mode,
// mirrored to a full symmetric matrix (reference matrix.c:341-385).
template <typename Mat>
int dot_mul_seq_triangle(ptiIndex mode, ptiIndex nmodes, Mat **mats) {
    const ptiIndex n = mats[0]->nrows, R = mats[0]->ncols, S = mats[0]->stride;
    for (ptiIndex m = 1; m <= nmodes; ++m)
        if (mats[m]->ncols != R || mats[m]->nrows != n || mats[m]->stride != S)
            PTI_FAIL(PTIERR_SHAPE_MISMATCH, "Mtx DotMulSeqTriangle", "Gram shapes differ");
    ptiValue *o = mats[nmodes]->values;
    for (ptiIndex i = 0; i < n; ++i)
        for (ptiIndex j = 0; j < R; ++j) o[(size_t)i * S + j] = 1.0f;
    for (ptiIndex m = 1; m < nmodes; ++m) {
        const ptiValue *v = mats[(mode + m) % nmodes]->values;
        for (ptiIndex i = 0; i < n; ++i)
            for (ptiIndex j = i; j < R; ++j) o[(size_t)i * S + j] *= v[(size_t)i * S + j];
    }
    for (ptiIndex i = 0; i < n; ++i)
        for (ptiIndex j = 0; j < i && j < R; ++j) o[(size_t)i * S + j] = o[(size_t)j * S + i];
    return 0;
}

// ||K||^2 = sum_pq l_p l_q prod_k ata_k[p][q] over the full symmetric
// Hadamard product built from the upper triangles (FP64).
double kruskal_norm_sq(ptiIndex nmodes, const ptiValue *lambda, const ptiMatrix *const *ata, ptiValue *scratch) {
    const ptiIndex R = ata[0]->ncols, S = ata[0]->stride;
    double norm = 0.0;
    for (ptiIndex i = 0; i < R; ++i)
        for (ptiIndex j = i; j < R; ++j) {
            double h = 1.0;
            for (ptiIndex m = 0; m < nmodes; ++m) h *= ata[m]->values[(size_t)i * S + j];
            if (scratch) {
                scratch[(size_t)i * S + j] = (ptiValue)h;
                scratch[(size_t)j * S + i] = (ptiValue)h;
            }
            const double t = h * lambda[i] * lambda[j];
            norm += (i == j) ? t : 2 * t;
        }
    return std::fabs(norm);
}

// <X,K> = sum_r lambda_r sum_i last[i][r] * mttkrp[i][r] (FP64).
double inner_product(ptiIndex nmodes, const ptiValue *lambda, const ptiMatrix *last, const ptiMatrix *mt) {
    const ptiIndex R = last->ncols, S = last->stride;
    double inner = 0.0;
    for (ptiIndex r = 0; r < R; ++r) {
        double acc = 0.0;
        for (ptiIndex i = 0; i < last->nrows; ++i)
            acc += (double)last->values[(size_t)i * S + r] * (double)mt->values[(size_t)i * mt->stride + r];
        inner += acc * lambda[r];
    }
    (void)nmodes;
    return inner;
}

double fit_from(double normX2, double normK, double inner) {
    double residual = normX2 + normK - 2 * inner;
    if (residual > 0.0) residual = std::sqrt(residual);
    return 1 - residual / std::sqrt(normX2);
}

}  // namespace

extern "C" {

int ptiMatrixDotMulSeqTriangle(ptiIndex const mode, ptiIndex const nmodes, ptiMatrix **mats) {
    return dot_mul_seq_triangle(mode, nmodes, mats);
}

int ptiRankMatrixDotMulSeqTriangle(ptiIndex const mode, ptiIndex const nmodes, ptiRankMatrix **mats) {
    return dot_mul_seq_triangle(mode, nmodes, mats);
}

void GetFinalLambda(ptiIndex const rank, ptiIndex const nmodes, ptiMatrix **mats, ptiValue *const lambda) {
    std::vector<ptiValue> tmp(rank);
    for (ptiIndex m = 0; m < nmodes; ++m) {
        ptiMatrix2Norm(mats[m], tmp.data());
        for (ptiIndex r = 0; r < rank; ++r) lambda[r] *= tmp[r];
    }
}

void GetRankFinalLambda(ptiElementIndex const rank, ptiIndex const nmodes, ptiRankMatrix **mats,
                        ptiValue *const lambda) {
    std::vector<ptiValue> tmp(rank);
    for (ptiIndex m = 0; m < nmodes; ++m) {
        ptiRankMatrix2Norm(mats[m], tmp.data());
        for (ptiIndex r = 0; r < rank; ++r) lambda[r] *= tmp[r];
    }
}

double KruskalTensorFrobeniusNormSquared(ptiIndex const nmodes, ptiValue const *const lambda, ptiMatrix **ata) {
    return kruskal_norm_sq(nmodes, lambda, ata, ata[nmodes]->values);
}

double SparseKruskalTensorInnerProduct(ptiIndex const nmodes, ptiValue const *const lambda, ptiMatrix **mats) {
    return inner_product(nmodes, lambda, mats[nmodes - 1], mats[nmodes]);
}

double KruskalTensorFit(ptiSparseTensor const *const ptien, ptiValue const *const lambda, ptiMatrix **mats,
                        ptiMatrix **ata) {
    const ptiIndex N = ptien->nmodes;
    const double normX2 = SparseTensorFrobeniusNormSquared(ptien);
    const double normK = KruskalTensorFrobeniusNormSquared(N, lambda, ata);
    const double inner = SparseKruskalTensorInnerProduct(N, lambda, mats);
    return fit_from(normX2, normK, inner);
}

double KruskalTensorFrobeniusNormSquaredRank(ptiIndex const nmodes, ptiValue const *const lambda,
                                             ptiRankMatrix **ata) {
    std::vector<ptiMatrix> v(nmodes + 1);
    std::vector<const ptiMatrix *> p(nmodes + 1);
    for (ptiIndex m = 0; m <= nmodes; ++m) {
        v[m] = view_of(ata[m]);
        p[m] = &v[m];
    }
    return kruskal_norm_sq(nmodes, lambda, p.data(), ata[nmodes]->values);
}

double SparseKruskalTensorInnerProductRank(ptiIndex const nmodes, ptiValue const *const lambda,
                                           ptiRankMatrix **mats) {
    const ptiMatrix last = view_of(mats[nmodes - 1]), mt = view_of(mats[nmodes]);
    return inner_product(nmodes, lambda, &last, &mt);
}

double KruskalTensorFitHiCOO(ptiSparseTensorHiCOO const *const hitsr, ptiValue const *const lambda,
                             ptiRankMatrix **mats, ptiRankMatrix **ata) {
    const ptiIndex N = hitsr->nmodes;
    const double normX2 = SparseTensorFrobeniusNormSquaredHiCOO(hitsr);
    const double normK = KruskalTensorFrobeniusNormSquaredRank(N, lambda, ata);
    const double inner = SparseKruskalTensorInnerProductRank(N, lambda, mats);
    return fit_from(normX2, normK, inner);
}

// ---- ptiRankMatrix (reference src/matrix/rankmatrix.c) --------------------

int ptiNewRankMatrix(ptiRankMatrix *mtx, ptiIndex const nrows, ptiElementIndex const ncols) {
    if (ncols == 0 || ncols > 248) PTI_FAIL(PTIERR_VALUE_ERROR, "RankMtx New", "ncols must be in [1, 248]");
    mtx->nrows = nrows;
    mtx->ncols = ncols;
    mtx->cap = nrows != 0 ? nrows : 1;
    mtx->stride = (ptiElementIndex)(((ncols - 1) / 8 + 1) * 8);
    void *p = nullptr;
    const size_t bytes = (size_t)mtx->cap * mtx->stride * sizeof(ptiValue);
    if (posix_memalign(&p, 8 * sizeof(ptiValue), bytes) != 0) p = nullptr;
    mtx->values = static_cast<ptiValue *>(p);
    if (!mtx->values) PTI_FAIL(PTIERR_OS_ERROR + ENOMEM, "RankMtx New", "out of memory");
    return 0;
}

int ptiRandomizeRankMatrix(ptiRankMatrix *mtx, ptiIndex const nrows, ptiElementIndex const ncols) {
    for (ptiIndex i = 0; i < nrows; ++i)
        for (ptiElementIndex j = 0; j < ncols; ++j) mtx->values[(size_t)i * mtx->stride + j] = (ptiValue)(i + j + 1);
    return 0;
}

int ptiConstantRankMatrix(ptiRankMatrix *mtx, ptiValue const val) {
    for (ptiIndex i = 0; i < mtx->nrows; ++i)
        for (ptiElementIndex j = 0; j < mtx->ncols; ++j) mtx->values[(size_t)i * mtx->stride + j] = val;
    return 0;
}

void ptiFreeRankMatrix(ptiRankMatrix *mtx) {
    std::free(mtx->values);
    mtx->values = nullptr;
    mtx->nrows = 0;
    mtx->ncols = 0;
    mtx->cap = 0;
    mtx->stride = 0;
}

int ptiDumpRankMatrix(ptiRankMatrix *mtx, FILE *fp) {
    ptiMatrix v = view_of(mtx);
    return ptiDumpMatrix(&v, fp);
}

int ptiRankMatrix2Norm(ptiRankMatrix *const A, ptiValue *const lambda) {
    ptiMatrix v = view_of(A);
    return ptiMatrix2Norm(&v, lambda);
}

int ptiRankMatrixMaxNorm(ptiRankMatrix *const A, ptiValue *const lambda) {
    ptiMatrix v = view_of(A);
    return ptiMatrixMaxNorm(&v, lambda);
}

// ---- ptiRankKruskalTensor (reference src/ktensor/rankktensor.c) ----------

int ptiNewRankKruskalTensor(ptiRankKruskalTensor *ktsr, ptiIndex nmodes, const ptiIndex ndims[],
                            ptiElementIndex rank) {
    ktsr->nmodes = nmodes;
    ktsr->rank = rank;
    ktsr->ndims = static_cast<ptiIndex *>(std::malloc((nmodes ? nmodes : 1) * sizeof(ptiIndex)));
    ktsr->lambda = static_cast<ptiValue *>(std::calloc(rank ? rank : 1, sizeof(ptiValue)));
    PTI_CHECK_ALLOC(ktsr->ndims, "RankKruskalTns New");
    PTI_CHECK_ALLOC(ktsr->lambda, "RankKruskalTns New");
    for (ptiIndex i = 0; i < nmodes; ++i) ktsr->ndims[i] = ndims[i];
    ktsr->fit = 0.0;
    ktsr->factors = nullptr;
    return 0;
}

void ptiFreeRankKruskalTensor(ptiRankKruskalTensor *ktsr) {
    if (ktsr->factors) {
        for (ptiIndex i = 0; i < ktsr->nmodes; ++i)
            if (ktsr->factors[i]) {
                ptiFreeRankMatrix(ktsr->factors[i]);
                std::free(ktsr->factors[i]);
            }
        std::free(ktsr->factors);
    }
    std::free(ktsr->ndims);
    std::free(ktsr->lambda);
    ktsr->factors = nullptr;
    ktsr->ndims = nullptr;
    ktsr->lambda = nullptr;
    ktsr->rank = 0;
    ktsr->fit = 0.0;
    ktsr->nmodes = 0;
}

int ptiDumpRankKruskalTensor(ptiRankKruskalTensor *ktsr, FILE *fp) {
    if (std::fprintf(fp, "nmodes: %" PRIu32 ", rank: %u\n", ktsr->nmodes, (unsigned)ktsr->rank) < 0)
        PTI_FAIL(PTIERR_OS_ERROR + errno, "RankKruskalTns Dump", std::strerror(errno));
    for (ptiIndex m = 0; m < ktsr->nmodes; ++m) std::fprintf(fp, m ? " %" PRIu32 : "%" PRIu32, ktsr->ndims[m]);
    std::fputs("\n", fp);
    std::fprintf(fp, "fit: %lf\n", ktsr->fit);
    std::fprintf(fp, "lambda:\n");
    for (ptiIndex r = 0; r < ktsr->rank; ++r) std::fprintf(fp, "%f ", (double)ktsr->lambda[r]);
    std::fprintf(fp, "\n");
    std::fprintf(fp, "Factor matrices:\n");
    for (ptiIndex m = 0; m < ktsr->nmodes; ++m)
        PTI_CHECK(ptiDumpRankMatrix(ktsr->factors[m], fp), "RankKruskalTns Dump");
    return 0;
}

}  // extern "C"
