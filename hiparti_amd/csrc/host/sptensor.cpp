// COO sparse tensors: allocation, copy, text IO, lexicographic sort, norm.
//
// Layout and semantics follow the reference (src/sptensor/sptensor.c:30-109,
// src/sptensor/load.c:63-118 and 277-297, src/sptensor/dump.c:29-53).  The
// text loader reads "nmodes", the mode sizes, then one nonzero per line as
// nmodes indices (start_index based) and a value; it rejects indices below
// start_index with PTIERR_VALUE_ERROR and returns PTIERR_OS_ERROR+errno when
// the file cannot be opened (the reference aborts there).
#include "internal.h"
#include "keys.h"

#include <cstdio>
#include <cstdlib>

extern "C" {

int ptiNewSparseTensor(ptiSparseTensor *tsr, ptiIndex nmodes, const ptiIndex ndims[]) {
    tsr->nmodes = nmodes;
    tsr->sortorder = static_cast<ptiIndex *>(std::malloc((nmodes ? nmodes : 1) * sizeof(ptiIndex)));
    tsr->ndims = static_cast<ptiIndex *>(std::malloc((nmodes ? nmodes : 1) * sizeof(ptiIndex)));
    tsr->inds = static_cast<ptiIndexVector *>(std::malloc((nmodes ? nmodes : 1) * sizeof(ptiIndexVector)));
    if (!tsr->sortorder || !tsr->ndims || !tsr->inds)
        PTI_FAIL(PTIERR_OS_ERROR + ENOMEM, "SpTns New", "out of memory");
    for (ptiIndex i = 0; i < nmodes; ++i) {
        tsr->sortorder[i] = i;
        tsr->ndims[i] = ndims[i];
        PTI_CHECK(ptiNewIndexVector(&tsr->inds[i], 0, 0), "SpTns New");
    }
    tsr->nnz = 0;
    PTI_CHECK(ptiNewValueVector(&tsr->values, 0, 0), "SpTns New");
    return 0;
}

int ptiCopySparseTensor(ptiSparseTensor *dest, const ptiSparseTensor *src, int const nt) {
    (void)nt;
    PTI_CHECK(ptiNewSparseTensor(dest, src->nmodes, src->ndims), "SpTns Copy");
    std::memcpy(dest->sortorder, src->sortorder, src->nmodes * sizeof(ptiIndex));
    dest->nnz = src->nnz;
    for (ptiIndex m = 0; m < src->nmodes; ++m) {
        PTI_CHECK(ptiResizeIndexVector(&dest->inds[m], src->nnz), "SpTns Copy");
        if (src->nnz) std::memcpy(dest->inds[m].data, src->inds[m].data, src->nnz * sizeof(ptiIndex));
    }
    PTI_CHECK(ptiResizeValueVector(&dest->values, src->nnz), "SpTns Copy");
    if (src->nnz) std::memcpy(dest->values.data, src->values.data, src->nnz * sizeof(ptiValue));
    return 0;
}

void ptiFreeSparseTensor(ptiSparseTensor *tsr) {
    for (ptiIndex i = 0; i < tsr->nmodes; ++i) ptiFreeIndexVector(&tsr->inds[i]);
    std::free(tsr->sortorder);
    std::free(tsr->ndims);
    std::free(tsr->inds);
    ptiFreeValueVector(&tsr->values);
    tsr->sortorder = nullptr;
    tsr->ndims = nullptr;
    tsr->inds = nullptr;
    tsr->nmodes = 0;
    tsr->nnz = 0;
}

double SparseTensorFrobeniusNormSquared(ptiSparseTensor const *const ptien) {
    double norm = 0;
    const ptiValue *v = ptien->values.data;
    for (ptiNnzIndex n = 0; n < ptien->nnz; ++n) norm += (double)v[n] * (double)v[n];
    return norm;
}

int ptiLoadSparseTensor(ptiSparseTensor *tsr, ptiIndex start_index, char const *const fname) {
    FILE *fp = std::fopen(fname, "r");
    if (!fp) PTI_FAIL(PTIERR_OS_ERROR + errno, "SpTns Load", std::strerror(errno));
    unsigned nmodes = 0;
    if (std::fscanf(fp, "%u", &nmodes) != 1 || nmodes == 0 || nmodes > 64) {
        std::fclose(fp);
        PTI_FAIL(PTIERR_VALUE_ERROR, "SpTns Load", "bad header: nmodes");
    }
    std::vector<ptiIndex> ndims(nmodes);
    for (unsigned m = 0; m < nmodes; ++m) {
        if (std::fscanf(fp, "%u", &ndims[m]) != 1) {
            std::fclose(fp);
            PTI_FAIL(PTIERR_VALUE_ERROR, "SpTns Load", "bad header: ndims");
        }
    }
    int rc = ptiNewSparseTensor(tsr, nmodes, ndims.data());
    if (rc) {
        std::fclose(fp);
        return rc;
    }
    // The reference leaves sortorder zero-filled after loading (load.c:70-72).
    std::memset(tsr->sortorder, 0, nmodes * sizeof(ptiIndex));
    for (;;) {
        unsigned idx[64];
        unsigned m = 0;
        for (; m < nmodes; ++m)
            if (std::fscanf(fp, "%u", &idx[m]) != 1) break;
        if (m < nmodes) break;
        double value;
        if (std::fscanf(fp, "%lf", &value) != 1) break;
        for (m = 0; m < nmodes; ++m) {
            if (idx[m] < start_index) {
                std::fclose(fp);
                PTI_FAIL(PTIERR_VALUE_ERROR, "SpTns Load", "index < start_index");
            }
            ptiAppendIndexVector(&tsr->inds[m], idx[m] - start_index);
        }
        ptiAppendValueVector(&tsr->values, (ptiValue)value);
        ++tsr->nnz;
    }
    std::fclose(fp);
    return 0;
}

int ptiDumpSparseTensor(const ptiSparseTensor *tsr, ptiIndex start_index, FILE *fp) {
    if (std::fprintf(fp, "%" PRIu32 "\n", tsr->nmodes) < 0)
        PTI_FAIL(PTIERR_OS_ERROR + errno, "SpTns Dump", std::strerror(errno));
    for (ptiIndex m = 0; m < tsr->nmodes; ++m)
        std::fprintf(fp, m ? " %" PRIu32 : "%" PRIu32, tsr->ndims[m]);
    std::fputs("\n", fp);
    for (ptiNnzIndex i = 0; i < tsr->nnz; ++i) {
        for (ptiIndex m = 0; m < tsr->nmodes; ++m)
            std::fprintf(fp, "%" PRIu32 "\t", tsr->inds[m].data[i] + start_index);
        std::fprintf(fp, "%f\n", (double)tsr->values.data[i]);
    }
    return 0;
}

// Lexicographic sort with mode 0 slowest (reference: sort.c:1570-1600 with the
// compare of sort.c, equal coordinates ordered as the reference's quicksort
// leaves them).
void ptiSparseTensorSortIndex(ptiSparseTensor *tsr, int force, int tk) {
    int needsort = 0;
    for (ptiIndex m = 0; m < tsr->nmodes; ++m) {
        if (tsr->sortorder[m] != m) {
            tsr->sortorder[m] = m;
            needsort = 1;
        }
    }
    if (!(needsort || force) || tsr->nnz < 2) return;
    int nthreads = tk > 0 ? tk : pti::default_threads();
    std::vector<uint32_t> perm = pti::exact_tuple_sort(tsr, 0, nthreads);
    pti::apply_permutation(tsr, perm, nthreads);
}

}  // extern "C"
