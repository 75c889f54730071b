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

// Text coordinate format (reference load.c:63-118): "nmodes", nmodes
// dimensions, then records of nmodes indices and one value, all separated by
// whitespace; reading stops at the first incomplete or malformed record and
// keeps the records before it.  The file is read whole and tokenised with
// strtoul / strtod, which accept what the reference's fscanf("%u") /
// fscanf("%lf") accept.
static int load_text(ptiSparseTensor *tsr, ptiIndex start_index, FILE *fp) {
    std::vector<char> buf;
    {
        char tmp[1 << 16];
        size_t n;
        while ((n = std::fread(tmp, 1, sizeof tmp, fp)) > 0) buf.insert(buf.end(), tmp, tmp + n);
        if (std::ferror(fp)) PTI_FAIL(PTIERR_OS_ERROR + errno, "SpTns Load", std::strerror(errno));
    }
    buf.push_back('\0');
    const char *p = buf.data();
    auto read_u = [&](unsigned long &v) {
        char *e;
        errno = 0;
        v = std::strtoul(p, &e, 10);
        if (e == p) return false;
        p = e;
        return true;
    };
    unsigned long nm = 0;
    if (!read_u(nm) || nm == 0 || nm > 64) PTI_FAIL(PTIERR_VALUE_ERROR, "SpTns Load", "bad header: nmodes");
    const unsigned nmodes = (unsigned)nm;
    std::vector<ptiIndex> ndims(nmodes);
    for (unsigned m = 0; m < nmodes; ++m) {
        unsigned long d;
        if (!read_u(d)) PTI_FAIL(PTIERR_VALUE_ERROR, "SpTns Load", "bad header: ndims");
        ndims[m] = (ptiIndex)d;
    }
    PTI_CHECK(ptiNewSparseTensor(tsr, nmodes, ndims.data()), "SpTns Load");
    // The reference leaves sortorder zero-filled after loading (load.c:70-72).
    std::memset(tsr->sortorder, 0, nmodes * sizeof(ptiIndex));
    std::vector<std::vector<ptiIndex>> inds(nmodes);
    std::vector<ptiValue> vals;
    for (;;) {
        ptiIndex idx[64];
        unsigned m = 0;
        for (; m < nmodes; ++m) {
            unsigned long v;
            if (!read_u(v)) break;
            idx[m] = (ptiIndex)v;
        }
        if (m < nmodes) break;
        char *e;
        const double value = std::strtod(p, &e);
        if (e == p) break;
        p = e;
        for (m = 0; m < nmodes; ++m) {
            if (idx[m] < start_index) PTI_FAIL(PTIERR_VALUE_ERROR, "SpTns Load", "index < start_index");
            inds[m].push_back(idx[m] - start_index);
        }
        vals.push_back((ptiValue)value);
    }
    const ptiNnzIndex nnz = vals.size();
    for (unsigned m = 0; m < nmodes; ++m) {
        PTI_CHECK(ptiResizeIndexVector(&tsr->inds[m], nnz), "SpTns Load");
        if (nnz) std::memcpy(tsr->inds[m].data, inds[m].data(), nnz * sizeof(ptiIndex));
    }
    PTI_CHECK(ptiResizeValueVector(&tsr->values, nnz), "SpTns Load");
    if (nnz) std::memcpy(tsr->values.data, vals.data(), nnz * sizeof(ptiValue));
    tsr->nnz = nnz;
    return 0;
}

// SPLATT binary coordinate format (reference load.c:121-250): header
// int32 magic, u64 idx_width, u64 val_width (read field by field, 20 bytes);
// then nmodes and the dimensions as 32-bit integers, nnz as a 64-bit integer
// when idx_width is 8 and a 32-bit one otherwise, the indices mode by mode
// (32-bit), and the values (float when val_width is 4, else double).  Index
// widths above 32 bits are rejected, as the reference does (it exits).
static int load_binary(ptiSparseTensor *tsr, FILE *fp) {
    static const char *M = "SpTns Load";
    int32_t magic;
    uint64_t idx_width, val_width;
    if (std::fread(&magic, sizeof magic, 1, fp) != 1 || std::fread(&idx_width, sizeof idx_width, 1, fp) != 1 ||
        std::fread(&val_width, sizeof val_width, 1, fp) != 1)
        PTI_FAIL(PTIERR_VALUE_ERROR, M, "truncated binary header");
    if (idx_width > sizeof(ptiIndex)) PTI_FAIL(PTIERR_VALUE_ERROR, M, "binary tensor has indices wider than 32 bits");
    uint32_t nmodes = 0;
    if (std::fread(&nmodes, sizeof nmodes, 1, fp) != 1 || nmodes == 0 || nmodes > 64)
        PTI_FAIL(PTIERR_VALUE_ERROR, M, "bad binary header: nmodes");
    std::vector<ptiIndex> dims(nmodes);
    if (std::fread(dims.data(), sizeof(ptiIndex), nmodes, fp) != nmodes)
        PTI_FAIL(PTIERR_VALUE_ERROR, M, "bad binary header: ndims");
    uint64_t nnz = 0;
    if (idx_width == sizeof(ptiNnzIndex)) {
        if (std::fread(&nnz, sizeof nnz, 1, fp) != 1) PTI_FAIL(PTIERR_VALUE_ERROR, M, "bad binary header: nnz");
    } else {
        uint32_t n32;
        if (std::fread(&n32, sizeof n32, 1, fp) != 1) PTI_FAIL(PTIERR_VALUE_ERROR, M, "bad binary header: nnz");
        nnz = n32;
    }
    PTI_CHECK(ptiNewSparseTensor(tsr, nmodes, dims.data()), M);
    std::memset(tsr->sortorder, 0, nmodes * sizeof(ptiIndex));
    for (uint32_t m = 0; m < nmodes; ++m) {
        PTI_CHECK(ptiResizeIndexVector(&tsr->inds[m], nnz), M);
        if (std::fread(tsr->inds[m].data, sizeof(ptiIndex), nnz, fp) != nnz)
            PTI_FAIL(PTIERR_VALUE_ERROR, M, "truncated binary indices");
    }
    PTI_CHECK(ptiResizeValueVector(&tsr->values, nnz), M);
    if (val_width == sizeof(ptiValue)) {
        if (std::fread(tsr->values.data, sizeof(ptiValue), nnz, fp) != nnz)
            PTI_FAIL(PTIERR_VALUE_ERROR, M, "truncated binary values");
    } else {
        std::vector<double> d(nnz);
        if (std::fread(d.data(), sizeof(double), nnz, fp) != nnz)
            PTI_FAIL(PTIERR_VALUE_ERROR, M, "truncated binary values");
        for (uint64_t i = 0; i < nnz; ++i) tsr->values.data[i] = (ptiValue)d[i];
    }
    tsr->nnz = nnz;
    return 0;
}

// File type by extension (reference load.c:25-60): .tns / .coo text, .bin
// binary; anything else is read as text.
int ptiLoadSparseTensor(ptiSparseTensor *tsr, ptiIndex start_index, char const *const fname) {
    FILE *fp = std::fopen(fname, "rb");
    if (!fp) PTI_FAIL(PTIERR_OS_ERROR + errno, "SpTns Load", std::strerror(errno));
    const char *dot = std::strrchr(fname, '.');
    const bool binary = dot && std::strcmp(dot, ".bin") == 0;
    if (!binary && !(dot && (std::strcmp(dot, ".tns") == 0 || std::strcmp(dot, ".coo") == 0)))
        std::fprintf(stderr, "SPLATT: extension for '%s' not recognized. Defaulting to ASCII coordinate form.\n",
                     fname);
    const int rc = binary ? load_binary(tsr, fp) : load_text(tsr, start_index, fp);
    std::fclose(fp);
    return rc;
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
