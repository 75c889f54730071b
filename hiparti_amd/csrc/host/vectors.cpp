// Growable arrays with the reference's observable behaviour
// (src/vector/vector.c): new vectors are zero-filled with at least two
// allocated slots, append grows the capacity by half when full, resize sets
// len = cap = size.  Storage comes from malloc/realloc so callers may free()
// vector data the way reference-linked code does.
#include "internal.h"

#include <cerrno>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace {

template <typename V, typename T>
int vec_new(V *vec, uint64_t len, uint64_t cap, const char *module) {
    if (cap < len) cap = len;
    if (cap < 2) cap = 2;
    vec->len = len;
    vec->cap = cap;
    vec->data = static_cast<T *>(std::calloc(cap, sizeof(T)));
    if (!vec->data) PTI_FAIL(PTIERR_OS_ERROR + ENOMEM, module, "out of memory");
    return 0;
}

template <typename V, typename T>
int vec_append(V *vec, T value, const char *module) {
    if (vec->cap <= vec->len) {
        uint64_t newcap = vec->cap + vec->cap / 2;
        if (newcap <= vec->len) newcap = vec->len + 1;
        T *nd = static_cast<T *>(std::realloc(vec->data, newcap * sizeof(T)));
        if (!nd) PTI_FAIL(PTIERR_OS_ERROR + ENOMEM, module, "out of memory");
        vec->cap = newcap;
        vec->data = nd;
    }
    vec->data[vec->len++] = value;
    return 0;
}

template <typename V, typename T>
int vec_resize(V *vec, uint64_t size, const char *module) {
    T *nd = static_cast<T *>(std::realloc(vec->data, (size ? size : 1) * sizeof(T)));
    if (!nd) PTI_FAIL(PTIERR_OS_ERROR + ENOMEM, module, "out of memory");
    vec->len = size;
    vec->cap = size;
    vec->data = nd;
    return 0;
}

template <typename V>
void vec_free(V *vec) {
    std::free(vec->data);
    vec->data = nullptr;
    vec->len = 0;
    vec->cap = 0;
}

}  // namespace

extern "C" {

int ptiNewValueVector(ptiValueVector *v, uint64_t len, uint64_t cap) {
    return vec_new<ptiValueVector, ptiValue>(v, len, cap, "ValVec New");
}
int ptiConstantValueVector(ptiValueVector *const v, ptiValue const val) {
    for (ptiNnzIndex i = 0; i < v->len; ++i) v->data[i] = val;
    return 0;
}
int ptiAppendValueVector(ptiValueVector *v, ptiValue const value) {
    return vec_append<ptiValueVector, ptiValue>(v, value, "ValVec Append");
}
int ptiResizeValueVector(ptiValueVector *v, ptiNnzIndex const size) {
    return vec_resize<ptiValueVector, ptiValue>(v, size, "ValVec Resize");
}
void ptiFreeValueVector(ptiValueVector *v) { vec_free(v); }

int ptiNewIndexVector(ptiIndexVector *v, uint64_t len, uint64_t cap) {
    return vec_new<ptiIndexVector, ptiIndex>(v, len, cap, "IdxVec New");
}
int ptiAppendIndexVector(ptiIndexVector *v, ptiIndex const value) {
    return vec_append<ptiIndexVector, ptiIndex>(v, value, "IdxVec Append");
}
int ptiResizeIndexVector(ptiIndexVector *v, ptiNnzIndex const size) {
    return vec_resize<ptiIndexVector, ptiIndex>(v, size, "IdxVec Resize");
}
void ptiFreeIndexVector(ptiIndexVector *v) { vec_free(v); }

int ptiNewElementIndexVector(ptiElementIndexVector *v, uint64_t len, uint64_t cap) {
    return vec_new<ptiElementIndexVector, ptiElementIndex>(v, len, cap, "EleIdxVec New");
}
int ptiAppendElementIndexVector(ptiElementIndexVector *v, ptiElementIndex const value) {
    return vec_append<ptiElementIndexVector, ptiElementIndex>(v, value, "EleIdxVec Append");
}
int ptiResizeElementIndexVector(ptiElementIndexVector *v, ptiNnzIndex const size) {
    return vec_resize<ptiElementIndexVector, ptiElementIndex>(v, size, "EleIdxVec Resize");
}
void ptiFreeElementIndexVector(ptiElementIndexVector *v) { vec_free(v); }

int ptiNewBlockIndexVector(ptiBlockIndexVector *v, uint64_t len, uint64_t cap) {
    return vec_new<ptiBlockIndexVector, ptiBlockIndex>(v, len, cap, "BlkIdxVec New");
}
int ptiAppendBlockIndexVector(ptiBlockIndexVector *v, ptiBlockIndex const value) {
    return vec_append<ptiBlockIndexVector, ptiBlockIndex>(v, value, "BlkIdxVec Append");
}
int ptiResizeBlockIndexVector(ptiBlockIndexVector *v, ptiNnzIndex const size) {
    return vec_resize<ptiBlockIndexVector, ptiBlockIndex>(v, size, "BlkIdxVec Resize");
}
void ptiFreeBlockIndexVector(ptiBlockIndexVector *v) { vec_free(v); }

int ptiNewNnzIndexVector(ptiNnzIndexVector *v, uint64_t len, uint64_t cap) {
    return vec_new<ptiNnzIndexVector, ptiNnzIndex>(v, len, cap, "NnzIdxVec New");
}
int ptiAppendNnzIndexVector(ptiNnzIndexVector *v, ptiNnzIndex const value) {
    return vec_append<ptiNnzIndexVector, ptiNnzIndex>(v, value, "NnzIdxVec Append");
}
int ptiResizeNnzIndexVector(ptiNnzIndexVector *v, ptiNnzIndex const size) {
    return vec_resize<ptiNnzIndexVector, ptiNnzIndex>(v, size, "NnzIdxVec Resize");
}
void ptiFreeNnzIndexVector(ptiNnzIndexVector *v) { vec_free(v); }

// Text dumps (src/vector/dump.c:31-175): a "<type> length: n" line, then
// the entries, each followed by a tab, then a newline.
}  // extern "C"

namespace {
template <typename T>
int dump_entries(const char *title, const T *data, uint64_t n, const char *fmt, const char *module, FILE *fp) {
    if (std::fprintf(fp, "%s length: %" PRIu64 "\n", title, n) < 0)
        PTI_FAIL(PTIERR_OS_ERROR + errno, module, std::strerror(errno));
    for (uint64_t i = 0; i < n; ++i)
        if (std::fprintf(fp, fmt, data[i]) < 0) PTI_FAIL(PTIERR_OS_ERROR + errno, module, std::strerror(errno));
    std::fprintf(fp, "\n");
    return 0;
}
}  // namespace

extern "C" {

int ptiDumpElementIndexVector(ptiElementIndexVector *vec, FILE *fp) {
    return dump_entries("ptiElementIndexVector", vec->data, vec->len, "%" PRIu8 "\t", "EleIdxVec Dump", fp);
}
int ptiDumpIndexVector(ptiIndexVector *vec, FILE *fp) {
    return dump_entries("ptiIndexVector", vec->data, vec->len, "%" PRIu32 "\t", "IdxVec Dump", fp);
}
int ptiDumpBlockIndexVector(ptiBlockIndexVector *vec, FILE *fp) {
    return dump_entries("ptiBlockIndexVector", vec->data, vec->len, "%" PRIu32 "\t", "BlkIdxVec Dump", fp);
}
int ptiDumpNnzIndexVector(ptiNnzIndexVector *vec, FILE *fp) {
    return dump_entries("ptiNnzIndexVector", vec->data, vec->len, "%" PRIu64 "\t", "NnzIdxVec Dump", fp);
}
int ptiDumpValueVector(ptiValueVector *vec, FILE *fp) {
    const ptiValue *d = vec->data;
    if (std::fprintf(fp, "ptiValueVector length: %" PRIu64 "\n", vec->len) < 0)
        PTI_FAIL(PTIERR_OS_ERROR + errno, "ValVec Dump", std::strerror(errno));
    for (uint64_t i = 0; i < vec->len; ++i) std::fprintf(fp, "%f\t", (double)d[i]);
    std::fprintf(fp, "\n");
    return 0;
}
int ptiDumpIndexArray(ptiIndex const *array, ptiNnzIndex const n, FILE *fp) {
    return dump_entries("ptiIndex array", array, n, "%" PRIu32 "\t", "IdxArray Dump", fp);
}
int ptiDumpNnzIndexArray(ptiNnzIndex const *array, ptiNnzIndex const n, FILE *fp) {
    return dump_entries("ptiNnzIndex array", array, n, "%" PRIu64 "\t", "NnzIdxArray Dump", fp);
}

ptiIndex ptiMaxIndexArray(ptiIndex const *const indices, ptiNnzIndex const size) {
    ptiIndex m = indices[0];
    for (ptiNnzIndex i = 1; i < size; ++i)
        if (indices[i] > m) m = indices[i];
    return m;
}

}  // extern "C"
