// Growable arrays with the reference's observable behaviour
// (src/vector/vector.c): new vectors are zero-filled with at least two
// allocated slots, append grows the capacity by half when full, resize sets
// len = cap = size.  Storage comes from malloc/realloc so callers may free()
// vector data the way reference-linked code does.
#include "internal.h"

#include <cstdlib>

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

ptiIndex ptiMaxIndexArray(ptiIndex const *const indices, ptiNnzIndex const size) {
    ptiIndex m = indices[0];
    for (ptiNnzIndex i = 1; i < size; ++i)
        if (indices[i] > m) m = indices[i];
    return m;
}

}  // extern "C"
