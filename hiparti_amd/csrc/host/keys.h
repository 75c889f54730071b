// Order-preserving packing of index tuples into sort keys.
#pragma once

#include "sorting.h"

namespace pti {

inline int bit_width_u32(uint32_t v) {
    int b = 0;
    while (v) {
        ++b;
        v >>= 1;
    }
    return b;
}

// Maximum of each mode's indices (0 for an empty tensor).
inline std::vector<uint32_t> mode_maxima(const ptiSparseTensor *tsr, int nthreads) {
    const ptiIndex N = tsr->nmodes;
    std::vector<uint32_t> mx(N, 0);
    // parallel_for resolves nthreads <= 0 to default_threads(); size the
    // per-thread slots for the count it will actually run.
    if (nthreads <= 0) nthreads = default_threads();
    for (ptiIndex m = 0; m < N; ++m) {
        const uint32_t *d = tsr->inds[m].data;
        std::vector<uint32_t> part(nthreads, 0);
        parallel_for(tsr->nnz, nthreads, [&](uint64_t b, uint64_t e, int t) {
            uint32_t v = 0;
            for (uint64_t z = b; z < e; ++z) v = d[z] > v ? d[z] : v;
            part[t] = v;
        });
        for (uint32_t v : part) mx[m] = std::max(mx[m], v);
    }
    return mx;
}

// Build keys key[z] = concat over m (mode 0 most significant) of
// (inds[m][z] >> shift) using widths[m] bits each.  Returns false when the
// total width exceeds the key type.
template <typename K>
bool pack_shifted_tuple(const ptiSparseTensor *tsr, const std::vector<int> &widths, int shift,
                        std::vector<K> &key, int nthreads) {
    int total = 0;
    for (int w : widths) total += w;
    if (total > (int)(8 * sizeof(K))) return false;
    const ptiIndex N = tsr->nmodes;
    key.resize(tsr->nnz);
    parallel_for(tsr->nnz, nthreads, [&](uint64_t b, uint64_t e, int) {
        for (uint64_t z = b; z < e; ++z) {
            K k = 0;
            for (ptiIndex m = 0; m < N; ++m) {
                k = (widths[m] ? (k << widths[m]) : k) | (K)(tsr->inds[m].data[z] >> shift);
            }
            key[z] = k;
        }
    });
    return true;
}

inline void pack_tuple8(const ptiSparseTensor *tsr, int shift, std::vector<Tuple8> &key, int nthreads) {
    const ptiIndex N = tsr->nmodes;
    key.resize(tsr->nnz);
    parallel_for(tsr->nnz, nthreads, [&](uint64_t b, uint64_t e, int) {
        for (uint64_t z = b; z < e; ++z) {
            Tuple8 t;
            std::memset(&t, 0, sizeof(t));
            t.n = N;
            for (ptiIndex m = 0; m < N; ++m) t.w[m] = tsr->inds[m].data[z] >> shift;
            key[z] = t;
        }
    });
}

// Apply a gather permutation to all coordinate arrays and values:
// new[x] = old[perm[x]].
inline void apply_permutation(ptiSparseTensor *tsr, const std::vector<uint32_t> &perm, int nthreads) {
    const uint64_t n = tsr->nnz;
    if (n < 2) return;
    std::vector<uint32_t> tmp(n);
    for (ptiIndex m = 0; m < tsr->nmodes; ++m) {
        uint32_t *d = tsr->inds[m].data;
        parallel_for(n, nthreads, [&](uint64_t b, uint64_t e, int) {
            for (uint64_t x = b; x < e; ++x) tmp[x] = d[perm[x]];
        });
        std::memcpy(d, tmp.data(), n * sizeof(uint32_t));
    }
    float *v = tsr->values.data;
    std::vector<float> tv(n);
    parallel_for(n, nthreads, [&](uint64_t b, uint64_t e, int) {
        for (uint64_t x = b; x < e; ++x) tv[x] = v[perm[x]];
    });
    std::memcpy(v, tv.data(), n * sizeof(float));
}

// Sort positions [0, nnz) of a COO tensor by the lexicographic order of the
// tuples (inds[m] >> shift), reproducing the reference quicksort's order of
// equal tuples.  Returns the gather permutation.
inline std::vector<uint32_t> exact_tuple_sort(const ptiSparseTensor *tsr, int shift, int nthreads) {
    const uint64_t n = tsr->nnz;
    std::vector<uint32_t> id(n);
    for (uint64_t z = 0; z < n; ++z) id[z] = (uint32_t)z;
    if (n < 2) return id;
    std::vector<uint32_t> mx = mode_maxima(tsr, nthreads);
    std::vector<int> widths(tsr->nmodes);
    for (ptiIndex m = 0; m < tsr->nmodes; ++m) widths[m] = bit_width_u32(mx[m] >> shift);
    std::vector<uint64_t> k64;
    if (pack_shifted_tuple<uint64_t>(tsr, widths, shift, k64, nthreads)) {
        hoare_sort_pairs<uint64_t>(k64.data(), id.data(), 0, n, nthreads);
        return id;
    }
    std::vector<u128> k128;
    if (pack_shifted_tuple<u128>(tsr, widths, shift, k128, nthreads)) {
        hoare_sort_pairs<u128>(k128.data(), id.data(), 0, n, nthreads);
        return id;
    }
    std::vector<Tuple8> kt;
    pack_tuple8(tsr, shift, kt, nthreads);
    hoare_sort_pairs<Tuple8>(kt.data(), id.data(), 0, n, nthreads);
    return id;
}

}  // namespace pti
