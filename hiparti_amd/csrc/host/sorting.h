// Host sorting engines used by COO sorting and COO->HiCOO conversion.
//
// Two engines, both operating on (key, id) pairs so the N+1 coordinate and
// value arrays are permuted once at the end instead of on every swap:
//
//  * radix_sort_pairs: parallel LSD radix sort (8-bit digits).  Used when the
//    sort keys are pairwise distinct; then any correct sort yields the same
//    order, so the result equals the reference's quicksort output.
//
//  * hoare_sort_pairs: the reference's in-place quicksort scheme, restated on
//    keys (src/sptensor/sort.c:1434-1465 and its siblings): pivot position
//    p = (l+r)/2 tracked through swaps, scans i upward while key[i] < key[p]
//    and j downward while key[p] < key[j], swap, recurse on [l,i) and [i,r).
//    Its control flow depends only on key comparisons, so applied to the same
//    starting order it reproduces the reference's order exactly, including
//    the order of equal keys.  Sub-ranges are disjoint, so they are sorted in
//    parallel without changing the result (the reference uses OpenMP tasks
//    the same way).
#pragma once

#include <cstdint>
#include <cstring>
#include <future>
#include <thread>
#include <vector>

#include "internal.h"

namespace pti {

typedef unsigned __int128 u128;

// Fixed-width lexicographic tuple of up to 8 x 32-bit words; used when a key
// does not fit into 128 bits.
struct Tuple8 {
    uint32_t w[8];
    uint32_t n;
    bool operator<(const Tuple8 &o) const {
        for (uint32_t k = 0; k < n; ++k) {
            if (w[k] < o.w[k]) return true;
            if (w[k] > o.w[k]) return false;
        }
        return false;
    }
    bool operator==(const Tuple8 &o) const {
        for (uint32_t k = 0; k < n; ++k)
            if (w[k] != o.w[k]) return false;
        return true;
    }
    bool operator!=(const Tuple8 &o) const { return !(*this == o); }
};

template <typename K>
void hoare_sort_rec(K *key, uint32_t *id, uint64_t l, uint64_t r, int par_depth) {
    while (r - l >= 2) {
        uint64_t p = (l + r) / 2;
        uint64_t i = l, j = r - 1;
        for (;; ++i, --j) {
            while (key[i] < key[p]) ++i;
            while (key[p] < key[j]) --j;
            if (i >= j) break;
            K tk = key[i];
            key[i] = key[j];
            key[j] = tk;
            uint32_t ti = id[i];
            id[i] = id[j];
            id[j] = ti;
            if (i == p)
                p = j;
            else if (j == p)
                p = i;
        }
        // Left part [l, i), right part [i, r).
        if (par_depth > 0 && (i - l) > (1u << 16) && (r - i) > (1u << 16)) {
            auto fut = std::async(std::launch::async, hoare_sort_rec<K>, key, id, l, i, par_depth - 1);
            hoare_sort_rec<K>(key, id, i, r, par_depth - 1);
            fut.get();
            return;
        }
        // Recurse into the smaller side, loop on the larger one.
        if (i - l < r - i) {
            hoare_sort_rec<K>(key, id, l, i, 0);
            l = i;
        } else {
            hoare_sort_rec<K>(key, id, i, r, 0);
            r = i;
        }
    }
}

template <typename K>
void hoare_sort_pairs(K *key, uint32_t *id, uint64_t l, uint64_t r, int nthreads) {
    int depth = 0;
    while ((1 << depth) < nthreads * 2 && depth < 12) ++depth;
    hoare_sort_rec<K>(key, id, l, r, nthreads > 1 ? depth : 0);
}

template <typename K>
inline uint32_t radix_digit(const K &k, int shift) {
    return (uint32_t)(k >> shift) & 0xFFu;
}

// LSD radix sort of (key, id) pairs on the low `key_bits` bits of the keys.
// Stable; deterministic for any thread count.
template <typename K>
void radix_sort_pairs(std::vector<K> &key, std::vector<uint32_t> &id, int key_bits, int nthreads) {
    const uint64_t n = key.size();
    if (n < 2 || key_bits <= 0) return;
    if (nthreads <= 0) nthreads = default_threads();
    if ((uint64_t)nthreads * 4096 > n) nthreads = (int)std::max<uint64_t>(1, n / 4096);
    std::vector<K> key2(n);
    std::vector<uint32_t> id2(n);
    std::vector<uint64_t> hist((size_t)nthreads * 256);
    K *src_k = key.data(), *dst_k = key2.data();
    uint32_t *src_i = id.data(), *dst_i = id2.data();
    int passes = 0;
    for (int shift = 0; shift < key_bits; shift += 8) {
        std::fill(hist.begin(), hist.end(), 0);
        parallel_for(n, nthreads, [&](uint64_t b, uint64_t e, int t) {
            uint64_t *h = &hist[(size_t)t * 256];
            for (uint64_t x = b; x < e; ++x) ++h[radix_digit(src_k[x], shift)];
        });
        // Exclusive scan in (digit, thread) order keeps the sort stable.
        uint64_t sum = 0;
        for (int d = 0; d < 256; ++d)
            for (int t = 0; t < nthreads; ++t) {
                uint64_t c = hist[(size_t)t * 256 + d];
                hist[(size_t)t * 256 + d] = sum;
                sum += c;
            }
        parallel_for(n, nthreads, [&](uint64_t b, uint64_t e, int t) {
            uint64_t *h = &hist[(size_t)t * 256];
            for (uint64_t x = b; x < e; ++x) {
                uint64_t pos = h[radix_digit(src_k[x], shift)]++;
                dst_k[pos] = src_k[x];
                dst_i[pos] = src_i[x];
            }
        });
        std::swap(src_k, dst_k);
        std::swap(src_i, dst_i);
        ++passes;
    }
    if (passes & 1) {
        key.swap(key2);
        id.swap(id2);
    }
}

}  // namespace pti
