// HiCOO containers and the COO -> HiCOO conversion.
//
// The conversion produces the reference's arrays bit for bit
// (src/sptensor/hicoo/convert.c:430-860 with sort.c:397-501, 929-1288,
// 1331-1465):
//   1. nonzeros are ordered by superblock tuple (idx >> sk_bits, mode 0 most
//      significant), then, for 3-D and 4-D tensors, by Morton order inside each
//      superblock (bit b of mode 0 above bit b of mode 1 above ..., 3-D keys
//      only see the low 24 bits of each index);
//   2. superblock pointers kptr/knnzs, the per-mode superblock lists kschr and
//      nkiters, and the balanced schedule are derived from that order;
//   3. blocks start wherever the tuple (idx >> sb_bits) changes between
//      consecutive nonzeros; einds = idx & (2^sb_bits - 1); chunk pointers
//      follow the reference's nnz-count rule; kptr is rewritten to block
//      offsets.
// Where the order is unique (3-D/4-D, distinct coordinates, 3-D indices below
// 2^24) a parallel radix sort on a composite (superblock, Morton) key gives it
// directly.  Otherwise the reference's quicksort is replayed on keys
// (sorting.h), which reproduces its order of equal keys.
//
// Unlike the reference, the caller's COO is permuted once at the end of the
// sort instead of swap by swap; the final COO order is the same.
#include "internal.h"
#include "keys.h"
#include "hicoo_build.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cerrno>
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace {

using pti::CompositeLayout;
using pti::u128;

int new_vec_array(ptiIndexVector **out, ptiIndex n) {
    *out = static_cast<ptiIndexVector *>(std::malloc((n ? n : 1) * sizeof(ptiIndexVector)));
    if (!*out) return PTIERR_OS_ERROR + ENOMEM;
    for (ptiIndex i = 0; i < n; ++i) {
        int rc = ptiNewIndexVector(&(*out)[i], 0, 0);
        if (rc) return rc;
    }
    return 0;
}

// Spread the 8 bits of b to positions 0, stride, 2*stride, ...
inline uint64_t spread_bits(uint32_t b, int nbits, int stride) {
    uint64_t r = 0;
    for (int i = 0; i < nbits; ++i) r |= (uint64_t)((b >> i) & 1u) << (i * stride);
    return r;
}

struct MortonTables {
    uint32_t s3[256];  // 8 bits spread with stride 3
    MortonTables() {
        for (uint32_t b = 0; b < 256; ++b) s3[b] = (uint32_t)spread_bits(b, 8, 3);
    }
};
const MortonTables &morton_tables() {
    static MortonTables t;
    return t;
}

// Order-equivalent form of the reference's 3-D Morton key: interleave of the
// low 24 bits of (x, y, z) with x the most significant bit of each triple.
inline u128 morton3_key(uint32_t x, uint32_t y, uint32_t z) {
    const MortonTables &t = morton_tables();
    u128 k = 0;
    for (int byte = 2; byte >= 0; --byte) {
        uint32_t bx = (x >> (8 * byte)) & 0xFF, by = (y >> (8 * byte)) & 0xFF, bz = (z >> (8 * byte)) & 0xFF;
        uint32_t part = (t.s3[bx] << 2) | (t.s3[by] << 1) | t.s3[bz];
        k = (k << 24) | part;
    }
    return k;
}

// The reference's 4-D Morton key: 32-bit interleave, x most significant.
inline u128 morton4_key(uint32_t x, uint32_t y, uint32_t z, uint32_t w) {
    u128 k = 0;
    for (int i = 31; i >= 0; --i) {
        k = (k << 4) | (u128)((((x >> i) & 1u) << 3) | (((y >> i) & 1u) << 2) | (((z >> i) & 1u) << 1) |
                              ((w >> i) & 1u));
    }
    return k;
}

template <typename K>
void build_composite_keys(const ptiSparseTensor *tsr, const CompositeLayout &L, std::vector<K> &key,
                          int nthreads) {
    const int N = L.N;
    key.resize(tsr->nnz);
    pti::parallel_for(tsr->nnz, nthreads, [&](uint64_t b, uint64_t e, int) {
        uint32_t c[8];
        for (uint64_t z = b; z < e; ++z) {
            for (int m = 0; m < N; ++m) c[m] = tsr->inds[m].data[z];
            K k = 0;
            for (int m = 0; m < N; ++m)
                if (L.hb[m]) k = (k << L.hb[m]) | (K)(c[m] >> L.sk);
            for (int bit = L.max_lb - 1; bit >= 0; --bit)
                for (int m = 0; m < N; ++m)
                    if (bit < L.lb[m]) k = (k << 1) | (K)((c[m] >> bit) & 1u);
            key[z] = k;
        }
    });
}

template <typename K>
bool try_unique_order(const ptiSparseTensor *tsr, const CompositeLayout &L, std::vector<uint32_t> &perm,
                      int nthreads) {
    std::vector<K> key;
    build_composite_keys<K>(tsr, L, key, nthreads);
    perm.resize(tsr->nnz);
    for (uint64_t z = 0; z < tsr->nnz; ++z) perm[z] = (uint32_t)z;
    pti::radix_sort_pairs<K>(key, perm, L.total_bits, nthreads);
    std::atomic<bool> dup(false);
    pti::parallel_for(tsr->nnz > 0 ? tsr->nnz - 1 : 0, nthreads, [&](uint64_t b, uint64_t e, int) {
        for (uint64_t z = b; z < e; ++z)
            if (key[z] == key[z + 1]) {
                dup = true;
                return;
            }
    });
    return !dup.load();
}

// Superblock boundaries of a tensor already ordered by superblock tuple
// (reference: convert.c:314-369).  kptr gets nk+1 nnz offsets, knnzs nk sizes.
int set_kernel_pointers(ptiNnzIndexVector *kptr, ptiNnzIndexVector *knnzs, const ptiSparseTensor *tsr,
                        int sk) {
    const ptiIndex N = tsr->nmodes;
    const uint64_t nnz = tsr->nnz;
    PTI_CHECK(ptiAppendNnzIndexVector(kptr, 0), "HiSpTns Convert");
    // An empty tensor behaves as one superblock at coordinate 0.
    uint64_t knnz = 0;
    for (uint64_t z = 0; z < nnz; ++z) {
        bool same = true;
        if (z > 0)
            for (ptiIndex m = 0; m < N; ++m)
                if ((tsr->inds[m].data[z] >> sk) != (tsr->inds[m].data[z - 1] >> sk)) {
                    same = false;
                    break;
                }
        if (same) {
            ++knnz;
        } else {
            PTI_CHECK(ptiAppendNnzIndexVector(kptr, knnz + kptr->data[kptr->len - 1]), "HiSpTns Convert");
            PTI_CHECK(ptiAppendNnzIndexVector(knnzs, knnz), "HiSpTns Convert");
            knnz = 1;
        }
    }
    PTI_CHECK(ptiAppendNnzIndexVector(kptr, nnz), "HiSpTns Convert");
    PTI_CHECK(ptiAppendNnzIndexVector(knnzs, knnz), "HiSpTns Convert");
    return 0;
}

// Exact Morton ordering inside each superblock for tensors whose order is not
// unique; superblocks are independent and are processed in parallel.
void exact_morton_within_superblocks(ptiSparseTensor *tsr, const ptiNnzIndexVector &kptr, int nthreads) {
    const ptiIndex N = tsr->nmodes;
    const uint64_t nnz = tsr->nnz;
    std::vector<u128> key(nnz);
    std::vector<uint32_t> id(nnz);
    pti::parallel_for(nnz, nthreads, [&](uint64_t b, uint64_t e, int) {
        for (uint64_t z = b; z < e; ++z) {
            const uint32_t x = tsr->inds[0].data[z], y = tsr->inds[1].data[z], zz = tsr->inds[2].data[z];
            key[z] = N == 3 ? morton3_key(x, y, zz) : morton4_key(x, y, zz, tsr->inds[3].data[z]);
            id[z] = (uint32_t)z;
        }
    });
    const uint64_t nk = kptr.len - 1;
    std::atomic<uint64_t> next(0);
    int nt = nthreads > 0 ? nthreads : pti::default_threads();
    std::vector<std::thread> ts;
    for (int t = 0; t < nt; ++t)
        ts.emplace_back([&] {
            for (;;) {
                uint64_t k = next.fetch_add(1);
                if (k >= nk) break;
                uint64_t b = kptr.data[k], e = kptr.data[k + 1];
                if (e - b >= 2) pti::hoare_sort_pairs<u128>(key.data(), id.data(), b, e, 1);
            }
        });
    for (auto &th : ts) th.join();
    pti::apply_permutation(tsr, id, nthreads);
}

}  // namespace

namespace pti {

// Composite (superblock tuple, compressed Morton) key for the unique-order
// fast path.  Bits of mode m below lb[m] take part in the Morton interleave;
// bits from sk upward form the superblock tuple.  Returns false if the key
// needs more than 128 bits.
bool composite_layout(const std::vector<uint32_t> &mx, int sk, CompositeLayout &L) {
    L.N = (int)mx.size();
    L.sk = sk;
    L.lb.resize(L.N);
    L.hb.resize(L.N);
    L.morton_bits = 0;
    L.max_lb = 0;
    int hbits = 0;
    for (int m = 0; m < L.N; ++m) {
        int b = pti::bit_width_u32(mx[m]);
        L.lb[m] = std::min(sk, b);
        L.hb[m] = std::max(0, b - sk);
        L.morton_bits += L.lb[m];
        hbits += L.hb[m];
        L.max_lb = std::max(L.max_lb, L.lb[m]);
    }
    L.total_bits = L.morton_bits + hbits;
    return L.total_bits <= 128;
}

inline uint32_t coord_or_zero(const ptiSparseTensor *tsr, ptiIndex m, uint64_t z) {
    return z < tsr->nnz ? tsr->inds[m].data[z] : 0u;
}

// Per-mode superblock lists (reference: convert.c:378-421).
int set_kernel_scheduler(ptiSparseTensorHiCOO *h, const ptiSparseTensor *tsr, int sk) {
    const ptiIndex N = tsr->nmodes;
    const uint64_t nk = h->kptr.len - 1;
    for (uint64_t k = 0; k < nk; ++k) {
        uint64_t z = h->kptr.data[k];
        for (ptiIndex m = 0; m < N; ++m)
            PTI_CHECK(ptiAppendIndexVector(&h->kschr[m][coord_or_zero(tsr, m, z) >> sk], (ptiIndex)k),
                      "HiSpTns Convert");
    }
    const uint64_t skv = 1ull << sk;
    for (ptiIndex m = 0; m < N; ++m) {
        ptiIndex kernel_ndim = (ptiIndex)((h->ndims[m] + skv - 1) / skv);
        uint64_t mx = 0;
        for (ptiIndex i = 0; i < kernel_ndim; ++i) mx = std::max<uint64_t>(mx, h->kschr[m][i].len);
        h->nkiters[m] = (ptiIndex)mx;
    }
    return 0;
}

// Balanced superblock partitions per mode (reference: convert.c:474-571),
// restated with the reference's integer widths: the partition threshold is a
// 32-bit value, per-row sums are 64-bit.
int set_balanced_schedule(ptiSparseTensorHiCOO *h, int sk) {
    const ptiIndex N = h->nmodes;
    const uint64_t nk = h->kptr.len - 1;
    uint64_t max_nnzk = 0;
    for (uint64_t k = 0; k < nk; ++k) max_nnzk = std::max(max_nnzk, h->knnzs.data[k]);
    const uint64_t par_nnzk_th = 5 * max_nnzk;
    const uint64_t skv = 1ull << sk;
    for (ptiIndex m = 0; m < N; ++m) {
        ptiIndexVector *rows = h->kschr[m];
        ptiIndexVector *bal = h->kschr_balanced[m];
        ptiIndexVector *pos = h->kschr_balanced_pos[m];
        const ptiIndex kernel_ndim = (ptiIndex)((h->ndims[m] + skv - 1) / skv);
        for (ptiIndex i = 0; i < kernel_ndim; ++i) PTI_CHECK(ptiAppendIndexVector(&pos[i], 0), "HiSpTns Convert");
        const ptiIndex empty_rows_th = kernel_ndim > 1 ? kernel_ndim : 1;
        ptiIndex npars = 0;
        bool rest = false;
        uint64_t count_nk = 0;
        while (!rest && count_nk < nk) {
            uint32_t max_col = 0;
            ptiIndex empty_rows = 0;
            for (ptiIndex i = 0; i < kernel_ndim; ++i) {
                if (empty_rows > empty_rows_th) {
                    rest = true;
                    break;
                }
                if (npars >= pos[i].len) {
                    ++empty_rows;
                    continue;
                }
                ptiIndex j = pos[i].data[npars];
                if (j >= rows[i].len) {
                    ++empty_rows;
                    continue;
                }
                uint64_t knz = h->knnzs.data[rows[i].data[j]];
                if (max_col < knz) max_col = (uint32_t)knz;
            }
            if (rest) {
                for (ptiIndex i = 0; i < kernel_ndim; ++i) {
                    if (npars >= pos[i].len) continue;
                    for (ptiIndex j2 = pos[i].data[npars]; j2 < rows[i].len; ++j2) {
                        PTI_CHECK(ptiAppendIndexVector(&h->kschr_rest[m], rows[i].data[j2]), "HiSpTns Convert");
                        ++count_nk;
                    }
                }
            } else {
                const uint32_t par_nnzk = (uint32_t)((uint64_t)max_col > par_nnzk_th ? (uint64_t)max_col : par_nnzk_th);
                for (ptiIndex i = 0; i < kernel_ndim; ++i) {
                    if (npars >= pos[i].len) continue;
                    ptiIndex tj = pos[i].data[npars];
                    if (tj >= rows[i].len) continue;
                    ptiIndex kn = rows[i].data[tj];
                    uint64_t sum = h->knnzs.data[kn];
                    while (sum <= par_nnzk) {
                        PTI_CHECK(ptiAppendIndexVector(&bal[i], kn), "HiSpTns Convert");
                        ++count_nk;
                        ++tj;
                        if (tj < rows[i].len) {
                            kn = rows[i].data[tj];
                            sum += h->knnzs.data[kn];
                        } else {
                            break;
                        }
                    }
                    PTI_CHECK(ptiAppendIndexVector(&pos[i], tj), "HiSpTns Convert");
                }
                ++npars;
            }
        }
        h->nkpars[m] = npars;
    }
    return 0;
}

// Chunk pointers and block-offset superblock pointers, per block
// (reference: convert.c:728-835).  On entry h->bptr holds the block starts
// and h->kptr the superblock starts as nonzero offsets; on return kptr holds
// block offsets and cptr the chunk boundaries.  The reference walks the
// nonzeros; at the start of block b >= 1 its counter of the open block equals
// the size of block b-1, so the same rule is evaluated once per block: a new
// chunk starts at block b when the finished blocks of the current chunk plus
// block b-1 reach 2^sc_bits nonzeros, and every superblock starts a chunk.
int set_chunks_from_blocks(ptiSparseTensorHiCOO *h) {
    const uint64_t nb = h->bptr.len - 1;
    const uint64_t nk = h->kptr.len - 1;
    const uint64_t sc = 1ull << h->sc_bits;
    const uint64_t *bp = h->bptr.data;
    // First block of each superblock (superblock starts are block starts).
    std::vector<uint64_t> fb(nk + 1);
    for (uint64_t k = 0; k < nk; ++k) {
        const uint64_t z = h->kptr.data[k];
        fb[k] = (uint64_t)(std::lower_bound(bp, bp + nb, z) - bp);
    }
    fb[nk] = nb;
    for (uint64_t k = 0; k < nk; ++k) {
        h->kptr.data[k] = k == 0 ? 0 : fb[k];
        PTI_CHECK(ptiAppendNnzIndexVector(&h->cptr, h->kptr.data[k]), "HiSpTns Convert");
        uint64_t chunk_size = 0;
        for (uint64_t b = fb[k]; b < fb[k + 1]; ++b) {
            if (b == 0) continue;
            const uint64_t prev = bp[b] - bp[b - 1];
            if (chunk_size + prev >= sc) {
                PTI_CHECK(ptiAppendNnzIndexVector(&h->cptr, b), "HiSpTns Convert");
                chunk_size = 0;
            } else {
                chunk_size += prev;
            }
        }
    }
    h->kptr.data[nk] = nb;
    PTI_CHECK(ptiAppendNnzIndexVector(&h->cptr, nb), "HiSpTns Convert");
    return 0;
}

}  // namespace pti

extern "C" {

int ptiNewSparseTensorHiCOO(ptiSparseTensorHiCOO *hitsr, const ptiIndex nmodes, const ptiIndex ndims[],
                            const ptiNnzIndex nnz, const ptiElementIndex sb_bits, const ptiElementIndex sk_bits,
                            const ptiElementIndex sc_bits) {
    std::memset(hitsr, 0, sizeof(*hitsr));
    hitsr->nmodes = nmodes;
    hitsr->sortorder = static_cast<ptiIndex *>(std::malloc((nmodes ? nmodes : 1) * sizeof(ptiIndex)));
    hitsr->ndims = static_cast<ptiIndex *>(std::malloc((nmodes ? nmodes : 1) * sizeof(ptiIndex)));
    PTI_CHECK_ALLOC(hitsr->sortorder, "HiSpTns New");
    PTI_CHECK_ALLOC(hitsr->ndims, "HiSpTns New");
    for (ptiIndex i = 0; i < nmodes; ++i) {
        hitsr->sortorder[i] = i;
        hitsr->ndims[i] = ndims[i];
    }
    hitsr->nnz = nnz;
    hitsr->sb_bits = sb_bits;
    hitsr->sk_bits = sk_bits;
    hitsr->sc_bits = sc_bits;
    const uint64_t sk = 1ull << sk_bits;
    hitsr->kschr = static_cast<ptiIndexVector **>(std::malloc((nmodes ? nmodes : 1) * sizeof(ptiIndexVector *)));
    hitsr->kschr_balanced = static_cast<ptiIndexVector **>(std::malloc((nmodes ? nmodes : 1) * sizeof(ptiIndexVector *)));
    hitsr->kschr_balanced_pos = static_cast<ptiIndexVector **>(std::malloc((nmodes ? nmodes : 1) * sizeof(ptiIndexVector *)));
    hitsr->nkiters = static_cast<ptiIndex *>(std::calloc(nmodes ? nmodes : 1, sizeof(ptiIndex)));
    hitsr->nkpars = static_cast<ptiIndex *>(std::calloc(nmodes ? nmodes : 1, sizeof(ptiIndex)));
    PTI_CHECK_ALLOC(hitsr->kschr, "HiSpTns New");
    PTI_CHECK_ALLOC(hitsr->kschr_balanced, "HiSpTns New");
    PTI_CHECK_ALLOC(hitsr->kschr_balanced_pos, "HiSpTns New");
    PTI_CHECK_ALLOC(hitsr->nkiters, "HiSpTns New");
    PTI_CHECK_ALLOC(hitsr->nkpars, "HiSpTns New");
    for (ptiIndex m = 0; m < nmodes; ++m) {
        ptiIndex kernel_ndim = (ptiIndex)((ndims[m] + sk - 1) / sk);
        PTI_CHECK(new_vec_array(&hitsr->kschr[m], kernel_ndim), "HiSpTns New");
        PTI_CHECK(new_vec_array(&hitsr->kschr_balanced[m], kernel_ndim), "HiSpTns New");
        PTI_CHECK(new_vec_array(&hitsr->kschr_balanced_pos[m], kernel_ndim), "HiSpTns New");
    }
    PTI_CHECK(new_vec_array(&hitsr->kschr_rest, nmodes), "HiSpTns New");
    PTI_CHECK(ptiNewNnzIndexVector(&hitsr->kptr, 0, 0), "HiSpTns New");
    PTI_CHECK(ptiNewNnzIndexVector(&hitsr->cptr, 0, 0), "HiSpTns New");
    PTI_CHECK(ptiNewNnzIndexVector(&hitsr->knnzs, 0, 0), "HiSpTns New");
    PTI_CHECK(ptiNewNnzIndexVector(&hitsr->bptr, 0, 0), "HiSpTns New");
    hitsr->binds = static_cast<ptiBlockIndexVector *>(std::malloc((nmodes ? nmodes : 1) * sizeof(ptiBlockIndexVector)));
    hitsr->einds = static_cast<ptiElementIndexVector *>(std::malloc((nmodes ? nmodes : 1) * sizeof(ptiElementIndexVector)));
    PTI_CHECK_ALLOC(hitsr->binds, "HiSpTns New");
    PTI_CHECK_ALLOC(hitsr->einds, "HiSpTns New");
    for (ptiIndex m = 0; m < nmodes; ++m) {
        PTI_CHECK(ptiNewBlockIndexVector(&hitsr->binds[m], 0, 0), "HiSpTns New");
        PTI_CHECK(ptiNewElementIndexVector(&hitsr->einds[m], 0, 0), "HiSpTns New");
    }
    PTI_CHECK(ptiNewValueVector(&hitsr->values, 0, 0), "HiSpTns New");
    return 0;
}

void ptiFreeSparseTensorHiCOO(ptiSparseTensorHiCOO *hitsr) {
    const ptiIndex N = hitsr->nmodes;
    const uint64_t sk = 1ull << hitsr->sk_bits;
    for (ptiIndex m = 0; m < N; ++m) {
        ptiIndex kernel_ndim = (ptiIndex)((hitsr->ndims[m] + sk - 1) / sk);
        for (ptiIndex i = 0; i < kernel_ndim; ++i) {
            if (hitsr->kschr) ptiFreeIndexVector(&hitsr->kschr[m][i]);
            if (hitsr->kschr_balanced) ptiFreeIndexVector(&hitsr->kschr_balanced[m][i]);
            if (hitsr->kschr_balanced_pos) ptiFreeIndexVector(&hitsr->kschr_balanced_pos[m][i]);
        }
        if (hitsr->kschr) std::free(hitsr->kschr[m]);
        if (hitsr->kschr_balanced) std::free(hitsr->kschr_balanced[m]);
        if (hitsr->kschr_balanced_pos) std::free(hitsr->kschr_balanced_pos[m]);
        if (hitsr->kschr_rest) ptiFreeIndexVector(&hitsr->kschr_rest[m]);
        if (hitsr->binds) ptiFreeBlockIndexVector(&hitsr->binds[m]);
        if (hitsr->einds) ptiFreeElementIndexVector(&hitsr->einds[m]);
    }
    std::free(hitsr->kschr);
    std::free(hitsr->kschr_balanced);
    std::free(hitsr->kschr_balanced_pos);
    std::free(hitsr->kschr_rest);
    std::free(hitsr->nkiters);
    std::free(hitsr->nkpars);
    std::free(hitsr->binds);
    std::free(hitsr->einds);
    ptiFreeNnzIndexVector(&hitsr->kptr);
    ptiFreeNnzIndexVector(&hitsr->cptr);
    ptiFreeNnzIndexVector(&hitsr->knnzs);
    ptiFreeNnzIndexVector(&hitsr->bptr);
    ptiFreeValueVector(&hitsr->values);
    std::free(hitsr->sortorder);
    std::free(hitsr->ndims);
    std::memset(hitsr, 0, sizeof(*hitsr));
}

double SparseTensorFrobeniusNormSquaredHiCOO(ptiSparseTensorHiCOO const *const hitsr) {
    double norm = 0;
    for (ptiNnzIndex n = 0; n < hitsr->nnz; ++n)
        norm += (double)hitsr->values.data[n] * (double)hitsr->values.data[n];
    return norm;
}

int ptiSparseTensorToHiCOO(ptiSparseTensorHiCOO *hitsr, ptiNnzIndex *max_nnzb, ptiSparseTensor *tsr,
                           const ptiElementIndex sb_bits, const ptiElementIndex sk_bits,
                           const ptiElementIndex sc_bits, int const tk) {
    const ptiIndex N = tsr->nmodes;
    const uint64_t nnz = tsr->nnz;
    if (sk_bits < sb_bits || sc_bits < sb_bits)
        PTI_FAIL(PTIERR_VALUE_ERROR, "HiSpTns Convert", "requires sk_bits >= sb_bits and sc_bits >= sb_bits");
    if (sb_bits > 7)
        PTI_FAIL(PTIERR_VALUE_ERROR, "HiSpTns Convert", "sb_bits must be <= 7 (8-bit element indices)");
    if (sk_bits > 31 || sc_bits > 63)
        PTI_FAIL(PTIERR_VALUE_ERROR, "HiSpTns Convert", "sk_bits must be < 32");
    if (N == 0 || N > 8) PTI_FAIL(PTIERR_VALUE_ERROR, "HiSpTns Convert", "nmodes must be in [1, 8]");
    if (nnz >= (1ull << 32)) PTI_FAIL(PTIERR_VALUE_ERROR, "HiSpTns Convert", "nnz must be < 2^32");
    const int nthreads = tk > 0 ? tk : pti::default_threads();
    auto t_start = std::chrono::steady_clock::now();
    auto phase = [&](const char *what) {
        if (!pti::verbose()) return;
        const auto now = std::chrono::steady_clock::now();
        std::fprintf(stderr, "[HiSpTns Convert] %s: %.3f s\n", what, std::chrono::duration<double>(now - t_start).count());
        t_start = now;
    };

    PTI_CHECK(ptiNewSparseTensorHiCOO(hitsr, N, tsr->ndims, nnz, sb_bits, sk_bits, sc_bits), "HiSpTns Convert");
    for (ptiIndex m = 0; m < N; ++m) tsr->sortorder[m] = m;

    // ---- 1. order the nonzeros -------------------------------------------
    bool ordered = false;
    if ((N == 3 || N == 4) && nnz >= 2) {
        std::vector<uint32_t> mx = pti::mode_maxima(tsr, nthreads);
        bool quirk_free = true;
        if (N == 3)
            for (uint32_t v : mx) quirk_free = quirk_free && v < (1u << 24);
        CompositeLayout L;
        if (quirk_free && pti::composite_layout(mx, sk_bits, L)) {
            std::vector<uint32_t> perm;
            bool unique = L.total_bits <= 64 ? try_unique_order<uint64_t>(tsr, L, perm, nthreads)
                                             : try_unique_order<u128>(tsr, L, perm, nthreads);
            if (unique) {
                pti::apply_permutation(tsr, perm, nthreads);
                ordered = true;
            }
        }
    }
    if (!ordered && nnz >= 2) {
        std::vector<uint32_t> perm = pti::exact_tuple_sort(tsr, sk_bits, nthreads);
        pti::apply_permutation(tsr, perm, nthreads);
    }

    phase(ordered ? "order (composite-key radix sort)" : "order (exact quicksort replay)");
    // ---- 2. superblocks and schedules --------------------------------------
    PTI_CHECK(set_kernel_pointers(&hitsr->kptr, &hitsr->knnzs, tsr, sk_bits), "HiSpTns Convert");
    PTI_CHECK(pti::set_kernel_scheduler(hitsr, tsr, sk_bits), "HiSpTns Convert");
    PTI_CHECK(pti::set_balanced_schedule(hitsr, sk_bits), "HiSpTns Convert");
    if (!ordered && (N == 3 || N == 4) && nnz >= 2) exact_morton_within_superblocks(tsr, hitsr->kptr, nthreads);

    phase("superblocks, schedules");
    // ---- 3. blocks, element offsets, chunks --------------------------------
    const uint32_t emask = (1u << sb_bits) - 1u;
    std::vector<uint8_t> is_new(nnz, 0);
    pti::parallel_for(nnz, nthreads, [&](uint64_t b, uint64_t e, int) {
        for (uint64_t z = b; z < e; ++z) {
            if (z == 0) continue;
            for (ptiIndex m = 0; m < N; ++m)
                if ((tsr->inds[m].data[z] >> sb_bits) != (tsr->inds[m].data[z - 1] >> sb_bits)) {
                    is_new[z] = 1;
                    break;
                }
        }
    });
    uint64_t nb = 1;
    for (uint64_t z = 1; z < nnz; ++z) nb += is_new[z];

    PTI_CHECK(ptiResizeNnzIndexVector(&hitsr->bptr, nb + 1), "HiSpTns Convert");
    for (ptiIndex m = 0; m < N; ++m) {
        PTI_CHECK(ptiResizeBlockIndexVector(&hitsr->binds[m], nb), "HiSpTns Convert");
        PTI_CHECK(ptiResizeElementIndexVector(&hitsr->einds[m], nnz), "HiSpTns Convert");
    }
    PTI_CHECK(ptiResizeValueVector(&hitsr->values, nnz), "HiSpTns Convert");
    if (nnz) std::memcpy(hitsr->values.data, tsr->values.data, nnz * sizeof(ptiValue));
    for (ptiIndex m = 0; m < N; ++m) {
        const uint32_t *src = tsr->inds[m].data;
        uint8_t *dst = hitsr->einds[m].data;
        pti::parallel_for(nnz, nthreads, [&](uint64_t b, uint64_t e, int) {
            for (uint64_t z = b; z < e; ++z) dst[z] = (uint8_t)(src[z] & emask);
        });
    }
    // Block starts and coordinates.
    hitsr->bptr.data[0] = 0;
    for (ptiIndex m = 0; m < N; ++m) hitsr->binds[m].data[0] = pti::coord_or_zero(tsr, m, 0) >> sb_bits;
    {
        uint64_t b = 0;
        for (uint64_t z = 1; z < nnz; ++z) {
            if (!is_new[z]) continue;
            ++b;
            hitsr->bptr.data[b] = (uint64_t)(uint32_t)z;
            for (ptiIndex m = 0; m < N; ++m) hitsr->binds[m].data[b] = tsr->inds[m].data[z] >> sb_bits;
        }
    }
    hitsr->bptr.data[nb] = nnz;

    PTI_CHECK(pti::set_chunks_from_blocks(hitsr), "HiSpTns Convert");

    phase("blocks, element offsets, chunks");
    uint64_t mx_nnzb = hitsr->bptr.data[1] - hitsr->bptr.data[0];
    for (uint64_t b = 0; b < nb; ++b) mx_nnzb = std::max(mx_nnzb, hitsr->bptr.data[b + 1] - hitsr->bptr.data[b]);
    if (max_nnzb) *max_nnzb = mx_nnzb;
    return 0;
}

// Text dump of a HiCOO tensor in the reference's layout (src/sptensor/hicoo/
// dump.c:30-74): order and sizes, nkiters, the superblock-row schedule of
// every mode, kptr, cptr, bptr, the block coordinates, the element offsets
// and the values, each array in the vector dump format.
int ptiDumpSparseTensorHiCOO(ptiSparseTensorHiCOO *const hitsr, FILE *fp) {
    static const char *M = "HiSpTns Dump";
    const ptiIndex N = hitsr->nmodes;
    if (std::fprintf(fp, "%" PRIu32 "\n", N) < 0) PTI_FAIL(PTIERR_OS_ERROR + errno, M, std::strerror(errno));
    for (ptiIndex m = 0; m < N; ++m) std::fprintf(fp, m ? " %" PRIu32 : "%" PRIu32, hitsr->ndims[m]);
    std::fputs("\n", fp);
    std::fprintf(fp, "nkiters:\n");
    PTI_CHECK(ptiDumpIndexArray(hitsr->nkiters, N, fp), M);
    std::fprintf(fp, "kschr:\n");
    const uint64_t sk = 1ull << hitsr->sk_bits;
    for (ptiIndex m = 0; m < N; ++m) {
        std::fprintf(fp, "mode %" PRIu32 "\n", m);
        const uint64_t nrows = (hitsr->ndims[m] + sk - 1) / sk;
        for (uint64_t i = 0; i < nrows; ++i) PTI_CHECK(ptiDumpIndexVector(&hitsr->kschr[m][i], fp), M);
    }
    std::fprintf(fp, "kptr:\n");
    PTI_CHECK(ptiDumpNnzIndexVector(&hitsr->kptr, fp), M);
    std::fprintf(fp, "cptr:\n");
    PTI_CHECK(ptiDumpNnzIndexVector(&hitsr->cptr, fp), M);
    std::fprintf(fp, "bptr:\n");
    PTI_CHECK(ptiDumpNnzIndexVector(&hitsr->bptr, fp), M);
    std::fprintf(fp, "binds:\n");
    for (ptiIndex m = 0; m < N; ++m) PTI_CHECK(ptiDumpBlockIndexVector(&hitsr->binds[m], fp), M);
    std::fprintf(fp, "einds:\n");
    for (ptiIndex m = 0; m < N; ++m) PTI_CHECK(ptiDumpElementIndexVector(&hitsr->einds[m], fp), M);
    std::fprintf(fp, "values:\n");
    PTI_CHECK(ptiDumpValueVector(&hitsr->values, fp), M);
    return 0;
}

}  // extern "C"
