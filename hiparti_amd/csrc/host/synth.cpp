// Seeded synthetic sparse tensors and factor matrices.
//
// Every random number is a pure function of (seed, stream tag, counter):
// r = mix(mix(seed ^ tag) + counter), with mix the SplitMix64 finaliser.  The
// output therefore does not depend on the number of threads.
//
// Zipf(s) draws use Vose's alias method over P(k) ~ (k+1)^-s, k in [0, dim).
// Coordinates are packed into an order-preserving integer key (mode 0 most
// significant); duplicates are removed by sorting (key, draw number) pairs and
// keeping the first draw of each key; further rounds are drawn until enough
// unique coordinates exist, and the nnz unique keys with the lowest draw
// numbers are kept, in key (lexicographic) order.
#include "../../../include/hiparti_synth.h"
#include "internal.h"
#include "keys.h"
#include "sorting.h"

#include <cmath>
#include <cstdlib>

namespace {

using pti::u128;

inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

inline uint64_t stream_base(uint64_t seed, uint64_t tag) { return mix64(seed ^ mix64(tag)); }
inline uint64_t rnd(uint64_t base, uint64_t counter) { return mix64(base + mix64(counter)); }

enum : uint64_t { TAG_COORD = 0x1000, TAG_PERM = 0x2000, TAG_VALUE = 0x3000, TAG_MATRIX = 0x4000 };

struct Alias {
    std::vector<double> prob;
    std::vector<uint32_t> alias;
    uint32_t n = 0;
    void build_zipf(uint32_t dim, double s) {
        n = dim;
        prob.assign(dim, 0.0);
        alias.assign(dim, 0);
        std::vector<double> w(dim);
        double sum = 0;
        for (uint32_t k = 0; k < dim; ++k) {
            w[k] = std::pow((double)(k + 1), -s);
            sum += w[k];
        }
        std::vector<uint32_t> small, large;
        small.reserve(dim);
        large.reserve(dim);
        for (uint32_t k = 0; k < dim; ++k) {
            w[k] = w[k] * dim / sum;
            (w[k] < 1.0 ? small : large).push_back(k);
        }
        while (!small.empty() && !large.empty()) {
            uint32_t l = small.back();
            small.pop_back();
            uint32_t g = large.back();
            prob[l] = w[l];
            alias[l] = g;
            w[g] = (w[g] + w[l]) - 1.0;
            if (w[g] < 1.0) {
                large.pop_back();
                small.push_back(g);
            }
        }
        for (uint32_t k : large) prob[k] = 1.0, alias[k] = k;
        for (uint32_t k : small) prob[k] = 1.0, alias[k] = k;
    }
    inline uint32_t sample(uint64_t r) const {
        uint32_t col = (uint32_t)(((r >> 32) * (uint64_t)n) >> 32);
        double u = (double)(r & 0xFFFFFFFFull) * (1.0 / 4294967296.0);
        return u < prob[col] ? col : alias[col];
    }
};

struct ModeSampler {
    int dist = PTI_SYNTH_UNIFORM;
    uint32_t dim = 1;
    Alias alias;
    std::vector<uint32_t> relabel;  // empty = identity
    inline uint32_t draw(uint64_t r) const {
        uint32_t k = dist == PTI_SYNTH_ZIPF ? alias.sample(r) : (uint32_t)(((r >> 32) * (uint64_t)dim) >> 32);
        return relabel.empty() ? k : relabel[k];
    }
};

template <typename K>
int generate(ptiSparseTensor *tsr, ptiIndex N, const ptiIndex *ndims, uint64_t nnz,
             const std::vector<ModeSampler> &ms, const std::vector<int> &bits, int total_bits, int value_dist,
             uint64_t seed, int nthreads) {
    const uint64_t base = stream_base(seed, TAG_COORD);
    std::vector<K> ukey;         // unique keys, sorted
    std::vector<uint32_t> udraw; // first draw number of each unique key
    uint64_t drawn = 0;
    double ratio = 1.25;
    while (ukey.size() < nnz) {
        uint64_t need = nnz - ukey.size();
        uint64_t D = (uint64_t)(need * ratio) + 1024;
        if (drawn + D >= (1ull << 32)) PTI_FAIL(PTIERR_VALUE_ERROR, "Synth", "too many draws");
        size_t old = ukey.size();
        ukey.resize(old + D);
        udraw.resize(old + D);
        pti::parallel_for(D, nthreads, [&](uint64_t b, uint64_t e, int) {
            for (uint64_t x = b; x < e; ++x) {
                uint64_t d = drawn + x;
                K k = 0;
                for (ptiIndex m = 0; m < N; ++m) {
                    uint32_t idx = ms[m].draw(rnd(base, d * 16 + m));
                    k = (bits[m] ? (k << bits[m]) : k) | (K)idx;
                }
                ukey[old + x] = k;
                udraw[old + x] = (uint32_t)d;
            }
        });
        drawn += D;
        // Stable sort keeps the earliest draw first among equal keys.
        pti::radix_sort_pairs<K>(ukey, udraw, total_bits, nthreads);
        size_t w = 0;
        for (size_t x = 0; x < ukey.size(); ++x) {
            if (w > 0 && ukey[w - 1] == ukey[x]) continue;
            ukey[w] = ukey[x];
            udraw[w] = udraw[x];
            ++w;
        }
        ukey.resize(w);
        udraw.resize(w);
        double got = (double)w;
        ratio = std::max(1.05, std::min(4.0, (double)drawn / std::max(got, 1.0)));
        if (drawn > 64 * (nnz + 1024)) PTI_FAIL(PTIERR_VALUE_ERROR, "Synth", "cannot reach nnz unique coordinates");
    }
    if (ukey.size() > nnz) {
        // Keep the nnz keys with the lowest draw numbers (draw numbers are distinct).
        std::vector<uint32_t> d(udraw);
        std::nth_element(d.begin(), d.begin() + (nnz - 1), d.end());
        uint32_t cutoff = d[nnz - 1];
        size_t w = 0;
        for (size_t x = 0; x < ukey.size(); ++x)
            if (udraw[x] <= cutoff) ukey[w++] = ukey[x];
        ukey.resize(w);
    }
    PTI_CHECK(ptiNewSparseTensor(tsr, N, ndims), "Synth");
    for (ptiIndex m = 0; m < N; ++m) PTI_CHECK(ptiResizeIndexVector(&tsr->inds[m], nnz), "Synth");
    PTI_CHECK(ptiResizeValueVector(&tsr->values, nnz), "Synth");
    tsr->nnz = nnz;
    const uint64_t vbase = stream_base(seed, TAG_VALUE);
    pti::parallel_for(nnz, nthreads, [&](uint64_t b, uint64_t e, int) {
        for (uint64_t z = b; z < e; ++z) {
            K k = ukey[z];
            for (int m = (int)N - 1; m >= 0; --m) {
                tsr->inds[m].data[z] = bits[m] ? (uint32_t)(k & (((K)1 << bits[m]) - 1)) : 0u;
                k = bits[m] ? (k >> bits[m]) : k;
            }
            uint64_t r1 = rnd(vbase, 2 * z);
            if (value_dist == PTI_SYNTH_VAL_LOGNORMAL) {
                uint64_t r2 = rnd(vbase, 2 * z + 1);
                double u1 = ((double)(r1 >> 11) + 0.5) * (1.0 / 9007199254740992.0);
                double u2 = ((double)(r2 >> 11) + 0.5) * (1.0 / 9007199254740992.0);
                double g = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
                tsr->values.data[z] = (float)std::exp(g);
            } else {
                tsr->values.data[z] = (float)(((double)(r1 >> 40) + 0.5) * (1.0 / 16777216.0));
            }
        }
    });
    return 0;
}

}  // namespace

extern "C" {

int ptiSynthSparseTensor(ptiSparseTensor *tsr, ptiIndex nmodes, const ptiIndex ndims[], ptiNnzIndex nnz,
                         const int dist[], const double zipf_s[], const int permute[], int value_dist,
                         uint64_t seed, int nthreads) {
    if (nmodes == 0 || nmodes > 8) PTI_FAIL(PTIERR_VALUE_ERROR, "Synth", "nmodes must be in [1, 8]");
    long double cells = 1;
    for (ptiIndex m = 0; m < nmodes; ++m) {
        if (ndims[m] == 0) PTI_FAIL(PTIERR_VALUE_ERROR, "Synth", "empty mode");
        cells *= ndims[m];
    }
    if ((long double)nnz > cells) PTI_FAIL(PTIERR_VALUE_ERROR, "Synth", "nnz exceeds the number of cells");
    if (nnz >= (1ull << 31)) PTI_FAIL(PTIERR_VALUE_ERROR, "Synth", "nnz must be < 2^31");
    if (nthreads <= 0) nthreads = pti::default_threads();
    std::vector<ModeSampler> ms(nmodes);
    std::vector<int> bits(nmodes);
    int total = 0;
    for (ptiIndex m = 0; m < nmodes; ++m) {
        ms[m].dist = dist ? dist[m] : PTI_SYNTH_UNIFORM;
        ms[m].dim = ndims[m];
        if (ms[m].dist == PTI_SYNTH_ZIPF) ms[m].alias.build_zipf(ndims[m], zipf_s ? zipf_s[m] : 1.0);
        if (permute && permute[m]) {
            std::vector<uint32_t> &p = ms[m].relabel;
            p.resize(ndims[m]);
            for (uint32_t i = 0; i < ndims[m]; ++i) p[i] = i;
            const uint64_t pb = stream_base(seed, TAG_PERM + m);
            for (uint32_t i = ndims[m] - 1; i > 0; --i) {
                uint32_t j = (uint32_t)(((rnd(pb, i) >> 32) * (uint64_t)(i + 1)) >> 32);
                std::swap(p[i], p[j]);
            }
        }
        bits[m] = pti::bit_width_u32(ndims[m] - 1);
        total += bits[m];
    }
    if (total <= 64)
        return generate<uint64_t>(tsr, nmodes, ndims, nnz, ms, bits, total, value_dist, seed, nthreads);
    if (total <= 128)
        return generate<u128>(tsr, nmodes, ndims, nnz, ms, bits, total, value_dist, seed, nthreads);
    PTI_FAIL(PTIERR_VALUE_ERROR, "Synth", "coordinates need more than 128 key bits");
}

int ptiSynthUniformMatrix(ptiMatrix *mtx, uint64_t seed) {
    const uint64_t base = stream_base(seed, TAG_MATRIX);
    const uint64_t nr = mtx->nrows, nc = mtx->ncols, st = mtx->stride;
    pti::parallel_for(nr, 0, [&](uint64_t b, uint64_t e, int) {
        for (uint64_t i = b; i < e; ++i)
            for (uint64_t j = 0; j < nc; ++j)
                mtx->values[i * st + j] = (float)(((double)(rnd(base, i * nc + j) >> 40) + 0.5) * (1.0 / 16777216.0));
    });
    return 0;
}

}  // extern "C"
