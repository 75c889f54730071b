// Host-only regression driver for the COO -> HiCOO conversion, built with
// AddressSanitizer from the library's host sources (tests/test_host_asan.py).
//
// pti::mode_maxima once sized its per-thread slots for the thread count it
// was given (0 = "default") rather than for the count parallel_for runs, so
// with OMP_NUM_THREADS > 1 every worker past the first wrote past the end of
// a one-slot buffer.  This driver calls it, and the whole host conversion,
// with thread counts 0 (default), 1 and 3 under OMP_NUM_THREADS=8, and checks
// that every count gives the same maxima and the same HiCOO arrays.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/HiParTI.h"
#include "../../include/hiparti_synth.h"
#include "../../hiparti_amd/csrc/host/keys.h"

// This CPU build links no device code: the device-event timer hooks report
// that no device exists.
namespace pti {
int dev_timer_create(void **) { return PTIERR_CUDA_ERROR + 100; }
int dev_timer_start(void *) { return PTIERR_CUDA_ERROR + 100; }
int dev_timer_stop(void *) { return PTIERR_CUDA_ERROR + 100; }
double dev_timer_elapsed(void *) { return -1.0; }
void dev_timer_destroy(void *) {}
}  // namespace pti

static int fail(const char *what) {
    std::fprintf(stderr, "FAILED: %s\n", what);
    return 1;
}

template <typename V>
static bool same_vec(const V &a, const V &b, size_t elem) {
    return a.len == b.len && (a.len == 0 || std::memcmp(a.data, b.data, a.len * elem) == 0);
}

static int check_order(ptiIndex N, const ptiIndex *ndims, ptiNnzIndex nnz, int dist, uint64_t seed) {
    std::vector<int> d(N, dist);
    std::vector<double> s(N, 1.1);
    std::vector<int> perm(N, 0);
    ptiSparseTensor base;
    if (ptiSynthSparseTensor(&base, N, ndims, nnz, d.data(), s.data(), perm.data(), 0, seed, 0)) return fail("synth");
    // mode_maxima with every thread count against a serial scan.
    for (int nt : {0, 1, 3}) {
        std::vector<uint32_t> mx = pti::mode_maxima(&base, nt);
        for (ptiIndex m = 0; m < N; ++m) {
            uint32_t ref = 0;
            for (ptiNnzIndex z = 0; z < base.nnz; ++z) ref = base.inds[m].data[z] > ref ? base.inds[m].data[z] : ref;
            if (mx[m] != ref) return fail("mode_maxima");
        }
    }
    // The conversion's arrays do not depend on the thread count.
    ptiSparseTensorHiCOO h[3];
    int tks[3] = {1, 0, 3};
    for (int i = 0; i < 3; ++i) {
        ptiSparseTensor t;
        if (ptiCopySparseTensor(&t, &base, 1)) return fail("copy");
        ptiNnzIndex mxb = 0;
        if (ptiSparseTensorToHiCOO(&h[i], &mxb, &t, 3, 5, 5, tks[i])) return fail("convert");
        ptiFreeSparseTensor(&t);
    }
    for (int i = 1; i < 3; ++i) {
        if (!same_vec(h[0].bptr, h[i].bptr, sizeof(ptiNnzIndex))) return fail("bptr");
        if (!same_vec(h[0].values, h[i].values, sizeof(ptiValue))) return fail("values");
        if (!same_vec(h[0].kptr, h[i].kptr, sizeof(ptiNnzIndex))) return fail("kptr");
        for (ptiIndex m = 0; m < N; ++m) {
            if (!same_vec(h[0].binds[m], h[i].binds[m], sizeof(ptiBlockIndex))) return fail("binds");
            if (!same_vec(h[0].einds[m], h[i].einds[m], sizeof(ptiElementIndex))) return fail("einds");
        }
    }
    for (auto &x : h) ptiFreeSparseTensorHiCOO(&x);
    ptiFreeSparseTensor(&base);
    return 0;
}

int main() {
    const ptiIndex d3[3] = {5000, 4000, 300};
    const ptiIndex d5[5] = {300, 200, 100, 30, 12};
    if (check_order(3, d3, 60000, 1, 7)) return 1;
    if (check_order(5, d5, 20000, 0, 8)) return 1;
    if (check_order(3, d3, 5, 0, 9)) return 1;   // fewer nonzeros than threads
    std::printf("ok\n");
    return 0;
}
