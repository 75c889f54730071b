// COO -> HiCOO conversion with the ordering and the block emission on the GPU
// (reference: src/sptensor/hicoo/convert.c:430-860, sort.c:397-501,
// 929-1288).
//
// When the reference's order is unique -- 3-D or 4-D tensors with distinct
// coordinates, 3-D indices below 2^24 (where its Morton key keeps every
// bit), and a composite key of at most 64 bits -- that order is the sort
// order of the composite (superblock tuple, compressed Morton) key:
//   1. keys on the device, rocPRIM radix sort of (key, position) pairs;
//   2. a duplicate key means the order is not unique: the host conversion
//      (ptiSparseTensorToHiCOO, which replays the reference's quicksort) is
//      used instead;
//   3. the COO is gathered into the sorted order; block and superblock
//      starts are flagged where the (idx >> sb_bits) / (idx >> sk_bits) tuple
//      changes, scanned into block / superblock ids, and bptr, binds, einds
//      and the superblock starts are scattered;
//   4. the host copies the sorted COO back into the caller's tensor (which
//      the reference also sorts in place) and builds the small per-superblock
//      structures (kschr, nkiters, the balanced schedule, cptr) with the same
//      code as the host conversion.
// Every other case (orders 2 and 5-8, duplicates, wide keys) runs the host
// conversion, so the result is always the reference's, bit for bit.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "../../../include/hiparti_dev.h"
#include "../host/hicoo_build.h"
#include "../host/keys.h"
#include "runtime.hpp"

namespace pti {
namespace dev {
namespace {

struct KeyLayout {
    uint32_t N, sk, max_lb;
    uint32_t lb[4], hb[4];
};

struct Modes4 {
    const uint32_t *p[4];
};
struct MutModes4 {
    uint32_t *p[4];
};

inline dim3 grid_for(uint64_t n) {
    uint64_t b = (n + 255) / 256;
    if (b > 65536) b = 65536;
    if (b == 0) b = 1;
    return dim3((uint32_t)b);
}

__global__ void __launch_bounds__(256) composite_keys_kernel(Modes4 in, uint64_t nnz, KeyLayout L,
                                                             uint64_t *__restrict__ keys, uint32_t *__restrict__ ids) {
    for (uint64_t z = (uint64_t)blockIdx.x * 256 + threadIdx.x; z < nnz; z += (uint64_t)gridDim.x * 256) {
        uint32_t c[4] = {0, 0, 0, 0};
        for (uint32_t m = 0; m < L.N; ++m) c[m] = in.p[m][z];
        uint64_t k = 0;
        for (uint32_t m = 0; m < L.N; ++m)
            if (L.hb[m]) k = (k << L.hb[m]) | (uint64_t)(c[m] >> L.sk);
        for (int bit = (int)L.max_lb - 1; bit >= 0; --bit)
            for (uint32_t m = 0; m < L.N; ++m)
                if ((uint32_t)bit < L.lb[m]) k = (k << 1) | (uint64_t)((c[m] >> bit) & 1u);
        keys[z] = k;
        ids[z] = (uint32_t)z;
    }
}

__global__ void __launch_bounds__(256) duplicate_kernel(const uint64_t *__restrict__ keys, uint64_t nnz, int *flag) {
    for (uint64_t z = (uint64_t)blockIdx.x * 256 + threadIdx.x; z + 1 < nnz; z += (uint64_t)gridDim.x * 256)
        if (keys[z] == keys[z + 1]) atomicOr(flag, 1);
}

__global__ void __launch_bounds__(256) gather_sorted_kernel(Modes4 in, const float *__restrict__ vin,
                                                            const uint32_t *__restrict__ ids, uint64_t nnz,
                                                            uint32_t N, MutModes4 out, float *__restrict__ vout) {
    for (uint64_t z = (uint64_t)blockIdx.x * 256 + threadIdx.x; z < nnz; z += (uint64_t)gridDim.x * 256) {
        const uint32_t s = ids[z];
        for (uint32_t m = 0; m < N; ++m) out.p[m][z] = in.p[m][s];
        vout[z] = vin[s];
    }
}

// flag[z] bit 0: a block starts at z (z > 0), bit 1: a superblock starts.
__global__ void __launch_bounds__(256) start_flags_kernel(Modes4 s, uint64_t nnz, uint32_t N, uint32_t sb,
                                                          uint32_t sk, uint32_t *__restrict__ bflag,
                                                          uint32_t *__restrict__ kflag) {
    for (uint64_t z = (uint64_t)blockIdx.x * 256 + threadIdx.x; z < nnz; z += (uint64_t)gridDim.x * 256) {
        uint32_t b = 0, k = 0;
        if (z > 0)
            for (uint32_t m = 0; m < N; ++m) {
                const uint32_t x = s.p[m][z], y = s.p[m][z - 1];
                b |= (x >> sb) != (y >> sb);
                k |= (x >> sk) != (y >> sk);
            }
        bflag[z] = b;
        kflag[z] = k;
    }
}

__global__ void __launch_bounds__(256) emit_blocks_kernel(Modes4 s, uint64_t nnz, uint32_t N, uint32_t sb,
                                                          const uint32_t *__restrict__ bflag,
                                                          const uint32_t *__restrict__ bid,
                                                          const uint32_t *__restrict__ kflag,
                                                          const uint32_t *__restrict__ kid,
                                                          uint64_t *__restrict__ bptr, MutModes4 binds,
                                                          uint64_t *__restrict__ kstart, uint8_t *__restrict__ e0,
                                                          uint8_t *__restrict__ e1, uint8_t *__restrict__ e2,
                                                          uint8_t *__restrict__ e3) {
    const uint32_t mask = (1u << sb) - 1u;
    uint8_t *eo[4] = {e0, e1, e2, e3};
    for (uint64_t z = (uint64_t)blockIdx.x * 256 + threadIdx.x; z < nnz; z += (uint64_t)gridDim.x * 256) {
        if (z == 0 || bflag[z]) {
            const uint32_t b = bid[z];
            bptr[b] = z;
            for (uint32_t m = 0; m < N; ++m) binds.p[m][b] = s.p[m][z] >> sb;
        }
        if (z == 0 || kflag[z]) kstart[kid[z]] = z;
        for (uint32_t m = 0; m < N; ++m) eo[m][z] = (uint8_t)(s.p[m][z] & mask);
    }
}

template <typename T>
int put(Buffer &b, const T *src, size_t n, hipStream_t s, const char *M) {
    PTI_HIP(b.alloc(n * sizeof(T)), M);
    if (n) PTI_HIP(hipMemcpyAsync(b.ptr, src, n * sizeof(T), hipMemcpyHostToDevice, s), M);
    return 0;
}

template <typename T>
int get(T *dst, const Buffer &b, size_t n, hipStream_t s, const char *M) {
    if (n) PTI_HIP(hipMemcpyAsync(dst, b.ptr, n * sizeof(T), hipMemcpyDeviceToHost, s), M);
    return 0;
}

// Returns 1 when the GPU path does not apply (the caller runs the host
// conversion), 0 on success, a PTIERR code on failure.
int gpu_convert(ptiSparseTensorHiCOO *hitsr, ptiNnzIndex *max_nnzb, ptiSparseTensor *tsr, ptiElementIndex sb,
                ptiElementIndex sk, ptiElementIndex sc, hipStream_t s) {
    static const char *M = "GPU HiSpTns Convert";
    const uint32_t N = tsr->nmodes;
    const uint64_t nnz = tsr->nnz;
    if ((N != 3 && N != 4) || nnz < 2) return 1;
    const std::vector<uint32_t> mx = pti::mode_maxima(tsr, pti::default_threads());
    if (N == 3)
        for (uint32_t v : mx)
            if (v >= (1u << 24)) return 1;
    CompositeLayout L;
    if (!pti::composite_layout(mx, sk, L) || L.total_bits > 64) return 1;
    PTI_CHECK(require_device(M), M);
    KeyLayout kl{};
    kl.N = N;
    kl.sk = sk;
    kl.max_lb = (uint32_t)L.max_lb;
    for (uint32_t m = 0; m < N; ++m) {
        kl.lb[m] = (uint32_t)L.lb[m];
        kl.hb[m] = (uint32_t)L.hb[m];
    }
    Buffer din[4], dvals, keys0, keys1, ids0, ids1, dflag, dsorted[4], dsv, bflag, kflag, bid, kid, tmp;
    Modes4 in{};
    for (uint32_t m = 0; m < N; ++m) {
        PTI_CHECK(put(din[m], tsr->inds[m].data, nnz, s, M), M);
        in.p[m] = din[m].as<uint32_t>();
    }
    PTI_CHECK(put(dvals, tsr->values.data, nnz, s, M), M);
    PTI_HIP(keys0.alloc(nnz * 8), M);
    PTI_HIP(keys1.alloc(nnz * 8), M);
    PTI_HIP(ids0.alloc(nnz * 4), M);
    PTI_HIP(ids1.alloc(nnz * 4), M);
    hipLaunchKernelGGL(composite_keys_kernel, grid_for(nnz), dim3(256), 0, s, in, nnz, kl, keys0.as<uint64_t>(),
                       ids0.as<uint32_t>());
    PTI_HIP(hipGetLastError(), M);
    size_t tbytes = 0;
    PTI_HIP(rocprim::radix_sort_pairs(nullptr, tbytes, keys0.as<uint64_t>(), keys1.as<uint64_t>(), ids0.as<uint32_t>(),
                                      ids1.as<uint32_t>(), (size_t)nnz, 0, (unsigned)L.total_bits, s),
            M);
    PTI_HIP(tmp.alloc(tbytes), M);
    PTI_HIP(rocprim::radix_sort_pairs(tmp.ptr, tbytes, keys0.as<uint64_t>(), keys1.as<uint64_t>(), ids0.as<uint32_t>(),
                                      ids1.as<uint32_t>(), (size_t)nnz, 0, (unsigned)L.total_bits, s),
            M);
    PTI_HIP(dflag.alloc(sizeof(int)), M);
    PTI_HIP(hipMemsetAsync(dflag.ptr, 0, sizeof(int), s), M);
    hipLaunchKernelGGL(duplicate_kernel, grid_for(nnz), dim3(256), 0, s, keys1.as<uint64_t>(), nnz, dflag.as<int>());
    PTI_HIP(hipGetLastError(), M);
    int dup = 0;
    PTI_CHECK(get(&dup, dflag, 1, s, M), M);
    PTI_HIP(hipStreamSynchronize(s), M);
    if (dup) return 1;
    keys0.release();
    keys1.release();
    ids0.release();

    // Sorted COO, block and superblock starts.
    MutModes4 sorted{};
    Modes4 sorted_c{};
    for (uint32_t m = 0; m < N; ++m) {
        PTI_HIP(dsorted[m].alloc(nnz * 4), M);
        sorted.p[m] = dsorted[m].as<uint32_t>();
        sorted_c.p[m] = sorted.p[m];
    }
    PTI_HIP(dsv.alloc(nnz * 4), M);
    hipLaunchKernelGGL(gather_sorted_kernel, grid_for(nnz), dim3(256), 0, s, in, dvals.as<float>(),
                       ids1.as<uint32_t>(), nnz, N, sorted, dsv.as<float>());
    PTI_HIP(hipGetLastError(), M);
    // The gather must finish before its inputs are released.
    PTI_HIP(hipStreamSynchronize(s), M);
    for (uint32_t m = 0; m < N; ++m) din[m].release();
    dvals.release();
    ids1.release();
    PTI_HIP(bflag.alloc(nnz * 4), M);
    PTI_HIP(kflag.alloc(nnz * 4), M);
    PTI_HIP(bid.alloc(nnz * 4), M);
    PTI_HIP(kid.alloc(nnz * 4), M);
    hipLaunchKernelGGL(start_flags_kernel, grid_for(nnz), dim3(256), 0, s, sorted_c, nnz, N, (uint32_t)sb,
                       (uint32_t)sk, bflag.as<uint32_t>(), kflag.as<uint32_t>());
    PTI_HIP(hipGetLastError(), M);
    size_t sbytes = 0;
    PTI_HIP(rocprim::inclusive_scan(nullptr, sbytes, bflag.as<uint32_t>(), bid.as<uint32_t>(), (size_t)nnz,
                                    rocprim::plus<uint32_t>(), s),
            M);
    PTI_HIP(tmp.alloc(sbytes), M);   // the sort's storage is idle: the stream was synchronised
    PTI_HIP(rocprim::inclusive_scan(tmp.ptr, sbytes, bflag.as<uint32_t>(), bid.as<uint32_t>(), (size_t)nnz,
                                    rocprim::plus<uint32_t>(), s),
            M);
    PTI_HIP(rocprim::inclusive_scan(tmp.ptr, sbytes, kflag.as<uint32_t>(), kid.as<uint32_t>(), (size_t)nnz,
                                    rocprim::plus<uint32_t>(), s),
            M);
    uint32_t last_b = 0, last_k = 0;
    PTI_HIP(hipMemcpyAsync(&last_b, bid.as<uint32_t>() + nnz - 1, 4, hipMemcpyDeviceToHost, s), M);
    PTI_HIP(hipMemcpyAsync(&last_k, kid.as<uint32_t>() + nnz - 1, 4, hipMemcpyDeviceToHost, s), M);
    PTI_HIP(hipStreamSynchronize(s), M);
    const uint64_t nb = (uint64_t)last_b + 1, nk = (uint64_t)last_k + 1;
    Buffer dbptr, dbinds[4], dkstart, de[4];
    MutModes4 binds{};
    PTI_HIP(dbptr.alloc(nb * 8), M);
    PTI_HIP(dkstart.alloc(nk * 8), M);
    uint8_t *eo[4] = {nullptr, nullptr, nullptr, nullptr};
    for (uint32_t m = 0; m < N; ++m) {
        PTI_HIP(dbinds[m].alloc(nb * 4), M);
        binds.p[m] = dbinds[m].as<uint32_t>();
        PTI_HIP(de[m].alloc(nnz), M);
        eo[m] = de[m].as<uint8_t>();
    }
    for (uint32_t m = N; m < 4; ++m) eo[m] = eo[0];  // unused slots (N == 3)
    hipLaunchKernelGGL(emit_blocks_kernel, grid_for(nnz), dim3(256), 0, s, sorted_c, nnz, N, (uint32_t)sb,
                       bflag.as<uint32_t>(), bid.as<uint32_t>(), kflag.as<uint32_t>(), kid.as<uint32_t>(),
                       dbptr.as<uint64_t>(), binds, dkstart.as<uint64_t>(), eo[0], eo[1], eo[2], eo[3]);
    PTI_HIP(hipGetLastError(), M);

    // Host structures.
    PTI_CHECK(ptiNewSparseTensorHiCOO(hitsr, N, tsr->ndims, nnz, sb, sk, sc), M);
    for (uint32_t m = 0; m < N; ++m) tsr->sortorder[m] = m;
    for (uint32_t m = 0; m < N; ++m) PTI_CHECK(get(tsr->inds[m].data, dsorted[m], nnz, s, M), M);
    PTI_CHECK(get(tsr->values.data, dsv, nnz, s, M), M);
    PTI_CHECK(ptiResizeNnzIndexVector(&hitsr->bptr, nb + 1), M);
    for (uint32_t m = 0; m < N; ++m) {
        PTI_CHECK(ptiResizeBlockIndexVector(&hitsr->binds[m], nb), M);
        PTI_CHECK(ptiResizeElementIndexVector(&hitsr->einds[m], nnz), M);
        PTI_CHECK(get(hitsr->binds[m].data, dbinds[m], nb, s, M), M);
        PTI_CHECK(get(hitsr->einds[m].data, de[m], nnz, s, M), M);
    }
    PTI_CHECK(get(hitsr->bptr.data, dbptr, nb, s, M), M);
    PTI_CHECK(ptiResizeValueVector(&hitsr->values, nnz), M);
    PTI_CHECK(get(hitsr->values.data, dsv, nnz, s, M), M);
    PTI_CHECK(ptiResizeNnzIndexVector(&hitsr->kptr, nk + 1), M);
    PTI_CHECK(get(hitsr->kptr.data, dkstart, nk, s, M), M);
    PTI_HIP(hipStreamSynchronize(s), M);
    hitsr->bptr.data[nb] = nnz;
    hitsr->kptr.data[nk] = nnz;
    PTI_CHECK(ptiResizeNnzIndexVector(&hitsr->knnzs, nk), M);
    for (uint64_t k = 0; k < nk; ++k) hitsr->knnzs.data[k] = hitsr->kptr.data[k + 1] - hitsr->kptr.data[k];
    PTI_CHECK(pti::set_kernel_scheduler(hitsr, tsr, sk), M);
    PTI_CHECK(pti::set_balanced_schedule(hitsr, sk), M);
    PTI_CHECK(pti::set_chunks_from_blocks(hitsr), M);
    uint64_t mxb = 0;
    for (uint64_t b = 0; b < nb; ++b) mxb = std::max<uint64_t>(mxb, hitsr->bptr.data[b + 1] - hitsr->bptr.data[b]);
    if (max_nnzb) *max_nnzb = mxb;
    return 0;
}

}  // namespace
}  // namespace dev
}  // namespace pti

extern "C" int ptiDevSparseTensorToHiCOO(ptiSparseTensorHiCOO *hitsr, ptiNnzIndex *max_nnzb, ptiSparseTensor *tsr,
                                         ptiElementIndex const sb_bits, ptiElementIndex const sk_bits,
                                         ptiElementIndex const sc_bits, void *stream) {
    static const char *M = "GPU HiSpTns Convert";
    if (sk_bits < sb_bits || sc_bits < sb_bits)
        PTI_FAIL(PTIERR_VALUE_ERROR, M, "requires sk_bits >= sb_bits and sc_bits >= sb_bits");
    if (sb_bits > 7) PTI_FAIL(PTIERR_VALUE_ERROR, M, "sb_bits must be <= 7 (8-bit element indices)");
    if (sk_bits > 31 || sc_bits > 63) PTI_FAIL(PTIERR_VALUE_ERROR, M, "sk_bits must be < 32");
    if (tsr->nnz >= (1ull << 32)) PTI_FAIL(PTIERR_VALUE_ERROR, M, "nnz must be < 2^32");
    const int rc = pti::dev::gpu_convert(hitsr, max_nnzb, tsr, sb_bits, sk_bits, sc_bits,
                                         static_cast<hipStream_t>(stream));
    if (rc == 1) return ptiSparseTensorToHiCOO(hitsr, max_nnzb, tsr, sb_bits, sk_bits, sc_bits, 0);
    return rc;
}
