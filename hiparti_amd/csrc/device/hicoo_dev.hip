// HiCOO tensors resident in HBM and the HiCOO MTTKRP entry points.
//
// Device layout (built once by ptiDevHiCOOUpload from the host HiCOO arrays):
//   bptr      u32[nb + 1]        nonzero offset of each block (nnz < 2^32)
//   binds     u32[nb * N]        block coordinates, block-major (one 16-byte
//                                read per block for N = 3 with bptr)
//   einds     u32[nnz * EW]      the N 8-bit element offsets of a nonzero
//                                packed into EW = ceil(N/4) words
//   vals      f32[nnz]
//   chunk_blk u32[nchunks + 1]   block holding nonzero c * 256; last = nb
//   perm[m]   u8[nchunks * 256]  per mode: each 256-nonzero chunk's local
//                                indices sorted by the mode-m coordinate
// The host HiCOO struct itself is left untouched.
#include <atomic>
#include <mutex>
#include <unordered_map>

#include "../../../include/hiparti_dev.h"
#include "dropin.hpp"
#include "launch.hpp"

struct ptiDevHiCOO {
    uint32_t nmodes = 0;
    uint32_t sb_bits = 0;
    uint64_t nnz = 0;
    uint64_t nb = 0;
    uint32_t nchunks = 0;
    ptiIndex ndims[8] = {0};
    pti::dev::Buffer bptr, binds, einds, vals, chunk_blk;
    pti::dev::Buffer perm[8];
    uint64_t bytes = 0;
};

namespace pti {
namespace dev {

HiCOOArgs hicoo_args(const ptiDevHiCOO *t, int mode_perm) {
    HiCOOArgs h;
    h.bptr = t->bptr.as<uint32_t>();
    h.binds = t->binds.as<uint32_t>();
    h.einds = t->einds.as<uint32_t>();
    h.vals = t->vals.as<float>();
    h.chunk_blk = t->chunk_blk.as<uint32_t>();
    h.perm = mode_perm >= 0 ? t->perm[mode_perm].as<uint8_t>() : nullptr;
    h.nnz = t->nnz;
    h.nb = (uint32_t)t->nb;
    h.nchunks = t->nchunks;
    h.sb_bits = t->sb_bits;
    return h;
}

namespace {

std::mutex g_reg_mu;
std::unordered_map<const void *, ptiDevHiCOO *> g_hicoo_reg;

int upload(ptiDevHiCOO **out, const ptiSparseTensorHiCOO *hitsr, hipStream_t s) {
    static const char *M = "DevHiCOO Upload";
    PTI_CHECK(require_device(M), M);
    const uint32_t N = hitsr->nmodes;
    if (N < 2 || N > 8) PTI_FAIL(PTIERR_VALUE_ERROR, M, "nmodes must be in [2, 8]");
    const uint64_t nnz = hitsr->nnz;
    const uint64_t nb = hitsr->bptr.len ? hitsr->bptr.len - 1 : 0;
    if (nnz >= (1ull << 32)) PTI_FAIL(PTIERR_VALUE_ERROR, M, "nnz must be < 2^32");
    if (hitsr->values.len < nnz || (nnz && nb == 0)) PTI_FAIL(PTIERR_VALUE_ERROR, M, "inconsistent HiCOO arrays");
    for (uint32_t m = 0; m < N; ++m)
        if (hitsr->einds[m].len < nnz || hitsr->binds[m].len < nb)
            PTI_FAIL(PTIERR_VALUE_ERROR, M, "inconsistent HiCOO arrays");
    if (nb && hitsr->bptr.data[nb] != nnz) PTI_FAIL(PTIERR_VALUE_ERROR, M, "bptr does not end at nnz");
    // The kernels index factor rows and LDS staging buffers with these
    // arrays: every block must be non-empty (a 256-nonzero chunk then spans
    // at most 257 blocks) and every coordinate must lie inside its mode.
    if (nnz) {
        std::atomic<bool> bad(false);
        pti::parallel_for(nb, 0, [&](uint64_t b, uint64_t e, int) {
            for (uint64_t x = b; x < e && !bad.load(std::memory_order_relaxed); ++x) {
                const uint64_t z0 = hitsr->bptr.data[x], z1 = hitsr->bptr.data[x + 1];
                if (z1 <= z0 || z1 > nnz) {
                    bad = true;
                    break;
                }
                for (uint32_t m = 0; m < N; ++m) {
                    const uint64_t base = (uint64_t)hitsr->binds[m].data[x] << hitsr->sb_bits;
                    for (uint64_t z = z0; z < z1; ++z)
                        if (base + hitsr->einds[m].data[z] >= hitsr->ndims[m]) bad = true;
                }
            }
        });
        if (bad.load()) PTI_FAIL(PTIERR_VALUE_ERROR, M, "empty block or coordinate outside its mode");
    }

    ptiDevHiCOO *t = new ptiDevHiCOO();
    t->nmodes = N;
    t->sb_bits = hitsr->sb_bits;
    t->nnz = nnz;
    t->nb = nb;
    for (uint32_t m = 0; m < N; ++m) t->ndims[m] = hitsr->ndims[m];
    const int EW = (N + 3) / 4;
    const uint64_t nchunks = (nnz + CHUNK - 1) / CHUNK;
    t->nchunks = (uint32_t)nchunks;

    // Host-side repacking.
    std::vector<uint32_t> hb(nb + 1), hbi(nb * N), he(nnz * EW, 0u), hc(nchunks + 1);
    pti::parallel_for(nb + 1, 0, [&](uint64_t b, uint64_t e, int) {
        for (uint64_t x = b; x < e; ++x) hb[x] = (uint32_t)hitsr->bptr.data[x];
    });
    pti::parallel_for(nb, 0, [&](uint64_t b, uint64_t e, int) {
        for (uint64_t x = b; x < e; ++x)
            for (uint32_t m = 0; m < N; ++m) hbi[x * N + m] = hitsr->binds[m].data[x];
    });
    pti::parallel_for(nnz, 0, [&](uint64_t b, uint64_t e, int) {
        for (uint64_t z = b; z < e; ++z)
            for (uint32_t m = 0; m < N; ++m) he[z * EW + (m >> 2)] |= (uint32_t)hitsr->einds[m].data[z] << ((m & 3) * 8);
    });
    {
        uint64_t blk = 0;
        for (uint64_t c = 0; c < nchunks; ++c) {
            const uint64_t z = c * CHUNK;
            while (hitsr->bptr.data[blk + 1] <= z) ++blk;
            hc[c] = (uint32_t)blk;
        }
        hc[nchunks] = (uint32_t)nb;
    }
    auto put = [&](Buffer &buf, const void *src, size_t bytes) -> int {
        PTI_HIP(buf.alloc(bytes), M);
        if (bytes) PTI_HIP(hipMemcpyAsync(buf.ptr, src, bytes, hipMemcpyHostToDevice, s), M);
        t->bytes += bytes;
        return 0;
    };
    int rc = 0;
    if (!rc) rc = put(t->bptr, hb.data(), hb.size() * 4);
    if (!rc) rc = put(t->binds, hbi.data(), hbi.size() * 4);
    if (!rc) rc = put(t->einds, he.data(), he.size() * 4);
    if (!rc) rc = put(t->vals, hitsr->values.data, nnz * 4);
    if (!rc) rc = put(t->chunk_blk, hc.data(), hc.size() * 4);
    if (!rc) {
        HiCOOArgs h = hicoo_args(t, -1);
        for (uint32_t m = 0; m < N && !rc; ++m) {
            hipError_t e = t->perm[m].alloc(nchunks * CHUNK);
            if (e == hipSuccess) e = launch_sort_chunks(N, h, m, t->perm[m].as<uint8_t>(), s);
            if (e != hipSuccess) {
                pti_ComplainError(M, hip_code(e), __FILE__, __LINE__, hipGetErrorString(e));
                rc = hip_code(e);
            }
            t->bytes += nchunks * CHUNK;
        }
    }
    if (!rc) {
        hipError_t e = hipStreamSynchronize(s);
        if (e != hipSuccess) {
            pti_ComplainError(M, hip_code(e), __FILE__, __LINE__, hipGetErrorString(e));
            rc = hip_code(e);
        }
    }
    if (rc) {
        (void)hipStreamSynchronize(s);  // pending copies read the host staging vectors
        delete t;
        return rc;
    }
    *out = t;
    return 0;
}

}  // namespace

// Shared by the FP32 and BF16 entry points: zero the output rows, then run the
// kernel once per 64*V-rank column tile.
int run_hicoo(const ptiDevHiCOO *t, const void *const facs[], float *out, ptiIndex R, ptiIndex stride,
              ptiIndex out_stride, const ptiIndex mats_order[], ptiIndex mode, bool bf16, hipStream_t s,
              const char *M) {
    const uint32_t N = t->nmodes;
    if (mode >= N) PTI_FAIL(PTIERR_VALUE_ERROR, M, "mode out of range");
    if (R == 0 || stride < R || out_stride < R) PTI_FAIL(PTIERR_SHAPE_MISMATCH, M, "bad rank or stride");
    for (uint32_t i = 1; i < N; ++i)
        if (mats_order[i] >= N) PTI_FAIL(PTIERR_VALUE_ERROR, M, "mats_order entry out of range");
    PTI_HIP(zero_floats(out, (uint64_t)t->ndims[mode] * out_stride, s), M);
    if (t->nnz == 0) return 0;
    const ptiIndex tile = 256;
    for (ptiIndex c0 = 0; c0 < R; c0 += tile) {
        const ptiIndex Rt = std::min<ptiIndex>(tile, R - c0);
        Shape sh;
        if (!choose_shape(Rt, stride, out_stride, sh)) PTI_FAIL(PTIERR_VALUE_ERROR, M, "unsupported rank tile");
        FactorArgs a;
        const size_t esz = bf16 ? 2 : 4;
        for (uint32_t m = 0; m < N; ++m) a.fac[m] = static_cast<const char *>(facs[m]) + (size_t)c0 * esz;
        a.out = out + c0;
        a.R = Rt;
        a.stride = stride;
        a.out_stride = out_stride;
        a.mode = mode;
        for (uint32_t i = 0; i < N; ++i) a.order[i] = mats_order[i];
        PTI_HIP(launch_hicoo(N, hicoo_args(t, (int)mode), a, sh, bf16, s), M);
    }
    return 0;
}

}  // namespace dev
}  // namespace pti

using namespace pti::dev;

extern "C" {

int ptiDevHiCOOUpload(ptiDevHiCOO **out, ptiSparseTensorHiCOO const *hitsr, void *stream) {
    return upload(out, hitsr, static_cast<hipStream_t>(stream));
}

void ptiDevHiCOOFree(ptiDevHiCOO *t) { delete t; }

int ptiDevHiCOOInfo(ptiDevHiCOO const *t, uint64_t *nnz, uint64_t *nblocks, uint32_t *nmodes, uint64_t *bytes) {
    if (nnz) *nnz = t->nnz;
    if (nblocks) *nblocks = t->nb;
    if (nmodes) *nmodes = t->nmodes;
    if (bytes) *bytes = t->bytes;
    return 0;
}

int ptiDevMTTKRPHiCOO(ptiDevHiCOO const *t, float *const dmats[], ptiIndex R, ptiIndex stride,
                      ptiIndex const mats_order[], ptiIndex mode, void *stream) {
    const void *facs[8];
    for (uint32_t m = 0; m < t->nmodes; ++m) facs[m] = dmats[m];
    return run_hicoo(t, facs, dmats[t->nmodes], R, stride, stride, mats_order, mode, false,
                     static_cast<hipStream_t>(stream), "DevHiCOO MTTKRP");
}

int ptiDevMTTKRPHiCOO_BF16(ptiDevHiCOO const *t, uint16_t const *const dfacs[], float *dout, ptiIndex R,
                           ptiIndex stride, ptiIndex out_stride, ptiIndex const mats_order[], ptiIndex mode,
                           void *stream) {
    const void *facs[8];
    for (uint32_t m = 0; m < t->nmodes; ++m) facs[m] = dfacs[m] ? (const void *)dfacs[m] : (const void *)dout;
    return run_hicoo(t, facs, dout, R, stride, out_stride, mats_order, mode, true, static_cast<hipStream_t>(stream),
                     "DevHiCOO MTTKRP BF16");
}

int ptiRegisterHiCOO(ptiSparseTensorHiCOO const *hitsr) {
    ptiDevHiCOO *t = nullptr;
    PTI_CHECK(upload(&t, hitsr, nullptr), "Register HiCOO");
    std::lock_guard<std::mutex> lk(g_reg_mu);
    auto it = g_hicoo_reg.find(hitsr);
    if (it != g_hicoo_reg.end()) delete it->second;
    g_hicoo_reg[hitsr] = t;
    return 0;
}

int ptiUnregisterHiCOO(ptiSparseTensorHiCOO const *hitsr) {
    std::lock_guard<std::mutex> lk(g_reg_mu);
    auto it = g_hicoo_reg.find(hitsr);
    if (it == g_hicoo_reg.end()) PTI_FAIL(PTIERR_VALUE_ERROR, "Unregister HiCOO", "tensor is not registered");
    delete it->second;
    g_hicoo_reg.erase(it);
    return 0;
}

int ptiMTTKRPHiCOO(ptiSparseTensorHiCOO const *const hitsr, ptiMatrix *mats[], ptiIndex const mats_order[],
                   ptiIndex const mode) {
    static const char *M = "GPU  HiCOO SpTns MTTKRP";
    const ptiIndex N = hitsr->nmodes;
    PTI_CHECK(check_mttkrp_args(N, hitsr->ndims, mats, mats_order, mode, M), M);
    PTI_CHECK(require_device(M), M);
    hipStream_t s = nullptr;
    ptiDevHiCOO *t = nullptr;
    ptiDevHiCOO *owned = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_reg_mu);
        auto it = g_hicoo_reg.find(hitsr);
        if (it != g_hicoo_reg.end()) t = it->second;
    }
    if (!t) {
        PTI_CHECK(upload(&owned, hitsr, s), M);
        t = owned;
    }
    StagedFactors sf;
    int rc = stage_factors(N, mats, mats_order, mode, sf, s, M);
    if (!rc) {
        const void *facs[8];
        for (uint32_t m = 0; m < N; ++m) facs[m] = sf.ptrs[m];
        rc = run_hicoo(t, facs, sf.ptrs[N], mats[N]->ncols, mats[0]->stride, mats[0]->stride, mats_order, mode,
                       false, s, M);
    }
    if (!rc) rc = fetch_output(N, mats, mode, sf, s, M);
    delete owned;
    return rc;
}

int ptiCudaMTTKRPHiCOO(ptiSparseTensorHiCOO const *const hitsr, ptiMatrix **const mats, ptiIndex *const mats_order,
                       ptiIndex const mode, ptiNnzIndex const max_nnzb, int const impl_num) {
    (void)max_nnzb;
    (void)impl_num;
    return ptiMTTKRPHiCOO(hitsr, mats, mats_order, mode);
}

}  // extern "C"
