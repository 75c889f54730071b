// HiCOO tensors resident in HBM and the HiCOO MTTKRP entry points.
//
// Device layout.  For every target mode m the upload keeps one copy of the
// HiCOO blocks, stably reordered so that blocks with the same mode-m block
// coordinate (a "block-row": 2^sb_bits output rows) are contiguous.  Block
// coordinates, 8-bit element offsets and values are the host tensor's.  The
// nonzeros of each block-row are cut into chunks of <= 256; inside a chunk
// they are ordered (stably) by their mode-m element offset, i.e. by output
// row, and each carries its block's index relative to the chunk's first
// block.  Per mode:
//   binds      u32[nb * N]       block coordinates, block-major, reordered
//   einds      u32[nnz * EW]     the N element offsets of a nonzero packed
//                                into EW = ceil(N/4) words
//   vals       f32[nnz]
//   lblk       u8[nnz]           block index - chunk_blk[chunk]  (< 256)
//   chunk_z    u32[nchunks + 1]  first nonzero of each chunk
//   chunk_blk  u32[nchunks]      first block of each chunk
//   unit_chunk u32[nunits + 1]   units: runs of <= UNIT_CHUNKS chunks of one
//   unit_br    u32[nunits]       block-row, each processed by one workgroup
//   unit_excl  u32[nunits]       1 if the block-row has a single unit
// N copies cost N times the memory of one (about 6 GB for a 100M-nonzero
// order-3 tensor with 73M blocks), which HBM holds easily; in exchange each
// MTTKRP streams its copy sequentially, reduces equal rows in registers and
// accumulates outputs in LDS.  The host HiCOO struct is not referenced after
// the upload.
#include <atomic>
#include <mutex>
#include <unordered_map>

#include "../../../include/hiparti_dev.h"
#include "../host/sorting.h"
#include "dropin.hpp"
#include "launch.hpp"

namespace {
constexpr uint32_t UNIT_CHUNKS = 64;  // chunks (of 256 nonzeros) per workgroup unit
}

struct ptiDevHiCOOMode {
    pti::dev::Buffer binds, einds, vals, lblk, chunk_z, chunk_blk, unit_chunk, unit_br, unit_excl;
    uint32_t nchunks = 0, nunits = 0;
    bool ready = false;
};

struct ptiDevHiCOO {
    uint32_t nmodes = 0;
    uint32_t sb_bits = 0;
    uint64_t nnz = 0;
    uint64_t nb = 0;
    ptiIndex ndims[8] = {0};
    ptiDevHiCOOMode mc[8];
    uint64_t bytes = 0;
};

namespace pti {
namespace dev {

HiCOOArgs hicoo_args(const ptiDevHiCOO *t, uint32_t mode) {
    const ptiDevHiCOOMode &c = t->mc[mode];
    HiCOOArgs h;
    h.binds = c.binds.as<uint32_t>();
    h.einds = c.einds.as<uint32_t>();
    h.vals = c.vals.as<float>();
    h.lblk = c.lblk.as<uint8_t>();
    h.chunk_z = c.chunk_z.as<uint32_t>();
    h.chunk_blk = c.chunk_blk.as<uint32_t>();
    h.unit_chunk = c.unit_chunk.as<uint32_t>();
    h.unit_br = c.unit_br.as<uint32_t>();
    h.unit_excl = c.unit_excl.as<uint32_t>();
    h.nunits = c.nunits;
    h.sb_bits = t->sb_bits;
    h.nrows = t->ndims[mode];
    return h;
}

namespace {

std::mutex g_reg_mu;
std::unordered_map<const void *, ptiDevHiCOO *> g_hicoo_reg;

int validate(const ptiSparseTensorHiCOO *hitsr, const char *M) {
    const uint32_t N = hitsr->nmodes;
    if (N < 2 || N > 8) PTI_FAIL(PTIERR_VALUE_ERROR, M, "nmodes must be in [2, 8]");
    if (hitsr->sb_bits > 7) PTI_FAIL(PTIERR_VALUE_ERROR, M, "sb_bits must be <= 7");
    const uint64_t nnz = hitsr->nnz;
    const uint64_t nb = hitsr->bptr.len ? hitsr->bptr.len - 1 : 0;
    if (nnz >= (1ull << 32) || nb >= (1ull << 32)) PTI_FAIL(PTIERR_VALUE_ERROR, M, "nnz must be < 2^32");
    if (hitsr->values.len < nnz || (nnz && nb == 0)) PTI_FAIL(PTIERR_VALUE_ERROR, M, "inconsistent HiCOO arrays");
    for (uint32_t m = 0; m < N; ++m)
        if (hitsr->einds[m].len < nnz || hitsr->binds[m].len < nb)
            PTI_FAIL(PTIERR_VALUE_ERROR, M, "inconsistent HiCOO arrays");
    if (nb && hitsr->bptr.data[nb] != nnz) PTI_FAIL(PTIERR_VALUE_ERROR, M, "bptr does not end at nnz");
    // The kernels index factor rows, output rows and LDS staging buffers with
    // these arrays: every block must be non-empty (a 256-nonzero chunk then
    // spans at most 257 blocks) and every coordinate must lie inside its mode.
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
    return 0;
}

// Builds and uploads the copy of target mode `mode`.
int build_mode(ptiDevHiCOO *t, const ptiSparseTensorHiCOO *hitsr, uint32_t mode, hipStream_t s, const char *M) {
    const uint32_t N = t->nmodes;
    const uint64_t nnz = t->nnz, nb = t->nb;
    const int EW = (N + 3) / 4;
    ptiDevHiCOOMode &c = t->mc[mode];

    // Stable order of the blocks by their mode-m block coordinate.
    std::vector<uint32_t> key(nb), ids(nb);
    uint32_t kmax = 0;
    for (uint64_t b = 0; b < nb; ++b) {
        key[b] = hitsr->binds[mode].data[b];
        ids[b] = (uint32_t)b;
        kmax = std::max(kmax, key[b]);
    }
    int kbits = 0;
    while (kbits < 32 && (kmax >> kbits)) ++kbits;
    pti::radix_sort_pairs<uint32_t>(key, ids, kbits, 0);

    // Reordered blocks: nonzero offsets (block order) and coordinates.
    std::vector<uint32_t> hb(nb + 1), hbi(nb * N);
    hb[0] = 0;
    for (uint64_t i = 0; i < nb; ++i) {
        const uint64_t b = ids[i];
        hb[i + 1] = hb[i] + (uint32_t)(hitsr->bptr.data[b + 1] - hitsr->bptr.data[b]);
    }
    pti::parallel_for(nb, 0, [&](uint64_t lo, uint64_t hi, int) {
        for (uint64_t i = lo; i < hi; ++i)
            for (uint32_t m = 0; m < N; ++m) hbi[i * N + m] = hitsr->binds[m].data[ids[i]];
    });

    // Chunks (never crossing a block-row) and units.
    std::vector<uint32_t> cz, cb, uc, ubr, uex;
    cz.reserve(nnz / pti::dev::CHUNK + 1024);
    cb.reserve(nnz / pti::dev::CHUNK + 1024);
    for (uint64_t i = 0; i < nb;) {
        const uint32_t br = key[i];
        uint64_t j = i;
        while (j < nb && key[j] == br) ++j;
        const uint64_t gz1 = hb[j];
        const uint64_t first = cz.size();
        uint64_t blk = i;
        for (uint64_t z = hb[i]; z < gz1; z += pti::dev::CHUNK) {
            while (hb[blk + 1] <= z) ++blk;
            cz.push_back((uint32_t)z);
            cb.push_back((uint32_t)blk);
        }
        const uint64_t nch = cz.size() - first;
        const uint64_t nu = (nch + UNIT_CHUNKS - 1) / UNIT_CHUNKS;
        for (uint64_t u = 0; u < nu; ++u) {
            uc.push_back((uint32_t)(first + u * nch / nu));
            ubr.push_back(br);
            uex.push_back(nu == 1 ? 1u : 0u);
        }
        i = j;
    }
    c.nchunks = (uint32_t)cz.size();
    c.nunits = (uint32_t)ubr.size();
    cz.push_back((uint32_t)nnz);
    uc.push_back(c.nchunks);

    // Nonzeros: per chunk, a stable counting sort on the mode-m element
    // offset; each nonzero records its block relative to the chunk's first.
    std::vector<uint32_t> he(nnz * EW, 0u);
    std::vector<float> hv(nnz);
    std::vector<uint8_t> hl(nnz);
    const uint32_t nrow = 1u << t->sb_bits;
    pti::parallel_for(c.nchunks, 0, [&](uint64_t lo, uint64_t hi, int) {
        std::vector<uint32_t> src(pti::dev::CHUNK), bl(pti::dev::CHUNK), cnt(nrow + 1);
        for (uint64_t ch = lo; ch < hi; ++ch) {
            const uint32_t z0 = cz[ch], n = cz[ch + 1] - z0;
            // Source nonzero and block (in block order) of each position.
            uint64_t blk = cb[ch];
            for (uint32_t k = 0; k < n; ++k) {
                while (hb[blk + 1] <= z0 + k) ++blk;
                src[k] = (uint32_t)(hitsr->bptr.data[ids[blk]] + (z0 + k - hb[blk]));
                bl[k] = (uint32_t)(blk - cb[ch]);
            }
            std::fill(cnt.begin(), cnt.end(), 0u);
            for (uint32_t k = 0; k < n; ++k) ++cnt[hitsr->einds[mode].data[src[k]] + 1];
            for (uint32_t r = 0; r < nrow; ++r) cnt[r + 1] += cnt[r];
            for (uint32_t k = 0; k < n; ++k) {
                const uint32_t sz = src[k];
                const uint64_t z = z0 + cnt[hitsr->einds[mode].data[sz]]++;
                for (uint32_t m = 0; m < N; ++m)
                    he[z * EW + (m >> 2)] |= (uint32_t)hitsr->einds[m].data[sz] << ((m & 3) * 8);
                hv[z] = hitsr->values.data[sz];
                hl[z] = (uint8_t)bl[k];
            }
        }
    });

    auto put = [&](Buffer &buf, const void *src, size_t bytes) -> int {
        PTI_HIP(buf.alloc(bytes), M);
        if (bytes) PTI_HIP(hipMemcpyAsync(buf.ptr, src, bytes, hipMemcpyHostToDevice, s), M);
        t->bytes += bytes;
        return 0;
    };
    int rc = 0;
    if (!rc) rc = put(c.binds, hbi.data(), hbi.size() * 4);
    if (!rc) rc = put(c.einds, he.data(), he.size() * 4);
    if (!rc) rc = put(c.vals, hv.data(), hv.size() * 4);
    if (!rc) rc = put(c.lblk, hl.data(), hl.size());
    if (!rc) rc = put(c.chunk_z, cz.data(), cz.size() * 4);
    if (!rc) rc = put(c.chunk_blk, cb.data(), cb.size() * 4);
    if (!rc) rc = put(c.unit_chunk, uc.data(), uc.size() * 4);
    if (!rc) rc = put(c.unit_br, ubr.data(), ubr.size() * 4);
    if (!rc) rc = put(c.unit_excl, uex.data(), uex.size() * 4);
    // The copies read the host vectors above: finish them before they go.
    hipError_t e = hipStreamSynchronize(s);
    if (!rc && e != hipSuccess) {
        pti_ComplainError(M, hip_code(e), __FILE__, __LINE__, hipGetErrorString(e));
        rc = hip_code(e);
    }
    if (!rc) c.ready = true;
    return rc;
}

// Upload the copies of the modes in `mode_mask` (bit m = target mode m).
int upload(ptiDevHiCOO **out, const ptiSparseTensorHiCOO *hitsr, uint32_t mode_mask, hipStream_t s) {
    static const char *M = "DevHiCOO Upload";
    PTI_CHECK(require_device(M), M);
    PTI_CHECK(validate(hitsr, M), M);
    ptiDevHiCOO *t = new ptiDevHiCOO();
    t->nmodes = hitsr->nmodes;
    t->sb_bits = hitsr->sb_bits;
    t->nnz = hitsr->nnz;
    t->nb = hitsr->nnz ? hitsr->bptr.len - 1 : 0;
    for (uint32_t m = 0; m < t->nmodes; ++m) t->ndims[m] = hitsr->ndims[m];
    for (uint32_t m = 0; m < t->nmodes; ++m) {
        if (!(mode_mask >> m & 1u)) continue;
        int rc = build_mode(t, hitsr, m, s, M);
        if (rc) {
            delete t;
            return rc;
        }
    }
    *out = t;
    return 0;
}

}  // namespace

// Shared by the FP32 and BF16 entry points: zero the output rows, then run the
// tile kernel once per 64-rank column tile.
int run_hicoo(const ptiDevHiCOO *t, const void *const facs[], float *out, ptiIndex R, ptiIndex stride,
              ptiIndex out_stride, const ptiIndex mats_order[], ptiIndex mode, bool bf16, hipStream_t s,
              const char *M) {
    const uint32_t N = t->nmodes;
    if (mode >= N) PTI_FAIL(PTIERR_VALUE_ERROR, M, "mode out of range");
    if (R == 0 || stride < R || out_stride < R) PTI_FAIL(PTIERR_SHAPE_MISMATCH, M, "bad rank or stride");
    for (uint32_t i = 1; i < N; ++i)
        if (mats_order[i] >= N) PTI_FAIL(PTIERR_VALUE_ERROR, M, "mats_order entry out of range");
    if (t->nnz && !t->mc[mode].ready) PTI_FAIL(PTIERR_VALUE_ERROR, M, "mode was not uploaded");
    PTI_HIP(zero_floats(out, (uint64_t)t->ndims[mode] * out_stride, s), M);
    if (t->nnz == 0) return 0;
    // 64 ranks per LDS tile; 32 for orders >= 6, whose larger staging area
    // would otherwise push a workgroup past 64 KB of LDS.
    const ptiIndex tile = N >= 6 ? 32 : 64;
    const size_t esz = bf16 ? 2 : 4;
    for (ptiIndex c0 = 0; c0 < R; c0 += tile) {
        const ptiIndex Rt = std::min<ptiIndex>(tile, R - c0);
        // Column tiles start at multiples of 32 elements, so they keep the
        // alignment the vector width V relies on.
        Shape sh;
        if (!choose_shape(Rt, stride, out_stride, sh)) PTI_FAIL(PTIERR_VALUE_ERROR, M, "unsupported rank tile");
        // The kernel stages factor-row offsets as 32-bit counts of V-wide vectors.
        for (uint32_t i = 1; i < N; ++i)
            if ((uint64_t)t->ndims[mats_order[i]] * (stride / sh.V) >= (1ull << 32))
                PTI_FAIL(PTIERR_VALUE_ERROR, M, "factor matrix exceeds 2^32 row-vector slots");
        FactorArgs a;
        for (uint32_t m = 0; m < N; ++m) a.fac[m] = static_cast<const char *>(facs[m]) + (size_t)c0 * esz;
        a.out = out + c0;
        a.R = Rt;
        a.stride = stride;
        a.out_stride = out_stride;
        a.mode = mode;
        for (uint32_t i = 0; i < N; ++i) a.order[i] = mats_order[i];
        PTI_HIP(launch_hicoo(N, hicoo_args(t, mode), a, sh, bf16, s), M);
    }
    return 0;
}

// Shape and values of a resident tensor, for callers that run several
// kernels over it (CP-ALS).  vals is NULL when no mode copy is resident.
void hicoo_view(const ptiDevHiCOO *t, uint32_t *nmodes, const ptiIndex **ndims, uint64_t *nnz, const float **vals) {
    *nmodes = t->nmodes;
    *ndims = t->ndims;
    *nnz = t->nnz;
    *vals = nullptr;
    for (uint32_t m = 0; m < t->nmodes && !*vals; ++m)
        if (t->mc[m].ready) *vals = t->mc[m].vals.as<float>();
}

}  // namespace dev
}  // namespace pti

using namespace pti::dev;

extern "C" {

int ptiDevHiCOOUpload(ptiDevHiCOO **out, ptiSparseTensorHiCOO const *hitsr, void *stream) {
    return upload(out, hitsr, 0xFFu, static_cast<hipStream_t>(stream));
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
    PTI_CHECK(upload(&t, hitsr, 0xFFu, nullptr), "Register HiCOO");
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
        // Unregistered tensors are uploaded for this call only, and only the
        // copy of the requested mode is built.
        PTI_CHECK(upload(&owned, hitsr, 1u << mode, s), M);
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
