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
//
// Each mode copy also gets a stream form on its first MTTKRP (ensure_stream):
// the nonzeros of each block-row sorted by output row and decoded into
// (N + 1)-word entries (four-word packed entries for orders 4 and 5 where
// they fit) laid out for the lane groups of the rank shape
// (StreamShape in mttkrp_kernels.hpp), 4 KB per chunk and 16-byte plane, read
// by hicoo_stream_kernel.
#include <algorithm>
#include <atomic>
#include <mutex>
#include <unordered_map>

#include "../../../include/hiparti_dev.h"
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "dropin.hpp"
#include "launch.hpp"

namespace {
struct DimArgs {
    uint32_t d[8];
};
}

struct ptiDevHiCOOMode {
    pti::dev::Buffer binds, einds, vals, lblk, chunk_z, chunk_blk, unit_chunk, unit_br, unit_excl;
    uint32_t nchunks = 0, nunits = 0;
    bool ready = false;
    // Stream forms of the copy (16 bytes per chunk position and plane), each
    // built on first use for the lane-group count `ng` of a rank shape.  Two
    // are kept, so that the FP32 and the wider BF16 lane shape of one rank
    // (CP-ALS alternates them) do not rebuild each other; a third shape
    // replaces the least recently used one.
    struct StreamForm {
        pti::dev::Buffer ent;
        uint32_t ng = 0;
        uint64_t last_use = 0;
        bool packed = false;   // one plane of packed entries (StreamShape<N, true>)
    };
    mutable std::mutex stream_mu;
    mutable StreamForm forms[2];
    mutable uint64_t use_clock = 0;
    mutable pti::dev::Buffer unit_order;   // units by decreasing chunk count
    // One byte per block-row: 1 when a single unit owns it (it then writes
    // the block-row's output rows whole), built on first use.
    mutable pti::dev::Buffer br_owned;
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

// The host tensor's arrays in device memory, shared by the per-mode builds.
struct RawHiCOO {
    Buffer bptr, vals, binds[8], einds[8];
};

struct Ptr8u32 {
    const uint32_t *p[8];
};
struct Ptr8u8 {
    const uint8_t *p[8];
};

inline dim3 grid_n(uint64_t n) {
    uint64_t b = (n + 255) / 256;
    if (b > 65536) b = 65536;
    if (b == 0) b = 1;
    return dim3((uint32_t)b);
}

// Every block must be non-empty and inside [0, nnz), every element offset
// below 2^sb_bits and every coordinate inside its mode: the kernels use them
// as memory and LDS indices.
__global__ void __launch_bounds__(256) validate_blocks_kernel(const uint64_t *__restrict__ bptr, uint64_t nb,
                                                              uint64_t nnz, Ptr8u32 binds, Ptr8u8 einds, uint32_t N,
                                                              uint32_t sb, DimArgs dims, int *bad) {
    for (uint64_t b = (uint64_t)blockIdx.x * 256 + threadIdx.x; b < nb; b += (uint64_t)gridDim.x * 256) {
        const uint64_t z0 = bptr[b], z1 = bptr[b + 1];
        if (z1 <= z0 || z1 > nnz) {
            atomicOr(bad, 1);
            continue;
        }
        for (uint32_t m = 0; m < N; ++m) {
            const uint64_t base = (uint64_t)binds.p[m][b] << sb;
            uint32_t emax = 0;
            for (uint64_t z = z0; z < z1; ++z) emax = max(emax, (uint32_t)einds.p[m][z]);
            if (emax >> sb || base + emax >= dims.d[m]) atomicOr(bad, 1);
        }
    }
}

__global__ void __launch_bounds__(256) iota_u32_kernel(uint32_t *p, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256)
        p[i] = (uint32_t)i;
}

__global__ void __launch_bounds__(256) max_u32_kernel(const uint32_t *p, uint64_t n, uint32_t *out) {
    uint32_t v = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256)
        v = max(v, p[i]);
    for (int off = 32; off > 0; off >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out, v);
}

// Per reordered block i: its size, its coordinates, and whether it starts a
// block-row (a new mode-m block coordinate).
__global__ void __launch_bounds__(256) block_info_kernel(const uint64_t *__restrict__ bptr,
                                                         const uint32_t *__restrict__ skey,
                                                         const uint32_t *__restrict__ sid, uint64_t nb, Ptr8u32 binds,
                                                         uint32_t N, uint32_t *__restrict__ bsz,
                                                         uint32_t *__restrict__ hbi, uint32_t *__restrict__ segflag) {
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < nb; i += (uint64_t)gridDim.x * 256) {
        const uint32_t b = sid[i];
        bsz[i] = (uint32_t)(bptr[b + 1] - bptr[b]);
        for (uint32_t m = 0; m < N; ++m) hbi[i * N + m] = binds.p[m][b];
        segflag[i] = (i == 0 || skey[i] != skey[i - 1]) ? 1u : 0u;
    }
}

// seg_first[s] = first reordered block of block-row s.
__global__ void __launch_bounds__(256) segment_starts_kernel(const uint32_t *__restrict__ segflag,
                                                             const uint32_t *__restrict__ segid, uint64_t nb,
                                                             uint32_t *__restrict__ seg_first) {
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < nb; i += (uint64_t)gridDim.x * 256)
        if (segflag[i]) seg_first[segid[i] - 1] = (uint32_t)i;
}

// Chunks and units per block-row: nch = ceil(nnz / CHUNK), nu = ceil(nch / UNIT_CHUNKS).
__global__ void __launch_bounds__(256) segment_counts_kernel(const uint32_t *__restrict__ seg_first,
                                                             const uint32_t *__restrict__ hb, uint32_t nseg,
                                                             uint32_t unit_chunks, uint32_t *__restrict__ nch,
                                                             uint32_t *__restrict__ nu) {
    for (uint32_t s = blockIdx.x * 256 + threadIdx.x; s < nseg; s += gridDim.x * 256) {
        const uint32_t len = hb[seg_first[s + 1]] - hb[seg_first[s]];
        const uint32_t c = (len + CHUNK - 1) / CHUNK;
        nch[s] = c;
        nu[s] = (c + unit_chunks - 1) / unit_chunks;
    }
}

__device__ inline uint32_t upper_bound_u32(const uint32_t *a, uint32_t lo, uint32_t hi, uint32_t v) {
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (a[mid] <= v)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

__global__ void __launch_bounds__(256) chunk_fill_kernel(const uint32_t *__restrict__ choff,
                                                         const uint32_t *__restrict__ seg_first,
                                                         const uint32_t *__restrict__ hb, uint32_t nseg,
                                                         uint32_t nchunks, uint32_t *__restrict__ chunk_z,
                                                         uint32_t *__restrict__ chunk_blk) {
    for (uint32_t c = blockIdx.x * 256 + threadIdx.x; c < nchunks; c += gridDim.x * 256) {
        const uint32_t s = upper_bound_u32(choff, 0, nseg + 1, c) - 1;
        const uint32_t first = seg_first[s], next = seg_first[s + 1];
        const uint32_t z = hb[first] + (c - choff[s]) * CHUNK;
        chunk_z[c] = z;
        chunk_blk[c] = upper_bound_u32(hb, first, next, z) - 1;
    }
}

__global__ void __launch_bounds__(256) unit_fill_kernel(const uint32_t *__restrict__ uoff,
                                                        const uint32_t *__restrict__ choff,
                                                        const uint32_t *__restrict__ seg_first,
                                                        const uint32_t *__restrict__ skey, uint32_t nseg,
                                                        uint32_t nunits, uint32_t *__restrict__ unit_chunk,
                                                        uint32_t *__restrict__ unit_br,
                                                        uint32_t *__restrict__ unit_excl) {
    for (uint32_t U = blockIdx.x * 256 + threadIdx.x; U < nunits; U += gridDim.x * 256) {
        const uint32_t s = upper_bound_u32(uoff, 0, nseg + 1, U) - 1;
        const uint64_t u = U - uoff[s], nu = uoff[s + 1] - uoff[s], nch = choff[s + 1] - choff[s];
        unit_chunk[U] = choff[s] + (uint32_t)(u * nch / nu);
        unit_br[U] = skey[seg_first[s]];
        unit_excl[U] = nu == 1 ? 1u : 0u;
    }
}

// One wave per chunk: the chunk's nonzeros (in reordered block order) are
// stably sorted by their mode-m element offset with a counting sort whose
// ranks come from ballots, and written out with their packed offsets, value
// and block tag.
__global__ void __launch_bounds__(256) chunk_sort_kernel(const uint64_t *__restrict__ bptr,
                                                         const uint32_t *__restrict__ sid,
                                                         const uint32_t *__restrict__ hb, uint64_t nb,
                                                         const uint32_t *__restrict__ chunk_z,
                                                         const uint32_t *__restrict__ chunk_blk, uint32_t nchunks,
                                                         Ptr8u8 einds, const float *__restrict__ vals, uint32_t N,
                                                         uint32_t mode, uint32_t sb, uint32_t *__restrict__ he,
                                                         float *__restrict__ hv, uint8_t *__restrict__ hl) {
    __shared__ uint32_t cnt_all[4][129];
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t c = blockIdx.x * 4 + wid;
    if (c >= nchunks) return;   // whole wave
    uint32_t *cnt = cnt_all[wid];
    const uint32_t nrow = 1u << sb;
    const uint32_t EW = (N + 3) / 4;
    const uint32_t z0 = chunk_z[c], n = chunk_z[c + 1] - z0, cb = chunk_blk[c];
    const uint32_t hi = (uint32_t)min<uint64_t>((uint64_t)cb + CHUNK + 1, nb + 1);
    for (uint32_t r = lane; r <= nrow; r += 64) cnt[r] = 0;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    uint32_t key[4], src[4], bl[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const uint32_t k = r * 64 + lane;
        key[r] = 0;
        src[r] = 0;
        bl[r] = 0;
        if (k < n) {
            const uint32_t z = z0 + k;
            const uint32_t blk = upper_bound_u32(hb, cb, hi, z) - 1;
            src[r] = (uint32_t)(bptr[sid[blk]] + (z - hb[blk]));
            bl[r] = blk - cb;
            key[r] = einds.p[mode][src[r]];
            atomicAdd(&cnt[key[r] + 1], 1u);
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (lane == 0)
        for (uint32_t v = 0; v < nrow; ++v) cnt[v + 1] += cnt[v];   // cnt[v] = first slot of key v
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const uint32_t k = r * 64 + lane;
        const bool valid = k < n;
        uint64_t same = __ballot(valid);
        for (uint32_t bit = 0; bit < sb; ++bit) {
            const bool one = (key[r] >> bit) & 1u;
            const uint64_t ones = __ballot(valid && one);
            same &= one ? ones : ~ones;
        }
        const uint32_t rank = (uint32_t)__popcll(same & below);
        uint32_t dest = 0;
        if (valid) dest = cnt[key[r]] + rank;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (valid && rank == 0) cnt[key[r]] += (uint32_t)__popcll(same);
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (valid) {
            const uint64_t z = (uint64_t)z0 + dest;
            for (uint32_t w = 0; w < EW; ++w) {
                uint32_t word = 0;
                for (uint32_t m = 4 * w; m < N && m < 4 * w + 4; ++m)
                    word |= (uint32_t)einds.p[m][src[r]] << ((m & 3) * 8);
                he[z * EW + w] = word;
            }
            hv[z] = vals[src[r]];
            hl[z] = (uint8_t)bl[r];
        }
    }
}

// Stream-form build, step 1 (one wave per chunk): for each nonzero of the
// mode copy its block and its sort key: output row (the target-mode
// coordinate) in the high word, the coordinate of mode `sec` in the low word;
// or, with wbits > 0 (windowed order), (block-row, window of `sec`'s
// coordinate, row inside the block-row, the coordinate's low WINDOW_BITS).
__global__ void __launch_bounds__(256) stream_keys_kernel(HiCOOArgs h, uint32_t nchunks, uint32_t N, uint32_t mode,
                                                          uint32_t sec, uint32_t wbits, uint64_t *__restrict__ key,
                                                          uint32_t *__restrict__ pos, uint32_t *__restrict__ blk_of) {
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t c = blockIdx.x * 4 + wid;
    if (c >= nchunks) return;   // whole wave
    const uint32_t EW = (N + 3) / 4;
    const uint32_t z0 = h.chunk_z[c], cnt = h.chunk_z[c + 1] - z0, b0 = h.chunk_blk[c];
    for (uint32_t zl = lane; zl < cnt; zl += 64) {
        const uint32_t z = z0 + zl;
        const uint32_t b = b0 + h.lblk[z];
        blk_of[z] = b;
        auto coord = [&](uint32_t m) {
            return (h.binds[(uint64_t)b * N + m] << h.sb_bits) +
                   ((h.einds[(uint64_t)z * EW + m / 4] >> ((m % 4) * 8)) & 0xFFu);
        };
        if (wbits) {
            const uint32_t row = coord(mode), x = coord(sec), sb = h.sb_bits;
            key[z] = ((uint64_t)(row >> sb) << (wbits + sb + WINDOW_BITS)) |
                     ((uint64_t)(x >> WINDOW_BITS) << (sb + WINDOW_BITS)) |
                     ((uint64_t)(row & ((1u << sb) - 1u)) << WINDOW_BITS) | (x & ((1u << WINDOW_BITS) - 1u));
        } else {
            key[z] = ((uint64_t)coord(mode) << 32) | coord(sec);
        }
        pos[z] = z;
    }
}

// Windowed order: nonzeros of the mode copy per window of mode m's
// coordinate (counts zeroed by the caller); one wave per chunk.
__global__ void __launch_bounds__(256) window_hist_kernel(HiCOOArgs h, uint32_t nchunks, uint32_t N, uint32_t m,
                                                          uint32_t *__restrict__ counts) {
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t c = blockIdx.x * 4 + wid;
    if (c >= nchunks) return;   // whole wave
    const uint32_t EW = (N + 3) / 4;
    const uint32_t z0 = h.chunk_z[c], cnt = h.chunk_z[c + 1] - z0, b0 = h.chunk_blk[c];
    for (uint32_t zb = 0; zb < cnt; zb += 64) {   // the same iterations in every lane
        const uint32_t zl = zb + lane;
        uint32_t x = 0;
        if (zl < cnt) {
            const uint32_t z = z0 + zl;
            const uint32_t b = b0 + h.lblk[z];
            x = (h.binds[(uint64_t)b * N + m] << h.sb_bits) +
                ((h.einds[(uint64_t)z * EW + m / 4] >> ((m % 4) * 8)) & 0xFFu);
        }
        wave_count(counts, x >> WINDOW_BITS, zl < cnt);
    }
}

// Step 2, after a stable sort of the nonzeros by (output row, coordinate of
// the largest non-target mode): each block-row keeps its place, since
// block-rows are in ascending block coordinate, so the chunk boundaries are
// unchanged.  Writes the entries of chunk c for NG lane groups in the layout
// of StreamShape (one wave per chunk).
// A chunk of at most INTERLEAVE_MAX_RUNS row runs is interleaved: slice g
// takes its sorted entries g, g + ng, g + 2 ng, ...  packed (orders 4 and 5):
// one plane of four words, StreamShape<N, true>.
__global__ void __launch_bounds__(256) stream_fill_kernel(HiCOOArgs h, uint32_t nchunks, uint32_t N, uint32_t mode,
                                                          uint32_t ng, uint32_t packed,
                                                          const uint32_t *__restrict__ spos,
                                                          const uint32_t *__restrict__ blk_of,
                                                          uint4 *__restrict__ ent) {
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t c = blockIdx.x * 4 + wid;
    if (c >= nchunks) return;   // whole wave
    const uint32_t EW = (N + 3) / 4, NP = packed ? 1u : (N + 4) / 4;
    const uint32_t z0 = h.chunk_z[c], cnt = h.chunk_z[c + 1] - z0;
    const uint32_t per = (cnt + ng - 1) / ng;
    const uint32_t G = 64 / ng;
    auto row_of = [&](uint32_t z) { return (h.einds[(uint64_t)z * EW + mode / 4] >> ((mode % 4) * 8)) & 0xFFu; };
    uint32_t breaks = 0;   // row changes between neighbouring sorted entries (the same in every lane)
    for (uint32_t kb = 1; kb < cnt; kb += 64) {
        const uint32_t k = kb + lane;
        const bool brk = k < cnt && row_of(spos[z0 + k]) != row_of(spos[z0 + k - 1]);
        breaks += (uint32_t)__builtin_popcountll(__builtin_amdgcn_ballot_w64(brk));
    }
    const bool interleave = breaks < INTERLEAVE_MAX_RUNS;
    for (uint32_t e = lane; e < CHUNK; e += 64) {
        const uint32_t g = (e % 64) / G, j = (e / 64) * G + e % G;   // e == stream_entry(G, g, j)
        uint32_t w[12];
        for (uint32_t k = 0; k < 12; ++k) w[k] = 0u;
        w[N] = DUMMY_ROW;
        const uint32_t k = interleave ? j * ng + g : g * per + j;
        if (j < per && k < cnt) {
            const uint32_t z = spos[z0 + k];
            const uint64_t b = blk_of[z];
            auto elem = [&](uint32_t m) { return (h.einds[(uint64_t)z * EW + m / 4] >> ((m % 4) * 8)) & 0xFFu; };
            for (uint32_t i = 0; i + 1 < N; ++i) {
                const uint32_t m = i < mode ? i : i + 1;
                w[i] = (h.binds[b * N + m] << h.sb_bits) + elem(m);
            }
            w[N - 1] = __float_as_uint(h.vals[z]);
            w[N] = elem(mode);
        }
        if (packed) {
            // {c0 | row << 24, c1, c2, value}; order 5 adds c3's low byte to
            // word 1 and its high byte to word 2.
            const uint32_t p1 = N == 5 ? w[1] | (w[3] << 24) : w[1];
            const uint32_t p2 = N == 5 ? w[2] | ((w[3] >> 8) << 24) : w[2];
            ent[(uint64_t)c * CHUNK + e] = make_uint4(w[0] | (w[N] << 24), p1, p2, w[N - 1]);
            continue;
        }
        for (uint32_t q = 0; q < NP; ++q)
            ent[((uint64_t)c * NP + q) * CHUNK + e] = make_uint4(w[4 * q], w[4 * q + 1], w[4 * q + 2], w[4 * q + 3]);
    }
}

__global__ void __launch_bounds__(256) unit_size_kernel(const uint32_t *__restrict__ unit_chunk, uint32_t n,
                                                        uint32_t *__restrict__ sz) {
    for (uint32_t u = blockIdx.x * 256 + threadIdx.x; u < n; u += gridDim.x * 256)
        sz[u] = unit_chunk[u + 1] - unit_chunk[u];
}

// Row blocks of nonzeros sorted by output row: flag[z] = 1 where row >> sb
// changes.
__global__ void __launch_bounds__(256) row_block_flag_kernel(const uint32_t *__restrict__ rows, uint64_t n, uint32_t sb,
                                                             uint32_t *__restrict__ flag) {
    for (uint64_t z = (uint64_t)blockIdx.x * 256 + threadIdx.x; z < n; z += (uint64_t)gridDim.x * 256)
        flag[z] = (z == 0 || (rows[z] >> sb) != (rows[z - 1] >> sb)) ? 1u : 0u;
}

// Block s starts at nonzero start[s] and holds rows key[s] << sb ...
__global__ void __launch_bounds__(256) row_block_start_kernel(const uint32_t *__restrict__ rows, uint64_t n,
                                                              uint32_t sb, const uint32_t *__restrict__ flag,
                                                              const uint32_t *__restrict__ segid,
                                                              uint32_t *__restrict__ start, uint32_t *__restrict__ key) {
    for (uint64_t z = (uint64_t)blockIdx.x * 256 + threadIdx.x; z < n; z += (uint64_t)gridDim.x * 256)
        if (flag[z]) {
            start[segid[z] - 1] = (uint32_t)z;
            key[segid[z] - 1] = rows[z] >> sb;
        }
}

template <typename T>
int dev_put(Buffer &b, const T *src, size_t n, hipStream_t s, const char *M) {
    PTI_HIP(b.alloc(n * sizeof(T)), M);
    if (n) PTI_HIP(hipMemcpyAsync(b.ptr, src, n * sizeof(T), hipMemcpyHostToDevice, s), M);
    return 0;
}

// Copies the host tensor's arrays to the device and checks them there.
int raw_upload(RawHiCOO &r, const ptiSparseTensorHiCOO *hitsr, hipStream_t s, const char *M) {
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
    if (!nnz) return 0;
    if (hitsr->bptr.data[0] != 0 || hitsr->bptr.data[nb] != nnz)
        PTI_FAIL(PTIERR_VALUE_ERROR, M, "bptr must start at 0 and end at nnz");
    PTI_CHECK(dev_put(r.bptr, hitsr->bptr.data, nb + 1, s, M), M);
    PTI_CHECK(dev_put(r.vals, hitsr->values.data, nnz, s, M), M);
    Ptr8u32 bp{};
    Ptr8u8 ep{};
    DimArgs dims{};
    for (uint32_t m = 0; m < N; ++m) {
        PTI_CHECK(dev_put(r.binds[m], hitsr->binds[m].data, nb, s, M), M);
        PTI_CHECK(dev_put(r.einds[m], hitsr->einds[m].data, nnz, s, M), M);
        bp.p[m] = r.binds[m].as<uint32_t>();
        ep.p[m] = r.einds[m].as<uint8_t>();
        dims.d[m] = hitsr->ndims[m];
    }
    Buffer bad;
    PTI_HIP(bad.alloc(sizeof(int)), M);
    PTI_HIP(hipMemsetAsync(bad.ptr, 0, sizeof(int), s), M);
    hipLaunchKernelGGL(validate_blocks_kernel, grid_n(nb), dim3(256), 0, s, r.bptr.as<uint64_t>(), nb, nnz, bp, ep, N,
                       (uint32_t)hitsr->sb_bits, dims, bad.as<int>());
    PTI_HIP(hipGetLastError(), M);
    int h_bad = 0;
    PTI_HIP(hipMemcpyAsync(&h_bad, bad.ptr, sizeof(int), hipMemcpyDeviceToHost, s), M);
    PTI_HIP(hipStreamSynchronize(s), M);
    if (h_bad) PTI_FAIL(PTIERR_VALUE_ERROR, M, "empty block or coordinate outside its mode");
    return 0;
}

// Builds the copy of target mode `mode` on the device from the raw arrays.
int build_mode(ptiDevHiCOO *t, const RawHiCOO &r, uint32_t mode, hipStream_t s, const char *M) {
    const uint32_t N = t->nmodes;
    const uint64_t nnz = t->nnz, nb = t->nb;
    const uint32_t EW = (N + 3) / 4;
    ptiDevHiCOOMode &c = t->mc[mode];
    Ptr8u32 bp{};
    Ptr8u8 ep{};
    for (uint32_t m = 0; m < N; ++m) {
        bp.p[m] = r.binds[m].as<uint32_t>();
        ep.p[m] = r.einds[m].as<uint8_t>();
    }
    // 1. Stable order of the blocks by their mode-m block coordinate.
    Buffer kmax, ids, skey, sid, tmp_sort, tmp_scan, tmp_seg;
    PTI_HIP(kmax.alloc(4), M);
    PTI_HIP(hipMemsetAsync(kmax.ptr, 0, 4, s), M);
    hipLaunchKernelGGL(max_u32_kernel, grid_n(nb), dim3(256), 0, s, bp.p[mode], nb, kmax.as<uint32_t>());
    PTI_HIP(hipGetLastError(), M);
    uint32_t h_kmax = 0;
    PTI_HIP(hipMemcpyAsync(&h_kmax, kmax.ptr, 4, hipMemcpyDeviceToHost, s), M);
    PTI_HIP(hipStreamSynchronize(s), M);
    unsigned kbits = 0;
    while (kbits < 32 && (h_kmax >> kbits)) ++kbits;
    PTI_HIP(ids.alloc(nb * 4), M);
    PTI_HIP(skey.alloc(nb * 4), M);
    PTI_HIP(sid.alloc(nb * 4), M);
    hipLaunchKernelGGL(iota_u32_kernel, grid_n(nb), dim3(256), 0, s, ids.as<uint32_t>(), nb);
    PTI_HIP(hipGetLastError(), M);
    size_t tb = 0;
    PTI_HIP(rocprim::radix_sort_pairs(nullptr, tb, bp.p[mode], skey.as<uint32_t>(), ids.as<uint32_t>(),
                                      sid.as<uint32_t>(), (size_t)nb, 0, std::max(kbits, 1u), s),
            M);
    PTI_HIP(tmp_sort.alloc(tb), M);
    PTI_HIP(rocprim::radix_sort_pairs(tmp_sort.ptr, tb, bp.p[mode], skey.as<uint32_t>(), ids.as<uint32_t>(),
                                      sid.as<uint32_t>(), (size_t)nb, 0, std::max(kbits, 1u), s),
            M);
    // 2. Block sizes -> nonzero offsets in the new order; coordinates; block-rows.
    Buffer bsz, hb, segflag, segid;
    PTI_HIP(bsz.alloc(nb * 4), M);
    PTI_HIP(hb.alloc((nb + 1) * 4), M);
    PTI_HIP(segflag.alloc(nb * 4), M);
    PTI_HIP(segid.alloc(nb * 4), M);
    PTI_HIP(c.binds.alloc(nb * N * 4), M);
    hipLaunchKernelGGL(block_info_kernel, grid_n(nb), dim3(256), 0, s, r.bptr.as<uint64_t>(), skey.as<uint32_t>(),
                       sid.as<uint32_t>(), nb, bp, N, bsz.as<uint32_t>(), c.binds.as<uint32_t>(),
                       segflag.as<uint32_t>());
    PTI_HIP(hipGetLastError(), M);
    size_t sb1 = 0, sb2 = 0;
    PTI_HIP(rocprim::exclusive_scan(nullptr, sb1, bsz.as<uint32_t>(), hb.as<uint32_t>(), 0u, (size_t)nb,
                                    rocprim::plus<uint32_t>(), s),
            M);
    PTI_HIP(rocprim::inclusive_scan(nullptr, sb2, segflag.as<uint32_t>(), segid.as<uint32_t>(), (size_t)nb,
                                    rocprim::plus<uint32_t>(), s),
            M);
    PTI_HIP(tmp_scan.alloc(std::max(sb1, sb2)), M);
    PTI_HIP(rocprim::exclusive_scan(tmp_scan.ptr, sb1, bsz.as<uint32_t>(), hb.as<uint32_t>(), 0u, (size_t)nb,
                                    rocprim::plus<uint32_t>(), s),
            M);
    const uint32_t nnz32 = (uint32_t)nnz;
    PTI_HIP(hipMemcpyAsync(hb.as<uint32_t>() + nb, &nnz32, 4, hipMemcpyHostToDevice, s), M);
    PTI_HIP(rocprim::inclusive_scan(tmp_scan.ptr, sb2, segflag.as<uint32_t>(), segid.as<uint32_t>(), (size_t)nb,
                                    rocprim::plus<uint32_t>(), s),
            M);
    uint32_t nseg = 0;
    PTI_HIP(hipMemcpyAsync(&nseg, segid.as<uint32_t>() + nb - 1, 4, hipMemcpyDeviceToHost, s), M);
    PTI_HIP(hipStreamSynchronize(s), M);
    Buffer seg_first, nch, nu, choff, uoff;
    PTI_HIP(seg_first.alloc((size_t)(nseg + 1) * 4), M);
    const uint32_t nb32 = (uint32_t)nb;
    PTI_HIP(hipMemcpyAsync(seg_first.as<uint32_t>() + nseg, &nb32, 4, hipMemcpyHostToDevice, s), M);
    hipLaunchKernelGGL(segment_starts_kernel, grid_n(nb), dim3(256), 0, s, segflag.as<uint32_t>(),
                       segid.as<uint32_t>(), nb, seg_first.as<uint32_t>());
    PTI_HIP(hipGetLastError(), M);
    // 3. Chunks and units per block-row.
    PTI_HIP(nch.alloc((size_t)(nseg + 1) * 4), M);
    PTI_HIP(nu.alloc((size_t)(nseg + 1) * 4), M);
    PTI_HIP(choff.alloc((size_t)(nseg + 1) * 4), M);
    PTI_HIP(uoff.alloc((size_t)(nseg + 1) * 4), M);
    hipLaunchKernelGGL(segment_counts_kernel, grid_n(nseg), dim3(256), 0, s, seg_first.as<uint32_t>(),
                       hb.as<uint32_t>(), nseg, UNIT_CHUNKS, nch.as<uint32_t>(), nu.as<uint32_t>());
    PTI_HIP(hipGetLastError(), M);
    const uint32_t zero = 0;
    PTI_HIP(hipMemcpyAsync(nch.as<uint32_t>() + nseg, &zero, 4, hipMemcpyHostToDevice, s), M);
    PTI_HIP(hipMemcpyAsync(nu.as<uint32_t>() + nseg, &zero, 4, hipMemcpyHostToDevice, s), M);
    size_t sb3 = 0;
    PTI_HIP(rocprim::exclusive_scan(nullptr, sb3, nch.as<uint32_t>(), choff.as<uint32_t>(), 0u, (size_t)nseg + 1,
                                    rocprim::plus<uint32_t>(), s),
            M);
    PTI_HIP(tmp_seg.alloc(sb3), M);
    PTI_HIP(rocprim::exclusive_scan(tmp_seg.ptr, sb3, nch.as<uint32_t>(), choff.as<uint32_t>(), 0u, (size_t)nseg + 1,
                                    rocprim::plus<uint32_t>(), s),
            M);
    PTI_HIP(rocprim::exclusive_scan(tmp_seg.ptr, sb3, nu.as<uint32_t>(), uoff.as<uint32_t>(), 0u, (size_t)nseg + 1,
                                    rocprim::plus<uint32_t>(), s),
            M);
    uint32_t tot[2] = {0, 0};
    PTI_HIP(hipMemcpyAsync(&tot[0], choff.as<uint32_t>() + nseg, 4, hipMemcpyDeviceToHost, s), M);
    PTI_HIP(hipMemcpyAsync(&tot[1], uoff.as<uint32_t>() + nseg, 4, hipMemcpyDeviceToHost, s), M);
    PTI_HIP(hipStreamSynchronize(s), M);
    c.nchunks = tot[0];
    c.nunits = tot[1];
    PTI_HIP(c.chunk_z.alloc((size_t)(c.nchunks + 1) * 4), M);
    PTI_HIP(c.chunk_blk.alloc((size_t)c.nchunks * 4), M);
    PTI_HIP(c.unit_chunk.alloc((size_t)(c.nunits + 1) * 4), M);
    PTI_HIP(c.unit_br.alloc((size_t)c.nunits * 4), M);
    PTI_HIP(c.unit_excl.alloc((size_t)c.nunits * 4), M);
    hipLaunchKernelGGL(chunk_fill_kernel, grid_n(c.nchunks), dim3(256), 0, s, choff.as<uint32_t>(),
                       seg_first.as<uint32_t>(), hb.as<uint32_t>(), nseg, c.nchunks, c.chunk_z.as<uint32_t>(),
                       c.chunk_blk.as<uint32_t>());
    PTI_HIP(hipGetLastError(), M);
    PTI_HIP(hipMemcpyAsync(c.chunk_z.as<uint32_t>() + c.nchunks, &nnz32, 4, hipMemcpyHostToDevice, s), M);
    hipLaunchKernelGGL(unit_fill_kernel, grid_n(c.nunits), dim3(256), 0, s, uoff.as<uint32_t>(), choff.as<uint32_t>(),
                       seg_first.as<uint32_t>(), skey.as<uint32_t>(), nseg, c.nunits, c.unit_chunk.as<uint32_t>(),
                       c.unit_br.as<uint32_t>(), c.unit_excl.as<uint32_t>());
    PTI_HIP(hipGetLastError(), M);
    PTI_HIP(hipMemcpyAsync(c.unit_chunk.as<uint32_t>() + c.nunits, &c.nchunks, 4, hipMemcpyHostToDevice, s), M);
    // 4. Nonzeros, sorted by output row inside each chunk.
    PTI_HIP(c.einds.alloc(nnz * EW * 4), M);
    PTI_HIP(c.vals.alloc(nnz * 4), M);
    PTI_HIP(c.lblk.alloc(nnz), M);
    hipLaunchKernelGGL(chunk_sort_kernel, dim3((c.nchunks + 3) / 4), dim3(256), 0, s, r.bptr.as<uint64_t>(),
                       sid.as<uint32_t>(), hb.as<uint32_t>(), nb, c.chunk_z.as<uint32_t>(), c.chunk_blk.as<uint32_t>(),
                       c.nchunks, ep, r.vals.as<float>(), N, mode, t->sb_bits, c.einds.as<uint32_t>(),
                       c.vals.as<float>(), c.lblk.as<uint8_t>());
    PTI_HIP(hipGetLastError(), M);
    // The host scalars copied above and every temporary buffer (released on
    // return) must outlive the queued work.
    PTI_HIP(hipStreamSynchronize(s), M);
    t->bytes += nb * N * 4 + nnz * (EW * 4 + 5) + (size_t)(2 * c.nchunks + 1) * 4 + (size_t)(3 * c.nunits + 1) * 4;
    c.ready = true;
    return 0;
}

// Upload the copies of the modes in `mode_mask` (bit m = target mode m).
int upload(ptiDevHiCOO **out, const ptiSparseTensorHiCOO *hitsr, uint32_t mode_mask, hipStream_t s) {
    static const char *M = "DevHiCOO Upload";
    PTI_CHECK(require_device(M), M);
    RawHiCOO raw;
    PTI_CHECK(raw_upload(raw, hitsr, s, M), M);
    ptiDevHiCOO *t = new ptiDevHiCOO();
    t->nmodes = hitsr->nmodes;
    t->sb_bits = hitsr->sb_bits;
    t->nnz = hitsr->nnz;
    t->nb = hitsr->nnz ? hitsr->bptr.len - 1 : 0;
    for (uint32_t m = 0; m < t->nmodes; ++m) t->ndims[m] = hitsr->ndims[m];
    for (uint32_t m = 0; m < t->nmodes; ++m) {
        if (!(mode_mask >> m & 1u)) continue;
        int rc = t->nnz ? build_mode(t, raw, m, s, M) : 0;
        if (rc) {
            delete t;
            return rc;
        }
    }
    *out = t;
    return 0;
}

}  // namespace

namespace {
// The stream form of the copy of `mode` for ng lane groups (*ent): a cached
// one, or one built now in the free or least recently used slot.
int ensure_stream(const ptiDevHiCOO *t, uint32_t mode, uint32_t ng, hipStream_t s, const char *M,
                  const uint4 **ent, bool *packed) {
    const ptiDevHiCOOMode &c = t->mc[mode];
    std::lock_guard<std::mutex> lk(c.stream_mu);
    for (auto &f : c.forms)
        if (f.ng == ng) {
            f.last_use = ++c.use_clock;
            *ent = f.ent.as<uint4>();
            *packed = f.packed;
            return 0;
        }
    ptiDevHiCOOMode::StreamForm &dst =
        c.forms[0].ng == 0 ? c.forms[0]
                           : (c.forms[1].ng == 0 || c.forms[1].last_use < c.forms[0].last_use ? c.forms[1] : c.forms[0]);
    // Work queued on the form being replaced must finish before it is released.
    if (dst.ent.ptr) PTI_HIP(hipDeviceSynchronize(), M);
    dst.ng = 0;
    const uint32_t N = t->nmodes;
    // Packed entries (StreamShape<N, true>): orders 4 and 5, rows of at most
    // 7 bits, gathered coordinates below 2^24, and for order 5 a last
    // gathered mode of at most 2^16 rows.
    bool pack = (N == 4 || N == 5) && t->sb_bits <= 7;
    for (uint32_t m = 0; m < N; ++m)
        if (m != mode && t->ndims[m] > (1u << 24)) pack = false;
    if (N == 5 && t->ndims[mode == 4 ? 3 : 4] > (1u << 16)) pack = false;
    const uint32_t NP = pack ? 1u : (N + 4) / 4;
    PTI_HIP(dst.ent.alloc((size_t)c.nchunks * NP * CHUNK * sizeof(uint4)), M);
    // Inside a row, nonzeros follow the coordinate of the largest non-target
    // mode, whose factor rows are the likeliest cache misses.
    uint32_t sec = mode == 0 ? 1 : 0;
    for (uint32_t m = 0; m < N; ++m)
        if (m != mode && t->ndims[m] >= t->ndims[sec]) sec = m;
    const uint64_t nnz = t->nnz;
    bool windowed = windowed_order(nnz, t->ndims[mode], t->ndims[sec], t->sb_bits);
    if (windowed) {
        const uint32_t nwin = (uint32_t)(((uint64_t)t->ndims[sec] + (1u << WINDOW_BITS) - 1) >> WINDOW_BITS);
        Buffer wc;
        PTI_HIP(wc.alloc((size_t)nwin * 4), M);
        PTI_HIP(zero_floats(wc.as<float>(), nwin, s), M);   // all-zero bits: count 0
        hipLaunchKernelGGL(window_hist_kernel, dim3((c.nchunks + 3) / 4), dim3(256), 0, s, hicoo_args(t, mode),
                           c.nchunks, N, sec, wc.as<uint32_t>());
        PTI_HIP(hipGetLastError(), M);
        PTI_CHECK(windows_even(wc.as<uint32_t>(), nwin, nnz, &windowed, s, M), M);
    }
    Buffer key, pos, skey, spos, blk_of, tmp;
    PTI_HIP(key.alloc(nnz * 8), M);
    PTI_HIP(pos.alloc(nnz * 4), M);
    PTI_HIP(skey.alloc(nnz * 8), M);
    PTI_HIP(spos.alloc(nnz * 4), M);
    PTI_HIP(blk_of.alloc(nnz * 4), M);
    const dim3 grid((c.nchunks + 3) / 4);
    unsigned rbits = 1, wbits = 0;
    while (rbits < 32 && ((uint64_t)t->ndims[mode] - 1) >> rbits) ++rbits;
    if (windowed)
        while (wbits < 32 && (((uint64_t)t->ndims[sec] - 1) >> WINDOW_BITS) >> wbits) ++wbits;
    hipLaunchKernelGGL(stream_keys_kernel, grid, dim3(256), 0, s, hicoo_args(t, mode), c.nchunks, N, mode, sec,
                       windowed ? wbits : 0u, key.as<uint64_t>(), pos.as<uint32_t>(), blk_of.as<uint32_t>());
    PTI_HIP(hipGetLastError(), M);
    const unsigned kbits = windowed ? std::max<unsigned>(rbits, t->sb_bits) + wbits + WINDOW_BITS : rbits + 32;
    size_t tb = 0;
    PTI_HIP(rocprim::radix_sort_pairs(nullptr, tb, key.as<uint64_t>(), skey.as<uint64_t>(), pos.as<uint32_t>(),
                                      spos.as<uint32_t>(), (size_t)nnz, 0, kbits, s),
            M);
    PTI_HIP(tmp.alloc(tb), M);
    PTI_HIP(rocprim::radix_sort_pairs(tmp.ptr, tb, key.as<uint64_t>(), skey.as<uint64_t>(), pos.as<uint32_t>(),
                                      spos.as<uint32_t>(), (size_t)nnz, 0, kbits, s),
            M);
    hipLaunchKernelGGL(stream_fill_kernel, grid, dim3(256), 0, s, hicoo_args(t, mode), c.nchunks, N, mode, ng,
                       pack ? 1u : 0u, spos.as<uint32_t>(), blk_of.as<uint32_t>(), dst.ent.as<uint4>());
    PTI_HIP(hipGetLastError(), M);
    if (!c.unit_order.ptr) {
        if (windowed)
            PTI_CHECK(build_rank_unit_order(c.unit_br.as<uint32_t>(), c.nunits, c.unit_order, s, M), M);
        else
            PTI_CHECK(build_unit_order(c.unit_chunk.as<uint32_t>(), c.nunits, c.unit_order, s, M), M);
    }
    // The temporaries are released on return: the queued work must finish.
    PTI_HIP(hipStreamSynchronize(s), M);
    dst.ng = ng;
    dst.last_use = ++c.use_clock;
    dst.packed = pack;
    *ent = dst.ent.as<uint4>();
    *packed = pack;
    return 0;
}
}  // namespace

// Block-rows owned by a single unit (flags zeroed by the caller).
__global__ void __launch_bounds__(256) owned_rows_kernel(const uint32_t *__restrict__ unit_br,
                                                         const uint32_t *__restrict__ unit_excl, uint32_t nunits,
                                                         uint8_t *__restrict__ owned) {
    const uint32_t u = blockIdx.x * 256 + threadIdx.x;
    if (u < nunits && unit_excl[u]) owned[unit_br[u]] = 1;
}

// Zeroes the output rows whose block-row no single unit owns: one wave per
// row, rows of `width` floats.
__global__ void __launch_bounds__(256) zero_shared_rows_kernel(float *__restrict__ out, uint32_t nrows,
                                                               uint32_t width, const uint8_t *__restrict__ owned,
                                                               uint32_t sb) {
    const uint32_t wave = (blockIdx.x * 256 + threadIdx.x) >> 6, lane = threadIdx.x & 63;
    const uint32_t nwaves = (gridDim.x * 256) >> 6;
    for (uint32_t r = wave; r < nrows; r += nwaves) {
        if (owned[r >> sb]) continue;
        float *row = out + (uint64_t)r * width;
        for (uint32_t c = lane; c < width; c += 64) row[c] = 0.f;
    }
}

int ensure_owned_rows(const ptiDevHiCOO *t, uint32_t mode, hipStream_t s, const char *M) {
    const ptiDevHiCOOMode &c = t->mc[mode];
    std::lock_guard<std::mutex> lk(c.stream_mu);
    if (c.br_owned.ptr) return 0;
    const uint64_t nbr = ((uint64_t)t->ndims[mode] + (1u << t->sb_bits) - 1) >> t->sb_bits;
    Buffer owned;
    PTI_HIP(owned.alloc((nbr + 3) / 4 * 4), M);
    PTI_HIP(zero_floats(owned.as<float>(), (nbr + 3) / 4, s), M);   // all-zero bits
    if (c.nunits) {
        hipLaunchKernelGGL(owned_rows_kernel, dim3((c.nunits + 255) / 256), dim3(256), 0, s, c.unit_br.as<uint32_t>(),
                           c.unit_excl.as<uint32_t>(), c.nunits, owned.as<uint8_t>());
        PTI_HIP(hipGetLastError(), M);
    }
    // Calls on other streams may read the flags as soon as they are published.
    PTI_HIP(hipStreamSynchronize(s), M);
    c.br_owned = std::move(owned);
    return 0;
}

hipError_t zero_shared_rows(float *out, uint32_t nrows, uint32_t width, const uint8_t *owned, uint32_t sb,
                            hipStream_t s) {
    uint64_t blocks = ((uint64_t)nrows + 3) / 4;   // four rows (waves) per workgroup
    if (blocks > 16384) blocks = 16384;
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(zero_shared_rows_kernel, dim3((uint32_t)blocks), dim3(256), 0, s, out, nrows, width, owned, sb);
    return hipGetLastError();
}

// LDS-resident factors for the stream kernel (orders 4 and 5, four ranks per
// lane): the last gathered modes, in ascending mode order, whose
// whole rank-tile copies (rows x G * V elements) fit beside the LDS tile in
// the 80 KB a workgroup may take at two workgroups (four waves per SIMD) per
// CU, at most one for order 4 and two for order 5 (the instantiated kernels),
// and only where a unit holds at least four
// nonzeros per resident row on average (each workgroup loads the copies
// once).  rows[j] receives the row counts, 0 where none.
void resident_factors(const ptiDevHiCOO *t, uint32_t mode, Shape sh, size_t esz, uint32_t nunits, uint32_t rows[2]) {
    constexpr size_t BUDGET = 80 * 1024;
    rows[0] = rows[1] = 0;
    const uint32_t N = t->nmodes;
    // Eight BF16 ranks per lane (16-byte lanes) gain nothing from the copies:
    // on the 200M-nonzero order-4 tensor at R = 64 that kernel is not bound
    // by its gathers, and the copies took modes 0-2 from 5.57 / 3.48 / 3.39
    // to 5.75 / 3.53 / 3.50 ms.
    if ((N != 4 && N != 5) || sh.V != 4 || nunits == 0) return;
    const size_t rowb = (size_t)sh.G * sh.V * esz;
    size_t used = res_lds_offset(1u << t->sb_bits, sh.G * sh.V + 1);
    uint32_t take[2] = {0, 0}, nres = 0;
    const uint32_t max_res = N == 4 ? 1 : 2;
    for (int i = (int)N - 2; i >= 0 && nres < max_res; --i) {
        const uint32_t m = (uint32_t)i < mode ? (uint32_t)i : (uint32_t)i + 1;
        const size_t bytes = (size_t)t->ndims[m] * rowb;
        if (used + bytes > BUDGET) break;
        used += bytes;
        take[nres++] = t->ndims[m];
    }
    uint64_t total = (uint64_t)take[0] + take[1];
    if (nres == 0 || 4 * total > t->nnz / nunits) return;
    // Positions K - nres .. K - 1 in ascending order: the last one found
    // first is the highest position.
    if (nres == 1) {
        rows[0] = take[0];
    } else {
        rows[0] = take[1];
        rows[1] = take[0];
    }
}

// Shared by the FP32 and BF16 entry points: zero the output rows, then run the
// stream kernel (factor matrices < 4 GiB) or the tile kernel once per rank
// column tile.
int run_hicoo(const ptiDevHiCOO *t, const void *const facs[], float *out, ptiIndex R, ptiIndex stride,
              ptiIndex out_stride, const ptiIndex mats_order[], ptiIndex mode, bool bf16, hipStream_t s,
              const char *M) {
    const uint32_t N = t->nmodes;
    if (mode >= N) PTI_FAIL(PTIERR_VALUE_ERROR, M, "mode out of range");
    if (R == 0 || stride < R || out_stride < R) PTI_FAIL(PTIERR_SHAPE_MISMATCH, M, "bad rank or stride");
    if (!order_is_permutation(N, mats_order, mode))
        PTI_FAIL(PTIERR_VALUE_ERROR, M, "mats_order[1..N-1] must be a permutation of the other modes");
    if (t->nnz && !t->mc[mode].ready) PTI_FAIL(PTIERR_VALUE_ERROR, M, "mode was not uploaded");
    if (t->nnz == 0) {
        PTI_HIP(zero_floats(out, (uint64_t)t->ndims[mode] * out_stride, s), M);
        return 0;
    }
    // 64 ranks per LDS tile; 32 for orders >= 6, whose larger staging area
    // would otherwise push a workgroup past 64 KB of LDS.
    const ptiIndex tile = N >= 6 ? 32 : 64;
    const size_t esz = bf16 ? 2 : 4;
    bool fits = true;   // the stream kernel gathers with 32-bit byte offsets
    for (uint32_t i = 1; i < N; ++i)
        if ((uint64_t)t->ndims[mats_order[i]] * stride * esz >= (1ull << 32)) fits = false;
    const bool coord24 = coords_fit_u24(t->ndims, N, mode, stride * esz);
    // Lane shape and kernel of the column tile starting at c0 (tiles in
    // order: `first` is the first tile's shape).
    Shape first{0, 0};
    auto plan_tile = [&](ptiIndex c0, Shape &sh, bool &use_stream) -> int {
        const ptiIndex Rt = std::min<ptiIndex>(tile, R - c0);
        // Column tiles start at multiples of 32 elements, so they keep the
        // alignment the vector width V relies on.
        if (!choose_shape(Rt, stride, out_stride, sh)) PTI_FAIL(PTIERR_VALUE_ERROR, M, "unsupported rank tile");
        // BF16 factors on the stream kernel: eight ranks (16 bytes) per lane
        // when the rank tile needs more than four such lanes (lane groups of
        // eight), so that one gather instruction covers twice the rows of the
        // 8-byte shape.  At 32 ranks and below the 16-byte shape would need
        // groups of four lanes, sixteen slices per chunk: measured on the
        // 100M-nonzero Zipf tensor at rank 32, their run-sum flushes into the
        // LDS tile (eight atomics per lane each) cost more than the halved
        // gathers save (1.26 against 1.07 ms per mode).
        if (bf16 && fits && Rt % 8 == 0 && stride % 8 == 0 && Rt > 32) {
            Shape w{8, 8};
            if (stream_supported(N, w.G)) sh = w;
        }
        // A narrower last tile keeps the first tile's lane-group width G, with
        // the widest V <= the first tile's V that divides its ranks, when its
        // ranks fit G * V (idle lanes past the rank): every tile then reads the
        // same stream form instead of rebuilding it for another lane-group
        // count.  A last tile that fits no such shape runs the tile kernel.
        bool keep_form = true;
        if (c0 == 0) {
            first = sh;
        } else {
            int v = first.V;
            while (v > 1 && Rt % (ptiIndex)v) v >>= 1;
            if (Rt <= (ptiIndex)(first.G * v))
                sh = Shape{v, first.G};
            else
                keep_form = sh.G == first.G;
        }
        use_stream = fits && stream_supported(N, sh.G) && keep_form;
        // The tile kernel takes at most four ranks per lane.
        if (!use_stream && sh.V == 8) choose_shape(Rt, stride, out_stride, sh);
        // The kernel stages factor-row offsets as 32-bit counts of V-wide vectors.
        for (uint32_t i = 1; i < N; ++i)
            if ((uint64_t)t->ndims[mats_order[i]] * (stride / sh.V) >= (1ull << 32))
                PTI_FAIL(PTIERR_VALUE_ERROR, M, "factor matrix exceeds 2^32 row-vector slots");
        return 0;
    };
    // Whole rows from single-unit block-rows: when every tile runs the stream
    // kernel and the last tile's lanes reach the end of the output row
    // (padding included), the unit that owns a block-row alone writes all
    // of its rows and columns, zeros included, so the zero fill only covers
    // block-rows that several units share or that hold no nonzero.  (A
    // packed form with a 16 MB row pitch or more would go to the tile kernel,
    // so orders 4 and 5 then keep the full fill.)
    bool whole_rows = coord24 || (N != 4 && N != 5);
    for (ptiIndex c0 = 0; c0 < R && whole_rows; c0 += tile) {
        Shape sh;
        bool use_stream = false;
        PTI_CHECK(plan_tile(c0, sh, use_stream), M);
        if (!use_stream || (c0 + tile >= R && out_stride > c0 + (ptiIndex)(sh.G * sh.V))) whole_rows = false;
    }
    if (whole_rows) {
        PTI_CHECK(ensure_owned_rows(t, mode, s, M), M);
        const ptiDevHiCOOMode &c = t->mc[mode];
        PTI_HIP(zero_shared_rows(out, t->ndims[mode], out_stride, c.br_owned.as<uint8_t>(), t->sb_bits, s), M);
    } else {
        PTI_HIP(zero_floats(out, (uint64_t)t->ndims[mode] * out_stride, s), M);
    }
    for (ptiIndex c0 = 0; c0 < R; c0 += tile) {
        const ptiIndex Rt = std::min<ptiIndex>(tile, R - c0);
        Shape sh;
        bool use_stream = false;
        PTI_CHECK(plan_tile(c0, sh, use_stream), M);
        FactorArgs a;
        for (uint32_t m = 0; m < N; ++m) a.fac[m] = static_cast<const char *>(facs[m]) + (size_t)c0 * esz;
        a.out = out + c0;
        a.R = Rt;
        a.stride = stride;
        a.out_stride = out_stride;
        a.mode = mode;
        a.out_cols = whole_rows ? out_stride - c0 : 0u;
        for (uint32_t i = 0; i < N; ++i) a.order[i] = mats_order[i];
        if (use_stream) {
            const uint4 *ent = nullptr;
            bool packed = false;
            PTI_CHECK(ensure_stream(t, mode, 64u / sh.G, s, M, &ent, &packed), M);
            const ptiDevHiCOOMode &c = t->mc[mode];
            StreamArgs h{};
            h.ent = ent;
            h.unit_chunk = c.unit_chunk.as<uint32_t>();
            h.unit_br = c.unit_br.as<uint32_t>();
            h.unit_excl = c.unit_excl.as<uint32_t>();
            h.nunits = c.nunits;
            h.sb_bits = t->sb_bits;
            h.nrows = t->ndims[mode];
            h.coord24 = coord24;
            h.unit_order = c.unit_order.as<uint32_t>();
            resident_factors(t, mode, sh, esz, c.nunits, h.res_rows);
            if (packed && !h.coord24) {
                // Packed entries need the 24-bit row offsets (row pitch below
                // 16 MB), which ignore the fields in the coordinates' high byte.
                PTI_HIP(launch_hicoo(N, hicoo_args(t, mode), a, sh, bf16, s), M);
                continue;
            }
            h.packed = packed ? 1u : 0u;
            PTI_HIP(launch_stream(N, h, a, sh, bf16, s), M);
        } else {
            PTI_HIP(launch_hicoo(N, hicoo_args(t, mode), a, sh, bf16, s), M);
        }
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

// order = the units' indices by decreasing chunk count (stable), for
// StreamArgs::unit_order.
int build_unit_order(const uint32_t *unit_chunk, uint32_t nunits, Buffer &order, hipStream_t s, const char *M) {
    PTI_HIP(order.alloc((size_t)nunits * 4), M);
    if (nunits == 0) return 0;
    Buffer sz, ssz, ids, tmp;
    PTI_HIP(sz.alloc((size_t)nunits * 4), M);
    PTI_HIP(ssz.alloc((size_t)nunits * 4), M);
    PTI_HIP(ids.alloc((size_t)nunits * 4), M);
    hipLaunchKernelGGL(unit_size_kernel, grid_n(nunits), dim3(256), 0, s, unit_chunk, nunits, sz.as<uint32_t>());
    hipLaunchKernelGGL(iota_u32_kernel, grid_n(nunits), dim3(256), 0, s, ids.as<uint32_t>(), (uint64_t)nunits);
    PTI_HIP(hipGetLastError(), M);
    size_t tb = 0;
    PTI_HIP(rocprim::radix_sort_pairs_desc(nullptr, tb, sz.as<uint32_t>(), ssz.as<uint32_t>(), ids.as<uint32_t>(),
                                           order.as<uint32_t>(), (size_t)nunits, 0, 8, s),
            M);
    PTI_HIP(tmp.alloc(tb), M);
    PTI_HIP(rocprim::radix_sort_pairs_desc(tmp.ptr, tb, sz.as<uint32_t>(), ssz.as<uint32_t>(), ids.as<uint32_t>(),
                                           order.as<uint32_t>(), (size_t)nunits, 0, 8, s),
            M);
    // The temporaries are released on return.
    PTI_HIP(hipStreamSynchronize(s), M);
    return 0;
}

int windows_even(const uint32_t *counts, uint32_t nwin, uint64_t nnz, bool *even, hipStream_t s, const char *M) {
    std::vector<uint32_t> h(nwin);
    PTI_HIP(hipMemcpyAsync(h.data(), counts, (size_t)nwin * 4, hipMemcpyDeviceToHost, s), M);
    PTI_HIP(hipStreamSynchronize(s), M);
    uint32_t mx = 0;
    for (uint32_t x : h) mx = std::max(mx, x);
    *even = nwin > 0 && (double)mx <= 4.0 * (double)nnz / (double)nwin;
    return 0;
}

// Units dealt by their rank inside their block-row, for stream forms whose
// block-rows walk a gathered factor window by window (unit r of every
// block-row covers about the same windows).  Workgroup b runs on XCD b mod 8
// (round-robin dealing), so the units of rank r go to the positions of XCD
// r mod 8, in rank order and then block-row order on each: each XCD's L2 then
// holds the windows of about one rank at a time instead of those of every
// rank in flight.  Positions left over by an XCD with fewer units take the
// next XCD's.
int build_rank_unit_order(const uint32_t *unit_br, uint32_t nunits, Buffer &order, hipStream_t s, const char *M) {
    constexpr uint32_t XCDS = 8;
    PTI_HIP(order.alloc((size_t)nunits * 4), M);
    if (nunits == 0) return 0;
    std::vector<uint32_t> br(nunits), rank(nunits), ord;
    PTI_HIP(hipMemcpyAsync(br.data(), unit_br, (size_t)nunits * 4, hipMemcpyDeviceToHost, s), M);
    PTI_HIP(hipStreamSynchronize(s), M);
    std::vector<uint32_t> by_rank(nunits);
    for (uint32_t i = 0; i < nunits; ++i) {
        rank[i] = (i > 0 && br[i] == br[i - 1]) ? rank[i - 1] + 1 : 0;
        by_rank[i] = i;
    }
    std::stable_sort(by_rank.begin(), by_rank.end(), [&](uint32_t a, uint32_t b) { return rank[a] < rank[b]; });
    std::vector<std::vector<uint32_t>> q(XCDS);
    for (uint32_t i : by_rank) q[rank[i] % XCDS].push_back(i);
    std::vector<size_t> head(XCDS, 0);
    ord.reserve(nunits);
    for (uint32_t b = 0; b < nunits; ++b) {
        for (uint32_t k = 0; k < XCDS; ++k) {
            const uint32_t x = (b + k) % XCDS;
            if (head[x] < q[x].size()) {
                ord.push_back(q[x][head[x]++]);
                break;
            }
        }
    }
    PTI_HIP(hipMemcpyAsync(order.ptr, ord.data(), (size_t)nunits * 4, hipMemcpyHostToDevice, s), M);
    // The host vector must outlive the copy.
    PTI_HIP(hipStreamSynchronize(s), M);
    return 0;
}

// Cuts n nonzeros sorted by output row (rows[]) into blocks of 2^sb rows,
// chunks of <= CHUNK nonzeros that never cross a block and units of <=
// UNIT_CHUNKS chunks of one block, the structure of a HiCOO mode copy (so the
// stream kernel can run a sorted COO copy).
int build_row_units(const uint32_t *rows, uint64_t n, uint32_t sb, RowUnits &u, hipStream_t s, const char *M) {
    u.nchunks = u.nunits = 0;
    if (n == 0) return 0;
    Buffer flag, segid, tmp_scan;
    PTI_HIP(flag.alloc(n * 4), M);
    PTI_HIP(segid.alloc(n * 4), M);
    hipLaunchKernelGGL(row_block_flag_kernel, grid_n(n), dim3(256), 0, s, rows, n, sb, flag.as<uint32_t>());
    PTI_HIP(hipGetLastError(), M);
    size_t tb = 0;
    PTI_HIP(rocprim::inclusive_scan(nullptr, tb, flag.as<uint32_t>(), segid.as<uint32_t>(), (size_t)n,
                                    rocprim::plus<uint32_t>(), s),
            M);
    PTI_HIP(tmp_scan.alloc(tb), M);
    PTI_HIP(rocprim::inclusive_scan(tmp_scan.ptr, tb, flag.as<uint32_t>(), segid.as<uint32_t>(), (size_t)n,
                                    rocprim::plus<uint32_t>(), s),
            M);
    uint32_t nseg = 0;
    PTI_HIP(hipMemcpyAsync(&nseg, segid.as<uint32_t>() + n - 1, 4, hipMemcpyDeviceToHost, s), M);
    PTI_HIP(hipStreamSynchronize(s), M);
    Buffer start, key, idx, nch, nu, choff, uoff, tmp_seg, cblk;
    PTI_HIP(start.alloc((size_t)(nseg + 1) * 4), M);
    PTI_HIP(key.alloc((size_t)(nseg + 1) * 4), M);
    PTI_HIP(idx.alloc((size_t)(nseg + 1) * 4), M);
    hipLaunchKernelGGL(row_block_start_kernel, grid_n(n), dim3(256), 0, s, rows, n, sb, flag.as<uint32_t>(),
                       segid.as<uint32_t>(), start.as<uint32_t>(), key.as<uint32_t>());
    PTI_HIP(hipGetLastError(), M);
    const uint32_t n32 = (uint32_t)n;
    PTI_HIP(hipMemcpyAsync(start.as<uint32_t>() + nseg, &n32, 4, hipMemcpyHostToDevice, s), M);
    hipLaunchKernelGGL(iota_u32_kernel, grid_n(nseg + 1), dim3(256), 0, s, idx.as<uint32_t>(), (uint64_t)nseg + 1);
    PTI_HIP(hipGetLastError(), M);
    // Segments here are row blocks; each is its own "first block" (idx) and
    // `start` plays the block offsets.
    PTI_HIP(nch.alloc((size_t)(nseg + 1) * 4), M);
    PTI_HIP(nu.alloc((size_t)(nseg + 1) * 4), M);
    PTI_HIP(choff.alloc((size_t)(nseg + 1) * 4), M);
    PTI_HIP(uoff.alloc((size_t)(nseg + 1) * 4), M);
    hipLaunchKernelGGL(segment_counts_kernel, grid_n(nseg), dim3(256), 0, s, idx.as<uint32_t>(), start.as<uint32_t>(),
                       nseg, UNIT_CHUNKS, nch.as<uint32_t>(), nu.as<uint32_t>());
    PTI_HIP(hipGetLastError(), M);
    const uint32_t zero = 0;
    PTI_HIP(hipMemcpyAsync(nch.as<uint32_t>() + nseg, &zero, 4, hipMemcpyHostToDevice, s), M);
    PTI_HIP(hipMemcpyAsync(nu.as<uint32_t>() + nseg, &zero, 4, hipMemcpyHostToDevice, s), M);
    size_t sb3 = 0;
    PTI_HIP(rocprim::exclusive_scan(nullptr, sb3, nch.as<uint32_t>(), choff.as<uint32_t>(), 0u, (size_t)nseg + 1,
                                    rocprim::plus<uint32_t>(), s),
            M);
    PTI_HIP(tmp_seg.alloc(sb3), M);
    PTI_HIP(rocprim::exclusive_scan(tmp_seg.ptr, sb3, nch.as<uint32_t>(), choff.as<uint32_t>(), 0u, (size_t)nseg + 1,
                                    rocprim::plus<uint32_t>(), s),
            M);
    PTI_HIP(rocprim::exclusive_scan(tmp_seg.ptr, sb3, nu.as<uint32_t>(), uoff.as<uint32_t>(), 0u, (size_t)nseg + 1,
                                    rocprim::plus<uint32_t>(), s),
            M);
    uint32_t tot[2] = {0, 0};
    PTI_HIP(hipMemcpyAsync(&tot[0], choff.as<uint32_t>() + nseg, 4, hipMemcpyDeviceToHost, s), M);
    PTI_HIP(hipMemcpyAsync(&tot[1], uoff.as<uint32_t>() + nseg, 4, hipMemcpyDeviceToHost, s), M);
    PTI_HIP(hipStreamSynchronize(s), M);
    u.nchunks = tot[0];
    u.nunits = tot[1];
    PTI_HIP(u.chunk_z.alloc((size_t)(u.nchunks + 1) * 4), M);
    PTI_HIP(cblk.alloc((size_t)u.nchunks * 4), M);
    PTI_HIP(u.unit_chunk.alloc((size_t)(u.nunits + 1) * 4), M);
    PTI_HIP(u.unit_br.alloc((size_t)u.nunits * 4), M);
    PTI_HIP(u.unit_excl.alloc((size_t)u.nunits * 4), M);
    hipLaunchKernelGGL(chunk_fill_kernel, grid_n(u.nchunks), dim3(256), 0, s, choff.as<uint32_t>(), idx.as<uint32_t>(),
                       start.as<uint32_t>(), nseg, u.nchunks, u.chunk_z.as<uint32_t>(), cblk.as<uint32_t>());
    PTI_HIP(hipGetLastError(), M);
    PTI_HIP(hipMemcpyAsync(u.chunk_z.as<uint32_t>() + u.nchunks, &n32, 4, hipMemcpyHostToDevice, s), M);
    hipLaunchKernelGGL(unit_fill_kernel, grid_n(u.nunits), dim3(256), 0, s, uoff.as<uint32_t>(), choff.as<uint32_t>(),
                       idx.as<uint32_t>(), key.as<uint32_t>(), nseg, u.nunits, u.unit_chunk.as<uint32_t>(),
                       u.unit_br.as<uint32_t>(), u.unit_excl.as<uint32_t>());
    PTI_HIP(hipGetLastError(), M);
    PTI_HIP(hipMemcpyAsync(u.unit_chunk.as<uint32_t>() + u.nunits, &u.nchunks, 4, hipMemcpyHostToDevice, s), M);
    // The host scalars and the temporaries must outlive the queued work.
    PTI_HIP(hipStreamSynchronize(s), M);
    return 0;
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
    if (bytes) {
        *bytes = t->bytes;
        for (uint32_t m = 0; m < t->nmodes; ++m) {
            std::lock_guard<std::mutex> lk(t->mc[m].stream_mu);
            for (const auto &f : t->mc[m].forms) *bytes += f.ent.bytes;
        }
    }
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
