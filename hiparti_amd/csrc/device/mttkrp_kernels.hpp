// MTTKRP kernels for gfx950 (wave64).
//
// Common pieces.  Nonzeros are processed in chunks of at most CHUNK = 256
// consecutive nonzeros, walked by a wavefront with its 64 lanes split into
// NG = 64/G groups of G lanes: a group handles one nonzero per step and each
// lane V consecutive ranks (float4 / float2 / float loads of the factor
// rows, or 4/2/1 bf16 values for the BF16 form).  The tile and COO kernels
// first decode a chunk into LDS (coordinates and values) and form the
// contribution v * prod_k A_k[i_k][r] in mats_order order; the stream kernel
// reads chunks already decoded and forms it in ascending mode order.  All in
// FP32.
//
// HiCOO kernel (hicoo_tile_kernel).  The device keeps, for each target mode
// m, a copy of the HiCOO blocks grouped by their mode-m block coordinate
// (block-row); block coordinates, 8-bit element offsets and values are the
// tensor's.  The copy is cut into chunks of <= 256 nonzeros that never cross
// a block-row; inside a chunk the nonzeros are ordered by their mode-m
// element offset (i.e. by output row) and each carries the index of its
// block relative to the chunk's first block (1 byte).  A workgroup owns one
// "unit": a run of chunks of one block-row.  Its waves add into an LDS tile
// of the block-row's 2^sb_bits output rows, indexed by the 8-bit element
// offset.  Each lane group walks a contiguous slice of the chunk and
// sums runs of equal rows in registers, so a run reaches the tile once per
// slice.  At the end the tile is written once: with plain stores
// when the unit is the block-row's only unit, otherwise with one FP32 atomic
// per element.  Global atomics thus scale with the number of units, not with
// the number of nonzeros, and a Zipf head row no longer serialises millions
// of atomics.
//
// Stream kernel (hicoo_stream_kernel): the same units and LDS tile, but each
// chunk is read already decoded from HBM (see StreamShape) and the nonzeros
// of a block-row are in output-row order across its chunks; lane groups take
// their entries from their own lanes with DPP moves.  It runs orders 2-5
// with lane groups of at most 16 lanes and factor matrices below 4 GiB; the
// tile kernel takes the rest.
//
// COO kernel (coo_mttkrp_kernel).  Per target mode the device keeps a copy
// sorted by the output index; each run of equal rows in a chunk is added to
// the output with one FP32 atomic per rank.  Where the stream kernel's shapes
// allow, the sorted copy is instead cut into 2^7-row blocks, chunks and units
// like a HiCOO mode copy and run by hicoo_stream_kernel (LDS tile per block);
// coo_mttkrp_kernel serves the rest and the replaced-values form
// (CP-completion residuals passed in).
//
// CP-completion (completion_forward_kernel, completion_grad_kernel).  The
// forward gathers every factor row of each observed entry in storage order
// and writes its residual; the gradient kernel is the stream kernel on a COO
// copy's stream form that forms each residual itself, from the output row's
// factor row staged in LDS, and accumulates residual * product.
//
// Run reduction (COO kernel).  Within a step the NG contributions are
// combined by a segmented inclusive scan over groups keyed on the output row;
// the last group of a segment passes the segment sum to the sink, and a
// segment still open at the end of a step is carried into the next step.
//
// Workgroup ids are remapped so that consecutive work lands on the same XCD
// (the hardware deals workgroups to the 8 XCDs round-robin), which keeps the
// rows used by neighbouring work in one L2.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

namespace pti {
namespace dev {

constexpr int CHUNK = 256;          // nonzeros per wave chunk
constexpr int WAVES_PER_BLOCK = 4;  // waves (chunks) per 256-thread workgroup
constexpr uint32_t INVALID_ROW = 0xFFFFFFFFu;
constexpr uint32_t DUMMY_ROW = 0xFFu;   // HiCOO staging: 8-bit rows, sb_bits <= 7
constexpr int MAX_MODES = 8;

struct FactorArgs {
    const void *fac[MAX_MODES];  // factor matrix base per mode (float* or bf16 bits)
    float *out;                  // output, rows of the target mode
    uint32_t R;                  // rank (columns used)
    uint32_t stride;             // factor row stride (elements)
    uint32_t out_stride;         // output row stride (elements)
    uint32_t mode;               // target mode
    uint32_t order[MAX_MODES];   // order[1..N-1]: factor modes in product order
    // > 0 (stream kernel): a unit that owns its block-row alone writes the
    // rows whole, columns [0, min(G * V, out_cols)) of out, zeros included
    // (columns past R as zeros), so the caller need not zero those rows.
    uint32_t out_cols = 0;
};

// One target mode's copy of a HiCOO tensor (see the header comment).
struct HiCOOArgs {
    const uint32_t *binds;       // nb * N block coordinates, block-major
    const uint32_t *einds;       // nnz * EW packed 8-bit element offsets
    const float *vals;           // nnz values
    const uint8_t *lblk;         // nnz: block of the nonzero minus its chunk's first block
    const uint32_t *chunk_z;     // nchunks + 1: first nonzero of each chunk; [nchunks] = nnz
    const uint32_t *chunk_blk;   // nchunks: first block of each chunk
    const uint32_t *unit_chunk;  // nunits + 1: first chunk of each unit
    const uint32_t *unit_br;     // nunits: block-row (mode-m block coordinate) of each unit
    const uint32_t *unit_excl;   // nunits: 1 if the unit is its block-row's only unit
    uint32_t nunits;
    uint32_t sb_bits;
    uint32_t nrows;              // ndims of the target mode
};

struct COOArgs {
    const uint32_t *inds[MAX_MODES];  // coordinates per mode (sorted by the target mode)
    const float *vals;
    // Optional replacement values in the tensor's storage order (for example
    // CP-completion residuals); value z of the sorted copy is then
    // vals_override[perm[z]].
    const float *vals_override;
    const uint32_t *perm;
    uint64_t nnz;
    uint32_t nchunks;
};

__device__ __forceinline__ uint32_t remap_block(uint32_t bid, uint32_t nblocks) {
    // Hardware assigns workgroup bid to XCD bid % 8.  Give XCD x a contiguous
    // range of logical workgroups so neighbouring chunks share an L2.
    const uint32_t nx = 8;
    uint32_t per = nblocks / nx, rem = nblocks % nx;
    uint32_t x = bid % nx, k = bid / nx;
    // XCD x holds logical ids [start_x, start_x + per + (x < rem)).
    uint32_t start = x * per + (x < rem ? x : rem);
    return start + k;
}

template <typename FT, int V>
struct FactorLoad;

template <>
struct FactorLoad<float, 4> {
    __device__ __forceinline__ static void load(const void *base, uint64_t off, float *o) {
        float4 t = *reinterpret_cast<const float4 *>(static_cast<const float *>(base) + off);
        o[0] = t.x; o[1] = t.y; o[2] = t.z; o[3] = t.w;
    }
};
template <>
struct FactorLoad<float, 2> {
    __device__ __forceinline__ static void load(const void *base, uint64_t off, float *o) {
        float2 t = *reinterpret_cast<const float2 *>(static_cast<const float *>(base) + off);
        o[0] = t.x; o[1] = t.y;
    }
};
template <>
struct FactorLoad<float, 1> {
    __device__ __forceinline__ static void load(const void *base, uint64_t off, float *o) {
        o[0] = static_cast<const float *>(base)[off];
    }
};
__device__ __forceinline__ float bf16_bits_to_float(uint32_t b) { return __uint_as_float(b << 16); }
template <>
struct FactorLoad<uint16_t, 4> {
    __device__ __forceinline__ static void load(const void *base, uint64_t off, float *o) {
        uint2 t = *reinterpret_cast<const uint2 *>(static_cast<const uint16_t *>(base) + off);
        o[0] = bf16_bits_to_float(t.x & 0xFFFFu); o[1] = bf16_bits_to_float(t.x >> 16);
        o[2] = bf16_bits_to_float(t.y & 0xFFFFu); o[3] = bf16_bits_to_float(t.y >> 16);
    }
};
template <>
struct FactorLoad<uint16_t, 2> {
    __device__ __forceinline__ static void load(const void *base, uint64_t off, float *o) {
        uint32_t t = *reinterpret_cast<const uint32_t *>(static_cast<const uint16_t *>(base) + off);
        o[0] = bf16_bits_to_float(t & 0xFFFFu); o[1] = bf16_bits_to_float(t >> 16);
    }
};
template <>
struct FactorLoad<uint16_t, 1> {
    __device__ __forceinline__ static void load(const void *base, uint64_t off, float *o) {
        o[0] = bf16_bits_to_float(static_cast<const uint16_t *>(base)[off]);
    }
};

// Staged chunk.  Position p is stored at p + p/32, so that the lane groups
// of a wave reading positions 32 apart hit different LDS banks.
__device__ __forceinline__ uint32_t skew(uint32_t p) { return p + (p >> 5); }
constexpr int CHUNK_SKEWED = CHUNK + CHUNK / 32;

template <int N>
struct Staging {
    uint32_t coord[N][CHUNK_SKEWED];
    float val[CHUNK_SKEWED];
};

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// HiCOO staging of one chunk, laid out for the slice walk.  The NG lane
// groups each take a contiguous slice of per = ceil(cnt / NG) positions;
// local position j of group g is stored at entry j * NG + g, so the groups of
// a wave read NG consecutive entries (distinct banks) and a group's next
// position is NG entries further on.  Per entry: the output row inside the
// block-row (the target mode's element offset) and the value in one 8-byte
// word, and for each factor in product order (mats_order[1..N-1]) the row's
// offset in units of the lane vector (coord * stride / V, checked on the
// host to fit 32 bits).  Entries past a group's slice, up to the last
// position the walk reads (hicoo_straight), are dummies (invalid row, offset
// 0) so that the pipelined walk can read and gather ahead without bounds
// tests; every entry it reads is written for the current chunk.
template <int NG, int D>
constexpr int hstage_entries() {
    return CHUNK + NG + 2 * D * NG;
}

// The slice walk runs straight-line over a full chunk's slice (PMAX = CHUNK
// / NG positions plus D of look-ahead, whatever the chunk's length) when
// PMAX <= 32, else it loops over the slice's own length.
template <int NG>
constexpr bool hicoo_straight() {
    return CHUNK / NG <= 32;
}

template <int N, int E>
struct HStage {
    uint32_t off[N - 1][E];
    float val[E];
    uint8_t row[E];             // output row inside the block-row; 0xFF: dummy
};

// Decode chunk `chunk` of a HiCOO mode copy (one wave): coordinate of mode m
// = (binds[first block + lblk] << sb_bits) + element offset.  Returns the
// number of nonzeros in the chunk (<= CHUNK).
template <int N, int V, int NG, int D, int E>
__device__ __forceinline__ uint32_t stage_hicoo(const HiCOOArgs &h, const FactorArgs &a, uint32_t chunk,
                                                HStage<N, E> &st, int lane) {
    constexpr int EW = (N + 3) / 4;
    static_assert((CHUNK / NG + D) * NG <= E && (CHUNK / NG + 2 * D) * NG <= E, "staging too small for the walk");
    const uint32_t z0 = h.chunk_z[chunk];
    const uint32_t cnt = h.chunk_z[chunk + 1] - z0;
    const uint32_t b0 = h.chunk_blk[chunk];
    const uint32_t vstride = a.stride / V;
    const uint32_t per = (cnt + NG - 1) / NG;
    wave_sync();
    for (uint32_t zl = lane; zl < cnt; zl += 64) {
        const uint32_t z = z0 + zl;
        const uint64_t b = (uint64_t)b0 + h.lblk[z];
        uint32_t ew[EW];
#pragma unroll
        for (int w = 0; w < EW; ++w) ew[w] = h.einds[(uint64_t)z * EW + w];
        auto elem = [&](uint32_t m) { return (ew[m >> 2] >> ((m & 3) * 8)) & 0xFFu; };
        const uint32_t g = zl / per;
        const uint32_t e = (zl - g * per) * NG + g;
        st.row[e] = (uint8_t)elem(a.mode);
        st.val[e] = h.vals[z];
#pragma unroll
        for (int i = 1; i < N; ++i) {
            const uint32_t m = a.order[i];
            st.off[i - 1][e] = ((h.binds[b * N + m] << h.sb_bits) + elem(m)) * vstride;
        }
    }
    // Dummies: local positions at or past each group's slice length, up to
    // the last position the walk reads (PMAX + D positions when it runs
    // straight-line, per + 2 * D when it loops).
    const uint32_t span = hicoo_straight<NG>() ? (uint32_t)(CHUNK / NG + D) * NG : (per + 2 * D) * NG;
    for (uint32_t e = lane; e < span; e += 64) {
        const uint32_t g = e % NG, j = e / NG;
        // Position j of slice g holds nonzero g * per + j when j < per and
        // that nonzero exists (no subtraction, so nothing can wrap).
        if (j >= per || g * per + j >= cnt) {
            st.row[e] = DUMMY_ROW;
            st.val[e] = 0.f;
#pragma unroll
            for (int i = 0; i < N - 1; ++i) st.off[i][e] = 0u;
        }
    }
    wave_sync();
    return cnt;
}

template <int N>
__device__ __forceinline__ uint32_t stage_coo(const COOArgs &c, uint32_t chunk, Staging<N> &st, int lane) {
    const uint64_t z0 = (uint64_t)chunk * CHUNK;
    const uint64_t rem = c.nnz - z0;
    const uint32_t cnt = rem < CHUNK ? (uint32_t)rem : (uint32_t)CHUNK;
    wave_sync();
    for (uint32_t zl = lane; zl < cnt; zl += 64) {
#pragma unroll
        for (int m = 0; m < N; ++m) st.coord[m][skew(zl)] = c.inds[m][z0 + zl];
        st.val[skew(zl)] = c.vals_override ? c.vals_override[c.perm[z0 + zl]] : c.vals[z0 + zl];
    }
    wave_sync();
    return cnt;
}

// Contribution of position p of the staged chunk for this lane's V ranks;
// returns the output row (INVALID_ROW past the end of the chunk).
template <int N, int V, typename FT>
__device__ __forceinline__ uint32_t contribution(const FactorArgs &a, const Staging<N> &st, uint32_t p, uint32_t cnt,
                                                 uint32_t col, float *out) {
#pragma unroll
    for (int k = 0; k < V; ++k) out[k] = 0.f;
    if (p >= cnt) return INVALID_ROW;
    const uint32_t li = skew(p);
    const uint32_t row = st.coord[a.mode][li];
    if (col < a.R) {
        const float v = st.val[li];
        float t[V];
        FactorLoad<FT, V>::load(a.fac[a.order[1]], (uint64_t)st.coord[a.order[1]][li] * a.stride + col, t);
#pragma unroll
        for (int k = 0; k < V; ++k) out[k] = v * t[k];
#pragma unroll
        for (int i = 2; i < N; ++i) {
            FactorLoad<FT, V>::load(a.fac[a.order[i]], (uint64_t)st.coord[a.order[i]][li] * a.stride + col, t);
#pragma unroll
            for (int k = 0; k < V; ++k) out[k] *= t[k];
        }
    }
    return row;
}

// Where a reduced run of one output row goes.
// GlobalSink: FP32 atomics into the output matrix (COO path).
struct GlobalSink {
    float *out;
    uint32_t out_stride;
    template <int V>
    __device__ __forceinline__ void add(uint32_t row, uint32_t col, const float *val) const {
        float *dst = out + (uint64_t)row * out_stride + col;
#pragma unroll
        for (int k = 0; k < V; ++k) unsafeAtomicAdd(dst + k, val[k]);
    }
};
// Walk a staged chunk (sorted by output row), reduce runs of equal output
// rows across the lane groups and hand each run's sum to the sink once.
template <int N, int V, int G, typename FT, typename Sink>
__device__ __forceinline__ void process_chunk(const FactorArgs &a, const Staging<N> &st, uint32_t cnt,
                                              const Sink &sink, int lane) {
    constexpr int NG = 64 / G;
    const int g = lane / G;
    const int s = lane % G;
    const uint32_t col = (uint32_t)s * V;

    float carry[V];
#pragma unroll
    for (int k = 0; k < V; ++k) carry[k] = 0.f;
    uint32_t carry_row = INVALID_ROW;

    float nxt[V];
    uint32_t nxt_row = contribution<N, V, FT>(a, st, (uint32_t)g, cnt, col, nxt);
    const int nsteps = (int)((cnt + NG - 1) / NG);
    for (int step = 0; step < nsteps; ++step) {
        float val[V];
#pragma unroll
        for (int k = 0; k < V; ++k) val[k] = nxt[k];
        uint32_t row = nxt_row;
        if (step + 1 < nsteps)
            nxt_row = contribution<N, V, FT>(a, st, (uint32_t)((step + 1) * NG + g), cnt, col, nxt);

        // Group 0 continues or closes the segment carried from the last step.
        if (g == 0 && carry_row != INVALID_ROW) {
            if (row == carry_row) {
#pragma unroll
                for (int k = 0; k < V; ++k) val[k] += carry[k];
            } else {
                if (col < a.R) sink.template add<V>(carry_row, col, carry);
            }
        }
        // Segment heads: first group, or row differs from the previous group.
        uint32_t prev_row = __shfl_up(row, G, 64);
        int f = (g == 0) || (prev_row != row);
#pragma unroll
        for (int d = 1; d < NG; d <<= 1) {
            int of = __shfl_up(f, d * G, 64);
            float ov[V];
#pragma unroll
            for (int k = 0; k < V; ++k) ov[k] = __shfl_up(val[k], d * G, 64);
            if (g >= d) {
                if (!f) {
#pragma unroll
                    for (int k = 0; k < V; ++k) val[k] += ov[k];
                }
                f |= of;
            }
        }
        uint32_t next_row = __shfl_down(row, G, 64);
        const bool last_group = (g == NG - 1);
        if (!last_group && next_row != row && row != INVALID_ROW && col < a.R) sink.template add<V>(row, col, val);
        // The last group's segment stays open: hand it to group 0.
        const int src = (NG - 1) * G + s;
        carry_row = __shfl(row, (NG - 1) * G, 64);
#pragma unroll
        for (int k = 0; k < V; ++k) carry[k] = __shfl(val[k], src, 64);
    }
    if (g == 0 && carry_row != INVALID_ROW && col < a.R) sink.template add<V>(carry_row, col, carry);
}

// Lane vector of V factor values as loaded from memory, and its conversion
// to FP32 (BF16 bits are the high half of an FP32 word).
template <typename FT, int V>
struct Vec;
template <> struct Vec<float, 4> { using T = float4; };
template <> struct Vec<float, 2> { using T = float2; };
template <> struct Vec<float, 1> { using T = float; };
template <> struct Vec<uint16_t, 8> { using T = uint4; };
template <> struct Vec<uint16_t, 4> { using T = uint2; };
template <> struct Vec<uint16_t, 2> { using T = uint32_t; };
template <> struct Vec<uint16_t, 1> { using T = uint16_t; };

__device__ __forceinline__ void to_float(const float4 &t, float *o) { o[0] = t.x; o[1] = t.y; o[2] = t.z; o[3] = t.w; }
__device__ __forceinline__ void to_float(const float2 &t, float *o) { o[0] = t.x; o[1] = t.y; }
__device__ __forceinline__ void to_float(const float &t, float *o) { o[0] = t; }
__device__ __forceinline__ void to_float(const uint4 &t, float *o) {
    o[0] = __uint_as_float(t.x << 16); o[1] = __uint_as_float(t.x & 0xFFFF0000u);
    o[2] = __uint_as_float(t.y << 16); o[3] = __uint_as_float(t.y & 0xFFFF0000u);
    o[4] = __uint_as_float(t.z << 16); o[5] = __uint_as_float(t.z & 0xFFFF0000u);
    o[6] = __uint_as_float(t.w << 16); o[7] = __uint_as_float(t.w & 0xFFFF0000u);
}
__device__ __forceinline__ void to_float(const uint2 &t, float *o) {
    o[0] = __uint_as_float(t.x << 16); o[1] = __uint_as_float(t.x & 0xFFFF0000u);
    o[2] = __uint_as_float(t.y << 16); o[3] = __uint_as_float(t.y & 0xFFFF0000u);
}
__device__ __forceinline__ void to_float(const uint32_t &t, float *o) {
    o[0] = __uint_as_float(t << 16); o[1] = __uint_as_float(t & 0xFFFF0000u);
}
__device__ __forceinline__ void to_float(const uint16_t &t, float *o) { o[0] = __uint_as_float((uint32_t)t << 16); }

// Walk a staged chunk (sorted by output row) with each lane group taking a
// contiguous slice of positions: equal rows are summed in registers and each
// run is added to the LDS tile once (a row spanning several slices is added
// once per slice).  The factor-row gathers are software-pipelined D
// positions deep and issued unconditionally (the staging pads every slice
// with dummy entries that gather row 0 and carry an invalid row), and lanes
// past the rank read column 0, so the loop has no branches around loads and
// keeps D * (N - 1) of them in flight.  Lanes past the rank add into tile
// columns >= R, which are never written out.
template <int N, int V, int G, int D, typename FT, int E>
__device__ __forceinline__ void hicoo_slices(const FactorArgs &a, const HStage<N, E> &st, uint32_t cnt, float *tile,
                                             uint32_t tp, int lane) {
    using VT = typename Vec<FT, V>::T;
    constexpr int NG = 64 / G;
    const uint32_t g = (uint32_t)lane / G;
    const uint32_t col = (uint32_t)(lane % G) * V;
    const uint32_t lcol = col < a.R ? col : 0u;
    const uint32_t per = (cnt + NG - 1) / NG;
    const VT *base[N - 1];
#pragma unroll
    for (int i = 0; i < N - 1; ++i)
        base[i] = reinterpret_cast<const VT *>(static_cast<const FT *>(a.fac[a.order[i + 1]]) + lcol);
    VT raw[D][N - 1];
    float vv[D];
    uint32_t rr[D];
    auto issue = [&](int j, uint32_t e) {
        rr[j] = st.row[e];
        vv[j] = st.val[e];
#pragma unroll
        for (int i = 0; i < N - 1; ++i) raw[j][i] = base[i][st.off[i][e]];
    };
#pragma unroll
    for (int j = 0; j < D; ++j) issue(j, j * NG + g);
    float acc[V];
#pragma unroll
    for (int k = 0; k < V; ++k) acc[k] = 0.f;
    uint32_t cur = DUMMY_ROW;
    // Consume slot j (position p), refill it with position p + D.
    auto step = [&](int j, uint32_t next) {
        float prod[V], t[V];
        to_float(raw[j][0], prod);
#pragma unroll
        for (int k = 0; k < V; ++k) prod[k] *= vv[j];
#pragma unroll
        for (int i = 1; i < N - 1; ++i) {
            to_float(raw[j][i], t);
#pragma unroll
            for (int k = 0; k < V; ++k) prod[k] *= t[k];
        }
        const uint32_t row = rr[j];
        issue(j, next);
        const bool brk = row != cur;
        if (brk && cur != DUMMY_ROW) {
            float *dst = tile + cur * tp + col;
#pragma unroll
            for (int k = 0; k < V; ++k) atomicAdd(dst + k, acc[k]);
        }
#pragma unroll
        for (int k = 0; k < V; ++k) acc[k] = brk ? prod[k] : acc[k] + prod[k];
        cur = row;
    };
    constexpr int PMAX = CHUNK / NG;   // slice length of a full chunk
    if constexpr (hicoo_straight<NG>()) {
        // Straight-line walk over a full chunk's slice (shorter slices run
        // into their dummy entries): with no loop back-edge the compiler's
        // wait counts follow the ring exactly, keeping D slots of gathers in
        // flight, and every LDS read has a constant offset.
#pragma unroll
        for (int p = 0; p < PMAX; ++p) step(p % D, (uint32_t)(p + D) * NG + g);
    } else {
        for (uint32_t b = 0; b < per; b += D) {
#pragma unroll
            for (int j = 0; j < D; ++j) step(j, (b + j + D) * NG + g);
        }
    }
    if (cur != DUMMY_ROW) {
        float *dst = tile + cur * tp + col;
#pragma unroll
        for (int k = 0; k < V; ++k) atomicAdd(dst + k, acc[k]);
    }
}

// Pipeline depth of the gathers: D * (N - 1) vector loads in flight per lane.
template <int N, int V>
struct PipeDepth {
    static constexpr int value = (N - 1) * V <= 8 ? 4 : ((N - 1) * V <= 16 ? 2 : 1);
};

// W waves per workgroup (W * 64 threads), one chunk per wave at a time.
template <int N, int V, int G, typename FT, int W>
__global__ void __launch_bounds__(W * 64, 6) hicoo_tile_kernel(HiCOOArgs h, FactorArgs a) {
    constexpr int NG = 64 / G;
    constexpr int D = PipeDepth<N, V>::value;
    constexpr int E = hstage_entries<NG, D>();
    extern __shared__ float tile[];  // (1 << sb_bits) rows x TP floats
    __shared__ HStage<N, E> st_all[W];
    constexpr int GV = G * V;
    constexpr int TP = GV + 1;
    constexpr int NT = W * 64;
    const int tid = threadIdx.x, wid = tid >> 6, lane = tid & 63;
    const uint32_t u = remap_block(blockIdx.x, gridDim.x);   // gridDim.x == nunits
    const uint32_t rows = 1u << h.sb_bits;
    for (uint32_t i = tid; i < rows * TP; i += NT) tile[i] = 0.f;
    __syncthreads();
    const uint32_t c0 = h.unit_chunk[u], c1 = h.unit_chunk[u + 1];
    for (uint32_t c = c0 + wid; c < c1; c += W) {
        const uint32_t cnt = stage_hicoo<N, V, NG, D, E>(h, a, c, st_all[wid], lane);
        if (cnt) hicoo_slices<N, V, G, D, FT, E>(a, st_all[wid], cnt, tile, TP, lane);
    }
    __syncthreads();
    const uint32_t base = h.unit_br[u] << h.sb_bits;
    const bool excl = h.unit_excl[u] != 0;
    const bool whole = excl && a.out_cols != 0;
    for (uint32_t i = tid; i < rows * GV; i += NT) {
        const uint32_t r = i / GV, c = i % GV;
        const uint32_t orow = base + r;
        if (orow >= h.nrows) continue;
        float *dst = a.out + (uint64_t)orow * a.out_stride + c;
        if (whole) {
            if (c < a.out_cols) *dst = c < a.R ? tile[r * TP + c] : 0.f;
            continue;
        }
        if (c >= a.R) continue;
        const float v = tile[r * TP + c];
        if (v == 0.f) continue;
        if (excl)
            *dst = v;
        else
            unsafeAtomicAdd(dst, v);
    }
}

// Static LDS of the HiCOO kernel's staging area for W waves.
template <int N, int V, int G>
constexpr size_t hicoo_stage_bytes(int W) {
    return (size_t)W * sizeof(HStage<N, hstage_entries<64 / G, PipeDepth<N, V>::value>()>);
}

// Stream form of a mode copy (hicoo_stream_kernel): decoded nonzeros kept in
// HBM, so that an MTTKRP neither decodes block coordinates nor stages chunks
// in LDS.  Each block-row's nonzeros are sorted (stably) by output row over
// the whole block-row, not only inside a chunk, so a row's nonzeros are
// consecutive and most of its contributions are summed in registers, and
// inside a row by the coordinate of the largest non-target mode, so repeated
// factor rows are gathered back to back; the chunk and unit boundaries are
// those of the copy.  An entry has N + 1 words: the coordinates of the N - 1
// non-target modes in ascending mode order, the value bits and the output row
// inside the block-row; positions past a slice's end are all-zero with row
// DUMMY_ROW (packed forms of orders 4 and 5: four words, StreamShape).
// Entries are stored in NP = ceil(words / 4) planes of 16-byte words.  Chunk c owns NP * CHUNK 16-byte words starting at c * NP * CHUNK,
// plane q first at offset q * CHUNK: its nonzeros are cut into NG = 64 / G
// slices of per = ceil(cnt / NG) positions (G lanes per slice) and position j
// of slice g is entry stream_entry(G, g, j) = 64 * (j / G) + G * g + j % G, so
// 64 consecutive entries of one plane are one wave-wide 16-byte-per-lane load
// in which each lane group holds G positions of its own slice.
__host__ __device__ constexpr uint32_t stream_entry(uint32_t G, uint32_t g, uint32_t j) {
    return 64u * (j / G) + G * g + j % G;
}

// Packed forms (orders 4 and 5, every gathered coordinate below 2^24, rows
// of at most 7 bits): four words, one plane, half the bytes of the two-plane
// entry.  Order 4: {c0 | row << 24, c1, c2, value}; order 5: {c0 | row << 24,
// c1 | (c3 & 255) << 24, c2 | (c3 >> 8) << 24, value}, which needs the last
// gathered mode to have at most 2^16 rows.  The walk's 24-bit multiplies
// ignore the high bytes, so only the row and c3 take shifts to unpack.
template <int N, bool PACK = false>
struct StreamShape {
    static_assert(!PACK || N == 4 || N == 5, "packed entries exist for orders 4 and 5");
    static constexpr int WORDS = PACK ? 4 : N + 1;
    static constexpr int NP = (WORDS + 3) / 4;
};

struct StreamArgs {
    const uint4 *ent;
    const uint32_t *unit_chunk;
    const uint32_t *unit_br;
    const uint32_t *unit_excl;
    uint32_t nunits;
    uint32_t sb_bits;
    uint32_t nrows;
    // 1 when every gathered coordinate and the factor row pitch in bytes are
    // below 2^24: row offsets then take one full-rate 24-bit multiply-add
    // instead of a quarter-rate 32-bit one.
    uint32_t coord24;
    // Optional: units in order of decreasing chunk count; workgroup b runs
    // unit_order[b], so the hardware's round-robin dealing gives every XCD
    // an even share and the small units run last.  Null: remap_block order.
    const uint32_t *unit_order;
    // LDS-resident factors (kernels instantiated with NRES > 0): the rows of
    // the last NRES gathered modes, which the workgroup copies whole into LDS.
    uint32_t res_rows[2];
    // 1 when the stream form holds packed entries (StreamShape<N, true>).
    uint32_t packed;
};

// Byte offset of the LDS-resident factors behind a tile of `rows` x tp floats.
__host__ __device__ constexpr uint32_t res_lds_offset(uint32_t rows, uint32_t tp) {
    return (rows * tp * 4u + 15u) & ~15u;
}

__device__ __forceinline__ uint32_t stream_unit(const StreamArgs &h) {
    return h.unit_order ? h.unit_order[blockIdx.x] : remap_block(blockIdx.x, gridDim.x);
}

// Broadcast, inside each group of G lanes, the word that lane J of the group
// holds.  Groups of 4, 8 and 16 lanes use DPP moves (VALU work: the LDS pipe,
// which also serves the tile's atomics, stays free).  4 lanes: a quad_perm
// copies lane J of each quad to the whole quad.  8 lanes: the quad_perm, then
// a 4-lane row shift written only to the group's other quad (bank mask)
// copies quad J / 4 there.  16 lanes: row_newbcast copies lane J of each
// 16-lane row to the whole row.  Wider groups use ds_bpermute.
template <int G, int J>
__device__ __forceinline__ uint32_t group_bcast(uint32_t x) {
    static_assert(G == 4 || G == 8 || G == 16, "DPP broadcast covers groups of 4, 8 and 16 lanes");
    if constexpr (G == 16) {
        return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x150 + J, 0xF, 0xF, false);
    } else {
        constexpr int QP = (J % 4) * 0x55;  // quad_perm [J%4, J%4, J%4, J%4]
        const int t = __builtin_amdgcn_mov_dpp((int)x, QP, 0xF, 0xF, false);
        if constexpr (G == 4) {
            return (uint32_t)t;
        } else {
            // row_shr:4 into lanes 4-7 and 12-15 of each row, or row_shl:4 into 0-3 and 8-11.
            constexpr int CTRL = J < 4 ? 0x114 : 0x104;
            constexpr int BANKS = J < 4 ? 0xA : 0x5;
            return (uint32_t)__builtin_amdgcn_update_dpp(t, t, CTRL, 0xF, BANKS, false);
        }
    }
}

// Lane J of each group's word x (J < G); gbase is the byte address of lane 0
// of the caller's group.
template <int G, int J>
__device__ __forceinline__ uint32_t group_word(uint32_t x, int gbase) {
    if constexpr (G <= 16)
        return group_bcast<G, J>(x);
    else
        return (uint32_t)__builtin_amdgcn_ds_bpermute(gbase + J * 4, (int)x);
}

// Byte offset of column byte colb in factor row w of pitch rowb bytes.
template <bool U24>
__device__ __forceinline__ uint32_t row_offset(uint32_t w, uint32_t rowb, uint32_t colb) {
    if constexpr (U24)
        return __umul24(w, rowb) + colb;
    else
        return w * rowb + colb;
}

// Sum of x over each group of G (4, 8 or 16) lanes, in every lane of the
// group, with DPP moves: swaps inside pairs and quads (quad_perm), then the
// mirrored 8-lane half-row (row_half_mirror) and the mirrored 16-lane row
// (row_mirror), each of which pairs every lane with one of the other half.
template <int G>
__device__ __forceinline__ float group_sum(float x) {
    static_assert(G == 4 || G == 8 || G == 16, "DPP sums cover groups of 4, 8 and 16 lanes");
    x += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0xB1, 0xF, 0xF, false));   // quad_perm [1,0,3,2]
    x += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x4E, 0xF, 0xF, false));   // quad_perm [2,3,0,1]
    if constexpr (G >= 8) x += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x141, 0xF, 0xF, false));
    if constexpr (G >= 16) x += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x140, 0xF, 0xF, false));
    return x;
}

// f(std::integral_constant<int, I>) for I = B .. E - 1, in order.
template <int B, int E, typename F>
__device__ __forceinline__ void static_for(F &&f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

// Walk one chunk of the stream form.  The chunk's entries are loaded once,
// one 16-byte word per lane per wave-wide block (NP * CHUNK / 64 blocks, read
// non-temporally: each is used once): lane G * g + j of block b of a plane
// holds position b * G + j of lane group g.  A lane group takes a word of
// position p from its own lane p % G (group_word), so entries cost NP
// vector-memory instructions per G positions and no LDS traffic for groups of
// up to 16 lanes.  Each lane group runs through its slice (PMAX positions,
// padded with dummies) in straight-line code, gathering the N - 1 factor rows
// D positions ahead, so every wait count is exact.  The product is formed in
// ascending mode order.  Runs of equal rows are summed in registers and added
// to the LDS tile once per slice.
// Where a slice's row runs go: the workgroup's LDS tile of the block-row,
// indexed by the row inside the block-row; lanes past the rank add into tile
// columns that are never written out.
struct TileSink {
    static constexpr uint32_t DUMMY = DUMMY_ROW;
    float *tile;
    uint32_t tp;
    template <int V>
    __device__ __forceinline__ void add(uint32_t row, uint32_t col, const float *v) const {
        float *dst = tile + (__umul24(row, tp) + col);  // row < 2^8, tp <= 65
#pragma unroll
        for (int k = 0; k < V; ++k) atomicAdd(dst + k, v[k]);
    }
};
// RESID (CP-completion gradient): each entry's value x is replaced by its
// residual e = x - sum_r prod_m A_m[i_m][r], the target row's factor row
// read from trow (the block-row's rows of A_mode in LDS, pitch TRP floats,
// rows past the block-row clamped to trow_last); dummies contribute 0.
// NRES > 0: the last NRES gathered factors (ascending mode order) are read
// from their whole copies in LDS at res (G * V elements per row, rows of the
// first resident factor, then of the second), every lane of every position,
// so the walk stays free of branches.
// PACK: the entries are packed (StreamShape<N, true>); needs U24.
template <int N, int V, int G, int D, typename FT, bool U24, bool RESID = false, int NRES = 0, bool PACK = false,
          typename Sink>
__device__ __forceinline__ void stream_walk(const FactorArgs &a, const uint4 *__restrict__ ent, const Sink &sink,
                                            int lane, const float *trow = nullptr, uint32_t trow_last = 0,
                                            double *esq_sum = nullptr, const char *res = nullptr,
                                            uint32_t res1_off = 0) {
    static_assert(!PACK || U24, "packed coordinate words carry other fields in their high byte");
    // Sum of squared residuals (RESID with esq_sum): kept in a register and
    // added to *esq_sum once at the end.
    double esq = 0.0;
    using VT = typename Vec<FT, V>::T;
    constexpr int K = N - 1;
    constexpr int NP = StreamShape<N, PACK>::NP;
    constexpr int NG = 64 / G;
    constexpr int PMAX = CHUNK / NG;
    constexpr int NB = CHUNK / 64;
    const uint32_t col = (uint32_t)(lane % G) * V;
    // Uniform bases and 32-bit byte offsets (the host checks that every
    // factor matrix spans < 4 GiB), so each load takes one offset register.
    const char *f[K];
#pragma unroll
    for (int i = 0; i < K; ++i) f[i] = static_cast<const char *>(a.fac[(uint32_t)i < a.mode ? i : i + 1]);
    const uint32_t rowb = a.stride * (uint32_t)sizeof(FT);
    const uint32_t colb = (col < a.R ? col : 0u) * (uint32_t)sizeof(FT);
    typedef unsigned v4u __attribute__((ext_vector_type(4)));
    const v4u *eb = reinterpret_cast<const v4u *>(ent) + lane;
    v4u blk[NP][NB];
#pragma unroll
    for (int q = 0; q < NP; ++q)
#pragma unroll
        for (int b = 0; b < NB; ++b) blk[q][b] = __builtin_nontemporal_load(eb + q * CHUNK + b * 64);
    const int gbase = (lane & ~(G - 1)) * 4;  // byte address of lane 0 of this lane's group
    // Positions are compile-time constants (static_for), so each hand-out
    // names its source lane in the instruction.
    auto word = [&](auto P, int w) -> uint32_t {
        constexpr int p = decltype(P)::value;
        return group_word<G, p % G>(blk[w / 4][p / G][w % 4], gbase);
    };
    // Coordinate word of gathered position i (packed order 5: c3 sits in the
    // high bytes of words 1 and 2; the other high bytes are ignored by the
    // 24-bit multiplies below).
    auto coord = [&](auto P, int i) -> uint32_t {
        if (PACK && K == 4 && i == 3) return (word(P, 1) >> 24) | ((word(P, 2) >> 24) << 8);
        return word(P, i);
    };
    VT raw[D][K];
    float vv[D];
    uint32_t rr[D];
    constexpr uint32_t RRB = G * V * (uint32_t)sizeof(FT);   // bytes per resident row in LDS
    auto issue = [&](auto P) {
        constexpr int p = decltype(P)::value;
#pragma unroll
        for (int i = 0; i < K; ++i) {
            if (i >= K - NRES) {
                const uint32_t off = (i == K - NRES ? 0u : res1_off) + __umul24(coord(P, i), RRB) + colb;
                raw[p % D][i] = *static_cast<const VT *>(__builtin_assume_aligned(res + off, sizeof(VT)));
            } else {
                raw[p % D][i] = *reinterpret_cast<const VT *>(f[i] + (uint64_t)row_offset<U24>(coord(P, i), rowb, colb));
            }
        }
        if constexpr (PACK) {
            vv[p % D] = __uint_as_float(word(P, 3));
            rr[p % D] = word(P, 0) >> 24;
        } else {
            vv[p % D] = __uint_as_float(word(P, K));
            rr[p % D] = word(P, K + 1);
        }
    };
    static_for<0, (D < PMAX ? D : PMAX)>(issue);
    float acc[V];
#pragma unroll
    for (int k = 0; k < V; ++k) acc[k] = 0.f;
    uint32_t cur = Sink::DUMMY;
    static_for<0, PMAX>([&](auto P) {
        constexpr int p = decltype(P)::value;
        if constexpr (RESID) {
            constexpr uint32_t TRP = G * V + 4;
            float pp[V], t[V];
            to_float(raw[p % D][0], pp);
#pragma unroll
            for (int i = 1; i < K; ++i) {
                to_float(raw[p % D][i], t);
#pragma unroll
                for (int k = 0; k < V; ++k) pp[k] *= t[k];
            }
            const uint32_t row = rr[p % D];
            const float *tr = trow + (row < trow_last ? row : trow_last) * TRP + col;
            float part = 0.f;
#pragma unroll
            for (int k = 0; k < V; ++k) part = __builtin_fmaf(pp[k], tr[k], part);
            part = group_sum<G>(part);
            const float e = row == Sink::DUMMY ? 0.f : vv[p % D] - part;
            if (col == 0) esq += (double)e * (double)e;   // one lane per entry
            const bool brk = row != cur;
            if (brk && cur != Sink::DUMMY) sink.template add<V>(cur, col, acc);
#pragma unroll
            for (int k = 0; k < V; ++k) {
                float base = brk ? 0.f : acc[k];
                asm volatile("" : "+v"(base));
                acc[k] = __builtin_fmaf(e, pp[k], base);
            }
            if constexpr (p + D < PMAX) issue(std::integral_constant<int, p + D>{});
            cur = row;
            return;
        }
        // prod = v * A_0 * ... * A_{K-2}; the last factor is fused into the
        // run sum: acc = fma(prod, A_{K-1}, acc), restarted at 0 on a new row.
        float prod[V], t[V];
        to_float(raw[p % D][0], prod);
        const float v = vv[p % D];
#pragma unroll
        for (int k = 0; k < V; ++k) prod[k] *= v;
#pragma unroll
        for (int i = 1; i < K - 1; ++i) {
            to_float(raw[p % D][i], t);
#pragma unroll
            for (int k = 0; k < V; ++k) prod[k] *= t[k];
        }
        const uint32_t row = rr[p % D];
        const bool brk = row != cur;
        if (brk && cur != Sink::DUMMY) sink.template add<V>(cur, col, acc);
        if constexpr (K > 1) {
            to_float(raw[p % D][K - 1], t);
#pragma unroll
            for (int k = 0; k < V; ++k) {
                float base = brk ? 0.f : acc[k];
                asm volatile("" : "+v"(base));  // one select and one fma, not fma(.., 0) plus fma plus select
                acc[k] = __builtin_fmaf(prod[k], t[k], base);
            }
        } else {
#pragma unroll
            for (int k = 0; k < V; ++k) acc[k] = (brk ? 0.f : acc[k]) + prod[k];
        }
        if constexpr (p + D < PMAX) issue(std::integral_constant<int, p + D>{});
        cur = row;
    });
    if (cur != Sink::DUMMY) sink.template add<V>(cur, col, acc);
    if constexpr (RESID)
        if (esq_sum) *esq_sum += esq;
}

// Pipeline of the stream walk: gather depth D (about 32 registers of rows in
// flight) and the minimum resident waves per SIMD the register budget is
// sized for.  Measured on the 100M-nonzero order-3 Zipf tensor at R = 32
// (FP32): with the DPP entry hand-out (80 VGPRs at D = 4) D = 4, 5, 6 and 8
// at four to six waves per SIMD were within 1% of each other, the kernel
// being limited by the vector-memory path's throughput rather than by
// latency; D = 4 keeps the smallest register budget.
// Up to 12 registers of gathered rows per position (order 4 at four FP32 or
// eight BF16 ranks per lane) still fit D = 4 without spills at four waves per
// SIMD (104-116 VGPRs); measured 3.5% faster than D = 2 on the 200M-nonzero
// order-4 tensor at R = 64.  The CP-completion gradient walk, which also
// holds the target row's sums, keeps D_RESID (D = 4 spills there).
// With NRES positions read from LDS-resident copies, only the other
// positions' rows come from memory: where those take at most 12 registers the
// walk gathers four positions ahead as well (order 5 at four FP32 ranks per
// lane with one or two resident factors: the CP-ALS leg's five MTTKRPs went
// from 1.42 / 1.36 / 1.33 / 1.63 / 1.42 to 1.37 / 1.30 / 1.26 / 1.59 / 1.37 ms).
template <int N, int V, typename FT>
struct StreamConfig {
    static constexpr int ROW_REGS = (N - 1) * V * (int)sizeof(FT) / 4;
    static constexpr int D = ROW_REGS <= 12 ? 4 : (ROW_REGS <= 16 ? 2 : 1);
    template <int NRES>
    static constexpr int depth() {
        return (N - 1 - NRES) * V * (int)sizeof(FT) / 4 <= 12 ? 4 : D;
    }
    static constexpr int D_RESID = ROW_REGS <= 8 ? 4 : (ROW_REGS <= 16 ? 2 : 1);
    static constexpr int OCC = N > 3 ? 4 : ((int)sizeof(FT) * V >= 16 ? 5 : 6);
};

// HiCOO MTTKRP on the stream form: one workgroup per unit, W waves taking the
// unit's chunks in turn, an LDS tile of the block-row's output rows written
// back once (stores when the unit owns the block-row, else atomics).  The
// factor gathers address a uniform base with 32-bit byte offsets: the host
// sends factor matrices of 4 GiB or more to the tile kernel.
// NRES > 0: the workgroup first copies the last NRES gathered factors whole
// (h.res_rows rows each, the rank tile's G * V columns, zero past the rank)
// into LDS behind the tile; the walk then reads them there.
// PACK: the form holds packed entries (orders 4 and 5; the host sends such
// forms here only with coord24).
template <int N, int V, int G, typename FT, int W, int NRES = 0, bool PACK = false>
__global__ void __launch_bounds__(W * 64, (StreamConfig<N, V, FT>::OCC))
    hicoo_stream_kernel(StreamArgs h, FactorArgs a) {
    constexpr int D = StreamConfig<N, V, FT>::template depth<NRES>();
    constexpr int NP = StreamShape<N, PACK>::NP;
    extern __shared__ float tile[];  // (1 << sb_bits) rows x TP floats [, resident factors]
    constexpr int GV = G * V;
    constexpr int TP = GV + 1;
    constexpr int NT = W * 64;
    const int tid = threadIdx.x, wid = tid >> 6, lane = tid & 63;
    const uint32_t u = stream_unit(h);   // gridDim.x == nunits
    const uint32_t rows = 1u << h.sb_bits;
    for (uint32_t i = tid; i < rows * TP; i += NT) tile[i] = 0.f;
    char *res = nullptr;
    uint32_t res1_off = 0;
    if constexpr (NRES > 0) {
        using VT = typename Vec<FT, V>::T;
        constexpr uint32_t RRB = GV * (uint32_t)sizeof(FT);
        constexpr int K = N - 1;
        res = reinterpret_cast<char *>(tile) + res_lds_offset(rows, TP);
        res1_off = h.res_rows[0] * RRB;
        const uint32_t rowb = a.stride * (uint32_t)sizeof(FT);
#pragma unroll
        for (int j = 0; j < NRES; ++j) {
            const int i = K - NRES + j;   // gathered position, ascending mode order
            const char *fi = static_cast<const char *>(a.fac[(uint32_t)i < a.mode ? i : i + 1]);
            char *dst = res + (j == 0 ? 0u : res1_off);
            const uint32_t n = h.res_rows[j] * G;
            for (uint32_t x = tid; x < n; x += NT) {
                const uint32_t r = x / G, v = x % G;
                VT val{};
                if (v * V < a.R) val = *reinterpret_cast<const VT *>(fi + (uint64_t)r * rowb + v * V * sizeof(FT));
                *static_cast<VT *>(__builtin_assume_aligned(dst + r * RRB + v * V * sizeof(FT), sizeof(VT))) = val;
            }
        }
    }
    __syncthreads();
    const uint32_t c0 = h.unit_chunk[u], c1 = h.unit_chunk[u + 1];
    const uint32_t w0 = __builtin_amdgcn_readfirstlane(wid);
    if (PACK || h.coord24) {
        for (uint32_t c = c0 + w0; c < c1; c += W)
            stream_walk<N, V, G, D, FT, true, false, NRES, PACK>(a, h.ent + (uint64_t)c * NP * CHUNK,
                                                                  TileSink{tile, TP}, lane, nullptr, 0, nullptr, res,
                                                                  res1_off);
    } else if constexpr (!PACK) {
        for (uint32_t c = c0 + w0; c < c1; c += W)
            stream_walk<N, V, G, D, FT, false, false, NRES>(a, h.ent + (uint64_t)c * NP * CHUNK, TileSink{tile, TP},
                                                             lane, nullptr, 0, nullptr, res, res1_off);
    }
    __syncthreads();
    const uint32_t base = h.unit_br[u] << h.sb_bits;
    const bool excl = h.unit_excl[u] != 0;
    const bool whole = excl && a.out_cols != 0;
    for (uint32_t i = tid; i < rows * GV; i += NT) {
        const uint32_t r = i / GV, c = i % GV;
        const uint32_t orow = base + r;
        if (orow >= h.nrows) continue;
        float *dst = a.out + (uint64_t)orow * a.out_stride + c;
        if (whole) {
            if (c < a.out_cols) *dst = c < a.R ? tile[r * TP + c] : 0.f;
            continue;
        }
        if (c >= a.R) continue;
        const float v = tile[r * TP + c];
        if (v == 0.f) continue;
        if (excl)
            *dst = v;
        else
            unsafeAtomicAdd(dst, v);
    }
}

// CP-completion gradient on a COO stream form: out[i_n] += e_z * prod_{m != n}
// A_m[i_m] with the residual e_z = x_z - sum_r prod_m A_m[i_m][r] formed in
// the walk, so the residuals need not be gathered through the sort
// permutation.  As hicoo_stream_kernel (FP32 factors), plus the unit's
// block-row of A_n (2^sb rows, zero past the tensor's rows and the rank) in
// LDS after the tile.
// loss_part (optional): per unit, the sum of e_z^2 over its entries in a
// fixed order (each wave's chunks in turn, then the waves in order).
// Four resident waves per SIMD: the residual's group sum and the FP64 loss
// term need registers beyond the plain walk's budget at 16-lane groups.
template <int N, int V, int G, int W>
__global__ void __launch_bounds__(W * 64, 4)
    completion_grad_kernel(StreamArgs h, FactorArgs a, double *loss_part) {
    constexpr int D = StreamConfig<N, V, float>::D_RESID;
    constexpr int NP = StreamShape<N>::NP;
    extern __shared__ float tile[];  // (1 << sb_bits) rows x TP floats, then as many rows x TRP
    constexpr int GV = G * V;
    constexpr int TP = GV + 1;
    constexpr int TRP = GV + 4;
    constexpr int NT = W * 64;
    const int tid = threadIdx.x, wid = tid >> 6, lane = tid & 63;
    const uint32_t u = stream_unit(h);   // gridDim.x == nunits
    const uint32_t rows = 1u << h.sb_bits;
    const uint32_t base = h.unit_br[u] << h.sb_bits;
    float *trow = tile + rows * TP;
    const float *An = static_cast<const float *>(a.fac[a.mode]);
    for (uint32_t i = tid; i < rows * TP; i += NT) tile[i] = 0.f;
    for (uint32_t i = tid; i < rows * GV; i += NT) {
        const uint32_t r = i / GV, c = i % GV;
        trow[r * TRP + c] = (c < a.R && base + r < h.nrows) ? An[(uint64_t)(base + r) * a.stride + c] : 0.f;
    }
    __syncthreads();
    const uint32_t c0 = h.unit_chunk[u], c1 = h.unit_chunk[u + 1];
    const uint32_t w0 = __builtin_amdgcn_readfirstlane(wid);
    const TileSink sink{tile, TP};
    double esq = 0.0;   // this lane's sum of squared residuals (used when loss_part is set)
    double *pe = &esq;
    if (h.coord24) {
        for (uint32_t c = c0 + w0; c < c1; c += W)
            stream_walk<N, V, G, D, float, true, true>(a, h.ent + (uint64_t)c * NP * CHUNK, sink, lane, trow, rows - 1,
                                                       pe);
    } else {
        for (uint32_t c = c0 + w0; c < c1; c += W)
            stream_walk<N, V, G, D, float, false, true>(a, h.ent + (uint64_t)c * NP * CHUNK, sink, lane, trow,
                                                        rows - 1, pe);
    }
    if (loss_part) {
        __shared__ double wsum[W];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) esq += __shfl_xor(esq, off, 64);
        if (lane == 0) wsum[wid] = esq;
        __syncthreads();
        if (tid == 0) {
            double t = 0.0;
            for (int w = 0; w < W; ++w) t += wsum[w];
            loss_part[u] = t;
        }
    }
    __syncthreads();
    const bool excl = h.unit_excl[u] != 0;
    for (uint32_t i = tid; i < rows * GV; i += NT) {
        const uint32_t r = i / GV, c = i % GV;
        const uint32_t orow = base + r;
        if (c >= a.R || orow >= h.nrows) continue;
        const float v = tile[r * TP + c];
        if (v == 0.f) continue;
        float *dst = a.out + (uint64_t)orow * a.out_stride + c;
        if (excl)
            *dst = v;
        else
            unsafeAtomicAdd(dst, v);
    }
}

template <int N, int V, int G, typename FT>
__global__ void __launch_bounds__(256) coo_mttkrp_kernel(COOArgs c, FactorArgs a) {
    __shared__ Staging<N> st_all[WAVES_PER_BLOCK];
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t chunk = remap_block(blockIdx.x, gridDim.x) * WAVES_PER_BLOCK + wid;
    if (chunk >= c.nchunks) return;
    Staging<N> &st = st_all[wid];
    const uint32_t cnt = stage_coo<N>(c, chunk, st, lane);
    process_chunk<N, V, G, FT>(a, st, cnt, GlobalSink{a.out, a.out_stride}, lane);
}

// Sum of x over each group of G lanes, in every lane of the group: DPP moves
// for groups of up to 16 lanes, a butterfly of shuffles for wider groups.
template <int G>
__device__ __forceinline__ float group_sum_any(float x) {
    if constexpr (G <= 16) {
        return group_sum<G>(x);
    } else {
#pragma unroll
        for (int off = G / 2; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
        return x;
    }
}

// Sampled reconstruction at the nonzeros (CP completion forward):
//   yhat_z = sum_r prod_m A_m[i_m][r],  resid_z = x_z - yhat_z,
// and each workgroup's partial sum of resid_z^2 in FP64.  Storage order, one
// wave per chunk of CHUNK nonzeros, taken as blocks of 64: lane l loads the
// coordinates and the value of nonzero l of the block (N + 1 coalesced loads
// per 64 nonzeros), and lane group g (G lanes x V ranks) then works through
// the G nonzeros its own lanes hold, taking each one's coordinates from its
// lane with group_word (DPP for groups of up to 16 lanes), gathering the N
// factor rows (stepping over column blocks of G * V ranks when R is wider)
// and summing over the ranks within the group.  The lane that holds a
// nonzero keeps its residual, so the block's residuals are written by one
// coalesced store.  Lanes past the chunk's end repeat the block's first
// nonzero; their results are neither stored nor counted.
template <int N, int V, int G>
__global__ void __launch_bounds__(256) completion_forward_kernel(COOArgs c, FactorArgs a, float *resid,
                                                                 double *partials) {
    __shared__ double wsum[WAVES_PER_BLOCK];
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t col = (uint32_t)(lane % G) * V;
    const int gbase = (lane & ~(G - 1)) * 4;  // byte address of lane 0 of this lane's group
    const uint64_t chunk = (uint64_t)blockIdx.x * WAVES_PER_BLOCK + wid;
    double sq = 0.0;
    if (chunk < c.nchunks) {
        const uint64_t z0 = chunk * CHUNK;
        const uint64_t rem = c.nnz - z0;
        const uint32_t cnt = rem < CHUNK ? (uint32_t)rem : (uint32_t)CHUNK;
        for (uint32_t b = 0; b < cnt; b += 64) {
            const bool mine = b + (uint32_t)lane < cnt;
            const uint64_t z = z0 + b + (mine ? (uint32_t)lane : 0u);
            uint32_t w[N];
#pragma unroll
            for (int m = 0; m < N; ++m) w[m] = c.inds[m][z];
            const float x = c.vals[z];
            float e = 0.f;
            static_for<0, G>([&](auto J) {
                constexpr int j = decltype(J)::value;
                uint32_t row[N];
#pragma unroll
                for (int m = 0; m < N; ++m) row[m] = group_word<G, j>(w[m], gbase);
                float part = 0.f;
                for (uint32_t cc = col; cc < a.R; cc += G * V) {
                    float prod[V], t[V];
                    FactorLoad<float, V>::load(a.fac[0], (uint64_t)row[0] * a.stride + cc, prod);
#pragma unroll
                    for (int m = 1; m < N; ++m) {
                        FactorLoad<float, V>::load(a.fac[m], (uint64_t)row[m] * a.stride + cc, t);
#pragma unroll
                        for (int k = 0; k < V; ++k) prod[k] *= t[k];
                    }
#pragma unroll
                    for (int k = 0; k < V; ++k) part += prod[k];
                }
                part = group_sum_any<G>(part);
                if (lane % G == j) e = x - part;
            });
            if (mine) {
                resid[z] = e;
                sq += (double)e * (double)e;
            }
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) sq += __shfl_xor(sq, off, 64);
    if (lane == 0) wsum[wid] = sq;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < WAVES_PER_BLOCK; ++w) t += wsum[w];
        partials[blockIdx.x] = t;
    }
}

}  // namespace dev
}  // namespace pti
