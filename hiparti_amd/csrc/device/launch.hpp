// Launch entry points of the MTTKRP kernel instantiations.  Each tensor order
// N in [2, 8] is compiled in its own object (mttkrp_inst.hip built with
// -DPTI_NMODES=N); this header declares the per-N launchers and a dispatcher.
#pragma once

#include <hip/hip_runtime.h>

#include "mttkrp_kernels.hpp"
#include "runtime.hpp"

namespace pti {
namespace dev {

// Lanes per nonzero G and ranks per lane V chosen for a rank R and strides.
struct Shape {
    int V;
    int G;
};

// Ranks per lane: 4 when R and every stride are multiples of 4 (16-byte row
// pieces), else 2 when they are even, else 1.
inline int vector_width(uint32_t R, uint32_t stride, uint32_t out_stride) {
    if (R % 4 == 0 && stride % 4 == 0 && out_stride % 4 == 0) return 4;
    if (R % 2 == 0 && stride % 2 == 0 && out_stride % 2 == 0) return 2;
    return 1;
}

// V = vector_width; G = ceil(R / V) rounded up to a power of two in [4, 64].
// Returns false when R exceeds 64 * V; callers split larger ranks into
// column tiles (64 * V columns for COO, 64 for the HiCOO LDS tile) or loop
// over column blocks (CP-completion forward).
inline bool choose_shape(uint32_t R, uint32_t stride, uint32_t out_stride, Shape &sh) {
    if (R == 0) return false;
    const int V = vector_width(R, stride, out_stride);
    uint32_t lanes = (R + V - 1) / V;
    int G = 4;
    while ((uint32_t)G < lanes) G <<= 1;
    if (G > 64) return false;
    sh.V = V;
    sh.G = G;
    return true;
}

// StreamArgs::coord24: every non-target mode has at most 2^24 rows and a
// factor row spans fewer than 2^24 bytes.
inline uint32_t coords_fit_u24(const uint32_t *ndims, uint32_t N, uint32_t mode, uint64_t row_bytes) {
    if (row_bytes >= (1ull << 24)) return 0;
    for (uint32_t m = 0; m < N; ++m)
        if (m != mode && ndims[m] > (1u << 24)) return 0;
    return 1;
}

constexpr uint32_t UNIT_CHUNKS = 64;  // chunks (of 256 nonzeros) per workgroup unit

// A sorted copy's rows cut into 2^sb-row blocks, chunks and units, the
// structure of a HiCOO mode copy (build_row_units in hicoo_dev.hip).
struct RowUnits {
    Buffer chunk_z, unit_chunk, unit_br, unit_excl;
    uint32_t nchunks = 0, nunits = 0;
};
int build_row_units(const uint32_t *rows, uint64_t n, uint32_t sb, RowUnits &u, hipStream_t s, const char *M);
int build_unit_order(const uint32_t *unit_chunk, uint32_t nunits, Buffer &order, hipStream_t s, const char *M);
int build_rank_unit_order(const uint32_t *unit_br, uint32_t nunits, Buffer &order, hipStream_t s, const char *M);

// Windowed entry order of a stream form: inside each block-row the nonzeros
// are ordered by windows of 2^WINDOW_BITS rows of the largest gathered mode
// (512 KB of an FP32 factor at rank 32), then by row, and units are dealt by
// their rank inside their block-row (build_rank_unit_order), so the units in
// flight on one XCD gather that factor from a few windows held in its L2.
// Used where block-rows hold two or more units' worth of nonzeros on average,
// rows keep at least 32 nonzeros per window on average (shorter runs would
// multiply the run-sum flushes into the LDS tile) and the gathered mode's
// coordinates are spread evenly over the windows (windows_even): with a
// skewed (Zipf) mode most (row, window) runs are far shorter than the
// average, and its frequent rows stay in L2 anyway.
constexpr uint32_t WINDOW_BITS = 12;
inline bool windowed_order(uint64_t nnz, uint32_t rows, uint32_t sec_rows, uint32_t row_bits) {
    const uint64_t blocks = ((uint64_t)rows + (1u << row_bits) - 1) >> row_bits;
    const uint64_t windows = ((uint64_t)sec_rows + (1u << WINDOW_BITS) - 1) >> WINDOW_BITS;
    return nnz >= blocks * 2 * UNIT_CHUNKS * 256 && windows >= 4 && nnz >= 32 * (uint64_t)rows * windows;
}

// Whether the per-window counts of a gathered mode's coordinates (counts[w]
// for the nwin windows, in device memory) are even: the largest at most four
// times the mean.  *even is written on success.
int windows_even(const uint32_t *counts, uint32_t nwin, uint64_t nnz, bool *even, hipStream_t s, const char *M);

// Interleaved slices: a stream-form chunk whose sorted entries hold at most
// two runs of equal rows hands them to the lane groups round-robin (slice g
// takes entries g, g + NG, g + 2 NG, ...) instead of in contiguous ranges, so
// that the lane groups of one position hold neighbouring entries of one row,
// which share gathered rows wherever the order by the largest gathered mode
// puts repeats together (a gather costs per distinct row).  In a chunk of
// many short rows every slice would then see every row and flush each one.
// Measured with a per-mode rule (rows of >= 256 nonzeros on average) before
// this per-chunk one: the 100M-nonzero Zipf tensor's mode 2 went from 1.217
// to 1.186 ms, the 200M-nonzero order-4 tensor's modes 1-3 by 2-4%, and the
// order-5 CP-ALS tensor's Zipf mode 1 (300 nonzeros per row on average, most
// rows short) from 1.55 to 1.92 ms.
constexpr uint32_t INTERLEAVE_MAX_RUNS = 2;

// Histogram update for a whole wave (every lane calls it): adds one to
// counts[x] for each lane with `valid`, as one atomic per distinct x in the
// wave, so that skewed (Zipf) coordinates do not queue millions of atomics
// on one counter.
__device__ __forceinline__ void wave_count(uint32_t *counts, uint32_t x, bool valid) {
    uint64_t left = __builtin_amdgcn_ballot_w64(valid);
    while (left) {
        const int leader = __builtin_ctzll(left);
        const uint32_t xl = (uint32_t)__builtin_amdgcn_readlane((int)x, leader);
        const uint64_t same = __builtin_amdgcn_ballot_w64(valid && x == xl);
        if ((int)(threadIdx.x & 63) == leader) atomicAdd(counts + xl, (uint32_t)__builtin_popcountll(same));
        valid = valid && x != xl;
        left &= ~same;
    }
}

// Shapes the stream kernel runs: orders up to 5 (beyond, the entry planes and
// the gathered rows no longer fit the register budget) and lane groups of at
// most 16 lanes (a slice of at most 64 positions, walked straight-line).
constexpr bool stream_supported(uint32_t N, int G) { return N <= 5 && G <= 16; }

#define PTI_DECLARE_LAUNCHERS(N)                                                                       \
    hipError_t launch_hicoo_##N(const HiCOOArgs &h, const FactorArgs &a, Shape sh, bool bf16, hipStream_t s); \
    hipError_t launch_stream_##N(const StreamArgs &h, const FactorArgs &a, Shape sh, bool bf16, hipStream_t s); \
    hipError_t launch_coo_##N(const COOArgs &c, const FactorArgs &a, Shape sh, hipStream_t s);        \
    hipError_t launch_completion_##N(const COOArgs &c, const FactorArgs &a, Shape sh, float *resid,    \
                                     double *partials, hipStream_t s);                                \
    hipError_t launch_completion_grad_##N(const StreamArgs &h, const FactorArgs &a, Shape sh, double *loss_part, \
                                          hipStream_t s);

PTI_DECLARE_LAUNCHERS(2)
PTI_DECLARE_LAUNCHERS(3)
PTI_DECLARE_LAUNCHERS(4)
PTI_DECLARE_LAUNCHERS(5)
PTI_DECLARE_LAUNCHERS(6)
PTI_DECLARE_LAUNCHERS(7)
PTI_DECLARE_LAUNCHERS(8)
#undef PTI_DECLARE_LAUNCHERS

#define PTI_DISPATCH_N(N, call)      \
    switch (N) {                     \
        case 2: return call(2);      \
        case 3: return call(3);      \
        case 4: return call(4);      \
        case 5: return call(5);      \
        case 6: return call(6);      \
        case 7: return call(7);      \
        case 8: return call(8);      \
        default: return hipErrorInvalidValue; \
    }

inline hipError_t launch_hicoo(uint32_t N, const HiCOOArgs &h, const FactorArgs &a, Shape sh, bool bf16,
                               hipStream_t s) {
#define PTI_CALL(n) launch_hicoo_##n(h, a, sh, bf16, s)
    PTI_DISPATCH_N(N, PTI_CALL)
#undef PTI_CALL
}
inline hipError_t launch_stream(uint32_t N, const StreamArgs &h, const FactorArgs &a, Shape sh, bool bf16,
                                hipStream_t s) {
#define PTI_CALL(n) launch_stream_##n(h, a, sh, bf16, s)
    PTI_DISPATCH_N(N, PTI_CALL)
#undef PTI_CALL
}
inline hipError_t launch_coo(uint32_t N, const COOArgs &c, const FactorArgs &a, Shape sh, hipStream_t s) {
#define PTI_CALL(n) launch_coo_##n(c, a, sh, s)
    PTI_DISPATCH_N(N, PTI_CALL)
#undef PTI_CALL
}

inline hipError_t launch_completion_grad(uint32_t N, const StreamArgs &h, const FactorArgs &a, Shape sh,
                                         double *loss_part, hipStream_t s) {
#define PTI_CALL(n) launch_completion_grad_##n(h, a, sh, loss_part, s)
    PTI_DISPATCH_N(N, PTI_CALL)
#undef PTI_CALL
}

inline hipError_t launch_completion(uint32_t N, const COOArgs &c, const FactorArgs &a, Shape sh, float *resid,
                                    double *partials, hipStream_t s) {
#define PTI_CALL(n) launch_completion_##n(c, a, sh, resid, partials, s)
    PTI_DISPATCH_N(N, PTI_CALL)
#undef PTI_CALL
}

}  // namespace dev
}  // namespace pti
