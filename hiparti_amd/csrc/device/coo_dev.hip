// COO tensors resident in HBM and the COO MTTKRP entry points.
//
// The uploaded tensor keeps its storage order (SoA: one u32 array per mode
// plus the values).  For each target mode a copy sorted by that mode's index
// is built on the device on first use (stable radix sort of (index, position)
// pairs with rocPRIM, then a gather), so that equal output rows are adjacent
// and the kernel's segmented reduction collapses them into one atomic per
// (chunk, row) run.  Where the stream kernel's shapes allow, the sorted copy
// is instead cut into blocks of 2^7 rows, chunks and units like a HiCOO mode
// copy and gets a stream form on its first MTTKRP (decoded entries in the
// layout of StreamShape, 16 bytes per nonzero and plane), which
// hicoo_stream_kernel runs with its LDS tile per block.
#include <atomic>
#include <mutex>
#include <unordered_map>

#include <rocprim/device/device_radix_sort.hpp>

#include "../../../include/hiparti_dev.h"
#include "dropin.hpp"
#include "launch.hpp"

// Scratch of ptiDevCOOCompletionForwardNoAlloc: the CP-completion forward's
// per-workgroup FP64 partial sums (one per 1024 nonzeros), so at most
// NOALLOC_MAX_PARTIALS * 1024 = 2^27 nonzeros.  Every GPU the library runs on
// holds its own copy (1 MiB) until the process ends.  In host code `scratch`
// names a host-side stand-in: its device address comes from
// hipGetSymbolAddress (noalloc_scratch).
constexpr uint32_t NOALLOC_MAX_PARTIALS = 1u << 17;
static __device__ __align__(256) double scratch[NOALLOC_MAX_PARTIALS];

struct ptiDevCOO {
    uint32_t nmodes = 0;
    uint64_t nnz = 0;
    ptiIndex ndims[8] = {0};
    pti::dev::Buffer inds[8], vals;           // storage order
    pti::dev::Buffer sinds[8][8], svals[8];   // [target mode][mode] sorted copies
    pti::dev::Buffer perm[8];                 // [target mode] sorted position -> storage position
    pti::dev::Buffer partials;                // completion forward: per-workgroup sums
    pti::dev::RowUnits units[8];              // [target mode] row blocks / chunks / units of the sorted copy
    pti::dev::Buffer stream[8];               // [target mode] stream form of the sorted copy
    pti::dev::Buffer unit_order[8];           // [target mode] its units by decreasing chunk count
    uint32_t stream_ng[8] = {0};              // its lane-group count (0: not built)
    bool ready[8] = {false};
    std::mutex mu;
};

namespace pti {
namespace dev {
namespace {

std::mutex g_coo_mu;
std::unordered_map<const void *, ptiDevCOO *> g_coo_reg;

__global__ void __launch_bounds__(256) iota_kernel(uint32_t *p, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = (uint32_t)i;
}

template <typename T>
__global__ void __launch_bounds__(256) gather_kernel(T *__restrict__ dst, const T *__restrict__ src,
                                                     const uint32_t *__restrict__ perm, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (uint64_t)gridDim.x * blockDim.x) dst[i] = src[perm[i]];
}

// 0.5 * sum of n partial sums, in a fixed order by one workgroup.
__global__ void __launch_bounds__(256) half_sum_kernel(const double *partials, uint32_t n, double *out) {
    __shared__ double s[256];
    double t = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += 256) t += partials[i];
    s[threadIdx.x] = t;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = 0.5 * s[0];
}

struct Ptr8 {
    const uint32_t *p[8];
};

// Stream form of the copy sorted by mode `mode`, cut by build_row_units into
// blocks of 2^ROW_BITS rows, chunks and units (one wave per chunk): entries in
// the StreamShape layout for ng lane groups, the coordinates of the other
// modes in ascending mode order, the value bits and the row inside its block.
// perm (optional): the entries of each block in another order (positions
// inside the block's range of the sorted copy).  A chunk of at most
// INTERLEAVE_MAX_RUNS row runs is interleaved: slice g takes its entries g,
// g + ng, g + 2 ng, ...
constexpr uint32_t ROW_BITS = 7;
__global__ void __launch_bounds__(256) coo_stream_fill_kernel(Ptr8 sinds, const float *__restrict__ svals,
                                                              const uint32_t *__restrict__ chunk_z, uint32_t nchunks,
                                                              uint32_t N, uint32_t mode, uint32_t ng,
                                                              const uint32_t *__restrict__ perm,
                                                              uint4 *__restrict__ ent) {
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t c = blockIdx.x * 4 + wid;
    if (c >= nchunks) return;   // whole wave
    const uint32_t NP = (N + 4) / 4;
    const uint32_t z0 = chunk_z[c], cnt = chunk_z[c + 1] - z0;
    const uint32_t per = (cnt + ng - 1) / ng;
    const uint32_t G = 64 / ng;
    auto row_of = [&](uint64_t zs) { return sinds.p[mode][perm ? perm[zs] : zs]; };
    uint32_t breaks = 0;   // row changes between neighbouring entries (the same in every lane)
    for (uint32_t kb = 1; kb < cnt; kb += 64) {
        const uint32_t k = kb + lane;
        const bool brk = k < cnt && row_of((uint64_t)z0 + k) != row_of((uint64_t)z0 + k - 1);
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
            const uint64_t zs = (uint64_t)z0 + k;
            const uint64_t z = perm ? perm[zs] : zs;
            for (uint32_t i = 0; i + 1 < N; ++i) w[i] = sinds.p[i < mode ? i : i + 1][z];
            w[N - 1] = __float_as_uint(svals[z]);
            w[N] = sinds.p[mode][z] & ((1u << ROW_BITS) - 1u);
        }
        for (uint32_t q = 0; q < NP; ++q)
            ent[((uint64_t)c * NP + q) * CHUNK + e] = make_uint4(w[4 * q], w[4 * q + 1], w[4 * q + 2], w[4 * q + 3]);
    }
}

// Sort keys of a stream form's order over the copy sorted by `mode`, as for
// the HiCOO stream forms: (row, coordinate of the largest gathered mode
// `sec`), whose factor rows are the likeliest cache misses; or, windowed
// (wbits > 0), (row block, window of `sec`, row inside the block, the low
// WINDOW_BITS of `sec`).
__global__ void __launch_bounds__(256) stream_keys_kernel(const uint32_t *__restrict__ rows,
                                                          const uint32_t *__restrict__ sec, uint64_t n, uint32_t wbits,
                                                          uint64_t *__restrict__ key, uint32_t *__restrict__ pos) {
    for (uint64_t z = (uint64_t)blockIdx.x * 256 + threadIdx.x; z < n; z += (uint64_t)gridDim.x * 256) {
        const uint32_t r = rows[z], x = sec[z];
        if (wbits)
            key[z] = ((uint64_t)(r >> ROW_BITS) << (wbits + ROW_BITS + WINDOW_BITS)) |
                     ((uint64_t)(x >> WINDOW_BITS) << (ROW_BITS + WINDOW_BITS)) |
                     ((uint64_t)(r & ((1u << ROW_BITS) - 1u)) << WINDOW_BITS) | (x & ((1u << WINDOW_BITS) - 1u));
        else
            key[z] = ((uint64_t)r << 32) | x;
        pos[z] = (uint32_t)z;
    }
}

// Windowed order: nonzeros per window of a mode's coordinates x[0..n).
__global__ void __launch_bounds__(256) window_count_kernel(const uint32_t *__restrict__ x, uint64_t n,
                                                           uint32_t *__restrict__ counts) {
    // Every lane of a wave runs the same iterations (wave_count is wave-wide).
    for (uint64_t z0 = (uint64_t)blockIdx.x * 256 + (threadIdx.x & ~63u); z0 < n; z0 += (uint64_t)gridDim.x * 256) {
        const uint64_t z = z0 + (threadIdx.x & 63u);
        wave_count(counts, z < n ? x[z] >> WINDOW_BITS : 0u, z < n);
    }
}

inline dim3 grid_1d(uint64_t n) {
    uint64_t b = (n + 255) / 256;
    if (b > 65536) b = 65536;
    if (b == 0) b = 1;
    return dim3((uint32_t)b);
}

int upload_coo(ptiDevCOO **out, const ptiSparseTensor *X, hipStream_t s) {
    static const char *M = "DevCOO Upload";
    PTI_CHECK(require_device(M), M);
    const uint32_t N = X->nmodes;
    if (N < 2 || N > 8) PTI_FAIL(PTIERR_VALUE_ERROR, M, "nmodes must be in [2, 8]");
    if (X->nnz >= (1ull << 32)) PTI_FAIL(PTIERR_VALUE_ERROR, M, "nnz must be < 2^32");
    // Coordinates index factor and output rows on the device: reject any that
    // lie outside their mode.
    for (uint32_t m = 0; m < N; ++m) {
        std::atomic<bool> bad(false);
        const uint32_t *d = X->inds[m].data;
        const uint32_t lim = X->ndims[m];
        pti::parallel_for(X->nnz, 0, [&](uint64_t b, uint64_t e, int) {
            for (uint64_t z = b; z < e; ++z)
                if (d[z] >= lim) {
                    bad = true;
                    return;
                }
        });
        if (bad.load()) PTI_FAIL(PTIERR_VALUE_ERROR, M, "coordinate outside its mode");
    }
    ptiDevCOO *t = new ptiDevCOO();
    t->nmodes = N;
    t->nnz = X->nnz;
    for (uint32_t m = 0; m < N; ++m) t->ndims[m] = X->ndims[m];
    int rc = 0;
    auto put = [&](Buffer &buf, const void *src, size_t bytes) -> int {
        PTI_HIP(buf.alloc(bytes), M);
        if (bytes) PTI_HIP(hipMemcpyAsync(buf.ptr, src, bytes, hipMemcpyHostToDevice, s), M);
        return 0;
    };
    for (uint32_t m = 0; m < N && !rc; ++m) rc = put(t->inds[m], X->inds[m].data, X->nnz * 4);
    if (!rc) rc = put(t->vals, X->values.data, X->nnz * 4);
    hipError_t e = hipStreamSynchronize(s);
    if (!rc && e != hipSuccess) {
        pti_ComplainError(M, hip_code(e), __FILE__, __LINE__, hipGetErrorString(e));
        rc = hip_code(e);
    }
    if (rc) {
        delete t;
        return rc;
    }
    *out = t;
    return 0;
}

int prepare_mode(ptiDevCOO *t, ptiIndex mode, hipStream_t s) {
    static const char *M = "DevCOO Prepare";
    if (mode >= t->nmodes) PTI_FAIL(PTIERR_VALUE_ERROR, M, "mode out of range");
    std::lock_guard<std::mutex> lk(t->mu);
    if (t->ready[mode]) return 0;
    const uint64_t n = t->nnz;
    const uint32_t N = t->nmodes;
    for (uint32_t m = 0; m < N; ++m) PTI_HIP(t->sinds[mode][m].alloc(n * 4), M);
    PTI_HIP(t->svals[mode].alloc(n * 4), M);
    if (n) {
        Buffer keys_out, ids_in, tmp;
        Buffer &ids_out = t->perm[mode];
        PTI_HIP(keys_out.alloc(n * 4), M);
        PTI_HIP(ids_in.alloc(n * 4), M);
        PTI_HIP(ids_out.alloc(n * 4), M);
        hipLaunchKernelGGL(iota_kernel, grid_1d(n), dim3(256), 0, s, ids_in.as<uint32_t>(), n);
        PTI_HIP(hipGetLastError(), M);
        unsigned end_bit = 1;
        while (end_bit < 32 && (1ull << end_bit) < (uint64_t)t->ndims[mode]) ++end_bit;
        size_t tmp_bytes = 0;
        PTI_HIP(rocprim::radix_sort_pairs(nullptr, tmp_bytes, t->inds[mode].as<uint32_t>(), keys_out.as<uint32_t>(),
                                          ids_in.as<uint32_t>(), ids_out.as<uint32_t>(), (size_t)n, 0u, end_bit, s),
                M);
        PTI_HIP(tmp.alloc(tmp_bytes), M);
        PTI_HIP(rocprim::radix_sort_pairs(tmp.ptr, tmp_bytes, t->inds[mode].as<uint32_t>(), keys_out.as<uint32_t>(),
                                          ids_in.as<uint32_t>(), ids_out.as<uint32_t>(), (size_t)n, 0u, end_bit, s),
                M);
        for (uint32_t m = 0; m < N; ++m) {
            hipLaunchKernelGGL(gather_kernel<uint32_t>, grid_1d(n), dim3(256), 0, s, t->sinds[mode][m].as<uint32_t>(),
                               t->inds[m].as<uint32_t>(), ids_out.as<uint32_t>(), n);
            PTI_HIP(hipGetLastError(), M);
        }
        hipLaunchKernelGGL(gather_kernel<float>, grid_1d(n), dim3(256), 0, s, t->svals[mode].as<float>(),
                           t->vals.as<float>(), ids_out.as<uint32_t>(), n);
        PTI_HIP(hipGetLastError(), M);
        // The temporaries are freed at scope exit: finish the work first.
        PTI_HIP(hipStreamSynchronize(s), M);
    }
    t->ready[mode] = true;
    return 0;
}

// Builds (or rebuilds, for another lane-group count) the stream form of the
// copy sorted by `mode`; the copy must exist.
int ensure_coo_stream(ptiDevCOO *t, ptiIndex mode, uint32_t ng, hipStream_t s, const char *M) {
    std::lock_guard<std::mutex> lk(t->mu);
    if (t->stream_ng[mode] == ng) return 0;
    // Work queued on the old form must finish before it is released.
    if (t->stream[mode].ptr) PTI_HIP(hipDeviceSynchronize(), M);
    const uint32_t N = t->nmodes;
    const uint32_t NP = (N + 4) / 4;
    RowUnits &u = t->units[mode];
    if (!u.chunk_z.ptr) PTI_CHECK(build_row_units(t->sinds[mode][mode].as<uint32_t>(), t->nnz, ROW_BITS, u, s, M), M);
    PTI_HIP(t->stream[mode].alloc((size_t)u.nchunks * NP * CHUNK * sizeof(uint4)), M);
    Ptr8 p{};
    for (uint32_t m = 0; m < N; ++m) p.p[m] = t->sinds[mode][m].as<uint32_t>();
    // Long row blocks take the windowed order (windowed_order).
    uint32_t sec = mode == 0 ? 1 : 0;
    for (uint32_t m = 0; m < N; ++m)
        if (m != mode && t->ndims[m] > t->ndims[sec]) sec = m;
    const uint64_t nblocks = ((uint64_t)t->ndims[mode] + (1u << ROW_BITS) - 1) >> ROW_BITS;
    bool windowed = windowed_order(t->nnz, t->ndims[mode], t->ndims[sec], ROW_BITS);
    if (windowed) {
        const uint32_t nwin = (uint32_t)(((uint64_t)t->ndims[sec] + (1u << WINDOW_BITS) - 1) >> WINDOW_BITS);
        Buffer wc;
        PTI_HIP(wc.alloc((size_t)nwin * 4), M);
        PTI_HIP(zero_floats(wc.as<float>(), nwin, s), M);   // all-zero bits: count 0
        hipLaunchKernelGGL(window_count_kernel, grid_1d(t->nnz), dim3(256), 0, s, t->inds[sec].as<uint32_t>(), t->nnz,
                           wc.as<uint32_t>());
        PTI_HIP(hipGetLastError(), M);
        PTI_CHECK(windows_even(wc.as<uint32_t>(), nwin, t->nnz, &windowed, s, M), M);
    }
    Buffer key, pos, skey, spos, tmp;
    const uint64_t n = t->nnz;
    if (n) {
        unsigned bbits = 0, wbits = 0, rbits = 1;
        while (bbits < 32 && ((nblocks - 1) >> bbits)) ++bbits;
        while (rbits < 32 && (((uint64_t)t->ndims[mode] - 1) >> rbits)) ++rbits;
        if (windowed)
            while (wbits < 32 && (((uint64_t)t->ndims[sec] - 1) >> WINDOW_BITS) >> wbits) ++wbits;
        PTI_HIP(key.alloc(n * 8), M);
        PTI_HIP(skey.alloc(n * 8), M);
        PTI_HIP(pos.alloc(n * 4), M);
        PTI_HIP(spos.alloc(n * 4), M);
        hipLaunchKernelGGL(stream_keys_kernel, grid_1d(n), dim3(256), 0, s, t->sinds[mode][mode].as<uint32_t>(),
                           t->sinds[mode][sec].as<uint32_t>(), n, windowed ? wbits : 0u, key.as<uint64_t>(),
                           pos.as<uint32_t>());
        PTI_HIP(hipGetLastError(), M);
        const unsigned end_bit = windowed ? bbits + wbits + ROW_BITS + WINDOW_BITS : rbits + 32;
        size_t tb = 0;
        PTI_HIP(rocprim::radix_sort_pairs(nullptr, tb, key.as<uint64_t>(), skey.as<uint64_t>(), pos.as<uint32_t>(),
                                          spos.as<uint32_t>(), (size_t)n, 0u, end_bit, s),
                M);
        PTI_HIP(tmp.alloc(tb), M);
        PTI_HIP(rocprim::radix_sort_pairs(tmp.ptr, tb, key.as<uint64_t>(), skey.as<uint64_t>(), pos.as<uint32_t>(),
                                          spos.as<uint32_t>(), (size_t)n, 0u, end_bit, s),
                M);
    }
    hipLaunchKernelGGL(coo_stream_fill_kernel, dim3((u.nchunks + 3) / 4), dim3(256), 0, s, p,
                       t->svals[mode].as<float>(), u.chunk_z.as<uint32_t>(), u.nchunks, N, (uint32_t)mode, ng,
                       n ? spos.as<uint32_t>() : nullptr, t->stream[mode].as<uint4>());
    PTI_HIP(hipGetLastError(), M);
    if (windowed)
        PTI_CHECK(build_rank_unit_order(u.unit_br.as<uint32_t>(), u.nunits, t->unit_order[mode], s, M), M);
    else
        PTI_CHECK(build_unit_order(u.unit_chunk.as<uint32_t>(), u.nunits, t->unit_order[mode], s, M), M);
    // The temporaries are released on return: the queued work must finish.
    PTI_HIP(hipStreamSynchronize(s), M);
    t->stream_ng[mode] = ng;
    return 0;
}

}  // namespace

int run_coo(ptiDevCOO *t, float *const dmats[], ptiIndex R, ptiIndex stride, const ptiIndex mats_order[],
            ptiIndex mode, hipStream_t s, const char *M, const float *vals_override) {
    const uint32_t N = t->nmodes;
    if (mode >= N) PTI_FAIL(PTIERR_VALUE_ERROR, M, "mode out of range");
    if (R == 0 || stride < R) PTI_FAIL(PTIERR_SHAPE_MISMATCH, M, "bad rank or stride");
    if (!order_is_permutation(N, mats_order, mode))
        PTI_FAIL(PTIERR_VALUE_ERROR, M, "mats_order[1..N-1] must be a permutation of the other modes");
    PTI_CHECK(prepare_mode(t, mode, s), M);
    float *out = dmats[N];
    PTI_HIP(zero_floats(out, (uint64_t)t->ndims[mode] * stride, s), M);
    if (t->nnz == 0) return 0;
    COOArgs c;
    for (uint32_t m = 0; m < N; ++m) c.inds[m] = t->sinds[mode][m].as<uint32_t>();
    c.vals = t->svals[mode].as<float>();
    c.vals_override = vals_override;
    c.perm = t->perm[mode].as<uint32_t>();
    c.nnz = t->nnz;
    c.nchunks = (uint32_t)((t->nnz + CHUNK - 1) / CHUNK);
    // The stream kernel takes ranks up to 64 (one launch) when its shapes
    // allow and every gathered factor matrix spans < 4 GiB (32-bit offsets).
    if (!vals_override && R <= 64) {
        Shape sh;
        bool ok = choose_shape(R, stride, stride, sh) && stream_supported(N, sh.G);
        for (uint32_t m = 0; m < N && ok; ++m)
            if (m != mode && (uint64_t)t->ndims[m] * stride * 4 >= (1ull << 32)) ok = false;
        if (ok) {
            PTI_CHECK(ensure_coo_stream(t, mode, 64u / sh.G, s, M), M);
            FactorArgs a;
            for (uint32_t m = 0; m < N; ++m) a.fac[m] = dmats[m];
            a.out = out;
            a.R = R;
            a.stride = stride;
            a.out_stride = stride;
            a.mode = mode;
            for (uint32_t i = 0; i < N; ++i) a.order[i] = mats_order[i];
            const RowUnits &u = t->units[mode];
            StreamArgs h{};
            h.ent = t->stream[mode].as<uint4>();
            h.unit_chunk = u.unit_chunk.as<uint32_t>();
            h.unit_br = u.unit_br.as<uint32_t>();
            h.unit_excl = u.unit_excl.as<uint32_t>();
            h.nunits = u.nunits;
            h.sb_bits = ROW_BITS;
            h.nrows = t->ndims[mode];
            h.coord24 = coords_fit_u24(t->ndims, N, mode, stride * 4);
            h.unit_order = t->unit_order[mode].as<uint32_t>();
            PTI_HIP(launch_stream(N, h, a, sh, false, s), M);
            return 0;
        }
    }
    // Column tiles of 64 lanes x V columns, V being the vector width the full
    // rank and stride allow: every tile then has a lane shape (G <= 64), and
    // tiles start at multiples of 64 columns, which keeps any V's alignment.
    const ptiIndex tile = 64u * (ptiIndex)vector_width(R, stride, stride);
    for (ptiIndex c0 = 0; c0 < R; c0 += tile) {
        const ptiIndex Rt = std::min<ptiIndex>(tile, R - c0);
        Shape sh;
        if (!choose_shape(Rt, stride, stride, sh)) PTI_FAIL(PTIERR_VALUE_ERROR, M, "unsupported rank tile");
        FactorArgs a;
        for (uint32_t m = 0; m < N; ++m) a.fac[m] = dmats[m] + c0;
        a.out = out + c0;
        a.R = Rt;
        a.stride = stride;
        a.out_stride = stride;
        a.mode = mode;
        for (uint32_t i = 0; i < N; ++i) a.order[i] = mats_order[i];
        PTI_HIP(launch_coo(N, c, a, sh, s), M);
    }
    return 0;
}

// Whether the CP-completion gradient of `mode` runs completion_grad_kernel:
// ranks <= 64 with the stream kernel's shapes (orders <= 5, lane groups of
// <= 16 lanes) and every gathered factor matrix < 4 GiB.
bool grad_fused(const ptiDevCOO *t, ptiIndex R, ptiIndex stride, ptiIndex mode, Shape &sh) {
    const uint32_t N = t->nmodes;
    bool ok = R >= 1 && R <= 64 && stride >= R && choose_shape(R, stride, stride, sh) && stream_supported(N, sh.G);
    for (uint32_t m = 0; m < N && ok; ++m)
        if (m != mode && (uint64_t)t->ndims[m] * stride * 4 >= (1ull << 32)) ok = false;
    return ok;
}

// CP-completion gradient of `mode`: out = MTTKRP of the residuals
// e_z = x_z - sum_r prod_m A_m[i_m][r] (dmats[0..N-1] = A_m, dmats[N] = out).
// On completion_grad_kernel (grad_fused) the residuals are formed inside the
// walk from A_mode's rows, and loss_part, when given, receives per unit the
// sum of e_z^2; otherwise this is the MTTKRP of `resid` (the forward's
// residuals in storage order) on the segmented-scan kernel.
int completion_grad(ptiDevCOO *t, const float *resid, float *const dmats[], ptiIndex R, ptiIndex stride,
                    ptiIndex mode, hipStream_t s, const char *M, double *loss_part = nullptr) {
    const uint32_t N = t->nmodes;
    if (mode >= N) PTI_FAIL(PTIERR_VALUE_ERROR, M, "mode out of range");
    if (R == 0 || stride < R) PTI_FAIL(PTIERR_SHAPE_MISMATCH, M, "bad rank or stride");
    Shape sh;
    if (grad_fused(t, R, stride, mode, sh)) {
        PTI_CHECK(prepare_mode(t, mode, s), M);
        float *out = dmats[N];
        PTI_HIP(zero_floats(out, (uint64_t)t->ndims[mode] * stride, s), M);
        if (t->nnz == 0) return 0;
        PTI_CHECK(ensure_coo_stream(t, mode, 64u / sh.G, s, M), M);
        FactorArgs a;
        for (uint32_t m = 0; m < N; ++m) a.fac[m] = dmats[m];
        a.out = out;
        a.R = R;
        a.stride = stride;
        a.out_stride = stride;
        a.mode = mode;
        a.order[0] = mode;
        for (uint32_t i = 1; i < N; ++i) a.order[i] = i - 1 < mode ? i - 1 : i;
        const RowUnits &u = t->units[mode];
        StreamArgs h{};
        h.ent = t->stream[mode].as<uint4>();
        h.unit_chunk = u.unit_chunk.as<uint32_t>();
        h.unit_br = u.unit_br.as<uint32_t>();
        h.unit_excl = u.unit_excl.as<uint32_t>();
        h.nunits = u.nunits;
        h.sb_bits = ROW_BITS;
        h.nrows = t->ndims[mode];
        h.coord24 = coords_fit_u24(t->ndims, N, mode, stride * 4);
        h.unit_order = t->unit_order[mode].as<uint32_t>();
        PTI_HIP(launch_completion_grad(N, h, a, sh, loss_part, s), M);
        return 0;
    }
    if (loss_part) PTI_FAIL(PTIERR_VALUE_ERROR, M, "no fused gradient for this shape");
    if (!resid && t->nnz) PTI_FAIL(PTIERR_VALUE_ERROR, M, "residuals must not be NULL for this shape");
    ptiIndex order[8];
    order[0] = mode;
    for (uint32_t i = 1; i < N; ++i) order[i] = i - 1 < mode ? i - 1 : i;
    return run_coo(t, dmats, R, stride, order, mode, s, M, resid);
}

// Loss and every mode's gradient MTTKRP in N passes of
// completion_grad_kernel; mode 0's pass also sums e_z^2 per unit, reduced in
// a fixed order.  Every mode must take the fused kernel (grad_fused).
int completion_loss_grad(ptiDevCOO *t, float *const dfacs[], float *const dgrads[], ptiIndex R, ptiIndex stride,
                         double *dloss, hipStream_t s, const char *M) {
    const uint32_t N = t->nmodes;
    Shape sh;
    for (uint32_t n = 0; n < N; ++n)
        if (!grad_fused(t, R, stride, n, sh)) PTI_FAIL(PTIERR_VALUE_ERROR, M, "no fused gradient for this shape");
    if (t->nnz == 0) {
        for (uint32_t n = 0; n < N; ++n) PTI_HIP(zero_floats(dgrads[n], (uint64_t)t->ndims[n] * stride, s), M);
        PTI_HIP(zero_floats(reinterpret_cast<float *>(dloss), 2, s), M);
        return 0;
    }
    for (uint32_t n = 0; n < N; ++n) {
        float *dm[9];
        for (uint32_t m = 0; m < N; ++m) dm[m] = dfacs[m];
        dm[N] = dgrads[n];
        double *part = nullptr;
        if (n == 0) {
            PTI_CHECK(prepare_mode(t, 0, s), M);
            PTI_CHECK(ensure_coo_stream(t, 0, 64u / sh.G, s, M), M);
            const uint32_t nu = t->units[0].nunits;
            if (t->partials.bytes < (size_t)std::max<uint32_t>(nu, 1) * sizeof(double))
                PTI_HIP(t->partials.alloc((size_t)std::max<uint32_t>(nu, 1) * sizeof(double)), M);
            part = t->partials.as<double>();
        }
        PTI_CHECK(completion_grad(t, nullptr, dm, R, stride, n, s, M, part), M);
        if (n == 0) {
            hipLaunchKernelGGL(half_sum_kernel, dim3(1), dim3(256), 0, s, part, t->units[0].nunits, dloss);
            PTI_HIP(hipGetLastError(), M);
        }
    }
    return 0;
}

// Workgroups of completion_forward_kernel for nnz entries (one FP64 partial
// sum each).
uint32_t completion_forward_workgroups(uint64_t nnz) {
    const uint64_t nchunks = (nnz + CHUNK - 1) / CHUNK;
    return (uint32_t)((nchunks + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK);
}

// CP-completion forward with the per-workgroup partial sums in `partials`
// (completion_forward_workgroups(nnz) doubles, supplied by the caller):
// residuals into dresid, 0.5 * sum of their squares into *dloss.  Ranks
// wider than 64 lanes x V run each lane group over column blocks.
int completion_forward(const ptiDevCOO *t, float *const dfacs[], ptiIndex R, ptiIndex stride, float *dresid,
                       double *dloss, double *partials, hipStream_t s, const char *M) {
    const uint32_t N = t->nmodes;
    if (R == 0 || R > 256 || stride < R) PTI_FAIL(PTIERR_VALUE_ERROR, M, "rank must be in [1, 256] and <= stride");
    Shape sh;
    sh.V = vector_width(R, stride, stride);
    sh.G = 4;
    while ((uint32_t)(sh.G * sh.V) < R && sh.G < 64) sh.G <<= 1;
    COOArgs c = {};
    for (uint32_t m = 0; m < N; ++m) c.inds[m] = t->inds[m].as<uint32_t>();
    c.vals = t->vals.as<float>();
    c.nnz = t->nnz;
    c.nchunks = (uint32_t)((t->nnz + CHUNK - 1) / CHUNK);
    const uint32_t nblk = completion_forward_workgroups(t->nnz);
    FactorArgs a = {};
    for (uint32_t m = 0; m < N; ++m) a.fac[m] = dfacs[m];
    a.R = R;
    a.stride = stride;
    a.out_stride = stride;
    PTI_HIP(launch_completion(N, c, a, sh, dresid, partials, s), M);
    hipLaunchKernelGGL(half_sum_kernel, dim3(1), dim3(256), 0, s, partials, nblk, dloss);
    PTI_HIP(hipGetLastError(), M);
    return 0;
}

// Device address of `scratch` on the current device (*out), looked up once
// per device with that device current: each GPU holds its own copy at its
// own address.
int noalloc_scratch(double **out, const char *M) {
    constexpr int MAX_DEVICES = 64;
    static std::mutex mu;
    static double *addr[MAX_DEVICES] = {};
    int dev = 0;
    PTI_HIP(hipGetDevice(&dev), M);
    if (dev < 0 || dev >= MAX_DEVICES) PTI_FAIL(PTIERR_VALUE_ERROR, M, "device ordinal beyond the scratch table");
    std::lock_guard<std::mutex> lk(mu);
    if (!addr[dev]) {
        void *p = nullptr;
        PTI_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), M);
        if (!p) PTI_FAIL(PTIERR_CUDA_ERROR, M, "no device address for the scratch array");
        addr[dev] = static_cast<double *>(p);
    }
    *out = addr[dev];
    return 0;
}

void coo_view(const ptiDevCOO *t, uint32_t *nmodes, const ptiIndex **ndims, uint64_t *nnz, const float **vals) {
    *nmodes = t->nmodes;
    *ndims = t->ndims;
    *nnz = t->nnz;
    *vals = t->vals.as<float>();
}

}  // namespace dev
}  // namespace pti

using namespace pti::dev;

extern "C" {

int ptiDevCOOUpload(ptiDevCOO **out, ptiSparseTensor const *X, void *stream) {
    return upload_coo(out, X, static_cast<hipStream_t>(stream));
}

int ptiDevCOOPrepareMode(ptiDevCOO *t, ptiIndex mode, void *stream) {
    return prepare_mode(t, mode, static_cast<hipStream_t>(stream));
}

void ptiDevCOOFree(ptiDevCOO *t) { delete t; }

int ptiDevMTTKRPCOO(ptiDevCOO *t, float *const dmats[], ptiIndex R, ptiIndex stride, ptiIndex const mats_order[],
                    ptiIndex mode, void *stream) {
    return run_coo(t, dmats, R, stride, mats_order, mode, static_cast<hipStream_t>(stream), "DevCOO MTTKRP",
                   nullptr);
}

int ptiDevMTTKRPCOOValues(ptiDevCOO *t, float const *dvals, float *const dmats[], ptiIndex R, ptiIndex stride,
                          ptiIndex const mats_order[], ptiIndex mode, void *stream) {
    if (!dvals && t->nnz) PTI_FAIL(PTIERR_VALUE_ERROR, "DevCOO MTTKRP values", "values must not be NULL");
    return run_coo(t, dmats, R, stride, mats_order, mode, static_cast<hipStream_t>(stream), "DevCOO MTTKRP values",
                   dvals);
}

int ptiDevCOOCompletionFused(ptiDevCOO const *t, ptiIndex R, ptiIndex stride) {
    Shape sh;
    for (uint32_t n = 0; n < t->nmodes; ++n)
        if (!grad_fused(t, R, stride, n, sh)) return 0;
    return 1;
}

int ptiDevCOOCompletionLossGrad(ptiDevCOO *t, float *const dfacs[], float *const dgrads[], ptiIndex R,
                                ptiIndex stride, double *dloss, void *stream) {
    return completion_loss_grad(t, dfacs, dgrads, R, stride, dloss, static_cast<hipStream_t>(stream),
                                "DevCOO Completion loss and gradient");
}

int ptiDevCOOCompletionGrad(ptiDevCOO *t, float const *dresid, float *const dmats[], ptiIndex R, ptiIndex stride,
                            ptiIndex mode, void *stream) {
    return completion_grad(t, dresid, dmats, R, stride, mode, static_cast<hipStream_t>(stream),
                           "DevCOO Completion gradient");
}

int ptiDevCOOCompletionForward(ptiDevCOO *t, float *const dfacs[], ptiIndex R, ptiIndex stride, float *dresid,
                               double *dloss, void *stream) {
    static const char *M = "DevCOO Completion";
    if (R == 0 || R > 256 || stride < R) PTI_FAIL(PTIERR_VALUE_ERROR, M, "rank must be in [1, 256] and <= stride");
    const uint32_t nblk = completion_forward_workgroups(t->nnz);
    if (t->partials.bytes < (size_t)std::max<uint32_t>(nblk, 1) * sizeof(double))
        PTI_HIP(t->partials.alloc((size_t)std::max<uint32_t>(nblk, 1) * sizeof(double)), M);
    return completion_forward(t, dfacs, R, stride, dresid, dloss, t->partials.as<double>(),
                              static_cast<hipStream_t>(stream), M);
}

int ptiDevCOOCompletionForwardNoAlloc(ptiDevCOO *t, float *const dfacs[], ptiIndex R, ptiIndex stride,
                                      float *dresid, double *dloss, void *stream) {
    static const char *M = "DevCOO Completion (no allocation)";
    if (R == 0 || R > 256 || stride < R) PTI_FAIL(PTIERR_VALUE_ERROR, M, "rank must be in [1, 256] and <= stride");
    if (completion_forward_workgroups(t->nnz) > NOALLOC_MAX_PARTIALS)
        PTI_FAIL(PTIERR_VALUE_ERROR, M, "more than 2^27 nonzeros: beyond the fixed scratch");
    double *part = nullptr;
    PTI_CHECK(noalloc_scratch(&part, M), M);
    // Every workgroup of the forward kernel writes its own partial before the
    // sum reads it, so the scratch needs no clearing.
    return completion_forward(t, dfacs, R, stride, dresid, dloss, part, static_cast<hipStream_t>(stream), M);
}

int ptiRegisterSparseTensor(ptiSparseTensor const *X) {
    ptiDevCOO *t = nullptr;
    PTI_CHECK(upload_coo(&t, X, nullptr), "Register SpTns");
    std::lock_guard<std::mutex> lk(g_coo_mu);
    auto it = g_coo_reg.find(X);
    if (it != g_coo_reg.end()) delete it->second;
    g_coo_reg[X] = t;
    return 0;
}

int ptiUnregisterSparseTensor(ptiSparseTensor const *X) {
    std::lock_guard<std::mutex> lk(g_coo_mu);
    auto it = g_coo_reg.find(X);
    if (it == g_coo_reg.end()) PTI_FAIL(PTIERR_VALUE_ERROR, "Unregister SpTns", "tensor is not registered");
    delete it->second;
    g_coo_reg.erase(it);
    return 0;
}

int ptiMTTKRP(ptiSparseTensor const *const X, ptiMatrix *mats[], ptiIndex const mats_order[], ptiIndex const mode) {
    static const char *M = "GPU  SpTns MTTKRP";
    const ptiIndex N = X->nmodes;
    PTI_CHECK(check_mttkrp_args(N, X->ndims, mats, mats_order, mode, M), M);
    PTI_CHECK(require_device(M), M);
    hipStream_t s = nullptr;
    ptiDevCOO *t = nullptr;
    ptiDevCOO *owned = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_coo_mu);
        auto it = g_coo_reg.find(X);
        if (it != g_coo_reg.end()) t = it->second;
    }
    if (!t) {
        PTI_CHECK(upload_coo(&owned, X, s), M);
        t = owned;
    }
    StagedFactors sf;
    int rc = stage_factors(N, mats, mats_order, mode, sf, s, M);
    if (!rc) rc = run_coo(t, sf.ptrs, mats[N]->ncols, mats[0]->stride, mats_order, mode, s, M, nullptr);
    if (!rc) rc = fetch_output(N, mats, mode, sf, s, M);
    delete owned;
    return rc;
}

int ptiCudaMTTKRP(ptiSparseTensor const *const X, ptiMatrix *mats[], ptiIndex *const mats_order, ptiIndex const mode,
                  ptiIndex const impl_num) {
    (void)impl_num;
    return ptiMTTKRP(X, mats, mats_order, mode);
}

}  // extern "C"
