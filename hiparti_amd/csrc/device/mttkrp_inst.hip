// Kernel instantiations for one tensor order.  The build compiles this file
// once per order with -DPTI_NMODES=N (N = 2..8) so the instantiations build in
// parallel.
#include "launch.hpp"

#ifndef PTI_NMODES
#error "compile with -DPTI_NMODES=<tensor order>"
#endif

#define PTI_CAT_(a, b) a##b
#define PTI_CAT(a, b) PTI_CAT_(a, b)

namespace pti {
namespace dev {

namespace {
constexpr int N = PTI_NMODES;

constexpr size_t LDS_PER_CU = 160 * 1024;   // gfx950
constexpr size_t MAX_WAVES_PER_CU = 32;

inline uint32_t grid_for_chunks(uint32_t nchunks) {
    return (nchunks + WAVES_PER_BLOCK - 1) / WAVES_PER_BLOCK;
}

template <int V, int G>
hipError_t hicoo_vg(const HiCOOArgs &h, const FactorArgs &a, bool bf16, hipStream_t s) {
    if constexpr (G * V > 64) {
        return hipErrorInvalidValue;  // the LDS tile holds at most 64 ranks
    } else {
        if (h.nunits == 0) return hipSuccess;
        const size_t lds = (size_t)(1u << h.sb_bits) * (G * V + 1) * sizeof(float);
        dim3 grid(h.nunits);
        // 8 or 4 waves per workgroup, whichever keeps more waves resident per
        // CU within gfx950's 160 KB of LDS (ties: 8).
        auto resident = [&](int W) -> size_t {
            const size_t per_wg = hicoo_stage_bytes<N, V, G>(W) + lds;
            if (per_wg > LDS_PER_CU) return 0;
            size_t wgs = LDS_PER_CU / per_wg;
            if (wgs * W > MAX_WAVES_PER_CU) wgs = MAX_WAVES_PER_CU / W;
            return wgs * W;
        };
        const size_t r8 = resident(8), r4 = resident(4);
        if (r8 == 0 && r4 == 0) return hipErrorInvalidValue;
        if (r8 >= r4) {
            if (bf16)
                hipLaunchKernelGGL((hicoo_tile_kernel<N, V, G, uint16_t, 8>), grid, dim3(512), lds, s, h, a);
            else
                hipLaunchKernelGGL((hicoo_tile_kernel<N, V, G, float, 8>), grid, dim3(512), lds, s, h, a);
        } else {
            if (bf16)
                hipLaunchKernelGGL((hicoo_tile_kernel<N, V, G, uint16_t, 4>), grid, dim3(256), lds, s, h, a);
            else
                hipLaunchKernelGGL((hicoo_tile_kernel<N, V, G, float, 4>), grid, dim3(256), lds, s, h, a);
        }
        return hipGetLastError();
    }
}

// The stream kernel with NRES LDS-resident factors (orders 4 and 5, four
// ranks per lane); dynamic LDS beyond 64 KB is allowed explicitly.
template <int V, int G, typename FT, int NRES, bool PACK>
hipError_t stream_res(const StreamArgs &h, const FactorArgs &a, size_t lds, hipStream_t s) {
    const void *fn = reinterpret_cast<const void *>(&hicoo_stream_kernel<N, V, G, FT, 8, NRES, PACK>);
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((hicoo_stream_kernel<N, V, G, FT, 8, NRES, PACK>), dim3(h.nunits), dim3(512), lds, s, h, a);
    return hipGetLastError();
}

template <int V, int G, bool PACK>
hipError_t stream_vgp(const StreamArgs &h, const FactorArgs &a, bool bf16, hipStream_t s);

// Packed forms exist for orders 4 and 5 only (StreamShape).
template <int V, int G>
hipError_t stream_vg(const StreamArgs &h, const FactorArgs &a, bool bf16, hipStream_t s) {
    if constexpr (N == 4 || N == 5) {
        if (h.packed) return stream_vgp<V, G, true>(h, a, bf16, s);
    } else {
        if (h.packed) return hipErrorInvalidValue;
    }
    return stream_vgp<V, G, false>(h, a, bf16, s);
}

template <int V, int G, bool PACK>
hipError_t stream_vgp(const StreamArgs &h, const FactorArgs &a, bool bf16, hipStream_t s) {
    if constexpr (G * V > 64 || !stream_supported(N, G)) {
        return hipErrorInvalidValue;  // the LDS tile holds at most 64 ranks
    } else {
        if (h.nunits == 0) return hipSuccess;
        const uint32_t rows = 1u << h.sb_bits;
        const int nres = h.res_rows[0] ? (h.res_rows[1] ? 2 : 1) : 0;
        if (nres) {
            if constexpr ((N == 4 || N == 5) && V == 4) {
                const size_t esz = bf16 ? 2 : 4;
                const size_t lds = res_lds_offset(rows, G * V + 1) +
                                   (size_t)(h.res_rows[0] + h.res_rows[1]) * G * V * esz;
                // Order 4 takes one resident factor, order 5 up to two.
                if (N == 4 && nres > 1) return hipErrorInvalidValue;
                if (bf16) {
                    if (nres == 1) return stream_res<V, G, uint16_t, 1, PACK>(h, a, lds, s);
                    if constexpr (N == 5) return stream_res<V, G, uint16_t, 2, PACK>(h, a, lds, s);
                }
                if (nres == 1) return stream_res<V, G, float, 1, PACK>(h, a, lds, s);
                if constexpr (N == 5) return stream_res<V, G, float, 2, PACK>(h, a, lds, s);
            }
            return hipErrorInvalidValue;
        }
        const size_t lds = (size_t)rows * (G * V + 1) * sizeof(float);
        dim3 grid(h.nunits);
        if constexpr (V == 8) {
            // Eight ranks per lane exist only for BF16 factors (16-byte pieces).
            if (!bf16) return hipErrorInvalidValue;
            hipLaunchKernelGGL((hicoo_stream_kernel<N, V, G, uint16_t, 8, 0, PACK>), grid, dim3(512), lds, s, h, a);
        } else {
            if (bf16)
                hipLaunchKernelGGL((hicoo_stream_kernel<N, V, G, uint16_t, 8, 0, PACK>), grid, dim3(512), lds, s, h, a);
            else
                hipLaunchKernelGGL((hicoo_stream_kernel<N, V, G, float, 8, 0, PACK>), grid, dim3(512), lds, s, h, a);
        }
        return hipGetLastError();
    }
}

template <int V, int G>
hipError_t grad_vg(const StreamArgs &h, const FactorArgs &a, double *loss_part, hipStream_t s) {
    if constexpr (G * V > 64 || !stream_supported(N, G)) {
        return hipErrorInvalidValue;
    } else {
        if (h.nunits == 0) return hipSuccess;
        const size_t rows = (size_t)1 << h.sb_bits;
        const size_t lds = rows * (G * V + 1) * sizeof(float) + rows * (G * V + 4) * sizeof(float);
        hipLaunchKernelGGL((completion_grad_kernel<N, V, G, 8>), dim3(h.nunits), dim3(512), lds, s, h, a, loss_part);
        return hipGetLastError();
    }
}

template <int V>
hipError_t grad_v(const StreamArgs &h, const FactorArgs &a, int G, double *loss_part, hipStream_t s) {
    switch (G) {
        case 4: return grad_vg<V, 4>(h, a, loss_part, s);
        case 8: return grad_vg<V, 8>(h, a, loss_part, s);
        case 16: return grad_vg<V, 16>(h, a, loss_part, s);
        default: return hipErrorInvalidValue;
    }
}

template <int V>
hipError_t stream_v(const StreamArgs &h, const FactorArgs &a, int G, bool bf16, hipStream_t s) {
    switch (G) {
        case 4: return stream_vg<V, 4>(h, a, bf16, s);
        case 8: return stream_vg<V, 8>(h, a, bf16, s);
        case 16: return stream_vg<V, 16>(h, a, bf16, s);
        case 32: return stream_vg<V, 32>(h, a, bf16, s);
        case 64: return stream_vg<V, 64>(h, a, bf16, s);
        default: return hipErrorInvalidValue;
    }
}

template <int V, int G>
hipError_t coo_vg(const COOArgs &c, const FactorArgs &a, hipStream_t s) {
    if (c.nchunks == 0) return hipSuccess;
    dim3 grid(grid_for_chunks(c.nchunks)), block(256);
    hipLaunchKernelGGL((coo_mttkrp_kernel<N, V, G, float>), grid, block, 0, s, c, a);
    return hipGetLastError();
}

template <int V, int G>
hipError_t completion_vg(const COOArgs &c, const FactorArgs &a, float *resid, double *partials, hipStream_t s) {
    if (c.nchunks == 0) return hipSuccess;
    dim3 grid(grid_for_chunks(c.nchunks)), block(256);
    hipLaunchKernelGGL((completion_forward_kernel<N, V, G>), grid, block, 0, s, c, a, resid, partials);
    return hipGetLastError();
}

template <int V>
hipError_t completion_v(const COOArgs &c, const FactorArgs &a, int G, float *resid, double *partials,
                        hipStream_t s) {
    switch (G) {
        case 4: return completion_vg<V, 4>(c, a, resid, partials, s);
        case 8: return completion_vg<V, 8>(c, a, resid, partials, s);
        case 16: return completion_vg<V, 16>(c, a, resid, partials, s);
        case 32: return completion_vg<V, 32>(c, a, resid, partials, s);
        case 64: return completion_vg<V, 64>(c, a, resid, partials, s);
        default: return hipErrorInvalidValue;
    }
}

template <int V>
hipError_t hicoo_v(const HiCOOArgs &h, const FactorArgs &a, int G, bool bf16, hipStream_t s) {
    switch (G) {
        case 4: return hicoo_vg<V, 4>(h, a, bf16, s);
        case 8: return hicoo_vg<V, 8>(h, a, bf16, s);
        case 16: return hicoo_vg<V, 16>(h, a, bf16, s);
        case 32: return hicoo_vg<V, 32>(h, a, bf16, s);
        case 64: return hicoo_vg<V, 64>(h, a, bf16, s);
        default: return hipErrorInvalidValue;
    }
}

template <int V>
hipError_t coo_v(const COOArgs &c, const FactorArgs &a, int G, hipStream_t s) {
    switch (G) {
        case 4: return coo_vg<V, 4>(c, a, s);
        case 8: return coo_vg<V, 8>(c, a, s);
        case 16: return coo_vg<V, 16>(c, a, s);
        case 32: return coo_vg<V, 32>(c, a, s);
        case 64: return coo_vg<V, 64>(c, a, s);
        default: return hipErrorInvalidValue;
    }
}
}  // namespace

hipError_t PTI_CAT(launch_hicoo_, PTI_NMODES)(const HiCOOArgs &h, const FactorArgs &a, Shape sh, bool bf16,
                                              hipStream_t s) {
    switch (sh.V) {
        case 4: return hicoo_v<4>(h, a, sh.G, bf16, s);
        case 2: return hicoo_v<2>(h, a, sh.G, bf16, s);
        case 1: return hicoo_v<1>(h, a, sh.G, bf16, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t PTI_CAT(launch_stream_, PTI_NMODES)(const StreamArgs &h, const FactorArgs &a, Shape sh, bool bf16,
                                               hipStream_t s) {
    switch (sh.V) {
        case 8:   // BF16 rank tiles above 32 ranks: lane groups of eight
            return sh.G == 8 ? stream_vg<8, 8>(h, a, bf16, s) : hipErrorInvalidValue;
        case 4: return stream_v<4>(h, a, sh.G, bf16, s);
        case 2: return stream_v<2>(h, a, sh.G, bf16, s);
        case 1: return stream_v<1>(h, a, sh.G, bf16, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t PTI_CAT(launch_completion_grad_, PTI_NMODES)(const StreamArgs &h, const FactorArgs &a, Shape sh,
                                                        double *loss_part, hipStream_t s) {
    switch (sh.V) {
        case 4: return grad_v<4>(h, a, sh.G, loss_part, s);
        case 2: return grad_v<2>(h, a, sh.G, loss_part, s);
        case 1: return grad_v<1>(h, a, sh.G, loss_part, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t PTI_CAT(launch_coo_, PTI_NMODES)(const COOArgs &c, const FactorArgs &a, Shape sh, hipStream_t s) {
    switch (sh.V) {
        case 4: return coo_v<4>(c, a, sh.G, s);
        case 2: return coo_v<2>(c, a, sh.G, s);
        case 1: return coo_v<1>(c, a, sh.G, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t PTI_CAT(launch_completion_, PTI_NMODES)(const COOArgs &c, const FactorArgs &a, Shape sh, float *resid,
                                                  double *partials, hipStream_t s) {
    switch (sh.V) {
        case 4: return completion_v<4>(c, a, sh.G, resid, partials, s);
        case 2: return completion_v<2>(c, a, sh.G, resid, partials, s);
        case 1: return completion_v<1>(c, a, sh.G, resid, partials, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace dev
}  // namespace pti
