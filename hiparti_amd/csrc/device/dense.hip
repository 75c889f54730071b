// FP64 dense step of CP-ALS on the device (reference: src/matrix/matrix.c:340-572,
// src/matrix/matrix_solver.c:32-83, src/ktensor/ktensor.c:77-240).
//
// Kernels:
//   gram_partial / gram_reduce   G = A^T A in 32 x 32 tiles of the upper
//                                triangle (one 16 x 16 tile for R <= 16),
//                                rows split over workgroups, partial tiles
//                                summed in split order.
//   solve_normals                one workgroup: V = Hadamard of the other
//                                modes' Grams, Cholesky (right-looking), else
//                                LU with partial pivoting (the reference's
//                                spotrf -> sgesv fallback), then W = V^-1 one
//                                column per thread.  Exactly singular V leaves
//                                the right-hand side unchanged, as gesv does.
//   apply                        A = M W, 64 x 64 output tiles (64 x 16 for
//                                R <= 16), FP64 sums.
//   colstat_partial / _reduce    column 2-norm, max (clamped >= 1) or dot.
//   scale                        A[:, r] /= lambda[r].
//   fit                          ||K||^2 from the full Grams, <X,K>, fit.
// Rank is limited to 256 (the reference's ptiRankMatrix bound).
#include <cmath>

#include "dense.hpp"

namespace pti {
namespace dev {
namespace {

constexpr int GT = 32;          // Gram tile edge
constexpr int GROWS = 64;       // rows staged per LDS pass in the Gram kernel
constexpr int G16ROWS = 256;    // rows staged per LDS pass in the R <= 16 Gram kernel
constexpr uint32_t MAX_PARTS = 1024;  // workgroups contributing partial sums

__host__ __device__ inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

inline uint32_t gram_tiles(uint32_t R) {
    const uint32_t nt = cdiv(R, GT);
    return nt * (nt + 1) / 2;
}

// Tile (tp, tq), tp <= tq, of linear index t in row-major order of the upper
// triangle of an nt x nt tile grid.
__device__ inline void tile_of(uint32_t t, uint32_t nt, uint32_t &tp, uint32_t &tq) {
    tp = 0;
    while (t >= nt - tp) {
        t -= nt - tp;
        ++tp;
    }
    tq = tp + t;
}

__global__ void __launch_bounds__(256) gram_partial_kernel(const float *__restrict__ A, uint64_t nrows, uint32_t R,
                                                           uint32_t stride, uint64_t rows_per_split,
                                                           uint32_t nsplit, double *__restrict__ partials) {
    __shared__ float sp[GROWS][GT + 1];
    __shared__ float sq[GROWS][GT + 1];
    uint32_t tp, tq;
    tile_of(blockIdx.y, cdiv(R, GT), tp, tq);
    const int tid = threadIdx.x, ty = tid >> 3, tx = tid & 7;  // entry row ty, columns 4*tx .. 4*tx+3
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    const uint64_t i0 = (uint64_t)blockIdx.x * rows_per_split;
    const uint64_t i1 = i0 + rows_per_split < nrows ? i0 + rows_per_split : nrows;
    // A diagonal tile (tp == tq, every tile for R <= 32) stages its rows once.
    const bool diag = tp == tq;
    float(*qs)[GT + 1] = diag ? sp : sq;
    for (uint64_t base = i0; base < i1; base += GROWS) {
        for (int e = tid; e < GROWS * GT; e += 256) {
            const int r = e / GT, c = e % GT;
            const uint64_t i = base + r;
            const uint32_t cp = tp * GT + c, cq = tq * GT + c;
            sp[r][c] = (i < i1 && cp < R) ? A[i * stride + cp] : 0.f;
            if (!diag) sq[r][c] = (i < i1 && cq < R) ? A[i * stride + cq] : 0.f;
        }
        __syncthreads();
#pragma unroll 4
        for (int r = 0; r < GROWS; ++r) {
            const double a = sp[r][ty];
#pragma unroll
            for (int k = 0; k < 4; ++k) acc[k] = fma(a, (double)qs[r][4 * tx + k], acc[k]);
        }
        __syncthreads();
    }
    double *out = partials + ((uint64_t)blockIdx.y * nsplit + blockIdx.x) * (GT * GT);
#pragma unroll
    for (int k = 0; k < 4; ++k) out[ty * GT + 4 * tx + k] = acc[k];
}

// Gram partials for R <= 16: the single tile holds at most 16 x 16 entries,
// so the 256 threads form 4 row phases x 64 entry groups (entry row p = e / 4,
// columns 4 (e % 4) .. 4 (e % 4) + 3), each phase taking every fourth staged
// row; the phases are added in a fixed order before the partial is written.
// Rows are staged G16ROWS at a time (16 independent loads per thread), so a
// workgroup's 1024 rows wait on four load latencies, not sixteen.
// Entry (p, q) goes where gram_reduce_kernel reads it (p * GT + q of tile 0);
// entries past the rank are never read there.
__global__ void __launch_bounds__(256) gram16_partial_kernel(const float *__restrict__ A, uint64_t nrows, uint32_t R,
                                                             uint32_t stride, uint64_t rows_per_split,
                                                             double *__restrict__ partials) {
    constexpr int E = 16;
    __shared__ float s[G16ROWS][E + 1];
    __shared__ double red[4][E * E];
    const int tid = threadIdx.x, ph = tid >> 6, e = tid & 63;
    const int p = e >> 2, q0 = (e & 3) * 4;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    const uint64_t i0 = (uint64_t)blockIdx.x * rows_per_split;
    const uint64_t i1 = i0 + rows_per_split < nrows ? i0 + rows_per_split : nrows;
    for (uint64_t base = i0; base < i1; base += G16ROWS) {
#pragma unroll
        for (int x = tid; x < G16ROWS * E; x += 256) {
            const int r = x / E, c = x % E;
            const uint64_t i = base + r;
            s[r][c] = (i < i1 && (uint32_t)c < R) ? A[i * stride + c] : 0.f;
        }
        __syncthreads();
#pragma unroll 4
        for (int r = ph; r < G16ROWS; r += 4) {
            const double a = s[r][p];
#pragma unroll
            for (int k = 0; k < 4; ++k) acc[k] = fma(a, (double)s[r][q0 + k], acc[k]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) red[ph][4 * e + k] = acc[k];   // entry p * 16 + q0 + k == 4 * e + k
    __syncthreads();
    const double v = ((red[0][tid] + red[1][tid]) + red[2][tid]) + red[3][tid];
    partials[(uint64_t)blockIdx.x * (GT * GT) + (tid / E) * GT + (tid % E)] = v;
}

// Sums the nsplit partial tiles: workgroup (tile, y) takes the tile's 64
// entries y*64 .. y*64+63 (two tile rows), entries outside R x R skipped; its
// 4 waves sum contiguous quarters of the partials, in order, then the
// quarters are added in order, so the result does not depend on timing.
__global__ void __launch_bounds__(256) gram_reduce_kernel(const double *__restrict__ partials, uint32_t nsplit,
                                                          uint32_t R, double *__restrict__ G) {
    __shared__ double red[4][64];
    uint32_t tp, tq;
    tile_of(blockIdx.x, cdiv(R, GT), tp, tq);
    const int tid = threadIdx.x, quarter = tid >> 6, lane = tid & 63;
    const uint32_t e = blockIdx.y * 64 + lane;
    const uint32_t p = tp * GT + e / GT, q = tq * GT + e % GT;
    if (tp * GT + (blockIdx.y * 64) / GT >= R) return;   // both tile rows past the rank: whole workgroup
    const uint32_t per = (nsplit + 3) / 4;
    const uint32_t k0 = quarter * per, k1 = k0 + per < nsplit ? k0 + per : nsplit;
    double sum = 0.0;
    if (p < R && q < R) {
        const double *src = partials + (uint64_t)blockIdx.x * nsplit * (GT * GT) + e;
#pragma unroll 8
        for (uint32_t k = k0; k < k1; ++k) sum += src[(uint64_t)k * (GT * GT)];
    }
    red[quarter][lane] = sum;
    __syncthreads();
    if (quarter == 0 && p < R && q < R && (tp < tq || p <= q)) {
        // Diagonal tiles hold both (p, q) and (q, p); take the upper one.
        const double t = ((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane];
        G[(size_t)p * R + q] = t;
        G[(size_t)q * R + p] = t;
    }
}

// One workgroup.  Vg, Wg: R x R scratch in device memory; piv: R ints;
// status: 1 int.  For R <= SOLVE_LDS_RANK the factorization and the inverse
// run in dynamic LDS (2 R^2 doubles, which the launch provides) and only W is
// written back: every elimination step is a workgroup barrier, and with the
// matrices in device memory each step also waited on L2 round trips (37 us
// per solve at R = 16 there, 25 us in LDS).
constexpr uint32_t SOLVE_LDS_RANK = 64;
__global__ void __launch_bounds__(256) solve_normals_kernel(const double *__restrict__ grams, uint32_t nmodes,
                                                            uint32_t mode, uint32_t R, double *Vg, double *Wg,
                                                            int *piv, int *status) {
    __shared__ int s_state;  // 0 running, 1 not SPD, 2 singular
    __shared__ int s_piv;
    extern __shared__ double s_mats[];
    const int tid = threadIdx.x;
    const uint32_t RR = R * R;
    const bool in_lds = R <= SOLVE_LDS_RANK;
    double *V = in_lds ? s_mats : Vg;
    double *W = in_lds ? s_mats + RR : Wg;
    // W back to device memory (from LDS), then the status.
    auto finish = [&](int st) {
        if (in_lds && st != 2) {
            __syncthreads();
            for (uint32_t x = tid; x < RR; x += 256) Wg[x] = W[x];
        }
        if (tid == 0) *status = st;
    };
    auto build = [&]() {
        for (uint32_t x = tid; x < RR; x += 256) {
            double v = 1.0;
            for (uint32_t k = 0; k < nmodes; ++k)
                if (k != mode) v *= grams[(size_t)k * RR + x];
            V[x] = v;
        }
    };
    build();
    if (tid == 0) s_state = 0;
    __syncthreads();
    // Cholesky V = L L^T in place (lower triangle), right-looking.
    for (uint32_t j = 0; j < R; ++j) {
        if (tid == 0) {
            const double d = V[(size_t)j * R + j];
            if (d > 0.0)
                V[(size_t)j * R + j] = sqrt(d);
            else
                s_state = 1;
        }
        __syncthreads();
        if (s_state) break;
        const double djj = V[(size_t)j * R + j];
        for (uint32_t i = j + 1 + tid; i < R; i += 256) V[(size_t)i * R + j] /= djj;
        __syncthreads();
        const uint32_t n = R - j - 1;
        for (uint32_t x = tid; x < n * n; x += 256) {
            const uint32_t i = j + 1 + x / n, k = j + 1 + x % n;
            if (k <= i) V[(size_t)i * R + k] -= V[(size_t)i * R + j] * V[(size_t)k * R + j];
        }
        __syncthreads();
    }
    if (s_state == 0) {
        // W = L^-T L^-1, column c by thread c: L y = e_c, then L^T x = y.
        for (uint32_t c = tid; c < R; c += 256) {
            for (uint32_t r = 0; r < R; ++r) {
                double s = (r == c) ? 1.0 : 0.0;
                for (uint32_t k = c; k < r; ++k) s -= V[(size_t)r * R + k] * W[(size_t)k * R + c];
                W[(size_t)r * R + c] = r < c ? 0.0 : s / V[(size_t)r * R + r];
            }
            for (uint32_t r = R; r-- > 0;) {
                double s = W[(size_t)r * R + c];
                for (uint32_t k = r + 1; k < R; ++k) s -= V[(size_t)k * R + r] * W[(size_t)k * R + c];
                W[(size_t)r * R + c] = s / V[(size_t)r * R + r];
            }
        }
        finish(0);
        return;
    }
    // Not positive definite: rebuild V and factor P V = L U.
    __syncthreads();
    build();
    if (tid == 0) s_state = 0;
    __syncthreads();
    for (uint32_t k = 0; k < R; ++k) {
        if (tid == 0) {
            uint32_t p = k;
            double best = fabs(V[(size_t)k * R + k]);
            for (uint32_t i = k + 1; i < R; ++i) {
                const double a = fabs(V[(size_t)i * R + k]);
                if (a > best) {
                    best = a;
                    p = i;
                }
            }
            piv[k] = (int)p;
            s_piv = (int)p;
            if (best == 0.0) s_state = 2;
        }
        __syncthreads();
        if (s_state) break;
        const uint32_t p = (uint32_t)s_piv;
        if (p != k)
            for (uint32_t j = tid; j < R; j += 256) {
                const double t = V[(size_t)k * R + j];
                V[(size_t)k * R + j] = V[(size_t)p * R + j];
                V[(size_t)p * R + j] = t;
            }
        __syncthreads();
        const double pk = V[(size_t)k * R + k];
        for (uint32_t i = k + 1 + tid; i < R; i += 256) V[(size_t)i * R + k] /= pk;
        __syncthreads();
        const uint32_t n = R - k - 1;
        for (uint32_t x = tid; x < n * n; x += 256) {
            const uint32_t i = k + 1 + x / n, j = k + 1 + x % n;
            V[(size_t)i * R + j] -= V[(size_t)i * R + k] * V[(size_t)k * R + j];
        }
        __syncthreads();
    }
    if (s_state == 2) {
        finish(2);
        return;
    }
    // W = V^-1: column c solves V x = e_c through the row swaps, L and U.
    for (uint32_t c = tid; c < R; c += 256) {
        for (uint32_t r = 0; r < R; ++r) W[(size_t)r * R + c] = (r == c) ? 1.0 : 0.0;
        for (uint32_t k = 0; k < R; ++k) {
            const uint32_t p = (uint32_t)piv[k];
            if (p != k) {
                const double t = W[(size_t)k * R + c];
                W[(size_t)k * R + c] = W[(size_t)p * R + c];
                W[(size_t)p * R + c] = t;
            }
        }
        for (uint32_t r = 0; r < R; ++r) {
            double s = W[(size_t)r * R + c];
            for (uint32_t k = 0; k < r; ++k) s -= V[(size_t)r * R + k] * W[(size_t)k * R + c];
            W[(size_t)r * R + c] = s;
        }
        for (uint32_t r = R; r-- > 0;) {
            double s = W[(size_t)r * R + c];
            for (uint32_t k = r + 1; k < R; ++k) s -= V[(size_t)r * R + k] * W[(size_t)k * R + c];
            W[(size_t)r * R + c] = s / V[(size_t)r * R + r];
        }
    }
    finish(1);
}

// A = M W for an nrows x R block of rows (64 x 16 NB output tiles, 16-deep
// k slices staged in LDS; NB = 4, or 1 for R <= 16).  Thread (ty, tx) owns
// rows ty + 16a, a in 0..3, and columns tx + 16b, b < NB.  When status == 2,
// A = M.
template <int NB>
__global__ void __launch_bounds__(256) apply_kernel(const float *__restrict__ Mrhs, float *__restrict__ A,
                                                    const double *__restrict__ W, const int *__restrict__ status,
                                                    uint64_t nrows, uint32_t R, uint32_t stride) {
    constexpr int TC = 16 * NB;   // output columns per tile
    __shared__ float sm[16][64 + 1];
    __shared__ double sw[16][TC];
    const int tid = threadIdx.x, ty = tid >> 4, tx = tid & 15;
    const uint64_t r0 = (uint64_t)blockIdx.x * 64;
    const uint32_t c0 = blockIdx.y * TC;
    const bool copy = *status == 2;
    double acc[4][NB] = {};
    if (!copy) {
        for (uint32_t k0 = 0; k0 < R; k0 += 16) {
            for (int e = tid; e < 64 * 16; e += 256) {
                const int r = e / 16, k = e % 16;
                const uint64_t i = r0 + r;
                sm[k][r] = (i < nrows && k0 + k < R) ? Mrhs[i * stride + k0 + k] : 0.f;
            }
            for (int e = tid; e < TC * 16; e += 256) {
                const int kk = e / TC, c = e % TC;
                sw[kk][c] = (k0 + kk < R && c0 + c < R) ? W[(size_t)(k0 + kk) * R + c0 + c] : 0.0;
            }
            __syncthreads();
#pragma unroll 4
            for (int k = 0; k < 16; ++k) {
                double m[4], w[NB];
#pragma unroll
                for (int a = 0; a < 4; ++a) m[a] = sm[k][ty + 16 * a];
#pragma unroll
                for (int b = 0; b < NB; ++b) w[b] = sw[k][tx + 16 * b];
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int b = 0; b < NB; ++b) acc[a][b] = fma(m[a], w[b], acc[a][b]);
            }
            __syncthreads();
        }
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const uint64_t i = r0 + ty + 16 * a;
        if (i >= nrows) continue;
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const uint32_t c = c0 + tx + 16 * b;
            if (c < R) A[i * stride + c] = copy ? Mrhs[i * stride + c] : (float)acc[a][b];
        }
    }
}

// Per-workgroup column partials over a row range: the 256 threads are CW
// column lanes (CW = 16, 32 or 64, the smallest covering R, columns done CW at
// a time) times 256 / CW row phases; phase f takes rows i0 + f, i0 + f + P,
// ...; the phases are combined in order.
__global__ void __launch_bounds__(256) colstat_partial_kernel(const float *__restrict__ A,
                                                              const float *__restrict__ B, uint64_t nrows,
                                                              uint32_t R, uint32_t stride, int kind,
                                                              uint64_t rows_per_split, double *__restrict__ partials) {
    __shared__ double red[256];
    const uint32_t CW = R > 32 ? 64u : (R > 16 ? 32u : 16u);
    const uint32_t P = 256 / CW;
    const uint32_t tid = threadIdx.x, phase = tid / CW, lane = tid % CW;
    const uint64_t i0 = (uint64_t)blockIdx.x * rows_per_split;
    const uint64_t i1 = i0 + rows_per_split < nrows ? i0 + rows_per_split : nrows;
    for (uint32_t g = 0; g < R; g += CW) {
        const uint32_t c = g + lane;
        double acc = 0.0;
        if (c < R) {
            // Unrolled so that eight loads are in flight; the accumulation
            // order stays the row order.
#pragma unroll 8
            for (uint64_t i = i0 + phase; i < i1; i += P) {
                const double a = A[i * stride + c];
                if (kind == COL_2NORM)
                    acc = fma(a, a, acc);
                else if (kind == COL_MAXNORM)
                    acc = a > acc ? a : acc;
                else
                    acc = fma(a, (double)B[i * stride + c], acc);
            }
        }
        red[tid] = acc;
        __syncthreads();
        if (phase == 0 && c < R) {
            double t = red[lane];
            for (uint32_t f = 1; f < P; ++f) {
                const double v = red[f * CW + lane];
                t = kind == COL_MAXNORM ? (v > t ? v : t) : t + v;
            }
            partials[(uint64_t)blockIdx.x * R + c] = t;
        }
        __syncthreads();
    }
}

// Combines the nsplit column partials: workgroup y takes columns
// y*64 .. y*64+63; its 4 waves reduce contiguous quarters of the partials in
// order, then the quarters are combined in order.
__global__ void __launch_bounds__(256) colstat_reduce_kernel(const double *__restrict__ partials, uint32_t nsplit,
                                                             uint32_t R, int kind, double *__restrict__ out,
                                                             double *__restrict__ mul) {
    __shared__ double red[4][64];
    const int tid = threadIdx.x, quarter = tid >> 6, lane = tid & 63;
    const uint32_t c = blockIdx.x * 64 + lane;
    const uint32_t per = (nsplit + 3) / 4;
    const uint32_t k0 = quarter * per, k1 = k0 + per < nsplit ? k0 + per : nsplit;
    double t = 0.0;
    if (c < R) {
#pragma unroll 8
        for (uint32_t k = k0; k < k1; ++k) {
            const double v = partials[(uint64_t)k * R + c];
            t = kind == COL_MAXNORM ? (v > t ? v : t) : t + v;
        }
    }
    red[quarter][lane] = t;
    __syncthreads();
    if (quarter == 0 && c < R) {
        t = red[0][lane];
        for (int f = 1; f < 4; ++f) t = kind == COL_MAXNORM ? (red[f][lane] > t ? red[f][lane] : t) : t + red[f][lane];
        if (kind == COL_2NORM) t = sqrt(t);
        if (kind == COL_MAXNORM && t < 1.0) t = 1.0;
        out[c] = t;
        if (mul) mul[c] *= t;
    }
}

__global__ void __launch_bounds__(256) scale_kernel(float *__restrict__ A, uint64_t nrows, uint32_t R,
                                                    uint32_t stride, const double *__restrict__ lambda) {
    const uint64_t n = nrows * R;
    if (n < (1ull << 31)) {
        // 32-bit index arithmetic (x plus the grid stride stays below 2^32):
        // a 64-bit division per element costs far more than the element's
        // load and store.
        for (uint32_t x = blockIdx.x * 256 + threadIdx.x; x < (uint32_t)n; x += gridDim.x * 256) {
            const uint32_t i = x / R, c = x % R;
            A[(uint64_t)i * stride + c] = (float)((double)A[(uint64_t)i * stride + c] / lambda[c]);
        }
        return;
    }
    for (uint64_t x = (uint64_t)blockIdx.x * 256 + threadIdx.x; x < n; x += (uint64_t)gridDim.x * 256) {
        const uint64_t i = x / R;
        const uint32_t c = (uint32_t)(x % R);
        A[i * stride + c] = (float)((double)A[i * stride + c] / lambda[c]);
    }
}

__global__ void __launch_bounds__(256) sumsq_partial_kernel(const float *__restrict__ v, uint64_t n,
                                                            uint64_t per_split, double *__restrict__ partials) {
    __shared__ double red[256];
    const uint64_t z0 = (uint64_t)blockIdx.x * per_split;
    const uint64_t z1 = z0 + per_split < n ? z0 + per_split : n;
    double acc = 0.0;
    for (uint64_t z = z0 + threadIdx.x; z < z1; z += 256) {
        const double a = v[z];
        acc = fma(a, a, acc);
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(256) sum_reduce_kernel(const double *__restrict__ partials, uint32_t n,
                                                         double *__restrict__ out) {
    __shared__ double red[256];
    double t = 0.0;
    for (uint32_t k = threadIdx.x; k < n; k += 256) t += partials[k];
    red[threadIdx.x] = t;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

// scal[0] = ||X||^2 (input); writes scal[1] = fit, scal[2] = ||K||^2,
// scal[3] = <X,K>.
__global__ void __launch_bounds__(256) fit_kernel(const double *__restrict__ grams, uint32_t nmodes, uint32_t R,
                                                  const double *__restrict__ lambda, const double *__restrict__ dot,
                                                  double *scal) {
    __shared__ double red[256];
    const uint32_t RR = R * R;
    double t = 0.0;
    for (uint32_t x = threadIdx.x; x < RR; x += 256) {
        double h = lambda[x / R] * lambda[x % R];
        for (uint32_t k = 0; k < nmodes; ++k) h *= grams[(size_t)k * RR + x];
        t += h;
    }
    red[threadIdx.x] = t;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double normK = fabs(red[0]);
        double inner = 0.0;
        for (uint32_t r = 0; r < R; ++r) inner += dot[r] * lambda[r];
        const double normX2 = scal[0];
        double residual = normX2 + normK - 2 * inner;
        if (residual > 0.0) residual = sqrt(residual);
        scal[1] = 1 - residual / sqrt(normX2);
        scal[2] = normK;
        scal[3] = inner;
    }
}

__global__ void __launch_bounds__(256) to_bf16_kernel(const float *__restrict__ A, uint16_t *__restrict__ B,
                                                      uint64_t nrows, uint32_t R, uint32_t stride) {
    const uint64_t n = nrows * R;
    const bool idx32 = n < (1ull << 31);   // 32-bit division, as in scale_kernel
    for (uint64_t x = (uint64_t)blockIdx.x * 256 + threadIdx.x; x < n; x += (uint64_t)gridDim.x * 256) {
        const uint64_t off = idx32 ? (uint64_t)((uint32_t)x / R) * stride + (uint32_t)x % R : (x / R) * stride + x % R;
        const uint32_t u = __float_as_uint(A[off]);
        uint16_t h;
        if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu))
            h = (uint16_t)((u >> 16) | 0x0040u);  // NaN stays a quiet NaN
        else
            h = (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);  // round to nearest, ties to even
        B[off] = h;
    }
}

inline dim3 grid_elems(uint64_t n) {
    uint64_t b = (n + 255) / 256;
    if (b > 8192) b = 8192;
    if (b == 0) b = 1;
    return dim3((uint32_t)b);
}

}  // namespace

int DenseWork::ensure(uint32_t r, const char *M) {
    if (r == 0 || r > DENSE_MAX_RANK) PTI_FAIL(PTIERR_VALUE_ERROR, M, "rank must be in [1, 256]");
    if (R >= r && partials.ptr) return 0;
    const size_t parts = (size_t)MAX_PARTS * (GT * GT);  // >= every kernel's partial count
    PTI_HIP(partials.alloc(parts * sizeof(double)), M);
    PTI_HIP(V.alloc((size_t)r * r * sizeof(double)), M);
    PTI_HIP(W.alloc((size_t)r * r * sizeof(double)), M);
    PTI_HIP(piv.alloc((size_t)r * sizeof(int)), M);
    PTI_HIP(status.alloc(sizeof(int)), M);
    PTI_HIP(colstat.alloc((size_t)r * sizeof(double)), M);
    PTI_HIP(scal.alloc(4 * sizeof(double)), M);
    R = r;
    return 0;
}

int dense_gram(const float *A, uint64_t nrows, uint32_t R, uint32_t stride, double *G, DenseWork &w, hipStream_t s,
               const char *M) {
    PTI_CHECK(w.ensure(R, M), M);
    if (stride < R) PTI_FAIL(PTIERR_SHAPE_MISMATCH, M, "stride < rank");
    const uint32_t ntiles = gram_tiles(R);
    // R <= 16: one 256-row stage per workgroup up to MAX_PARTS workgroups, so
    // that factors of a few thousand rows are not left to one workgroup.
    uint32_t nsplit = cdiv(nrows, R <= 16 ? G16ROWS : 1024);
    const uint32_t cap = MAX_PARTS / ntiles;
    if (nsplit > cap) nsplit = cap;
    if (nsplit == 0) nsplit = 1;
    uint64_t per = (nrows + nsplit - 1) / nsplit;
    const uint64_t stage = R <= 16 ? G16ROWS : GROWS;
    per = (per + stage - 1) / stage * stage;
    if (R <= 16)
        hipLaunchKernelGGL(gram16_partial_kernel, dim3(nsplit), dim3(256), 0, s, A, nrows, R, stride, per,
                           w.partials.as<double>());
    else
        hipLaunchKernelGGL(gram_partial_kernel, dim3(nsplit, ntiles), dim3(256), 0, s, A, nrows, R, stride, per,
                           nsplit, w.partials.as<double>());
    PTI_HIP(hipGetLastError(), M);
    hipLaunchKernelGGL(gram_reduce_kernel, dim3(ntiles, GT * GT / 64), dim3(256), 0, s, w.partials.as<double>(),
                       nsplit, R, G);
    PTI_HIP(hipGetLastError(), M);
    return 0;
}

int dense_solve_normals(const double *grams, uint32_t nmodes, uint32_t mode, uint32_t R, const float *Mrhs, float *A,
                        uint64_t nrows, uint32_t stride, DenseWork &w, hipStream_t s, const char *M) {
    PTI_CHECK(w.ensure(R, M), M);
    if (stride < R) PTI_FAIL(PTIERR_SHAPE_MISMATCH, M, "stride < rank");
    if (nrows && Mrhs == A) PTI_FAIL(PTIERR_VALUE_ERROR, M, "right-hand side and result must not alias");
    const size_t solve_lds = R <= SOLVE_LDS_RANK ? (size_t)2 * R * R * sizeof(double) : 0;
    hipLaunchKernelGGL(solve_normals_kernel, dim3(1), dim3(256), solve_lds, s, grams, nmodes, mode, R, w.V.as<double>(),
                       w.W.as<double>(), w.piv.as<int>(), w.status.as<int>());
    PTI_HIP(hipGetLastError(), M);
    if (nrows == 0) return 0;
    if (R <= 16)
        hipLaunchKernelGGL(apply_kernel<1>, dim3(cdiv(nrows, 64), 1), dim3(256), 0, s, Mrhs, A, w.W.as<double>(),
                           w.status.as<int>(), nrows, R, stride);
    else
        hipLaunchKernelGGL(apply_kernel<4>, dim3(cdiv(nrows, 64), cdiv(R, 64)), dim3(256), 0, s, Mrhs, A,
                           w.W.as<double>(), w.status.as<int>(), nrows, R, stride);
    PTI_HIP(hipGetLastError(), M);
    return 0;
}

int dense_colstat(float *A, const float *B, uint64_t nrows, uint32_t R, uint32_t stride, ColStat kind, double *out,
                  bool scale, double *mul, DenseWork &w, hipStream_t s, const char *M) {
    PTI_CHECK(w.ensure(R, M), M);
    if (stride < R) PTI_FAIL(PTIERR_SHAPE_MISMATCH, M, "stride < rank");
    uint32_t nsplit = cdiv(nrows, 256);   // 256 rows per workgroup up to MAX_PARTS workgroups
    if (nsplit > MAX_PARTS) nsplit = MAX_PARTS;
    if (nsplit == 0) nsplit = 1;
    const uint64_t per = (nrows + nsplit - 1) / nsplit;
    hipLaunchKernelGGL(colstat_partial_kernel, dim3(nsplit), dim3(256), 0, s, A, B, nrows, R, stride, (int)kind, per,
                       w.partials.as<double>());
    PTI_HIP(hipGetLastError(), M);
    hipLaunchKernelGGL(colstat_reduce_kernel, dim3(cdiv(R, 64)), dim3(256), 0, s, w.partials.as<double>(), nsplit, R,
                       (int)kind, out, mul);
    PTI_HIP(hipGetLastError(), M);
    if (scale && nrows) {
        hipLaunchKernelGGL(scale_kernel, grid_elems(nrows * R), dim3(256), 0, s, A, nrows, R, stride, out);
        PTI_HIP(hipGetLastError(), M);
    }
    return 0;
}

int dense_fit(const double *grams, uint32_t nmodes, uint32_t R, const double *lambda, const double *dot,
              DenseWork &w, hipStream_t s, const char *M) {
    PTI_CHECK(w.ensure(R, M), M);
    hipLaunchKernelGGL(fit_kernel, dim3(1), dim3(256), 0, s, grams, nmodes, R, lambda, dot, w.scal.as<double>());
    PTI_HIP(hipGetLastError(), M);
    return 0;
}

int dense_sumsq(const float *v, uint64_t n, DenseWork &w, hipStream_t s, const char *M) {
    if (!w.partials.ptr) PTI_FAIL(PTIERR_VALUE_ERROR, M, "dense workspace not initialised");
    uint32_t nsplit = cdiv(n, 65536);
    if (nsplit > MAX_PARTS) nsplit = MAX_PARTS;
    if (nsplit == 0) nsplit = 1;
    const uint64_t per = (n + nsplit - 1) / nsplit;
    hipLaunchKernelGGL(sumsq_partial_kernel, dim3(nsplit), dim3(256), 0, s, v, n, per, w.partials.as<double>());
    PTI_HIP(hipGetLastError(), M);
    hipLaunchKernelGGL(sum_reduce_kernel, dim3(1), dim3(256), 0, s, w.partials.as<double>(), nsplit,
                       w.scal.as<double>());
    PTI_HIP(hipGetLastError(), M);
    return 0;
}

int dense_to_bf16(const float *A, uint16_t *B, uint64_t nrows, uint32_t R, uint32_t stride, hipStream_t s,
                  const char *M) {
    if (nrows == 0) return 0;
    hipLaunchKernelGGL(to_bf16_kernel, grid_elems(nrows * R), dim3(256), 0, s, A, B, nrows, R, stride);
    PTI_HIP(hipGetLastError(), M);
    return 0;
}

}  // namespace dev
}  // namespace pti
