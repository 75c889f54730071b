// CP-ALS (reference: src/sptensor/cpd.c:32-208, src/matrix/matrix_solver.c,
// src/matrix/matrix.c:340-572, src/ktensor/ktensor.c:77-240).
//
// The tensor and the factor matrices stay in HBM for the whole run; each
// mode update is one MTTKRP on the device followed by the R x R normal
// equations in FP64:
//   V = Hadamard product of the Grams A_k^T A_k, k != m   (FP64)
//   A_m = M V^-1, via a Cholesky factorisation of V; when V is not positive
//         definite, LU with partial pivoting; when V is exactly singular the
//         rows keep the MTTKRP result, as LAPACK's gesv leaves its right-hand
//         side untouched in that case
//   column normalisation: 2-norm on the first iteration, then the column
//         maximum clamped to >= 1 (reference matrix.c:389-553)
// The fit is 1 - ||X - K|| / ||X|| with ||K||^2 = lambda^T (Hadamard of all
// Grams) lambda computed from the full Grams.  (The reference reads a
// triangle of the Gram product that it never writes, ktensor.c:128-143, so
// its printed fit depends on heap contents; this implementation uses the
// complete matrices.)
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "../../../include/hiparti_dev.h"
#include "dropin.hpp"

namespace pti {
namespace dev {
int run_coo(ptiDevCOO *t, float *const dmats[], ptiIndex R, ptiIndex stride, const ptiIndex mats_order[],
            ptiIndex mode, hipStream_t s, const char *M, const float *vals_override);
}
}  // namespace pti

namespace {

using pti::dev::Buffer;

// Gram G = A^T A of a row-major float matrix, accumulated in FP64.
void gram(const ptiMatrix *A, std::vector<double> &G) {
    const ptiIndex R = A->ncols, S = A->stride;
    G.assign((size_t)R * R, 0.0);
    const int nt = pti::default_threads();
    std::vector<std::vector<double>> part(nt, std::vector<double>((size_t)R * R, 0.0));
    pti::parallel_for(A->nrows, nt, [&](uint64_t b, uint64_t e, int t) {
        double *g = part[t].data();
        for (uint64_t i = b; i < e; ++i) {
            const float *row = A->values + i * S;
            for (ptiIndex p = 0; p < R; ++p) {
                const double vp = row[p];
                for (ptiIndex q = p; q < R; ++q) g[(size_t)p * R + q] += vp * row[q];
            }
        }
    });
    for (int t = 0; t < nt; ++t)
        for (size_t x = 0; x < G.size(); ++x) G[x] += part[t][x];
    for (ptiIndex p = 0; p < R; ++p)
        for (ptiIndex q = 0; q < p; ++q) G[(size_t)p * R + q] = G[(size_t)q * R + p];
}

// In-place Cholesky V = L L^T (lower, row-major).  Returns false when a pivot
// is not positive.
bool cholesky(std::vector<double> &L, ptiIndex R) {
    for (ptiIndex j = 0; j < R; ++j) {
        double d = L[(size_t)j * R + j];
        for (ptiIndex k = 0; k < j; ++k) d -= L[(size_t)j * R + k] * L[(size_t)j * R + k];
        if (!(d > 0.0)) return false;
        d = std::sqrt(d);
        L[(size_t)j * R + j] = d;
        for (ptiIndex i = j + 1; i < R; ++i) {
            double s = L[(size_t)i * R + j];
            for (ptiIndex k = 0; k < j; ++k) s -= L[(size_t)i * R + k] * L[(size_t)j * R + k];
            L[(size_t)i * R + j] = s / d;
        }
    }
    return true;
}

// LU with partial pivoting (Doolittle, row-major).  Returns false when a pivot
// is exactly zero.
bool lu(std::vector<double> &A, std::vector<ptiIndex> &piv, ptiIndex R) {
    piv.resize(R);
    for (ptiIndex k = 0; k < R; ++k) {
        ptiIndex p = k;
        for (ptiIndex i = k + 1; i < R; ++i)
            if (std::fabs(A[(size_t)i * R + k]) > std::fabs(A[(size_t)p * R + k])) p = i;
        piv[k] = p;
        if (A[(size_t)p * R + k] == 0.0) return false;
        if (p != k)
            for (ptiIndex j = 0; j < R; ++j) std::swap(A[(size_t)k * R + j], A[(size_t)p * R + j]);
        for (ptiIndex i = k + 1; i < R; ++i) {
            double f = A[(size_t)i * R + k] / A[(size_t)k * R + k];
            A[(size_t)i * R + k] = f;
            for (ptiIndex j = k + 1; j < R; ++j) A[(size_t)i * R + j] -= f * A[(size_t)k * R + j];
        }
    }
    return true;
}

// Solve x V = m for every row m of M (V symmetric), in place, FP64.
void solve_rows(ptiMatrix *M, const std::vector<double> &V, ptiIndex R) {
    std::vector<double> L(V);
    const bool spd = cholesky(L, R);
    std::vector<double> LU;
    std::vector<ptiIndex> piv;
    bool ok_lu = false;
    if (!spd) {
        LU = V;
        ok_lu = lu(LU, piv, R);
        if (!ok_lu) return;  // singular: rows keep the MTTKRP result
    }
    const ptiIndex S = M->stride;
    pti::parallel_for(M->nrows, 0, [&](uint64_t b, uint64_t e, int) {
        std::vector<double> x(R);
        for (uint64_t i = b; i < e; ++i) {
            float *row = M->values + i * S;
            for (ptiIndex r = 0; r < R; ++r) x[r] = row[r];
            // V symmetric: x V = m  <=>  V x^T = m^T.
            if (spd) {
                for (ptiIndex r = 0; r < R; ++r) {
                    double s = x[r];
                    for (ptiIndex k = 0; k < r; ++k) s -= L[(size_t)r * R + k] * x[k];
                    x[r] = s / L[(size_t)r * R + r];
                }
                for (ptiIndex r = R; r-- > 0;) {
                    double s = x[r];
                    for (ptiIndex k = r + 1; k < R; ++k) s -= L[(size_t)k * R + r] * x[k];
                    x[r] = s / L[(size_t)r * R + r];
                }
            } else {
                for (ptiIndex k = 0; k < R; ++k) std::swap(x[k], x[piv[k]]);
                for (ptiIndex r = 0; r < R; ++r)
                    for (ptiIndex k = 0; k < r; ++k) x[r] -= LU[(size_t)r * R + k] * x[k];
                for (ptiIndex r = R; r-- > 0;) {
                    double s = x[r];
                    for (ptiIndex k = r + 1; k < R; ++k) s -= LU[(size_t)r * R + k] * x[k];
                    x[r] = s / LU[(size_t)r * R + r];
                }
            }
            for (ptiIndex r = 0; r < R; ++r) row[r] = (float)x[r];
        }
    });
}

void normalize_columns(ptiMatrix *A, std::vector<double> &lambda, bool two_norm) {
    const ptiIndex R = A->ncols, S = A->stride;
    lambda.assign(R, 0.0);
    for (ptiIndex i = 0; i < A->nrows; ++i)
        for (ptiIndex r = 0; r < R; ++r) {
            double v = A->values[(size_t)i * S + r];
            if (two_norm)
                lambda[r] += v * v;
            else if (v > lambda[r])
                lambda[r] = v;
        }
    for (ptiIndex r = 0; r < R; ++r) {
        if (two_norm)
            lambda[r] = std::sqrt(lambda[r]);
        else if (lambda[r] < 1.0)
            lambda[r] = 1.0;
    }
    for (ptiIndex i = 0; i < A->nrows; ++i)
        for (ptiIndex r = 0; r < R; ++r) A->values[(size_t)i * S + r] = (float)(A->values[(size_t)i * S + r] / lambda[r]);
}

}  // namespace

extern "C" int ptiCpdAls(ptiSparseTensor const *const ptien, ptiIndex const rank, ptiIndex const niters,
                         double const tol, ptiKruskalTensor *ktensor) {
    static const char *M = "GPU  SpTns CPD-ALS";
    const ptiIndex N = ptien->nmodes;
    const ptiIndex R = rank;
    if (N < 2 || N > 8) PTI_FAIL(PTIERR_VALUE_ERROR, M, "nmodes must be in [2, 8]");
    if (R == 0) PTI_FAIL(PTIERR_VALUE_ERROR, M, "rank must be positive");
    PTI_CHECK(pti::dev::require_device(M), M);
    hipStream_t s = nullptr;

    ptiMatrix **mats = static_cast<ptiMatrix **>(std::calloc(N, sizeof(ptiMatrix *)));
    PTI_CHECK_ALLOC(mats, M);
    ptiIndex max_dim = 0;
    for (ptiIndex m = 0; m < N; ++m) {
        mats[m] = static_cast<ptiMatrix *>(std::malloc(sizeof(ptiMatrix)));
        PTI_CHECK_ALLOC(mats[m], M);
        PTI_CHECK(ptiNewMatrix(mats[m], ptien->ndims[m], R), M);
        PTI_CHECK(ptiRandomizeMatrix(mats[m]), M);
        max_dim = std::max(max_dim, ptien->ndims[m]);
    }
    const ptiIndex S = mats[0]->stride;
    ptiMatrix Mout;
    PTI_CHECK(ptiNewMatrix(&Mout, max_dim, R), M);

    ptiDevCOO *dt = nullptr;
    PTI_CHECK(ptiDevCOOUpload(&dt, ptien, s), M);
    Buffer dfac[8], dout;
    float *dptr[9];
    int rc = 0;
    for (ptiIndex m = 0; m < N && !rc; ++m) {
        size_t bytes = (size_t)ptien->ndims[m] * S * sizeof(float);
        if (dfac[m].alloc(bytes) != hipSuccess ||
            hipMemcpy(dfac[m].ptr, mats[m]->values, bytes, hipMemcpyHostToDevice) != hipSuccess)
            rc = PTIERR_CUDA_ERROR + (int)hipErrorOutOfMemory;
        dptr[m] = dfac[m].as<float>();
    }
    if (!rc && dout.alloc((size_t)max_dim * S * sizeof(float)) != hipSuccess) rc = PTIERR_CUDA_ERROR + (int)hipErrorOutOfMemory;
    if (rc) {
        ptiDevCOOFree(dt);
        ptiFreeMatrix(&Mout);
        PTI_FAIL(rc, M, "device allocation failed");
    }
    dptr[N] = dout.as<float>();

    std::vector<std::vector<double>> G(N);
    for (ptiIndex m = 0; m < N; ++m) gram(mats[m], G[m]);
    const double normX2 = SparseTensorFrobeniusNormSquared(ptien);
    std::vector<double> lambda(R, 1.0), V;
    std::vector<ptiIndex> order(N);
    double fit = 0.0, oldfit = 0.0;
    for (ptiIndex it = 0; it < niters && !rc; ++it) {
        for (ptiIndex m = 0; m < N && !rc; ++m) {
            order[0] = m;
            for (ptiIndex i = 1; i < N; ++i) order[i] = (m + i) % N;
            rc = pti::dev::run_coo(dt, dptr, R, S, order.data(), m, s, M, nullptr);
            if (rc) break;
            Mout.nrows = ptien->ndims[m];
            if (hipMemcpy(Mout.values, dout.ptr, (size_t)ptien->ndims[m] * S * sizeof(float), hipMemcpyDeviceToHost) !=
                hipSuccess) {
                rc = PTIERR_CUDA_ERROR + (int)hipErrorUnknown;
                break;
            }
            std::memcpy(mats[m]->values, Mout.values, (size_t)ptien->ndims[m] * S * sizeof(float));
            V.assign((size_t)R * R, 1.0);
            for (ptiIndex k = 0; k < N; ++k)
                if (k != m)
                    for (size_t x = 0; x < V.size(); ++x) V[x] *= G[k][x];
            solve_rows(mats[m], V, R);
            normalize_columns(mats[m], lambda, it == 0);
            gram(mats[m], G[m]);
            if (hipMemcpy(dfac[m].ptr, mats[m]->values, (size_t)ptien->ndims[m] * S * sizeof(float),
                          hipMemcpyHostToDevice) != hipSuccess)
                rc = PTIERR_CUDA_ERROR + (int)hipErrorUnknown;
        }
        if (rc) break;
        // Fit from the last MTTKRP output (mode N-1) and the full Grams.
        std::vector<double> H((size_t)R * R, 1.0);
        for (ptiIndex k = 0; k < N; ++k)
            for (size_t x = 0; x < H.size(); ++x) H[x] *= G[k][x];
        double normK = 0.0;
        for (ptiIndex p = 0; p < R; ++p)
            for (ptiIndex q = 0; q < R; ++q) normK += H[(size_t)p * R + q] * lambda[p] * lambda[q];
        normK = std::fabs(normK);
        const ptiMatrix *Al = mats[N - 1];
        double inner = 0.0;
        for (ptiIndex r = 0; r < R; ++r) {
            double acc = 0.0;
            for (ptiIndex i = 0; i < Al->nrows; ++i)
                acc += (double)Al->values[(size_t)i * S + r] * (double)Mout.values[(size_t)i * S + r];
            inner += acc * lambda[r];
        }
        double residual = normX2 + normK - 2 * inner;
        if (residual > 0.0) residual = std::sqrt(residual);
        fit = 1 - residual / std::sqrt(normX2);
        if (pti::verbose())
            std::printf("  its = %u fit = %0.5f  delta = %+0.4e\n", (unsigned)(it + 1), fit, fit - oldfit);
        if (it > 0 && std::fabs(fit - oldfit) < tol) break;
        oldfit = fit;
    }
    ptiDevCOOFree(dt);
    ptiFreeMatrix(&Mout);
    if (rc) {
        for (ptiIndex m = 0; m < N; ++m) {
            ptiFreeMatrix(mats[m]);
            std::free(mats[m]);
        }
        std::free(mats);
        PTI_FAIL(rc, M, "device MTTKRP failed");
    }
    // Final lambda: fold the 2-norms of every factor's columns into lambda.
    std::vector<double> nrm;
    for (ptiIndex m = 0; m < N; ++m) {
        normalize_columns(mats[m], nrm, true);
        for (ptiIndex r = 0; r < R; ++r) lambda[r] *= nrm[r];
    }
    for (ptiIndex r = 0; r < R; ++r) ktensor->lambda[r] = (ptiValue)lambda[r];
    ktensor->fit = fit;
    ktensor->factors = mats;
    return 0;
}
