// CP-ALS with the whole iteration on the device (reference: src/sptensor/cpd.c:32-208,
// src/sptensor/hicoo/cpd.c:31-204, src/matrix/matrix_solver.c:32-83,
// src/matrix/matrix.c:340-572, src/ktensor/ktensor.c:77-240).
//
// The tensor, the factor matrices, their FP64 Grams and lambda stay in HBM.
// Per mode m of an iteration:
//   M   = MTTKRP(X, m)                                   (COO or HiCOO kernel)
//   V   = Hadamard product of the Grams A_k^T A_k, k != m (FP64)
//   A_m = M V^-1  via Cholesky, else LU with partial pivoting; when V is
//         exactly singular A_m = M (LAPACK's gesv leaves the right-hand side
//         untouched in that case)
//   A_m /= lambda, lambda = column 2-norms on the first iteration, then the
//         column maxima clamped to >= 1                  (matrix.c:389-553)
//   G_m = A_m^T A_m
// After the modes: fit = 1 - ||X - K|| / ||X|| with ||K||^2 = l^T (Hadamard
// of all Grams) l from the full Grams and <X,K> from the last MTTKRP output
// (ktensor.c:77-240); one scalar is copied to the host for the tolerance
// test.  The reference reads a triangle of the Gram product that it never
// writes (ktensor.c:128-143), so its printed fit depends on heap contents;
// this implementation uses the complete matrices.  At the end every factor's
// columns are normalised to unit 2-norm and lambda absorbs the norms
// (GetFinalLambda, matrix.c:556-572).
//
// Optional precision schedule (HiCOO only): during the first bf16_iters
// iterations the MTTKRPs gather BF16 copies of the factors (refreshed after
// each update) with FP32 accumulation; the FP32 masters, the solve and the
// fit are unchanged.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <functional>

#include "../../../include/hiparti_dev.h"
#include "dense.hpp"
#include "dropin.hpp"

namespace pti {
namespace dev {
int run_coo(ptiDevCOO *t, float *const dmats[], ptiIndex R, ptiIndex stride, const ptiIndex mats_order[],
            ptiIndex mode, hipStream_t s, const char *M, const float *vals_override);
int run_hicoo(const ptiDevHiCOO *t, const void *const facs[], float *out, ptiIndex R, ptiIndex stride,
              ptiIndex out_stride, const ptiIndex mats_order[], ptiIndex mode, bool bf16, hipStream_t s,
              const char *M);
void hicoo_view(const ptiDevHiCOO *t, uint32_t *nmodes, const ptiIndex **ndims, uint64_t *nnz, const float **vals);
void coo_view(const ptiDevCOO *t, uint32_t *nmodes, const ptiIndex **ndims, uint64_t *nnz, const float **vals);

namespace {

// MTTKRP of mode `mode` into `out` (ndims[mode] x stride); `bf` selects the
// BF16 factor copies.
using MttkrpFn =
    std::function<int(uint32_t mode, bool bf, float *const facs[], uint16_t *const bfacs[], float *out)>;

struct CpdArgs {
    uint32_t N = 0, R = 0, stride = 0;
    const ptiIndex *ndims = nullptr;
    float *const *facs = nullptr;
    uint32_t niters = 0;
    double tol = 0.0;
    uint32_t bf16_iters = 0;
    double *dlambda = nullptr;
    double *fit = nullptr;
    double *fit_trace = nullptr;
    ptiIndex *iters_done = nullptr;
    const float *vals = nullptr;
    uint64_t nnz = 0;
    const char *name = "";
};

int cpd_core(const CpdArgs &a, const MttkrpFn &mttkrp, hipStream_t s) {
    const char *M = a.name;
    const uint32_t N = a.N, R = a.R, S = a.stride;
    if (N < 2 || N > 8) PTI_FAIL(PTIERR_VALUE_ERROR, M, "nmodes must be in [2, 8]");
    if (R == 0 || R > DENSE_MAX_RANK) PTI_FAIL(PTIERR_VALUE_ERROR, M, "rank must be in [1, 256]");
    if (S < R) PTI_FAIL(PTIERR_SHAPE_MISMATCH, M, "stride < rank");
    if (!a.vals && a.nnz) PTI_FAIL(PTIERR_VALUE_ERROR, M, "tensor values are not resident");
    DenseWork w;
    PTI_CHECK(w.ensure(R, M), M);
    const size_t RR = (size_t)R * R;
    uint64_t max_dim = 0;
    for (uint32_t m = 0; m < N; ++m) max_dim = std::max<uint64_t>(max_dim, a.ndims[m]);
    Buffer grams, mout, lam_own;
    Buffer bfb[8];
    uint16_t *bf[8] = {nullptr};
    PTI_HIP(grams.alloc(N * RR * sizeof(double)), M);
    PTI_HIP(mout.alloc(max_dim * S * sizeof(float)), M);
    double *lambda = a.dlambda;
    if (!lambda) {
        PTI_HIP(lam_own.alloc(R * sizeof(double)), M);
        lambda = lam_own.as<double>();
    }
    if (a.bf16_iters)
        for (uint32_t m = 0; m < N; ++m) {
            PTI_HIP(bfb[m].alloc((size_t)a.ndims[m] * S * sizeof(uint16_t)), M);
            bf[m] = bfb[m].as<uint16_t>();
        }
    double *G = grams.as<double>();
    float *out = mout.as<float>();
    double *colstat = w.colstat.as<double>();

    PTI_CHECK(dense_sumsq(a.vals, a.nnz, w, s, M), M);  // w.scal[0] = ||X||^2
    for (uint32_t m = 0; m < N; ++m) {
        PTI_CHECK(dense_gram(a.facs[m], a.ndims[m], R, S, G + m * RR, w, s, M), M);
        if (a.bf16_iters) PTI_CHECK(dense_to_bf16(a.facs[m], bf[m], a.ndims[m], R, S, s, M), M);
    }
    double fit = 0.0, oldfit = 0.0;
    uint32_t it = 0;
    while (it < a.niters) {
        const auto t0 = std::chrono::steady_clock::now();
        const bool use_bf = it < a.bf16_iters;
        for (uint32_t m = 0; m < N; ++m) {
            PTI_CHECK(mttkrp(m, use_bf, a.facs, bf, out), M);
            PTI_CHECK(dense_solve_normals(G, N, m, R, out, a.facs[m], a.ndims[m], S, w, s, M), M);
            PTI_CHECK(dense_colstat(a.facs[m], nullptr, a.ndims[m], R, S, it == 0 ? COL_2NORM : COL_MAXNORM, lambda,
                                    true, nullptr, w, s, M),
                      M);
            PTI_CHECK(dense_gram(a.facs[m], a.ndims[m], R, S, G + m * RR, w, s, M), M);
            if (use_bf) PTI_CHECK(dense_to_bf16(a.facs[m], bf[m], a.ndims[m], R, S, s, M), M);
        }
        // <X,K> uses the last mode's factor and its MTTKRP output.
        PTI_CHECK(dense_colstat(a.facs[N - 1], out, a.ndims[N - 1], R, S, COL_DOT, colstat, false, nullptr, w, s, M),
                  M);
        PTI_CHECK(dense_fit(G, N, R, lambda, colstat, w, s, M), M);
        PTI_HIP(hipMemcpyAsync(&fit, w.scal.as<double>() + 1, sizeof(double), hipMemcpyDeviceToHost, s), M);
        PTI_HIP(hipStreamSynchronize(s), M);
        if (a.fit_trace) a.fit_trace[it] = fit;
        ++it;
        if (verbose()) {
            const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            std::printf("  its = %u ( %.3lf s ) fit = %0.5f  delta = %+0.4e\n", it, sec, fit, fit - oldfit);
        }
        if (it > 1 && std::fabs(fit - oldfit) < a.tol) break;
        oldfit = fit;
    }
    // Final lambda: unit 2-norm columns, lambda *= norms.
    for (uint32_t m = 0; m < N; ++m)
        PTI_CHECK(dense_colstat(a.facs[m], nullptr, a.ndims[m], R, S, COL_2NORM, colstat, true, lambda, w, s, M), M);
    PTI_HIP(hipStreamSynchronize(s), M);
    if (a.fit) *a.fit = fit;
    if (a.iters_done) *a.iters_done = it;
    return 0;
}

void default_order(uint32_t N, uint32_t m, ptiIndex *order) {
    order[0] = m;
    for (uint32_t i = 1; i < N; ++i) order[i] = (m + i) % N;
}

CpdArgs make_args(uint32_t N, const ptiIndex *ndims, uint64_t nnz, const float *vals, float *const dfacs[],
                  ptiIndex R, ptiIndex stride, ptiIndex niters, double tol, ptiIndex bf16_iters, double *dlambda,
                  double *fit, double *fit_trace, ptiIndex *iters_done, const char *M) {
    CpdArgs a;
    a.N = N;
    a.ndims = ndims;
    a.nnz = nnz;
    a.vals = vals;
    a.R = R;
    a.stride = stride;
    a.facs = dfacs;
    a.niters = niters;
    a.tol = tol;
    a.bf16_iters = bf16_iters;
    a.dlambda = dlambda;
    a.fit = fit;
    a.fit_trace = fit_trace;
    a.iters_done = iters_done;
    a.name = M;
    return a;
}

int cpd_hicoo(const ptiDevHiCOO *t, float *const dfacs[], ptiIndex R, ptiIndex stride, ptiIndex niters, double tol,
              ptiIndex bf16_iters, double *dlambda, double *fit, double *fit_trace, ptiIndex *iters_done,
              hipStream_t s, const char *M) {
    uint32_t N;
    const ptiIndex *ndims;
    uint64_t nnz;
    const float *vals;
    hicoo_view(t, &N, &ndims, &nnz, &vals);
    const CpdArgs a = make_args(N, ndims, nnz, vals, dfacs, R, stride, niters, tol, bf16_iters, dlambda, fit,
                                fit_trace, iters_done, M);
    auto fn = [&](uint32_t m, bool bf, float *const facs[], uint16_t *const bfacs[], float *out) -> int {
        ptiIndex order[8];
        default_order(N, m, order);
        const void *f[8];
        for (uint32_t k = 0; k < N; ++k) f[k] = bf ? (const void *)bfacs[k] : (const void *)facs[k];
        return run_hicoo(t, f, out, R, stride, stride, order, m, bf, s, M);
    };
    return cpd_core(a, fn, s);
}

int cpd_coo(ptiDevCOO *t, float *const dfacs[], ptiIndex R, ptiIndex stride, ptiIndex niters, double tol,
            double *dlambda, double *fit, double *fit_trace, ptiIndex *iters_done, hipStream_t s, const char *M) {
    uint32_t N;
    const ptiIndex *ndims;
    uint64_t nnz;
    const float *vals;
    coo_view(t, &N, &ndims, &nnz, &vals);
    const CpdArgs a = make_args(N, ndims, nnz, vals, dfacs, R, stride, niters, tol, 0, dlambda, fit, fit_trace,
                                iters_done, M);
    auto fn = [&](uint32_t m, bool, float *const facs[], uint16_t *const *, float *out) -> int {
        ptiIndex order[8];
        default_order(N, m, order);
        float *mats[9];
        for (uint32_t k = 0; k < N; ++k) mats[k] = facs[k];
        mats[N] = out;
        return run_coo(t, mats, R, stride, order, m, s, M, nullptr);
    };
    return cpd_core(a, fn, s);
}

// Device copies of host factor matrices for the drop-in CP-ALS entry points.
struct DevFactors {
    Buffer dev[8];
    float *ptrs[8] = {nullptr};
    Buffer dlambda;
};

int upload_factors(uint32_t N, const ptiIndex *ndims, uint32_t R, uint32_t S, float *const host[], DevFactors &df,
                   const char *M) {
    for (uint32_t m = 0; m < N; ++m) {
        const size_t bytes = (size_t)ndims[m] * S * sizeof(float);
        PTI_HIP(df.dev[m].alloc(bytes), M);
        if (bytes) PTI_HIP(hipMemcpy(df.dev[m].ptr, host[m], bytes, hipMemcpyHostToDevice), M);
        df.ptrs[m] = df.dev[m].as<float>();
    }
    PTI_HIP(df.dlambda.alloc(R * sizeof(double)), M);
    return 0;
}

int download_factors(uint32_t N, const ptiIndex *ndims, uint32_t R, uint32_t S, float *const host[], DevFactors &df,
                     ptiValue *lambda, const char *M) {
    for (uint32_t m = 0; m < N; ++m) {
        const size_t bytes = (size_t)ndims[m] * S * sizeof(float);
        if (bytes) PTI_HIP(hipMemcpy(host[m], df.dev[m].ptr, bytes, hipMemcpyDeviceToHost), M);
    }
    std::vector<double> lam(R);
    PTI_HIP(hipMemcpy(lam.data(), df.dlambda.ptr, R * sizeof(double), hipMemcpyDeviceToHost), M);
    for (uint32_t r = 0; r < R; ++r) lambda[r] = (ptiValue)lam[r];
    return 0;
}

inline int hadamard_host(ptiIndex mode, ptiIndex nmodes, ptiMatrix **aTa) {
    return ptiMatrixDotMulSeqTriangle(mode, nmodes, aTa);
}
inline int hadamard_host(ptiIndex mode, ptiIndex nmodes, ptiRankMatrix **aTa) {
    return ptiRankMatrixDotMulSeqTriangle(mode, nmodes, aTa);
}

// Host-matrix normal-equation solve through the device kernels: the Grams'
// upper triangles become full FP64 matrices, the rows of rhs are solved on
// the GPU and copied back in place.
template <typename Mat>
int solve_normals_host(ptiIndex mode, ptiIndex nmodes, Mat **aTa, Mat *rhs, const char *M) {
    if (nmodes < 1 || nmodes > 8 || mode >= nmodes) PTI_FAIL(PTIERR_VALUE_ERROR, M, "bad nmodes or mode");
    const uint32_t R = aTa[0]->ncols, S = aTa[0]->stride, RS = rhs->stride;
    if (R == 0 || R > DENSE_MAX_RANK) PTI_FAIL(PTIERR_VALUE_ERROR, M, "rank must be in [1, 256]");
    if (rhs->ncols != R || RS < R) PTI_FAIL(PTIERR_SHAPE_MISMATCH, M, "rhs->ncols != rank");
    PTI_CHECK(hadamard_host(mode, nmodes, aTa), M);
    PTI_CHECK(require_device(M), M);
    const size_t RR = (size_t)R * R;
    std::vector<double> g(nmodes * RR);
    for (uint32_t k = 0; k < nmodes; ++k)
        for (uint32_t i = 0; i < R; ++i)
            for (uint32_t j = i; j < R; ++j) {
                const double v = aTa[k]->values[(size_t)i * S + j];
                g[k * RR + (size_t)i * R + j] = v;
                g[k * RR + (size_t)j * R + i] = v;
            }
    const uint64_t nrows = rhs->nrows;
    const size_t bytes = (size_t)nrows * RS * sizeof(float);
    Buffer dg, dm, da;
    DenseWork w;
    hipStream_t s = nullptr;
    PTI_HIP(dg.alloc(g.size() * sizeof(double)), M);
    PTI_HIP(dm.alloc(bytes), M);
    PTI_HIP(da.alloc(bytes), M);
    PTI_HIP(hipMemcpy(dg.ptr, g.data(), g.size() * sizeof(double), hipMemcpyHostToDevice), M);
    if (bytes) PTI_HIP(hipMemcpy(dm.ptr, rhs->values, bytes, hipMemcpyHostToDevice), M);
    PTI_CHECK(dense_solve_normals(dg.as<double>(), nmodes, mode, R, dm.as<float>(), da.as<float>(), nrows, RS, w, s,
                                  M),
              M);
    int status = 0;
    PTI_HIP(hipMemcpy(&status, w.status.ptr, sizeof(int), hipMemcpyDeviceToHost), M);
    if (status == 1 && verbose()) std::printf("Gram matrix is not SPD. Trying `gesv`.\n");
    if (bytes) PTI_HIP(hipMemcpy(rhs->values, da.ptr, bytes, hipMemcpyDeviceToHost), M);
    return 0;
}

}  // namespace
}  // namespace dev
}  // namespace pti

using namespace pti::dev;

extern "C" {

int ptiDevCpdAlsHiCOO(ptiDevHiCOO const *t, float *const dfacs[], ptiIndex R, ptiIndex stride, ptiIndex niters,
                      double tol, ptiIndex bf16_iters, double *dlambda, double *fit, double *fit_trace,
                      ptiIndex *iters_done, void *stream) {
    return cpd_hicoo(t, dfacs, R, stride, niters, tol, bf16_iters, dlambda, fit, fit_trace, iters_done,
                     static_cast<hipStream_t>(stream), "DevHiCOO CPD-ALS");
}

int ptiDevCpdAlsCOO(ptiDevCOO *t, float *const dfacs[], ptiIndex R, ptiIndex stride, ptiIndex niters, double tol,
                    double *dlambda, double *fit, double *fit_trace, ptiIndex *iters_done, void *stream) {
    return cpd_coo(t, dfacs, R, stride, niters, tol, dlambda, fit, fit_trace, iters_done,
                   static_cast<hipStream_t>(stream), "DevCOO CPD-ALS");
}

int ptiDevGram(float const *dA, ptiIndex nrows, ptiIndex R, ptiIndex stride, double *dG, void *stream) {
    static const char *M = "Dev Gram";
    PTI_CHECK(require_device(M), M);
    DenseWork w;
    hipStream_t s = static_cast<hipStream_t>(stream);
    PTI_CHECK(dense_gram(dA, nrows, R, stride, dG, w, s, M), M);
    PTI_HIP(hipStreamSynchronize(s), M);  // the workspace is released on return
    return 0;
}

int ptiDevSolveNormals(double const *dgrams, ptiIndex nmodes, ptiIndex mode, ptiIndex R, float const *dM, float *dA,
                       ptiIndex nrows, ptiIndex stride, int *status, void *stream) {
    static const char *M = "Dev SolveNormals";
    PTI_CHECK(require_device(M), M);
    if (nmodes < 1 || nmodes > 8 || mode >= nmodes) PTI_FAIL(PTIERR_VALUE_ERROR, M, "bad nmodes or mode");
    DenseWork w;
    hipStream_t s = static_cast<hipStream_t>(stream);
    PTI_CHECK(dense_solve_normals(dgrams, nmodes, mode, R, dM, dA, nrows, stride, w, s, M), M);
    int st = 0;
    PTI_HIP(hipMemcpyAsync(&st, w.status.ptr, sizeof(int), hipMemcpyDeviceToHost, s), M);
    PTI_HIP(hipStreamSynchronize(s), M);
    if (status) *status = st;
    return 0;
}

int ptiDevNormalizeColumns(float *dA, ptiIndex nrows, ptiIndex R, ptiIndex stride, int max_norm, double *dlambda,
                           void *stream) {
    static const char *M = "Dev NormalizeColumns";
    PTI_CHECK(require_device(M), M);
    DenseWork w;
    hipStream_t s = static_cast<hipStream_t>(stream);
    PTI_CHECK(dense_colstat(dA, nullptr, nrows, R, stride, max_norm ? COL_MAXNORM : COL_2NORM, dlambda, true,
                            nullptr, w, s, M),
              M);
    PTI_HIP(hipStreamSynchronize(s), M);
    return 0;
}

int ptiDevKruskalFit(double const *dgrams, ptiIndex nmodes, ptiIndex R, double const *dlambda, float const *dlast,
                     float const *dM, ptiIndex nrows_last, ptiIndex stride, double normX2, double *fit,
                     void *stream) {
    static const char *M = "Dev KruskalFit";
    PTI_CHECK(require_device(M), M);
    if (nmodes < 1 || nmodes > 8) PTI_FAIL(PTIERR_VALUE_ERROR, M, "bad nmodes");
    DenseWork w;
    hipStream_t s = static_cast<hipStream_t>(stream);
    PTI_CHECK(w.ensure(R, M), M);
    PTI_HIP(hipMemcpyAsync(w.scal.ptr, &normX2, sizeof(double), hipMemcpyHostToDevice, s), M);
    PTI_CHECK(dense_colstat(const_cast<float *>(dlast), dM, nrows_last, R, stride, COL_DOT, w.colstat.as<double>(),
                            false, nullptr, w, s, M),
              M);
    PTI_CHECK(dense_fit(dgrams, nmodes, R, dlambda, w.colstat.as<double>(), w, s, M), M);
    double f = 0.0;
    PTI_HIP(hipMemcpyAsync(&f, w.scal.as<double>() + 1, sizeof(double), hipMemcpyDeviceToHost, s), M);
    PTI_HIP(hipStreamSynchronize(s), M);
    if (fit) *fit = f;
    return 0;
}

int ptiMatrixSolveNormals(ptiIndex const mode, ptiIndex const nmodes, ptiMatrix **aTa, ptiMatrix *rhs) {
    return solve_normals_host(mode, nmodes, aTa, rhs, "Mtx SolveNormals");
}

int ptiRankMatrixSolveNormals(ptiIndex const mode, ptiIndex const nmodes, ptiRankMatrix **aTa, ptiRankMatrix *rhs) {
    return solve_normals_host(mode, nmodes, aTa, rhs, "RankMtx SolveNormals");
}

int ptiCpdAls(ptiSparseTensor const *const ptien, ptiIndex const rank, ptiIndex const niters, double const tol,
              ptiKruskalTensor *ktensor) {
    static const char *M = "GPU  SpTns CPD-ALS";
    const ptiIndex N = ptien->nmodes;
    const ptiIndex R = rank;
    if (N < 2 || N > 8) PTI_FAIL(PTIERR_VALUE_ERROR, M, "nmodes must be in [2, 8]");
    if (R == 0 || R > DENSE_MAX_RANK) PTI_FAIL(PTIERR_VALUE_ERROR, M, "rank must be in [1, 256]");
    PTI_CHECK(require_device(M), M);
    ptiMatrix **mats = static_cast<ptiMatrix **>(std::calloc(N + 1, sizeof(ptiMatrix *)));
    PTI_CHECK_ALLOC(mats, M);
    auto free_mats = [&]() {
        for (ptiIndex m = 0; m < N; ++m)
            if (mats[m]) {
                ptiFreeMatrix(mats[m]);
                std::free(mats[m]);
            }
        std::free(mats);
    };
    int rc = 0;
    for (ptiIndex m = 0; m < N && !rc; ++m) {
        mats[m] = static_cast<ptiMatrix *>(std::calloc(1, sizeof(ptiMatrix)));
        if (!mats[m]) rc = PTIERR_OS_ERROR + ENOMEM;
        if (!rc) rc = ptiNewMatrix(mats[m], ptien->ndims[m], R);
        if (!rc) rc = ptiRandomizeMatrix(mats[m]);
    }
    if (rc) {
        free_mats();
        PTI_FAIL(rc, M, "factor allocation failed");
    }
    const ptiIndex S = mats[0]->stride;
    float *host[8];
    for (ptiIndex m = 0; m < N; ++m) host[m] = mats[m]->values;
    ptiDevCOO *dt = nullptr;
    DevFactors df;
    double fit = 0.0;
    rc = ptiDevCOOUpload(&dt, ptien, nullptr);
    if (!rc) rc = upload_factors(N, ptien->ndims, R, S, host, df, M);
    if (!rc) rc = cpd_coo(dt, df.ptrs, R, S, niters, tol, df.dlambda.as<double>(), &fit, nullptr, nullptr, nullptr, M);
    if (!rc) rc = download_factors(N, ptien->ndims, R, S, host, df, ktensor->lambda, M);
    if (dt) ptiDevCOOFree(dt);
    if (rc) {
        free_mats();
        return rc;
    }
    ktensor->fit = fit;
    ktensor->factors = mats;
    return 0;
}

int ptiOmpCpdAls(ptiSparseTensor const *const ptien, ptiIndex const rank, ptiIndex const niters, double const tol,
                 const int tk, const int use_reduce, ptiKruskalTensor *ktensor) {
    (void)tk;
    (void)use_reduce;
    return ptiCpdAls(ptien, rank, niters, tol, ktensor);
}

int ptiCudaCpdAls(ptiSparseTensor const *const ptien, ptiIndex const rank, ptiIndex const niters, double const tol,
                  ptiKruskalTensor *ktensor) {
    return ptiCpdAls(ptien, rank, niters, tol, ktensor);
}

int ptiCpdAlsHiCOO(ptiSparseTensorHiCOO const *const hitsr, ptiIndex const rank, ptiIndex const niters,
                   double const tol, ptiRankKruskalTensor *ktensor) {
    static const char *M = "GPU  HiCOO SpTns CPD-ALS";
    const ptiIndex N = hitsr->nmodes;
    const ptiIndex R = rank;
    if (N < 2 || N > 8) PTI_FAIL(PTIERR_VALUE_ERROR, M, "nmodes must be in [2, 8]");
    // ptiRankMatrix keeps ncols and stride (ncols rounded up to 8) in 8 bits.
    if (R == 0 || R > 248) PTI_FAIL(PTIERR_VALUE_ERROR, M, "rank must be in [1, 248] for ptiRankMatrix factors");
    PTI_CHECK(require_device(M), M);
    ptiRankMatrix **mats = static_cast<ptiRankMatrix **>(std::calloc(N + 1, sizeof(ptiRankMatrix *)));
    PTI_CHECK_ALLOC(mats, M);
    auto free_mats = [&]() {
        for (ptiIndex m = 0; m < N; ++m)
            if (mats[m]) {
                ptiFreeRankMatrix(mats[m]);
                std::free(mats[m]);
            }
        std::free(mats);
    };
    int rc = 0;
    for (ptiIndex m = 0; m < N && !rc; ++m) {
        mats[m] = static_cast<ptiRankMatrix *>(std::calloc(1, sizeof(ptiRankMatrix)));
        if (!mats[m]) rc = PTIERR_OS_ERROR + ENOMEM;
        if (!rc) rc = ptiNewRankMatrix(mats[m], hitsr->ndims[m], (ptiElementIndex)R);
        if (!rc) rc = ptiRandomizeRankMatrix(mats[m], hitsr->ndims[m], (ptiElementIndex)R);
    }
    if (rc) {
        free_mats();
        PTI_FAIL(rc, M, "factor allocation failed");
    }
    const ptiIndex S = mats[0]->stride;
    float *host[8];
    for (ptiIndex m = 0; m < N; ++m) host[m] = mats[m]->values;
    ptiDevHiCOO *dt = nullptr;
    DevFactors df;
    double fit = 0.0;
    rc = ptiDevHiCOOUpload(&dt, hitsr, nullptr);
    if (!rc) rc = upload_factors(N, hitsr->ndims, R, S, host, df, M);
    if (!rc)
        rc = cpd_hicoo(dt, df.ptrs, R, S, niters, tol, 0, df.dlambda.as<double>(), &fit, nullptr, nullptr, nullptr, M);
    if (!rc) rc = download_factors(N, hitsr->ndims, R, S, host, df, ktensor->lambda, M);
    if (dt) ptiDevHiCOOFree(dt);
    if (rc) {
        free_mats();
        return rc;
    }
    ktensor->fit = fit;
    ktensor->factors = mats;
    return 0;
}

int ptiOmpCpdAlsHiCOO(ptiSparseTensorHiCOO const *const hitsr, ptiIndex const rank, ptiIndex const niters,
                      double const tol, const int tk, const int tb, int balanced, ptiRankKruskalTensor *ktensor) {
    (void)tk;
    (void)tb;
    (void)balanced;
    return ptiCpdAlsHiCOO(hitsr, rank, niters, tol, ktensor);
}

}  // extern "C"
