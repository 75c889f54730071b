/*
 * ORACLE -- test infrastructure only (see pti_oracle.h).
 *
 * CPU restatement of the reference HiParTI MTTKRP path: the FP64 closed form,
 * the sequential and OpenMP FP32 MTTKRPs in the reference's operation order,
 * and the COO -> HiCOO conversion with the reference's quicksort, Morton keys,
 * superblock scheduler, balanced scheduler and chunk rule.  Every function
 * names the reference lines it follows.  Compiled with -ffp-contract=off so
 * the FP32 expressions round exactly as written.
 */
#include "pti_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef unsigned __int128 u128;

/* ------------------------------------------------------------------------- */
/* MTTKRP                                                                     */
/* ------------------------------------------------------------------------- */

void orc_mttkrp_f64(uint32_t N, uint64_t nnz, const uint32_t *const *inds, const float *vals,
                    uint32_t R, uint32_t stride, const float *const *facs, const uint32_t *order,
                    uint32_t mode, uint64_t nrows_out, double *out) {
    memset(out, 0, nrows_out * R * sizeof(double));
    for (uint64_t z = 0; z < nnz; ++z) {
        double *row = out + (uint64_t)inds[mode][z] * R;
        for (uint32_t r = 0; r < R; ++r) {
            double p = vals[z];
            for (uint32_t k = 1; k < N; ++k) {
                uint32_t m = order[k];
                p *= (double)facs[m][(uint64_t)inds[m][z] * stride + r];
            }
            row[r] += p;
        }
    }
}

/* The FP64 closed form above restricted to a set of output rows, for
 * full-size accuracy checks: out[s][r] = sum over the nonzeros z whose mode
 * coordinate is sel[s] of v_z prod_{k != mode} A_k[i_k][r] (product in
 * ascending mode order).  OpenMP over the nonzeros with per-thread FP64
 * accumulators, summed in thread order.  Rows listed twice get the same sum.
 * Pinned against orc_mttkrp_f64 on the selected rows (tests). */
void orc_mttkrp_f64_rows(uint32_t N, uint64_t nnz, const uint32_t *const *inds, const float *vals, uint32_t R,
                         uint32_t stride, const float *const *facs, uint32_t mode, uint64_t nrows_out,
                         uint32_t nsel, const uint32_t *sel, double *out, int nthreads) {
    int32_t *slot = (int32_t *)malloc((nrows_out ? nrows_out : 1) * sizeof(int32_t));
    for (uint64_t i = 0; i < nrows_out; ++i) slot[i] = -1;
    for (uint32_t s = 0; s < nsel; ++s)
        if (sel[s] < nrows_out && slot[sel[s]] < 0) slot[sel[s]] = (int32_t)s;
    memset(out, 0, (size_t)nsel * R * sizeof(double));
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
    nthreads = 1;
#endif
    double *acc = (double *)calloc((size_t)nthreads * nsel * R + 1, sizeof(double));
#ifdef _OPENMP
#pragma omp parallel num_threads(nthreads)
#endif
    {
#ifdef _OPENMP
        const int t = omp_get_thread_num();
#else
        const int t = 0;
#endif
        double *my = acc + (size_t)t * nsel * R;
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (uint64_t z = 0; z < nnz; ++z) {
            const int32_t s = slot[inds[mode][z]];
            if (s < 0) continue;
            double *row = my + (size_t)s * R;
            for (uint32_t r = 0; r < R; ++r) {
                double p = vals[z];
                for (uint32_t m = 0; m < N; ++m)
                    if (m != mode) p *= (double)facs[m][(uint64_t)inds[m][z] * stride + r];
                row[r] += p;
            }
        }
    }
    for (uint32_t s = 0; s < nsel; ++s) {
        const int32_t src = slot[sel[s]];
        for (int t = 0; t < nthreads; ++t)
            for (uint32_t r = 0; r < R; ++r) out[(size_t)s * R + r] += acc[((size_t)t * nsel + src) * R + r];
    }
    free(acc);
    free(slot);
}

/* CP-completion loss and gradient rows in FP64 (the loss has no reference
 * counterpart; its gradient is the reference MTTKRP, src/sptensor/mttkrp.c:
 * 39-108, of the residual tensor): with e_z = v_z - sum_r prod_m A_m[i_m][r],
 *   *loss = 0.5 * sum_z e_z^2,
 *   out[s][r] = sum over z with i_mode = sel[s] of e_z prod_{m != mode} A_m[i_m][r]
 * (= -dL/dA_mode at row sel[s]).  nsel may be 0 (loss only).  OpenMP over the
 * nonzeros, per-thread FP64 sums added in thread order.  Pinned against
 * numpy FP64 and orc_mttkrp_f64 on the FP64 residuals (tests). */
void orc_completion_f64(uint32_t N, uint64_t nnz, const uint32_t *const *inds, const float *vals, uint32_t R,
                        uint32_t stride, const float *const *facs, uint32_t mode, uint64_t nrows_out,
                        uint32_t nsel, const uint32_t *sel, double *out, double *loss, int nthreads) {
    int32_t *slot = (int32_t *)malloc((nrows_out ? nrows_out : 1) * sizeof(int32_t));
    for (uint64_t i = 0; i < nrows_out; ++i) slot[i] = -1;
    for (uint32_t s = 0; s < nsel; ++s)
        if (sel[s] < nrows_out && slot[sel[s]] < 0) slot[sel[s]] = (int32_t)s;
    if (nsel) memset(out, 0, (size_t)nsel * R * sizeof(double));
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
    nthreads = 1;
#endif
    double *acc = (double *)calloc((size_t)nthreads * nsel * R + 1, sizeof(double));
    double *lsum = (double *)calloc((size_t)nthreads, sizeof(double));
#ifdef _OPENMP
#pragma omp parallel num_threads(nthreads)
#endif
    {
#ifdef _OPENMP
        const int t = omp_get_thread_num();
#else
        const int t = 0;
#endif
        double *my = acc + (size_t)t * nsel * R;
        double l = 0.0;
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (uint64_t z = 0; z < nnz; ++z) {
            double yhat = 0.0;
            for (uint32_t r = 0; r < R; ++r) {
                double p = 1.0;
                for (uint32_t m = 0; m < N; ++m) p *= (double)facs[m][(uint64_t)inds[m][z] * stride + r];
                yhat += p;
            }
            const double e = (double)vals[z] - yhat;
            l += e * e;
            const int32_t s = nsel ? slot[inds[mode][z]] : -1;
            if (s < 0) continue;
            double *row = my + (size_t)s * R;
            for (uint32_t r = 0; r < R; ++r) {
                double p = e;
                for (uint32_t m = 0; m < N; ++m)
                    if (m != mode) p *= (double)facs[m][(uint64_t)inds[m][z] * stride + r];
                row[r] += p;
            }
        }
        lsum[t] = l;
    }
    double l = 0.0;
    for (int t = 0; t < nthreads; ++t) l += lsum[t];
    *loss = 0.5 * l;
    for (uint32_t s = 0; s < nsel; ++s) {
        const int32_t src = slot[sel[s]];
        for (int t = 0; t < nthreads; ++t)
            for (uint32_t r = 0; r < R; ++r) out[(size_t)s * R + r] += acc[((size_t)t * nsel + src) * R + r];
    }
    free(lsum);
    free(acc);
    free(slot);
}

/* src/sptensor/mttkrp.c:111-163 (3-D) and :39-108 (N-D). */
void orc_mttkrp_coo_seq(uint32_t N, uint64_t nnz, const uint32_t *const *inds, const float *vals,
                        uint32_t R, uint32_t stride, const float *const *facs, const uint32_t *order,
                        uint32_t mode, uint64_t nrows_out, float *out) {
    memset(out, 0, nrows_out * stride * sizeof(float));
    if (N == 3) {
        const float *A1 = facs[order[1]], *A2 = facs[order[2]];
        const uint32_t *i1 = inds[order[1]], *i2 = inds[order[2]], *im = inds[mode];
        for (uint64_t x = 0; x < nnz; ++x) {
            float entry = vals[x];
            for (uint32_t r = 0; r < R; ++r)
                out[(uint64_t)im[x] * stride + r] +=
                    entry * A1[(uint64_t)i1[x] * stride + r] * A2[(uint64_t)i2[x] * stride + r];
        }
        return;
    }
    float *scratch = (float *)calloc(R ? R : 1, sizeof(float));
    for (uint64_t x = 0; x < nnz; ++x) {
        uint32_t m = order[1];
        float entry = vals[x];
        for (uint32_t r = 0; r < R; ++r) scratch[r] = entry * facs[m][(uint64_t)inds[m][x] * stride + r];
        for (uint32_t i = 2; i < N; ++i) {
            m = order[i];
            for (uint32_t r = 0; r < R; ++r) scratch[r] *= facs[m][(uint64_t)inds[m][x] * stride + r];
        }
        for (uint32_t r = 0; r < R; ++r) out[(uint64_t)inds[mode][x] * stride + r] += scratch[r];
    }
    free(scratch);
}

/* src/sptensor/mttkrp_omp.c:201-252 (3-D, one atomic per element) and
 * :125-198 (N-D, per-nonzero scratch). */
void orc_omp_mttkrp_coo(uint32_t N, uint64_t nnz, const uint32_t *const *inds, const float *vals,
                        uint32_t R, uint32_t stride, const float *const *facs, const uint32_t *order,
                        uint32_t mode, uint64_t nrows_out, float *out, int nthreads) {
    memset(out, 0, nrows_out * stride * sizeof(float));
    if (N == 3) {
        const float *A1 = facs[order[1]], *A2 = facs[order[2]];
        const uint32_t *i1 = inds[order[1]], *i2 = inds[order[2]], *im = inds[mode];
#pragma omp parallel for schedule(static) num_threads(nthreads)
        for (uint64_t x = 0; x < nnz; ++x) {
            float *row = out + (uint64_t)im[x] * stride;
            const float *a = A1 + (uint64_t)i1[x] * stride, *b = A2 + (uint64_t)i2[x] * stride;
            float entry = vals[x];
            for (uint32_t r = 0; r < R; ++r) {
                float p = entry * a[r] * b[r];
#pragma omp atomic update
                row[r] += p;
            }
        }
        return;
    }
#pragma omp parallel num_threads(nthreads)
    {
        float *scratch = (float *)calloc(R ? R : 1, sizeof(float));
#pragma omp for schedule(static)
        for (uint64_t x = 0; x < nnz; ++x) {
            uint32_t m = order[1];
            float entry = vals[x];
            for (uint32_t r = 0; r < R; ++r) scratch[r] = entry * facs[m][(uint64_t)inds[m][x] * stride + r];
            for (uint32_t i = 2; i < N; ++i) {
                m = order[i];
                for (uint32_t r = 0; r < R; ++r) scratch[r] *= facs[m][(uint64_t)inds[m][x] * stride + r];
            }
            float *row = out + (uint64_t)inds[mode][x] * stride;
            for (uint32_t r = 0; r < R; ++r) {
#pragma omp atomic update
                row[r] += scratch[r];
            }
        }
        free(scratch);
    }
}

/* ------------------------------------------------------------------------- */
/* CP-ALS on the host (the CPU baseline of the device CP-ALS)                */
/* ------------------------------------------------------------------------- */

static double orc_wtime(void) {
#ifdef _OPENMP
    return omp_get_wtime();
#else
    return 0.0;
#endif
}

/* G = A^T A (R x R, FP64, both triangles) of an I x R row-major FP32 matrix:
 * the role of ssyrk_ in the reference (src/sptensor/cpd_omp.c:75,146).
 * OpenMP over rows with per-thread partial sums added in thread order. */
static void orc_gram(const float *A, uint64_t I, uint32_t R, uint32_t stride, double *G, int nthreads) {
    double *part = (double *)calloc((size_t)nthreads * R * R, sizeof(double));
#pragma omp parallel num_threads(nthreads)
    {
#ifdef _OPENMP
        double *my = part + (size_t)omp_get_thread_num() * R * R;
#else
        double *my = part;
#endif
#pragma omp for schedule(static)
        for (uint64_t i = 0; i < I; ++i) {
            const float *a = A + i * stride;
            for (uint32_t p = 0; p < R; ++p) {
                const double ap = a[p];
                for (uint32_t q = p; q < R; ++q) my[p * R + q] += ap * a[q];
            }
        }
    }
    for (uint32_t p = 0; p < R; ++p)
        for (uint32_t q = p; q < R; ++q) {
            double t = 0.0;
            for (int th = 0; th < nthreads; ++th) t += part[(size_t)th * R * R + p * R + q];
            G[p * R + q] = G[q * R + p] = t;
        }
    free(part);
}

/* Rows of M (I x R, FP32) solved against V (R x R, FP64 SPD or not):
 * A[i] = M[i] V^-1, the work of ptiMatrixSolveNormals (src/matrix/
 * matrix_solver.c:32-83: spotrf/spotrs, on failure sgesv).  Cholesky by the
 * Cholesky-Banachiewicz recurrence (row by row, L L^T = V); if a pivot is
 * not positive, LU with partial pivoting by Doolittle's recurrence; if V is
 * exactly singular the rows are left as they are (as gesv leaves its
 * right-hand side).  Returns 0 Cholesky, 1 LU, 2 singular. */
static int orc_solve_rows(const double *V, uint32_t R, float *A, uint64_t I, uint32_t stride, int nthreads) {
    double *L = (double *)calloc((size_t)R * R, sizeof(double));
    int chol = 1;
    for (uint32_t i = 0; i < R && chol; ++i)
        for (uint32_t j = 0; j <= i; ++j) {
            double sum = V[i * R + j];
            for (uint32_t k = 0; k < j; ++k) sum -= L[i * R + k] * L[j * R + k];
            if (i == j) {
                if (!(sum > 0.0)) { chol = 0; break; }
                L[i * R + i] = sqrt(sum);
            } else {
                L[i * R + j] = sum / L[j * R + j];
            }
        }
    int status = 0;
    uint32_t *piv = NULL;
    if (!chol) {
        /* Doolittle LU with partial pivoting, in place in L (unit lower, upper). */
        status = 1;
        memcpy(L, V, (size_t)R * R * sizeof(double));
        piv = (uint32_t *)malloc((R ? R : 1) * sizeof(uint32_t));
        for (uint32_t i = 0; i < R; ++i) piv[i] = i;
        for (uint32_t k = 0; k < R && status == 1; ++k) {
            uint32_t p = k;
            for (uint32_t i = k + 1; i < R; ++i)
                if (fabs(L[i * R + k]) > fabs(L[p * R + k])) p = i;
            if (L[p * R + k] == 0.0) { status = 2; break; }
            if (p != k) {
                for (uint32_t j = 0; j < R; ++j) {
                    double t = L[k * R + j]; L[k * R + j] = L[p * R + j]; L[p * R + j] = t;
                }
                uint32_t t = piv[k]; piv[k] = piv[p]; piv[p] = t;
            }
            for (uint32_t i = k + 1; i < R; ++i) {
                L[i * R + k] /= L[k * R + k];
                for (uint32_t j = k + 1; j < R; ++j) L[i * R + j] -= L[i * R + k] * L[k * R + j];
            }
        }
    }
    if (status != 2) {
        /* V symmetric: A[i]^T = V^-1 M[i]^T. */
#pragma omp parallel num_threads(nthreads)
        {
            double *x = (double *)malloc((R ? R : 1) * sizeof(double));
#pragma omp for schedule(static)
            for (uint64_t i = 0; i < I; ++i) {
                float *a = A + i * stride;
                if (status == 0) {
                    for (uint32_t r = 0; r < R; ++r) {
                        double s = a[r];
                        for (uint32_t k = 0; k < r; ++k) s -= L[r * R + k] * x[k];
                        x[r] = s / L[r * R + r];
                    }
                    for (uint32_t r = R; r-- > 0;) {
                        double s = x[r];
                        for (uint32_t k = r + 1; k < R; ++k) s -= L[k * R + r] * x[k];
                        x[r] = s / L[r * R + r];
                    }
                } else {
                    for (uint32_t r = 0; r < R; ++r) {
                        double s = a[piv[r]];
                        for (uint32_t k = 0; k < r; ++k) s -= L[r * R + k] * x[k];
                        x[r] = s;
                    }
                    for (uint32_t r = R; r-- > 0;) {
                        double s = x[r];
                        for (uint32_t k = r + 1; k < R; ++k) s -= L[r * R + k] * x[k];
                        x[r] = s / L[r * R + r];
                    }
                }
                for (uint32_t r = 0; r < R; ++r) a[r] = (float)x[r];
            }
            free(x);
        }
    }
    free(piv);
    free(L);
    return status;
}

/* Column norms of A into lambda and division of the columns by them
 * (src/matrix/matrix.c:389-469: 2-norm; :472-553: the column maximum, not
 * the absolute value, floored at 0 and clamped to >= 1). */
static void orc_normalize(float *A, uint64_t I, uint32_t R, uint32_t stride, int maxnorm, float *lambda, int nthreads) {
    double *part = (double *)calloc((size_t)nthreads * R, sizeof(double));
#pragma omp parallel num_threads(nthreads)
    {
#ifdef _OPENMP
        double *my = part + (size_t)omp_get_thread_num() * R;
#else
        double *my = part;
#endif
#pragma omp for schedule(static)
        for (uint64_t i = 0; i < I; ++i)
            for (uint32_t r = 0; r < R; ++r) {
                const double v = A[i * stride + r];
                if (maxnorm)
                    my[r] = v > my[r] ? v : my[r];
                else
                    my[r] += v * v;
            }
    }
    for (uint32_t r = 0; r < R; ++r) {
        double t = 0.0;
        for (int th = 0; th < nthreads; ++th) {
            const double v = part[(size_t)th * R + r];
            t = maxnorm ? (v > t ? v : t) : t + v;
        }
        lambda[r] = maxnorm ? (float)(t < 1.0 ? 1.0 : t) : (float)sqrt(t);
    }
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (uint64_t i = 0; i < I; ++i)
        for (uint32_t r = 0; r < R; ++r) A[i * stride + r] /= lambda[r];
    free(part);
}

/* CP-ALS of src/sptensor/cpd_omp.c:32-244 (OmpCpdAlsStep, use_reduce = 0)
 * from the caller's factors facs[m] (ndims[m] x stride FP32, updated in
 * place): per mode the OpenMP COO MTTKRP (orc_omp_mttkrp_coo, mats_order
 * m, m+1, ...), the normal-equation solve against the Hadamard product of
 * the other modes' FP64 Grams (orc_solve_rows), the column norms (2-norm on
 * the first iteration, then the clamped max-norm) and the Gram refresh; then
 * the fit of src/ktensor/ktensor.c:77-240 with the full Grams (the
 * reference reads a triangle it never writes, ktensor.c:128-143); stop after
 * niters or when the fit changes by less than tol; finally GetFinalLambda
 * (matrix.c:556-572).  fit_trace (may be NULL) gets one fit per iteration;
 * t_mttkrp / t_other (may be NULL) the seconds spent in the MTTKRPs and in
 * everything else.  Returns the last fit. */
double orc_omp_cpd_als(uint32_t N, uint64_t nnz, const uint32_t *const *inds, const float *vals,
                       const uint32_t *ndims, uint32_t R, uint32_t stride, uint32_t niters, double tol,
                       float *const *facs, float *lambda, double *fit_trace, uint32_t *iters_done, double *t_mttkrp,
                       double *t_other, int nthreads) {
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
    nthreads = 1;
#endif
    const double t_start = orc_wtime();
    double tm = 0.0;
    uint64_t max_dim = 0;
    for (uint32_t m = 0; m < N; ++m) max_dim = ndims[m] > max_dim ? ndims[m] : max_dim;
    float *tmp = (float *)calloc(max_dim * stride + 1, sizeof(float));
    double *ata = (double *)calloc((size_t)(N + 1) * R * R, sizeof(double));
    double *V = (double *)malloc((size_t)R * R * sizeof(double));
    for (uint32_t m = 0; m < N; ++m) orc_gram(facs[m], ndims[m], R, stride, ata + (size_t)m * R * R, nthreads);
    double normX2 = 0.0;
#pragma omp parallel for reduction(+ : normX2) schedule(static) num_threads(nthreads)
    for (uint64_t z = 0; z < nnz; ++z) normX2 += (double)vals[z] * vals[z];
    uint32_t order[8];
    double fit = 0.0, oldfit = 0.0;
    uint32_t it = 0;
    while (it < niters) {
        for (uint32_t m = 0; m < N; ++m) {
            order[0] = m;
            for (uint32_t i = 1; i < N; ++i) order[i] = (m + i) % N;
            const double t0 = orc_wtime();
            orc_omp_mttkrp_coo(N, nnz, inds, vals, R, stride, (const float *const *)facs, order, m, ndims[m], tmp,
                               nthreads);
            tm += orc_wtime() - t0;
            memcpy(facs[m], tmp, (size_t)ndims[m] * stride * sizeof(float));
            for (uint32_t p = 0; p < R * R; ++p) V[p] = 1.0;
            for (uint32_t k = 0; k < N; ++k)
                if (k != m)
                    for (uint32_t p = 0; p < R * R; ++p) V[p] *= ata[(size_t)k * R * R + p];
            orc_solve_rows(V, R, facs[m], ndims[m], stride, nthreads);
            orc_normalize(facs[m], ndims[m], R, stride, it > 0, lambda, nthreads);
            orc_gram(facs[m], ndims[m], R, stride, ata + (size_t)m * R * R, nthreads);
        }
        /* ||K||^2 = lambda^T (Hadamard of all Grams) lambda; <X,K> from the
         * last mode's factor and its MTTKRP output (still in tmp). */
        double normK = 0.0;
        for (uint32_t p = 0; p < R; ++p)
            for (uint32_t q = 0; q < R; ++q) {
                double h = 1.0;
                for (uint32_t k = 0; k < N; ++k) h *= ata[(size_t)k * R * R + p * R + q];
                normK += (double)lambda[p] * lambda[q] * h;
            }
        normK = fabs(normK);
        double inner = 0.0;
        const float *last = facs[N - 1];
        for (uint32_t r = 0; r < R; ++r) {
            double col = 0.0;
#pragma omp parallel for reduction(+ : col) schedule(static) num_threads(nthreads)
            for (uint64_t i = 0; i < ndims[N - 1]; ++i) col += (double)last[i * stride + r] * tmp[i * stride + r];
            inner += lambda[r] * col;
        }
        double residual = normX2 + normK - 2 * inner;
        if (residual > 0) residual = sqrt(residual);
        fit = 1 - residual / sqrt(normX2);
        if (fit_trace) fit_trace[it] = fit;
        ++it;
        if (it > 1 && fabs(fit - oldfit) < tol) break;
        oldfit = fit;
    }
    /* GetFinalLambda: every factor to unit 2-norm columns, lambda *= norms. */
    float *nrm = (float *)malloc((R ? R : 1) * sizeof(float));
    for (uint32_t m = 0; m < N; ++m) {
        orc_normalize(facs[m], ndims[m], R, stride, 0, nrm, nthreads);
        for (uint32_t r = 0; r < R; ++r) lambda[r] *= nrm[r];
    }
    free(nrm);
    free(V);
    free(ata);
    free(tmp);
    if (iters_done) *iters_done = it;
    if (t_mttkrp) *t_mttkrp = tm;
    if (t_other) *t_other = orc_wtime() - t_start - tm;
    return fit;
}

/* src/sptensor/hicoo/mttkrp.c:232-310 (3-D blocked) and :74-151 (N-D). */
void orc_mttkrp_hicoo_seq(uint32_t N, uint64_t nk, const uint64_t *kptr, const uint64_t *bptr,
                          const uint32_t *const *binds, const uint8_t *const *einds, const float *vals,
                          uint32_t sb_bits, uint32_t R, uint32_t stride, const float *const *facs,
                          const uint32_t *order, uint32_t mode, uint64_t nrows_out, float *out) {
    memset(out, 0, nrows_out * stride * sizeof(float));
    float *scratch = (float *)calloc(R ? R : 1, sizeof(float));
    uint32_t coord[16];
    for (uint64_t k = 0; k < nk; ++k) {
        for (uint64_t b = kptr[k]; b < kptr[k + 1]; ++b) {
            if (N == 3) {
                float *bm = out + (uint64_t)(binds[mode][b] << sb_bits) * stride;
                const float *b1 = facs[order[1]] + (uint64_t)(binds[order[1]][b] << sb_bits) * stride;
                const float *b2 = facs[order[2]] + (uint64_t)(binds[order[2]][b] << sb_bits) * stride;
                for (uint64_t z = bptr[b]; z < bptr[b + 1]; ++z) {
                    float *orow = bm + (uint64_t)einds[mode][z] * stride;
                    const float *r1 = b1 + (uint64_t)einds[order[1]][z] * stride;
                    const float *r2 = b2 + (uint64_t)einds[order[2]][z] * stride;
                    float entry = vals[z];
                    for (uint32_t r = 0; r < R; ++r) orow[r] += entry * r1[r] * r2[r];
                }
                continue;
            }
            for (uint64_t z = bptr[b]; z < bptr[b + 1]; ++z) {
                for (uint32_t m = 0; m < N; ++m) coord[m] = (binds[m][b] << sb_bits) + einds[m][z];
                uint32_t m = order[1];
                float entry = vals[z];
                for (uint32_t r = 0; r < R; ++r) scratch[r] = entry * facs[m][(uint64_t)coord[m] * stride + r];
                for (uint32_t i = 2; i < N; ++i) {
                    m = order[i];
                    for (uint32_t r = 0; r < R; ++r) scratch[r] *= facs[m][(uint64_t)coord[m] * stride + r];
                }
                for (uint32_t r = 0; r < R; ++r) out[(uint64_t)coord[mode] * stride + r] += scratch[r];
            }
        }
    }
    free(scratch);
}

/* src/sptensor/hicoo/mttkrp_omp.c:297-397 (N-D) and :400-467 (3-D). */
void orc_omp_mttkrp_hicoo(uint32_t N, uint64_t nk, const uint64_t *kptr, const uint64_t *bptr,
                          const uint32_t *const *binds, const uint8_t *const *einds, const float *vals,
                          uint32_t sb_bits, uint32_t R, uint32_t stride, const float *const *facs,
                          const uint32_t *order, uint32_t mode, uint64_t nrows_out, float *out, int nthreads) {
    memset(out, 0, nrows_out * stride * sizeof(float));
#pragma omp parallel for num_threads(nthreads)
    for (uint64_t k = 0; k < nk; ++k) {
        float *scratch = (float *)calloc(R ? R : 1, sizeof(float));
        uint32_t coord[16];
        for (uint64_t b = kptr[k]; b < kptr[k + 1]; ++b) {
            if (N == 3) {
                uint32_t bm = binds[mode][b], b1 = binds[order[1]][b], b2 = binds[order[2]][b];
                const float *A1 = facs[order[1]], *A2 = facs[order[2]];
                for (uint64_t z = bptr[b]; z < bptr[b + 1]; ++z) {
                    uint32_t im = (bm << sb_bits) + einds[mode][z];
                    uint32_t t1 = (b1 << sb_bits) + einds[order[1]][z];
                    uint32_t t2 = (b2 << sb_bits) + einds[order[2]][z];
                    float entry = vals[z];
                    for (uint32_t r = 0; r < R; ++r) {
                        float p = entry * A1[(uint64_t)t1 * stride + r] * A2[(uint64_t)t2 * stride + r];
#pragma omp atomic update
                        out[(uint64_t)im * stride + r] += p;
                    }
                }
                continue;
            }
            for (uint64_t z = bptr[b]; z < bptr[b + 1]; ++z) {
                for (uint32_t m = 0; m < N; ++m) coord[m] = (binds[m][b] << sb_bits) + einds[m][z];
                uint32_t m = order[1];
                float entry = vals[z];
                for (uint32_t r = 0; r < R; ++r) scratch[r] = entry * facs[m][(uint64_t)coord[m] * stride + r];
                for (uint32_t i = 2; i < N; ++i) {
                    m = order[i];
                    for (uint32_t r = 0; r < R; ++r) scratch[r] *= facs[m][(uint64_t)coord[m] * stride + r];
                }
                for (uint32_t r = 0; r < R; ++r) {
#pragma omp atomic update
                    out[(uint64_t)coord[mode] * stride + r] += scratch[r];
                }
            }
        }
        free(scratch);
    }
}

/* ------------------------------------------------------------------------- */
/* Sorting (src/sptensor/sort.c)                                              */
/* ------------------------------------------------------------------------- */

typedef struct {
    uint32_t N;
    uint32_t **inds;
    float *vals;
    uint32_t bits; /* sk_bits for the row-block compare */
} sorter;

/* sort.c:150-160 */
static void swap_nz(sorter *s, uint64_t a, uint64_t b) {
    for (uint32_t m = 0; m < s->N; ++m) {
        uint32_t t = s->inds[m][a];
        s->inds[m][a] = s->inds[m][b];
        s->inds[m][b] = t;
    }
    float v = s->vals[a];
    s->vals[a] = s->vals[b];
    s->vals[b] = v;
}

/* sort.c:929-953: lexicographic on (idx >> sk_bits). */
static int cmp_rowblock(sorter *s, uint64_t a, uint64_t b) {
    for (uint32_t m = 0; m < s->N; ++m) {
        uint32_t x = s->inds[m][a] >> s->bits, y = s->inds[m][b] >> s->bits;
        if (x < y) return -1;
        if (x > y) return 1;
    }
    return 0;
}

static int cmp_lex(sorter *s, uint64_t a, uint64_t b) {
    for (uint32_t m = 0; m < s->N; ++m) {
        uint32_t x = s->inds[m][a], y = s->inds[m][b];
        if (x < y) return -1;
        if (x > y) return 1;
    }
    return 0;
}

/* Spread the 8 bits of v to every third bit (the reference keeps these as
 * 256-entry tables, sort.c:38-147; x is shifted by 2, y by 1, z by 0). */
static uint32_t spread3(uint32_t v) {
    uint32_t r = 0;
    for (int i = 0; i < 8; ++i) r |= ((v >> i) & 1u) << (3 * i);
    return r;
}

/* sort.c:1041-1098, including its 128-bit shifts that drop bits 24-31. */
static u128 morton3(uint32_t x, uint32_t y, uint32_t z) {
    u128 k = (u128)(spread3((z >> 24) & 0xFF) | (spread3((y >> 24) & 0xFF) << 1) | (spread3((x >> 24) & 0xFF) << 2));
    k = k << 72 | (spread3((z >> 16) & 0xFF) | (spread3((y >> 16) & 0xFF) << 1) | (spread3((x >> 16) & 0xFF) << 2));
    k = k << 48 | (spread3((z >> 8) & 0xFF) | (spread3((y >> 8) & 0xFF) << 1) | (spread3((x >> 8) & 0xFF) << 2));
    k = k << 24 | (spread3(z & 0xFF) | (spread3(y & 0xFF) << 1) | (spread3(x & 0xFF) << 2));
    return k;
}

static u128 spread4_32(uint32_t v) {
    u128 r = 0;
    for (int i = 0; i < 32; ++i) r |= (u128)((v >> i) & 1u) << (4 * i);
    return r;
}

/* sort.c:1109-1288: key = w | z<<1 | y<<2 | x<<3. */
static u128 morton4(uint32_t x, uint32_t y, uint32_t z, uint32_t w) {
    return spread4_32(w) | (spread4_32(z) << 1) | (spread4_32(y) << 2) | (spread4_32(x) << 3);
}

static int cmp_morton(sorter *s, uint64_t a, uint64_t b) {
    u128 ka, kb;
    if (s->N == 3) {
        ka = morton3(s->inds[0][a], s->inds[1][a], s->inds[2][a]);
        kb = morton3(s->inds[0][b], s->inds[1][b], s->inds[2][b]);
    } else {
        ka = morton4(s->inds[0][a], s->inds[1][a], s->inds[2][a], s->inds[3][a]);
        kb = morton4(s->inds[0][b], s->inds[1][b], s->inds[2][b], s->inds[3][b]);
    }
    return ka < kb ? -1 : (ka > kb ? 1 : 0);
}

/* The reference's quicksort scheme (sort.c:1434-1465 and siblings). */
static void quicksort(sorter *s, uint64_t l, uint64_t r, int (*cmp)(sorter *, uint64_t, uint64_t)) {
    uint64_t i, j, p;
    if (r - l < 2) return;
    p = (l + r) / 2;
    for (i = l, j = r - 1;; ++i, --j) {
        while (cmp(s, i, p) < 0) ++i;
        while (cmp(s, p, j) < 0) --j;
        if (i >= j) break;
        swap_nz(s, i, j);
        if (i == p)
            p = j;
        else if (j == p)
            p = i;
    }
    quicksort(s, l, i, cmp);
    quicksort(s, i, r, cmp);
}

void orc_sort_lexicographic(uint32_t N, uint64_t nnz, uint32_t **inds, float *vals) {
    sorter s = {N, inds, vals, 0};
    quicksort(&s, 0, nnz, cmp_lex);
}

/* ------------------------------------------------------------------------- */
/* Conversion (src/sptensor/hicoo/convert.c)                                  */
/* ------------------------------------------------------------------------- */

typedef struct {
    uint64_t len, cap;
    uint64_t *d;
} v64;
typedef struct {
    uint64_t len, cap;
    uint32_t *d;
} v32;

static void p64(v64 *v, uint64_t x) {
    if (v->len == v->cap) {
        v->cap = v->cap ? v->cap * 2 : 16;
        v->d = (uint64_t *)realloc(v->d, v->cap * sizeof(uint64_t));
    }
    v->d[v->len++] = x;
}
static void p32(v32 *v, uint32_t x) {
    if (v->len == v->cap) {
        v->cap = v->cap ? v->cap * 2 : 16;
        v->d = (uint32_t *)realloc(v->d, v->cap * sizeof(uint32_t));
    }
    v->d[v->len++] = x;
}

/* Flatten an array of per-row lists into (ptr, dat). */
static void flatten(v32 *rows, uint32_t n, uint64_t **ptr, uint32_t **dat) {
    uint64_t tot = 0;
    *ptr = (uint64_t *)malloc((n + 1) * sizeof(uint64_t));
    for (uint32_t i = 0; i < n; ++i) {
        (*ptr)[i] = tot;
        tot += rows[i].len;
    }
    (*ptr)[n] = tot;
    *dat = (uint32_t *)malloc((tot ? tot : 1) * sizeof(uint32_t));
    for (uint32_t i = 0; i < n; ++i) {
        if (rows[i].len) memcpy(*dat + (*ptr)[i], rows[i].d, rows[i].len * sizeof(uint32_t));
        free(rows[i].d);
    }
}

static uint32_t idx_or_zero(uint32_t **inds, uint64_t nnz, uint32_t m, uint64_t z) {
    /* The reference reads element 0 of an empty, zero-filled vector here. */
    return z < nnz ? inds[m][z] : 0u;
}

int orc_hicoo_convert(uint32_t N, const uint32_t *ndims, uint64_t nnz, uint32_t **inds, float *vals,
                      uint32_t sb_bits, uint32_t sk_bits, uint32_t sc_bits, orc_hicoo *out) {
    if (sk_bits < sb_bits || sc_bits < sb_bits || sb_bits > 7 || N < 1 || N > 16) return 3;
    memset(out, 0, sizeof(*out));
    out->nmodes = N;
    out->nnz = nnz;
    const uint64_t sk = 1ull << sk_bits, sc = 1ull << sc_bits;
    sorter s = {N, inds, vals, sk_bits};

    /* ptiPreprocessSparseTensor, convert.c:430-604 */
    quicksort(&s, 0, nnz, cmp_rowblock);

    /* ptiSetKernelPointers, convert.c:314-369 */
    v64 kptr = {0}, knnzs = {0};
    p64(&kptr, 0);
    {
        uint64_t knnz = 0;
        uint32_t prior[16];
        for (uint32_t m = 0; m < N; ++m) prior[m] = idx_or_zero(inds, nnz, m, 0) >> sk_bits;
        for (uint64_t z = 0; z < nnz; ++z) {
            int eq = 1;
            for (uint32_t m = 0; m < N; ++m)
                if ((inds[m][z] >> sk_bits) != prior[m]) eq = 0;
            if (eq) {
                ++knnz;
            } else {
                p64(&kptr, knnz + kptr.d[kptr.len - 1]);
                p64(&knnzs, knnz);
                for (uint32_t m = 0; m < N; ++m) prior[m] = inds[m][z] >> sk_bits;
                knnz = 1;
            }
        }
        p64(&kptr, nnz);
        p64(&knnzs, knnz);
    }
    const uint64_t nk = kptr.len - 1;

    /* ptiSetKernelScheduler, convert.c:378-421 */
    out->kernel_ndim = (uint32_t *)malloc(N * sizeof(uint32_t));
    out->nkiters = (uint32_t *)malloc(N * sizeof(uint32_t));
    out->nkpars = (uint32_t *)malloc(N * sizeof(uint32_t));
    v32 **kschr = (v32 **)malloc(N * sizeof(v32 *));
    v32 **bal = (v32 **)malloc(N * sizeof(v32 *));
    v32 **pos = (v32 **)malloc(N * sizeof(v32 *));
    v32 *rest = (v32 *)calloc(N, sizeof(v32));
    for (uint32_t m = 0; m < N; ++m) {
        uint32_t kd = (uint32_t)((ndims[m] + sk - 1) / sk);
        out->kernel_ndim[m] = kd;
        kschr[m] = (v32 *)calloc(kd ? kd : 1, sizeof(v32));
        bal[m] = (v32 *)calloc(kd ? kd : 1, sizeof(v32));
        pos[m] = (v32 *)calloc(kd ? kd : 1, sizeof(v32));
    }
    for (uint64_t k = 0; k < nk; ++k) {
        uint64_t z = kptr.d[k];
        for (uint32_t m = 0; m < N; ++m) p32(&kschr[m][idx_or_zero(inds, nnz, m, z) >> sk_bits], (uint32_t)k);
    }
    for (uint32_t m = 0; m < N; ++m) {
        uint32_t t = 0;
        for (uint32_t i = 0; i < out->kernel_ndim[m]; ++i)
            if (t < kschr[m][i].len) t = (uint32_t)kschr[m][i].len;
        out->nkiters[m] = t;
    }

    /* Balanced schedule, convert.c:474-571 */
    {
        uint64_t max_nnzk = 0;
        for (uint64_t k = 0; k < nk; ++k)
            if (max_nnzk < knnzs.d[k]) max_nnzk = knnzs.d[k];
        uint64_t par_nnzk_th = 5 * max_nnzk;
        for (uint32_t m = 0; m < N; ++m) {
            uint32_t kd = out->kernel_ndim[m];
            for (uint32_t i = 0; i < kd; ++i) p32(&pos[m][i], 0);
            uint32_t npars = 0;
            int tag_rest = 0;
            uint64_t count_nk = 0;
            uint32_t empty_th = 1.0 * kd > 1 ? (uint32_t)(1.0 * kd) : 1;
            while (tag_rest == 0 && count_nk < nk) {
                uint32_t max_col = 0, par_nnzk = 0, empty = 0;
                for (uint32_t i = 0; i < kd; ++i) {
                    if (empty > empty_th) {
                        tag_rest = 1;
                        break;
                    }
                    if (npars >= pos[m][i].len) {
                        ++empty;
                        continue;
                    }
                    uint32_t j = pos[m][i].d[npars];
                    if (j >= kschr[m][i].len) {
                        ++empty;
                        continue;
                    }
                    uint64_t knz = knnzs.d[kschr[m][i].d[j]];
                    if (max_col < knz) max_col = (uint32_t)knz;
                }
                if (tag_rest == 1) {
                    for (uint32_t i = 0; i < kd; ++i) {
                        if (npars >= pos[m][i].len) continue;
                        for (uint32_t j2 = pos[m][i].d[npars]; j2 < kschr[m][i].len; ++j2) {
                            p32(&rest[m], kschr[m][i].d[j2]);
                            ++count_nk;
                        }
                    }
                } else {
                    par_nnzk = (uint32_t)(max_col > par_nnzk_th ? (uint64_t)max_col : par_nnzk_th);
                    for (uint32_t i = 0; i < kd; ++i) {
                        if (npars >= pos[m][i].len) continue;
                        uint32_t tj = pos[m][i].d[npars];
                        if (tj >= kschr[m][i].len) continue;
                        uint32_t kn = kschr[m][i].d[tj];
                        uint64_t sum = knnzs.d[kn];
                        while (sum <= par_nnzk) {
                            p32(&bal[m][i], kn);
                            ++count_nk;
                            ++tj;
                            if (tj < kschr[m][i].len) {
                                kn = kschr[m][i].d[tj];
                                sum += knnzs.d[kn];
                            } else {
                                break;
                            }
                        }
                        p32(&pos[m][i], tj);
                    }
                    ++npars;
                }
            }
            out->nkpars[m] = npars;
        }
    }

    /* Morton order inside each superblock, convert.c:583-596 (tk = 1);
     * orders other than 3 and 4 are left as they are (sort.c:421-430). */
    if (N == 3 || N == 4)
        for (uint64_t k = 0; k < nk; ++k) quicksort(&s, kptr.d[k], kptr.d[k + 1], cmp_morton);

    /* Blocks, element offsets and chunks, convert.c:697-846 */
    v64 bptr = {0}, cptr = {0};
    v32 *binds = (v32 *)calloc(N, sizeof(v32));
    out->einds = (uint8_t **)malloc(N * sizeof(uint8_t *));
    for (uint32_t m = 0; m < N; ++m) out->einds[m] = (uint8_t *)malloc(nnz ? nnz : 1);
    out->values = (float *)malloc((nnz ? nnz : 1) * sizeof(float));
    uint32_t prior[16], cur[16];
    for (uint32_t m = 0; m < N; ++m) {
        prior[m] = idx_or_zero(inds, nnz, m, 0) >> sb_bits;
        p32(&binds[m], prior[m]);
    }
    p64(&bptr, 0);
    uint64_t nb = 1, ne = 0, chunk_size = 0;
    for (uint64_t k = 0; k < nk; ++k) {
        uint64_t kb = kptr.d[k], ke = kptr.d[k + 1];
        uint64_t nb_tmp = k == 0 ? 0 : nb;
        kptr.d[k] = nb_tmp;
        p64(&cptr, nb_tmp);
        chunk_size = 0;
        for (uint64_t z = kb; z < ke; ++z) {
            int same = 1;
            for (uint32_t m = 0; m < N; ++m) {
                cur[m] = inds[m][z] >> sb_bits;
                uint32_t base = cur[m] << sb_bits;
                uint32_t e = inds[m][z] < base ? inds[m][z] : inds[m][z] - base;
                out->einds[m][z] = (uint8_t)e;
                if (cur[m] != prior[m]) same = 0;
            }
            out->values[z] = vals[z];
            if (same) {
                ++ne;
            } else {
                p64(&bptr, (uint64_t)(uint32_t)z);
                for (uint32_t m = 0; m < N; ++m) {
                    p32(&binds[m], cur[m]);
                    prior[m] = cur[m];
                }
                if (chunk_size + ne >= sc) {
                    p64(&cptr, nb);
                    chunk_size = 0;
                } else {
                    chunk_size += ne;
                }
                ++nb;
                ne = 1;
            }
        }
    }
    kptr.d[kptr.len - 1] = bptr.len;
    p64(&cptr, bptr.len);
    p64(&bptr, nnz);

    out->nb = bptr.len - 1;
    out->nk = nk;
    out->nc = cptr.len;
    out->bptr = bptr.d;
    out->kptr = kptr.d;
    out->cptr = cptr.d;
    out->knnzs = knnzs.d;
    out->binds = (uint32_t **)malloc(N * sizeof(uint32_t *));
    for (uint32_t m = 0; m < N; ++m) out->binds[m] = binds[m].d;
    free(binds);
    out->max_nnzb = out->bptr[1] - out->bptr[0];
    for (uint64_t b = 0; b < out->nb; ++b) {
        uint64_t c = out->bptr[b + 1] - out->bptr[b];
        if (out->max_nnzb < c) out->max_nnzb = c;
    }
    out->kschr_ptr = (uint64_t **)malloc(N * sizeof(uint64_t *));
    out->kschr_dat = (uint32_t **)malloc(N * sizeof(uint32_t *));
    out->bal_ptr = (uint64_t **)malloc(N * sizeof(uint64_t *));
    out->bal_dat = (uint32_t **)malloc(N * sizeof(uint32_t *));
    out->pos_ptr = (uint64_t **)malloc(N * sizeof(uint64_t *));
    out->pos_dat = (uint32_t **)malloc(N * sizeof(uint32_t *));
    out->rest_len = (uint64_t *)malloc(N * sizeof(uint64_t));
    out->rest_dat = (uint32_t **)malloc(N * sizeof(uint32_t *));
    for (uint32_t m = 0; m < N; ++m) {
        flatten(kschr[m], out->kernel_ndim[m], &out->kschr_ptr[m], &out->kschr_dat[m]);
        flatten(bal[m], out->kernel_ndim[m], &out->bal_ptr[m], &out->bal_dat[m]);
        flatten(pos[m], out->kernel_ndim[m], &out->pos_ptr[m], &out->pos_dat[m]);
        out->rest_len[m] = rest[m].len;
        out->rest_dat[m] = rest[m].d ? rest[m].d : (uint32_t *)malloc(4);
        free(kschr[m]);
        free(bal[m]);
        free(pos[m]);
    }
    free(kschr);
    free(bal);
    free(pos);
    free(rest);
    return 0;
}

void orc_hicoo_free(orc_hicoo *h) {
    for (uint32_t m = 0; m < h->nmodes; ++m) {
        free(h->binds[m]);
        free(h->einds[m]);
        free(h->kschr_ptr[m]);
        free(h->kschr_dat[m]);
        free(h->bal_ptr[m]);
        free(h->bal_dat[m]);
        free(h->pos_ptr[m]);
        free(h->pos_dat[m]);
        free(h->rest_dat[m]);
    }
    free(h->binds);
    free(h->einds);
    free(h->kschr_ptr);
    free(h->kschr_dat);
    free(h->bal_ptr);
    free(h->bal_dat);
    free(h->pos_ptr);
    free(h->pos_dat);
    free(h->rest_dat);
    free(h->rest_len);
    free(h->bptr);
    free(h->kptr);
    free(h->cptr);
    free(h->knnzs);
    free(h->values);
    free(h->nkiters);
    free(h->nkpars);
    free(h->kernel_ndim);
    memset(h, 0, sizeof(*h));
}
