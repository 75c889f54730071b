// FP64 dense step of CP-ALS on the device: the small-matrix work on either
// side of every MTTKRP (reference: src/matrix/matrix.c:340-572,
// src/matrix/matrix_solver.c:32-83, src/ktensor/ktensor.c:77-240).
//
// All matrices are row-major.  Factor matrices are FP32 with a row stride;
// Grams, the normal-equation matrix, its inverse, lambda and the fit are
// FP64.  Every reduction runs in a fixed order (per-workgroup partials summed
// by one workgroup), so results do not depend on scheduling.
#pragma once

#include "runtime.hpp"

namespace pti {
namespace dev {

constexpr uint32_t DENSE_MAX_RANK = 256;

// Scratch shared by the dense kernels of one caller.  Sizes depend only on R.
struct DenseWork {
    Buffer partials;  // doubles: per-workgroup partial sums
    Buffer V, W;      // R x R doubles: normal-equation matrix and its inverse
    Buffer piv;       // R ints (LU pivots)
    Buffer status;    // 1 int: 0 Cholesky, 1 LU, 2 singular (rows unchanged)
    Buffer colstat;   // R doubles: column norms / dot products
    Buffer scal;      // 4 doubles: ||X||^2, fit, ||K||^2, <X,K>
    uint32_t R = 0;
    int ensure(uint32_t R, const char *M);
};

enum ColStat { COL_2NORM = 0, COL_MAXNORM = 1, COL_DOT = 2 };

// G = A^T A (R x R, full, FP64) of an nrows x R FP32 matrix.
int dense_gram(const float *A, uint64_t nrows, uint32_t R, uint32_t stride, double *G, DenseWork &w,
               hipStream_t s, const char *M);
// V = Hadamard product of grams[k], k != mode (grams: nmodes contiguous
// R x R FP64 matrices); W = V^-1 by Cholesky, else LU with partial pivoting;
// w.status says which (2: V exactly singular, W unset).  Then
// A = M W (rows of M solved against V, FP64), or A = M when singular.
int dense_solve_normals(const double *grams, uint32_t nmodes, uint32_t mode, uint32_t R, const float *Mrhs,
                        float *A, uint64_t nrows, uint32_t stride, DenseWork &w, hipStream_t s, const char *M);
// Column statistic over rows: 2-norm, max (not abs) clamped to >= 1, or the
// dot product of the columns of A and B.  out: R doubles.  When `scale` the
// columns of A are then divided by the statistic; when `mul` is non-null,
// mul[r] *= out[r].
int dense_colstat(float *A, const float *B, uint64_t nrows, uint32_t R, uint32_t stride, ColStat kind, double *out,
                  bool scale, double *mul, DenseWork &w, hipStream_t s, const char *M);
// fit = 1 - sqrt(||X||^2 + ||K||^2 - 2 <X,K>) / ||X|| with
// ||K||^2 = sum_pq l_p l_q prod_k G_k[p][q] and <X,K> = sum_r l_r dot[r].
// normX2 and the results live in w.scal (device).
int dense_fit(const double *grams, uint32_t nmodes, uint32_t R, const double *lambda, const double *dot,
              DenseWork &w, hipStream_t s, const char *M);
// w.scal[0] = sum of squares of n FP32 values (FP64).
int dense_sumsq(const float *v, uint64_t n, DenseWork &w, hipStream_t s, const char *M);
// BF16 copy (round to nearest even) of an nrows x R FP32 matrix, same stride.
int dense_to_bf16(const float *A, uint16_t *B, uint64_t nrows, uint32_t R, uint32_t stride, hipStream_t s,
                  const char *M);

}  // namespace dev
}  // namespace pti
