/*
 * Device-resident extension of the HiParTI C ABI (libHiParTI.so).
 *
 * None of these entry points has a counterpart in the reference: the
 * reference copies the whole tensor and every factor matrix to the device on
 * each MTTKRP call (src/sptensor/hicoo/mttkrp_cuda.cu:108-235,
 * src/sptensor/mttkrp_cuda.cu:96-150).  They exist for callers that keep the
 * tensor and the factors in HBM across calls (CPD-ALS, the benchmark, the
 * Python/torch layer), and for the drop-in pair ptiMTTKRP / ptiMTTKRPHiCOO,
 * which use a registered device copy instead of re-uploading the tensor.
 *
 * Conventions:
 *   - Device pointers are plain addresses in the current HIP device's memory.
 *   - `stream` is a hipStream_t passed as void*; NULL means the null stream.
 *     Calls only enqueue work; they do not synchronise unless stated.
 *   - Factor matrices are row-major with row stride `stride` (>= R) elements,
 *     the ptiMatrix layout.  Output rows are zeroed by the call.
 *   - Return values are PTIERR_* codes; device failures return
 *     PTIERR_CUDA_ERROR + hipError_t.
 */
#ifndef HIPARTI_DEV_H
#define HIPARTI_DEV_H

#include "HiParTI.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ptiDevHiCOO ptiDevHiCOO;  /* HiCOO tensor resident in HBM */
typedef struct ptiDevCOO   ptiDevCOO;    /* COO tensor resident in HBM   */

/* Number of visible HIP devices, or a negative PTIERR on failure. */
int ptiDevCount(void);

/* Upload a HiCOO tensor into the device layout used by the kernels.  The host
 * tensor is not referenced afterwards. */
int  ptiDevHiCOOUpload(ptiDevHiCOO **out, ptiSparseTensorHiCOO const *hitsr, void *stream);
void ptiDevHiCOOFree(ptiDevHiCOO *t);
/* Sizes of the resident tensor: nnz, number of blocks, nmodes, device bytes. */
int  ptiDevHiCOOInfo(ptiDevHiCOO const *t, uint64_t *nnz, uint64_t *nblocks,
                     uint32_t *nmodes, uint64_t *bytes);

/* Upload a COO tensor.  Per-mode copies sorted by the output index are built
 * on the device the first time a mode is used (or by ptiDevCOOPrepareMode). */
int  ptiDevCOOUpload(ptiDevCOO **out, ptiSparseTensor const *X, void *stream);
int  ptiDevCOOPrepareMode(ptiDevCOO *t, ptiIndex mode, void *stream);
void ptiDevCOOFree(ptiDevCOO *t);

/* MTTKRP on resident data, FP32 factors.  dmats[0..nmodes-1] are the factor
 * matrices (device), dmats[nmodes] the output with at least ndims[mode] rows.
 * Same mathematics and mats_order convention as ptiMTTKRP. */
int ptiDevMTTKRPHiCOO(ptiDevHiCOO const *t, float * const dmats[], ptiIndex R,
                      ptiIndex stride, ptiIndex const mats_order[], ptiIndex mode,
                      void *stream);
int ptiDevMTTKRPCOO(ptiDevCOO *t, float * const dmats[], ptiIndex R, ptiIndex stride,
                    ptiIndex const mats_order[], ptiIndex mode, void *stream);

/* Precision form: factors stored as bfloat16 (row stride `stride`), products
 * and accumulation in FP32, FP32 output dout (row stride `out_stride`). */
int ptiDevMTTKRPHiCOO_BF16(ptiDevHiCOO const *t, uint16_t const * const dfacs[],
                           float *dout, ptiIndex R, ptiIndex stride, ptiIndex out_stride,
                           ptiIndex const mats_order[], ptiIndex mode, void *stream);

/* MTTKRP of a COO tensor whose values are replaced by dvals (device, nnz
 * floats in the tensor's storage order), e.g. the residuals of a CP
 * completion model: the adjoint of the sampled reconstruction below. */
int ptiDevMTTKRPCOOValues(ptiDevCOO *t, float const *dvals, float * const dmats[], ptiIndex R,
                          ptiIndex stride, ptiIndex const mats_order[], ptiIndex mode, void *stream);

/* CP-completion forward at the observed entries (the tensor's nonzeros):
 *   resid_z = x_z - sum_r prod_m A_m[i_m][r]   (dresid: nnz floats, storage order)
 *   *dloss  = 0.5 * sum_z resid_z^2            (one double in device memory,
 *                                               summed in a fixed order)
 * dfacs: nmodes FP32 factor matrices with row stride `stride`; R <= 256. */
int ptiDevCOOCompletionForward(ptiDevCOO *t, float * const dfacs[], ptiIndex R, ptiIndex stride,
                               float *dresid, double *dloss, void *stream);

/* The same forward without any device allocation.  It has no counterpart in
 * the reference; it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * per-workgroup partial sums live in a fixed 1 MiB device array (one per
 * GPU, kept until the process ends), so the tensor may hold at most 2^27
 * nonzeros (PTIERR_VALUE_ERROR beyond).  Calls that overlap on one GPU share
 * that array and race; calls on one stream, or ordered across streams, do
 * not.  Results are bit-identical to ptiDevCOOCompletionForward. */
int ptiDevCOOCompletionForwardNoAlloc(ptiDevCOO *t, float * const dfacs[], ptiIndex R, ptiIndex stride,
                                      float *dresid, double *dloss, void *stream);

/* CP-completion gradient (no counterpart in the reference; the backward of
 * the loss above): dmats[nmodes] = MTTKRP of the residuals
 *   e_z = x_z - sum_r prod_m A_m[i_m][r],   A_m = dmats[m] (m < nmodes),
 * i.e. -dL/dA_mode.  Ranks <= 64 on orders <= 5 form e_z inside the MTTKRP
 * from the factors and ignore dresid; other shapes take the MTTKRP of dresid
 * (the forward's residuals, storage order).  Output row stride `stride`. */
int ptiDevCOOCompletionGrad(ptiDevCOO *t, float const *dresid, float * const dmats[], ptiIndex R,
                            ptiIndex stride, ptiIndex mode, void *stream);

/* CP-completion loss and every mode's gradient MTTKRP in one call (no
 * counterpart in the reference; for autograd callers that need both):
 * dgrads[n] = MTTKRP of the residuals for mode n (row stride `stride`) and
 * *dloss = 0.5 * sum_z e_z^2 (device double, fixed summation order), with
 * the residuals formed inside the MTTKRPs, so no forward pass is needed.
 * Only for shapes where ptiDevCOOCompletionFused returns 1 (every mode runs
 * the fused kernel); otherwise PTIERR_VALUE_ERROR and nothing is written. */
int ptiDevCOOCompletionFused(ptiDevCOO const *t, ptiIndex R, ptiIndex stride);
int ptiDevCOOCompletionLossGrad(ptiDevCOO *t, float * const dfacs[], float * const dgrads[], ptiIndex R,
                                ptiIndex stride, double *dloss, void *stream);

/* FP64 dense step of CP-ALS on device data.  The reference does this work on
 * the host with BLAS/LAPACK in FP32 (src/matrix/matrix_solver.c:32-83,
 * src/matrix/matrix.c:340-572, src/ktensor/ktensor.c:77-240).  Matrices are
 * row-major: factors FP32 with row stride `stride`, Grams FP64 R x R dense.
 * R <= 256.  Each call synchronises `stream` before returning. */
/* dG = A^T A (R x R, both triangles written), FP64 sums in a fixed order. */
int ptiDevGram(float const *dA, ptiIndex nrows, ptiIndex R, ptiIndex stride, double *dG, void *stream);
/* V = Hadamard product of dgrams[k] (nmodes contiguous R x R matrices),
 * k != mode; dA = dM V^-1 row by row (Cholesky, else LU with partial
 * pivoting; V exactly singular: dA = dM).  *status (host, may be NULL):
 * 0 Cholesky, 1 LU, 2 singular.  dA and dM must not overlap. */
int ptiDevSolveNormals(double const *dgrams, ptiIndex nmodes, ptiIndex mode, ptiIndex R, float const *dM,
                       float *dA, ptiIndex nrows, ptiIndex stride, int *status, void *stream);
/* Column norms into dlambda (R doubles) and division of the columns by them:
 * 2-norm, or (max_norm != 0) the column maximum clamped to >= 1. */
int ptiDevNormalizeColumns(float *dA, ptiIndex nrows, ptiIndex R, ptiIndex stride, int max_norm,
                           double *dlambda, void *stream);
/* fit = 1 - sqrt(normX2 + ||K||^2 - 2<X,K>) / sqrt(normX2), ||K||^2 from the
 * full Grams and dlambda, <X,K> = sum_r lambda_r sum_i dlast[i][r] dM[i][r]. */
int ptiDevKruskalFit(double const *dgrams, ptiIndex nmodes, ptiIndex R, double const *dlambda,
                     float const *dlast, float const *dM, ptiIndex nrows_last, ptiIndex stride,
                     double normX2, double *fit, void *stream);

/* CP-ALS entirely on the device, starting from the caller's factors
 * dfacs[0..nmodes-1] (FP32, row stride `stride`, updated in place).  The
 * iteration is the reference's (src/sptensor/cpd.c:32-160) with the dense
 * step above.  dlambda: R doubles (device, may be NULL).  *fit (host): last
 * fit; fit_trace (host, may be NULL): niters entries, one per iteration run;
 * *iters_done (host, may be NULL).  Stops after niters or when the fit
 * changes by less than tol.  For HiCOO, the first bf16_iters iterations
 * gather BF16 copies of the factors (FP32 accumulation, FP32 masters). */
int ptiDevCpdAlsHiCOO(ptiDevHiCOO const *t, float * const dfacs[], ptiIndex R, ptiIndex stride,
                      ptiIndex niters, double tol, ptiIndex bf16_iters, double *dlambda, double *fit,
                      double *fit_trace, ptiIndex *iters_done, void *stream);
int ptiDevCpdAlsCOO(ptiDevCOO *t, float * const dfacs[], ptiIndex R, ptiIndex stride, ptiIndex niters,
                    double tol, double *dlambda, double *fit, double *fit_trace, ptiIndex *iters_done,
                    void *stream);

/* ptiSparseTensorToHiCOO with the ordering and the block emission on the
 * GPU (the reference converts on the host, src/sptensor/hicoo/convert.c:
 * 646-860).  Same arguments, same output bit for bit, and like the
 * reference it sorts the caller's COO tensor in place.  The GPU path covers
 * tensors whose reference order is unique (3-D / 4-D, distinct coordinates,
 * 3-D indices < 2^24, a composite key of <= 64 bits); every other tensor is
 * converted by the host code of ptiSparseTensorToHiCOO. */
int ptiDevSparseTensorToHiCOO(ptiSparseTensorHiCOO *hitsr, ptiNnzIndex *max_nnzb, ptiSparseTensor *tsr,
                              ptiElementIndex const sb_bits, ptiElementIndex const sk_bits,
                              ptiElementIndex const sc_bits, void *stream);

/* Register a host HiCOO / COO tensor so that the drop-in ptiMTTKRPHiCOO /
 * ptiMTTKRP reuse a device copy instead of uploading the tensor per call.
 * The caller must unregister before modifying or freeing the host tensor. */
int ptiRegisterHiCOO(ptiSparseTensorHiCOO const *hitsr);
int ptiUnregisterHiCOO(ptiSparseTensorHiCOO const *hitsr);
int ptiRegisterSparseTensor(ptiSparseTensor const *X);
int ptiUnregisterSparseTensor(ptiSparseTensor const *X);

#ifdef __cplusplus
}
#endif

#endif /* HIPARTI_DEV_H */
