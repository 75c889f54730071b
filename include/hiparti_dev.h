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
