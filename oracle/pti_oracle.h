/*
 * ORACLE -- test infrastructure only.
 *
 * CPU restatement of the reference HiParTI algorithms on the MTTKRP path,
 * used to check libHiParTI.so's results.  Only tests/, __graft_entry__.smoke()
 * and bench.py's CPU-baseline leg may load it; the product never does.
 *
 * All arrays are plain C arrays so the functions can be driven from ctypes.
 * Coordinates are given per mode (inds[m][z]); factor matrices are row-major
 * with row stride `stride`; mats_order follows the reference convention
 * (entries 1..N-1 give the product order, entry 0 is ignored).
 */
#ifndef PTI_ORACLE_H
#define PTI_ORACLE_H

#include <stdint.h>

/* FP64 brute force:  out[i][r] = sum_z v_z prod_{k>=1} A_{order[k]}[i_{order[k]}][r]
 * (out has nrows_out x R doubles, zeroed here).  Summation order is the
 * storage order; no rounding to FP32 anywhere.  Pinned by closed forms in the
 * tests (all-ones factors, rank-1 tensors, linearity). */
void orc_mttkrp_f64(uint32_t N, uint64_t nnz, const uint32_t *const *inds, const float *vals,
                    uint32_t R, uint32_t stride, const float *const *facs, const uint32_t *order,
                    uint32_t mode, uint64_t nrows_out, double *out);

/* orc_mttkrp_f64 restricted to the output rows sel[0..nsel-1]: out is
 * nsel x R doubles (row s = output row sel[s]); OpenMP with nthreads threads
 * (<= 0: all).  For full-size accuracy checks. */
void orc_mttkrp_f64_rows(uint32_t N, uint64_t nnz, const uint32_t *const *inds, const float *vals, uint32_t R,
                         uint32_t stride, const float *const *facs, uint32_t mode, uint64_t nrows_out,
                         uint32_t nsel, const uint32_t *sel, double *out, int nthreads);

/* CP-completion in FP64: *loss = 0.5 sum_z e_z^2 with
 * e_z = v_z - sum_r prod_m A_m[i_m][r], and out (nsel x R) the MTTKRP of the
 * residuals at output rows sel[] of `mode` (-dL/dA_mode there). */
void orc_completion_f64(uint32_t N, uint64_t nnz, const uint32_t *const *inds, const float *vals, uint32_t R,
                        uint32_t stride, const float *const *facs, uint32_t mode, uint64_t nrows_out,
                        uint32_t nsel, const uint32_t *sel, double *out, double *loss, int nthreads);

/* Sequential FP32 MTTKRP in the reference's operation order
 * (src/sptensor/mttkrp.c:39-163): out zeroed over nrows_out*stride floats. */
void orc_mttkrp_coo_seq(uint32_t N, uint64_t nnz, const uint32_t *const *inds, const float *vals,
                        uint32_t R, uint32_t stride, const float *const *facs, const uint32_t *order,
                        uint32_t mode, uint64_t nrows_out, float *out);

/* OpenMP COO MTTKRP with per-element atomics (src/sptensor/mttkrp_omp.c:125-252). */
void orc_omp_mttkrp_coo(uint32_t N, uint64_t nnz, const uint32_t *const *inds, const float *vals,
                        uint32_t R, uint32_t stride, const float *const *facs, const uint32_t *order,
                        uint32_t mode, uint64_t nrows_out, float *out, int nthreads);

/* CP-ALS on the host: src/sptensor/cpd_omp.c:32-244 (OmpCpdAlsStep with
 * use_reduce = 0) with the OpenMP COO MTTKRP above and an FP64
 * Cholesky-Banachiewicz / Doolittle-LU normal-equation solve.  facs (ndims[m]
 * x stride FP32) are the initial factors and receive the final ones; lambda
 * gets R values.  Returns the last fit. */
double orc_omp_cpd_als(uint32_t N, uint64_t nnz, const uint32_t *const *inds, const float *vals,
                       const uint32_t *ndims, uint32_t R, uint32_t stride, uint32_t niters, double tol,
                       float *const *facs, float *lambda, double *fit_trace, uint32_t *iters_done, double *t_mttkrp,
                       double *t_other, int nthreads);

/* Sequential HiCOO MTTKRP (src/sptensor/hicoo/mttkrp.c:61-152, 232-310).
 * binds[m][b], einds[m][z]; kptr has nk+1 block offsets. */
void orc_mttkrp_hicoo_seq(uint32_t N, uint64_t nk, const uint64_t *kptr, const uint64_t *bptr,
                          const uint32_t *const *binds, const uint8_t *const *einds, const float *vals,
                          uint32_t sb_bits, uint32_t R, uint32_t stride, const float *const *facs,
                          const uint32_t *order, uint32_t mode, uint64_t nrows_out, float *out);

/* OpenMP HiCOO MTTKRP, superblock-parallel with atomics
 * (src/sptensor/hicoo/mttkrp_omp.c:184-204, 297-467, tk = nthreads, tb = 1). */
void orc_omp_mttkrp_hicoo(uint32_t N, uint64_t nk, const uint64_t *kptr, const uint64_t *bptr,
                          const uint32_t *const *binds, const uint8_t *const *einds, const float *vals,
                          uint32_t sb_bits, uint32_t R, uint32_t stride, const float *const *facs,
                          const uint32_t *order, uint32_t mode, uint64_t nrows_out, float *out, int nthreads);

/* COO -> HiCOO conversion restated from src/sptensor/hicoo/convert.c:190-860
 * and src/sptensor/sort.c (tk = 1).  Sorts inds/vals in place exactly as the
 * reference does and fills an orc_hicoo whose arrays are malloc'ed here
 * (release with orc_hicoo_free).  kschr-like lists are flattened per mode:
 * list i of mode m is dat[m][ptr[m][i] .. ptr[m][i+1]). */
typedef struct {
    uint32_t nmodes;
    uint64_t nnz, nb, nk, nc, max_nnzb;
    uint64_t *bptr;        /* nb + 1 */
    uint32_t **binds;      /* [m][nb] */
    uint8_t **einds;       /* [m][nnz] */
    float *values;         /* nnz */
    uint64_t *kptr;        /* nk + 1 (block offsets) */
    uint64_t *cptr;        /* nc */
    uint64_t *knnzs;       /* nk */
    uint32_t *nkiters;     /* [m] */
    uint32_t *nkpars;      /* [m] */
    uint32_t *kernel_ndim; /* [m] */
    uint64_t **kschr_ptr; uint32_t **kschr_dat;
    uint64_t **bal_ptr;   uint32_t **bal_dat;
    uint64_t **pos_ptr;   uint32_t **pos_dat;
    uint64_t *rest_len;   uint32_t **rest_dat;
} orc_hicoo;

int orc_hicoo_convert(uint32_t N, const uint32_t *ndims, uint64_t nnz, uint32_t **inds, float *vals,
                      uint32_t sb_bits, uint32_t sk_bits, uint32_t sc_bits, orc_hicoo *out);
void orc_hicoo_free(orc_hicoo *h);

/* Lexicographic quicksort (src/sptensor/sort.c:1570-1600), in place. */
void orc_sort_lexicographic(uint32_t N, uint64_t nnz, uint32_t **inds, float *vals);

#endif
