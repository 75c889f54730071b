/*
 * HiParTI-compatible C ABI for the MI355X sparse-MTTKRP path.
 *
 * The scalar widths and struct layouts below are byte-for-byte the ones the
 * reference library exposes (include/includes/types.h:27-103 and
 * include/includes/structs.h:27-312), because a drop-in replacement has to
 * accept the caller's existing ptiSparseTensor / ptiSparseTensorHiCOO /
 * ptiMatrix objects unchanged.  Every function declared here is exported by
 * libHiParTI.so built from hiparti_amd/csrc.  The comment above each
 * prototype names the reference declaration it replaces.
 *
 * Semantics that matter to callers:
 *   - All pti* functions are synchronous.  On return, output buffers that live
 *     in host memory hold the result.
 *   - Error codes follow the reference (PTIERR_*): 0 on success,
 *     PTIERR_SHAPE_MISMATCH for inconsistent matrices, PTIERR_OS_ERROR+errno
 *     for allocation / IO failures, PTIERR_CUDA_ERROR+hipError_t for device
 *     failures.  The last error is recorded per thread (ptiGetLastError).
 *   - ptiMTTKRP / ptiMTTKRPHiCOO run on the GPU.  There is no CPU fallback:
 *     without a usable HIP device they return PTIERR_CUDA_ERROR + code.
 */
#ifndef HIPARTI_H
#define HIPARTI_H

#include <stddef.h>
#include <stdio.h>
#include <stdint.h>
#include <inttypes.h>
#include <stdbool.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------------
 * Scalar types (reference: include/includes/types.h:27-103, 32-bit index,
 * 32-bit value, 8-bit element index build).
 * ------------------------------------------------------------------------- */
#define HIPARTI_INDEX_TYPEWIDTH 32
#define HIPARTI_VALUE_TYPEWIDTH 32
#define HIPARTI_ELEMENT_INDEX_TYPEWIDTH 8

typedef uint32_t ptiIndex;
typedef uint32_t ptiBlockIndex;
typedef float    ptiValue;
typedef uint8_t  ptiElementIndex;
typedef uint16_t ptiBlockMatrixIndex;
typedef ptiBlockIndex ptiBlockNnzIndex;
typedef uint64_t ptiNnzIndex;
typedef unsigned __int128 ptiMortonIndex;

#define HIPARTI_INDEX_MAX UINT32_MAX
#define HIPARTI_NNZ_INDEX_MAX UINT64_MAX
#define HIPARTI_PRI_INDEX PRIu32
#define HIPARTI_PRI_NNZ_INDEX PRIu64
#define HIPARTI_PRI_VALUE "f"

/* ---------------------------------------------------------------------------
 * Error codes (reference: include/includes/macros.h:47-56).
 * ------------------------------------------------------------------------- */
typedef enum {
    PTIERR_NO_ERROR       = 0,
    PTIERR_UNKNOWN        = 1,
    PTIERR_SHAPE_MISMATCH = 2,
    PTIERR_VALUE_ERROR    = 3,
    PTIERR_ZERO_DIVISION  = 4,
    PTIERR_NO_MORE        = 99,
    PTIERR_OS_ERROR       = 0x10000,
    PTIERR_CUDA_ERROR     = 0x20000,
} PtiError;

/* Opaque wall-clock timer (reference: include/includes/macros.h:45). */
typedef struct ptiTagTimer *ptiTimer;

/* ---------------------------------------------------------------------------
 * Containers (reference: include/includes/structs.h).  Field order, widths
 * and padding are those of the reference so existing objects can be passed.
 * ------------------------------------------------------------------------- */

/* Growable arrays: len used entries, cap allocated entries (structs.h:27-58). */
typedef struct { ptiNnzIndex len; ptiNnzIndex cap; ptiValue        *data; } ptiValueVector;
typedef struct { ptiNnzIndex len; ptiNnzIndex cap; ptiIndex        *data; } ptiIndexVector;
typedef struct { ptiNnzIndex len; ptiNnzIndex cap; ptiElementIndex *data; } ptiElementIndexVector;
typedef struct { ptiNnzIndex len; ptiNnzIndex cap; ptiBlockIndex   *data; } ptiBlockIndexVector;
typedef struct { ptiNnzIndex len; ptiNnzIndex cap; ptiNnzIndex     *data; } ptiNnzIndexVector;

/* Dense row-major matrix; stride = ncols rounded up to a multiple of 8
 * (structs.h:64-70, matrix.c:36-55). */
typedef struct {
    ptiIndex nrows;
    ptiIndex ncols;
    ptiIndex cap;
    ptiIndex stride;
    ptiValue *values;
} ptiMatrix;

/* Dense matrix with an 8-bit column count (structs.h:76-82). */
typedef struct {
    ptiIndex        nrows;
    ptiElementIndex ncols;
    ptiIndex        cap;
    ptiElementIndex stride;
    ptiValue       *values;
} ptiRankMatrix;

/* COO sparse tensor, structure of arrays (structs.h:142-149). */
typedef struct {
    ptiIndex        nmodes;
    ptiIndex       *sortorder;
    ptiIndex       *ndims;
    ptiNnzIndex     nnz;
    ptiIndexVector *inds;      /* inds[m].data[z] = index of nonzero z in mode m */
    ptiValueVector  values;
} ptiSparseTensor;

/* HiCOO sparse tensor (structs.h:155-185).
 *   bptr[b]..bptr[b+1]  nonzeros of block b;   binds[m].data[b] = block coord
 *   einds[m].data[z]    offset of nonzero z inside its block (< 2^sb_bits)
 *   kptr[k]..kptr[k+1]  blocks of superblock ("kernel") k
 *   cptr                chunk boundaries (block offsets)
 *   kschr[m][i]         superblocks whose mode-m superblock coordinate is i  */
typedef struct {
    ptiIndex            nmodes;
    ptiIndex           *sortorder;
    ptiIndex           *ndims;
    ptiNnzIndex         nnz;
    ptiElementIndex     sb_bits;
    ptiElementIndex     sk_bits;
    ptiElementIndex     sc_bits;
    ptiNnzIndexVector   kptr;
    ptiIndexVector    **kschr;
    ptiIndex           *nkiters;
    ptiNnzIndexVector   cptr;
    ptiIndexVector    **kschr_balanced;
    ptiIndexVector    **kschr_balanced_pos;
    ptiIndex           *nkpars;
    ptiIndexVector     *kschr_rest;
    ptiNnzIndexVector   knnzs;
    ptiNnzIndexVector   bptr;
    ptiBlockIndexVector   *binds;
    ptiElementIndexVector *einds;
    ptiValueVector      values;
} ptiSparseTensorHiCOO;

/* CP decomposition result (structs.h:249-256). */
typedef struct {
    ptiIndex    nmodes;
    ptiIndex    rank;
    ptiIndex   *ndims;
    ptiValue   *lambda;
    double      fit;
    ptiMatrix **factors;
} ptiKruskalTensor;

/* CP decomposition result with ptiRankMatrix factors (structs.h:262-275). */
typedef struct {
    ptiIndex         nmodes;
    ptiElementIndex  rank;
    ptiIndex        *ndims;
    ptiValue        *lambda;
    double           fit;
    ptiRankMatrix  **factors;
} ptiRankKruskalTensor;

/* OpenMP lock pool of the reference (structs.h:290-299).  Only passed
 * through by pointer here; never dereferenced. */
typedef struct ptiMutexPool ptiMutexPool;

/* ---------------------------------------------------------------------------
 * Errors and timers (reference: include/includes/helper_funcs.h:29-50,
 * src/helper_funcs/error.c:27-118, src/helper_funcs/timer.c:34-100).
 * ------------------------------------------------------------------------- */
int  ptiGetLastError(const char **module, const char **file, unsigned *line, const char **reason);
void ptiClearLastError(void);
void pti_ComplainError(const char *module, int errcode, const char *file, unsigned line, const char *reason);
void pti_Panic(const char *file, unsigned line, const char *expr);
#define ptiAssert(expr) ((expr) ? (void) 0 : pti_Panic(__FILE__, __LINE__, #expr))

/* use_cuda != 0 selects a device-event timer on the current HIP stream. */
int    ptiNewTimer(ptiTimer *timer, int use_cuda);
int    ptiStartTimer(ptiTimer timer);
int    ptiStopTimer(ptiTimer timer);
double ptiElapsedTime(const ptiTimer timer);
double ptiPrintElapsedTime(const ptiTimer timer, const char *name);
double ptiPrintAverageElapsedTime(const ptiTimer timer, const int niters, const char *name);
int    ptiFreeTimer(ptiTimer timer);

/* ---------------------------------------------------------------------------
 * Vectors (reference: include/includes/vectors.h:32-80).
 * ------------------------------------------------------------------------- */
int  ptiNewValueVector(ptiValueVector *vec, uint64_t len, uint64_t cap);
int  ptiConstantValueVector(ptiValueVector * const vec, ptiValue const val);
int  ptiAppendValueVector(ptiValueVector *vec, ptiValue const value);
int  ptiResizeValueVector(ptiValueVector *vec, ptiNnzIndex const size);
void ptiFreeValueVector(ptiValueVector *vec);
int  ptiNewIndexVector(ptiIndexVector *vec, uint64_t len, uint64_t cap);
int  ptiAppendIndexVector(ptiIndexVector *vec, ptiIndex const value);
int  ptiResizeIndexVector(ptiIndexVector *vec, ptiNnzIndex const size);
void ptiFreeIndexVector(ptiIndexVector *vec);
int  ptiNewElementIndexVector(ptiElementIndexVector *vec, uint64_t len, uint64_t cap);
int  ptiAppendElementIndexVector(ptiElementIndexVector *vec, ptiElementIndex const value);
int  ptiResizeElementIndexVector(ptiElementIndexVector *vec, ptiNnzIndex const size);
void ptiFreeElementIndexVector(ptiElementIndexVector *vec);
int  ptiNewBlockIndexVector(ptiBlockIndexVector *vec, uint64_t len, uint64_t cap);
int  ptiAppendBlockIndexVector(ptiBlockIndexVector *vec, ptiBlockIndex const value);
int  ptiResizeBlockIndexVector(ptiBlockIndexVector *vec, ptiNnzIndex const size);
void ptiFreeBlockIndexVector(ptiBlockIndexVector *vec);
int  ptiNewNnzIndexVector(ptiNnzIndexVector *vec, uint64_t len, uint64_t cap);
int  ptiAppendNnzIndexVector(ptiNnzIndexVector *vec, ptiNnzIndex const value);
int  ptiResizeNnzIndexVector(ptiNnzIndexVector *vec, ptiNnzIndex const size);
void ptiFreeNnzIndexVector(ptiNnzIndexVector *vec);
ptiIndex ptiMaxIndexArray(ptiIndex const * const indices, ptiNnzIndex const size);
/* Text dumps of vectors and arrays (reference: vectors.h, src/vector/dump.c:31-175). */
int ptiDumpElementIndexVector(ptiElementIndexVector *vec, FILE *fp);
int ptiDumpIndexVector(ptiIndexVector *vec, FILE *fp);
int ptiDumpBlockIndexVector(ptiBlockIndexVector *vec, FILE *fp);
int ptiDumpNnzIndexVector(ptiNnzIndexVector *vec, FILE *fp);
int ptiDumpValueVector(ptiValueVector *vec, FILE *fp);
int ptiDumpIndexArray(ptiIndex const *array, ptiNnzIndex const n, FILE *fp);
int ptiDumpNnzIndexArray(ptiNnzIndex const *array, ptiNnzIndex const n, FILE *fp);

/* ---------------------------------------------------------------------------
 * Dense matrices (reference: include/includes/matrices.h:26-65).
 * ------------------------------------------------------------------------- */
int  ptiNewMatrix(ptiMatrix *mtx, ptiIndex const nrows, ptiIndex const ncols);
/* Fills entry (i, j) with i + j + 1, as the reference does (matrix.c:67-74). */
int  ptiRandomizeMatrix(ptiMatrix *mtx);
int  ptiConstantMatrix(ptiMatrix * const mtx, ptiValue const val);
int  ptiCopyMatrix(ptiMatrix *dest, const ptiMatrix *src);
void ptiFreeMatrix(ptiMatrix *mtx);
int  ptiDumpMatrix(ptiMatrix *mtx, FILE *fp);
int  ptiMatrix2Norm(ptiMatrix * const A, ptiValue * const lambda);
int  ptiMatrixMaxNorm(ptiMatrix * const A, ptiValue * const lambda);

/* Normal-equation helpers of CP-ALS (reference: include/includes/matrices.h:47-65,
 * src/matrix/matrix.c:340-385,556-572, src/matrix/matrix_solver.c:32-83).
 * aTa[k] (k < nmodes) are R x R Grams of which only the row-major upper
 * triangle (entries [i][j], j >= i: the triangle a column-major ssyrk 'L'
 * writes) is read; aTa[nmodes] receives their Hadamard product over
 * k != mode, made symmetric.
 * ptiMatrixSolveNormals solves x V = m for every row m of rhs, in place, on
 * the GPU in FP64: Cholesky, else LU with partial pivoting, else (V exactly
 * singular) rhs is left unchanged.  Unlike the reference, aTa[nmodes] keeps
 * V itself rather than its factorisation. */
int  ptiMatrixDotMulSeqTriangle(ptiIndex const mode, ptiIndex const nmodes, ptiMatrix ** mats);
int  ptiMatrixSolveNormals(ptiIndex const mode, ptiIndex const nmodes, ptiMatrix ** aTa, ptiMatrix * rhs);
/* lambda[r] *= column-r 2-norm of every factor, which is normalised. */
void GetFinalLambda(ptiIndex const rank, ptiIndex const nmodes, ptiMatrix ** mats, ptiValue * const lambda);

/* Dense matrices with 8-bit ncols / stride (reference: matrices.h:68-89,
 * src/matrix/rankmatrix.c).  ncols <= 248 keeps the stride representable. */
int  ptiNewRankMatrix(ptiRankMatrix *mtx, ptiIndex const nrows, ptiElementIndex const ncols);
/* Fills entry (i, j) with i + j + 1 (rankmatrix.c:69-77). */
int  ptiRandomizeRankMatrix(ptiRankMatrix *mtx, ptiIndex const nrows, ptiElementIndex const ncols);
int  ptiConstantRankMatrix(ptiRankMatrix *mtx, ptiValue const val);
void ptiFreeRankMatrix(ptiRankMatrix *mtx);
int  ptiDumpRankMatrix(ptiRankMatrix *mtx, FILE *fp);
int  ptiRankMatrixDotMulSeqTriangle(ptiIndex const mode, ptiIndex const nmodes, ptiRankMatrix ** mats);
int  ptiRankMatrix2Norm(ptiRankMatrix * const A, ptiValue * const lambda);
int  ptiRankMatrixMaxNorm(ptiRankMatrix * const A, ptiValue * const lambda);
void GetRankFinalLambda(ptiElementIndex const rank, ptiIndex const nmodes, ptiRankMatrix ** mats,
                        ptiValue * const lambda);
int  ptiRankMatrixSolveNormals(ptiIndex const mode, ptiIndex const nmodes, ptiRankMatrix ** aTa,
                               ptiRankMatrix * rhs);

/* ---------------------------------------------------------------------------
 * COO sparse tensors (reference: include/includes/sptensors.h:22-50).
 * ------------------------------------------------------------------------- */
int    ptiNewSparseTensor(ptiSparseTensor *tsr, ptiIndex nmodes, const ptiIndex ndims[]);
int    ptiCopySparseTensor(ptiSparseTensor *dest, const ptiSparseTensor *src, int const nt);
void   ptiFreeSparseTensor(ptiSparseTensor *tsr);
double SparseTensorFrobeniusNormSquared(ptiSparseTensor const * const ptien);
/* Text coordinate format (.tns/.coo): nmodes, ndims, then "i_1 .. i_N value"
 * lines with start_index-based indices (reference: load.c:63-118,277-297). */
int    ptiLoadSparseTensor(ptiSparseTensor *tsr, ptiIndex start_index, char const * const fname);
int    ptiDumpSparseTensor(const ptiSparseTensor *tsr, ptiIndex start_index, FILE *fp);
/* Lexicographic sort, mode 0 slowest (reference: sptensors.h:80, sort.c:1570-1600). */
void   ptiSparseTensorSortIndex(ptiSparseTensor *tsr, int force, int tk);

/* ---------------------------------------------------------------------------
 * HiCOO sparse tensors (reference: include/includes/sptensors.h:122-160).
 * ------------------------------------------------------------------------- */
int  ptiNewSparseTensorHiCOO(ptiSparseTensorHiCOO *hitsr, const ptiIndex nmodes,
                             const ptiIndex ndims[], const ptiNnzIndex nnz,
                             const ptiElementIndex sb_bits, const ptiElementIndex sk_bits,
                             const ptiElementIndex sc_bits);
void ptiFreeSparseTensorHiCOO(ptiSparseTensorHiCOO *hitsr);
/* Converts COO -> HiCOO with the reference's element order and arrays
 * (reference: src/sptensor/hicoo/convert.c:646-860).  Like the reference it
 * sorts the caller's COO tensor in place.  tk is the host thread count used
 * for sorting; the result does not depend on it. */
int  ptiSparseTensorToHiCOO(ptiSparseTensorHiCOO *hitsr, ptiNnzIndex *max_nnzb,
                            ptiSparseTensor *tsr, const ptiElementIndex sb_bits,
                            const ptiElementIndex sk_bits, const ptiElementIndex sc_bits,
                            int const tk);
double SparseTensorFrobeniusNormSquaredHiCOO(ptiSparseTensorHiCOO const * const hitsr);
/* Text dump of every HiCOO array (reference: sptensors.h:128, src/sptensor/hicoo/dump.c:30-74). */
int  ptiDumpSparseTensorHiCOO(ptiSparseTensorHiCOO * const hitsr, FILE *fp);

/* ---------------------------------------------------------------------------
 * MTTKRP (the hot path).
 *
 * mats has nmodes+1 entries: mats[0..nmodes-1] are the factor matrices
 * (mats[m] has ndims[m] rows and R columns), mats[nmodes] is the output.  The
 * output is zeroed over mats[mode]->nrows rows and then receives
 *     M[i_mode][r] = sum_z  v_z * prod_{k=1..N-1} mats[mats_order[k]][i_{mats_order[k]}][r].
 * mats_order[0] is ignored.  Returns PTIERR_SHAPE_MISMATCH when a factor's
 * ncols differs from mats[nmodes]->ncols or its nrows differs from ndims.
 * ------------------------------------------------------------------------- */
/* reference: include/includes/sptensors.h:179-183, src/sptensor/mttkrp.c:39-108 */
int ptiMTTKRP(ptiSparseTensor const * const X, ptiMatrix * mats[],
              ptiIndex const mats_order[], ptiIndex const mode);
/* reference: include/includes/sptensors.h:253-257, src/sptensor/hicoo/mttkrp.c:61-152 */
int ptiMTTKRPHiCOO(ptiSparseTensorHiCOO const * const hitsr, ptiMatrix * mats[],
                   ptiIndex const mats_order[], ptiIndex const mode);
/* GPU-named entry points of the reference (sptensors.h:204-208 and
 * :306-312).  impl_num and max_nnzb are accepted and ignored: there is one
 * kernel family, tuned for gfx950. */
int ptiCudaMTTKRP(ptiSparseTensor const * const X, ptiMatrix * mats[],
                  ptiIndex * const mats_order, ptiIndex const mode, ptiIndex const impl_num);
int ptiCudaMTTKRPHiCOO(ptiSparseTensorHiCOO const * const hitsr, ptiMatrix ** const mats,
                       ptiIndex * const mats_order, ptiIndex const mode,
                       ptiNnzIndex const max_nnzb, int const impl_num);
/* The reference's OpenMP and matrix-tiling variants (sptensors.h:184-203,
 * :258-300).  They compute the same product and all run the GPU kernels:
 * tk, tb, balanced, copy_mats and lock_pool are accepted and not used.
 * ptiRankMatrix arguments share ptiMatrix's memory layout (row-major, 8-bit
 * ncols and stride). */
int ptiOmpMTTKRP(ptiSparseTensor const * const X, ptiMatrix * mats[],
                 ptiIndex const mats_order[], ptiIndex const mode, const int tk);
int ptiOmpMTTKRP_Reduce(ptiSparseTensor const * const X, ptiMatrix * mats[], ptiMatrix * copy_mats[],
                        ptiIndex const mats_order[], ptiIndex const mode, const int tk);
int ptiOmpMTTKRP_Lock(ptiSparseTensor const * const X, ptiMatrix * mats[],
                      ptiIndex const mats_order[], ptiIndex const mode, const int tk,
                      ptiMutexPool * lock_pool);
int ptiMTTKRPHiCOO_MatrixTiling(ptiSparseTensorHiCOO const * const hitsr, ptiRankMatrix * mats[],
                                ptiIndex const mats_order[], ptiIndex const mode);
int ptiOmpMTTKRPHiCOO(ptiSparseTensorHiCOO const * const hitsr, ptiMatrix * mats[],
                      ptiIndex const mats_order[], ptiIndex const mode, const int tk, const int tb);
int ptiOmpMTTKRPHiCOO_MatrixTiling(ptiSparseTensorHiCOO const * const hitsr, ptiRankMatrix * mats[],
                                   ptiIndex const mats_order[], ptiIndex const mode,
                                   const int tk, const int tb);
int ptiOmpMTTKRPHiCOO_MatrixTiling_Scheduled(ptiSparseTensorHiCOO const * const hitsr,
                                             ptiRankMatrix * mats[], ptiIndex const mats_order[],
                                             ptiIndex const mode, const int tk, const int tb,
                                             int balanced);
int ptiOmpMTTKRPHiCOO_MatrixTiling_Scheduled_Reduce(ptiSparseTensorHiCOO const * const hitsr,
                                                    ptiRankMatrix * mats[], ptiRankMatrix * copy_mats[],
                                                    ptiIndex const mats_order[], ptiIndex const mode,
                                                    const int tk, const int tb, int balanced);
int ptiOmpMTTKRPHiCOO_MatrixTiling_Scheduled_Reduce_Two(ptiSparseTensorHiCOO const * const hitsr,
                                                        ptiRankMatrix * mats[], ptiRankMatrix * copy_mats[],
                                                        ptiIndex const mats_order[], ptiIndex const mode,
                                                        const int tk, const int tb);

/* ---------------------------------------------------------------------------
 * Kruskal tensors and CPD-ALS (reference: include/includes/ktensors.h:23-39,
 * include/includes/cpds.h:26-31).
 * ------------------------------------------------------------------------- */
int    ptiNewKruskalTensor(ptiKruskalTensor *ktsr, ptiIndex nmodes, const ptiIndex ndims[], ptiIndex rank);
void   ptiFreeKruskalTensor(ptiKruskalTensor *ktsr);
int    ptiDumpKruskalTensor(ptiKruskalTensor *ktsr, FILE *fp);
/* CP-ALS on the GPU.  Factors start from the reference's deterministic
 * initialisation (entry (i, j) = i + j + 1); normal equations are formed and
 * solved in FP64 on the device.  Fills ktensor->lambda, ->factors and ->fit
 * (reference: src/sptensor/cpd.c:163-208). */
int    ptiCpdAls(ptiSparseTensor const * const ptien, ptiIndex const rank,
                 ptiIndex const niters, double const tol, ptiKruskalTensor * ktensor);
/* Same algorithm; tk and use_reduce are accepted and not used (cpds.h:32-45). */
int    ptiOmpCpdAls(ptiSparseTensor const * const ptien, ptiIndex const rank, ptiIndex const niters,
                    double const tol, const int tk, const int use_reduce, ptiKruskalTensor * ktensor);
int    ptiCudaCpdAls(ptiSparseTensor const * const ptien, ptiIndex const rank, ptiIndex const niters,
                     double const tol, ptiKruskalTensor * ktensor);
/* CP-ALS over the HiCOO MTTKRP, ptiRankMatrix factors, rank <= 248
 * (cpds.h:46-60, src/sptensor/hicoo/cpd.c:153-204). */
int    ptiCpdAlsHiCOO(ptiSparseTensorHiCOO const * const hitsr, ptiIndex const rank, ptiIndex const niters,
                      double const tol, ptiRankKruskalTensor * ktensor);
int    ptiOmpCpdAlsHiCOO(ptiSparseTensorHiCOO const * const hitsr, ptiIndex const rank,
                         ptiIndex const niters, double const tol, const int tk, const int tb,
                         int balanced, ptiRankKruskalTensor * ktensor);

/* Fit of a Kruskal tensor against X (ktensors.h:27-39, ktensor.c:77-240):
 * 1 - sqrt(||X||^2 + ||K||^2 - 2<X,K>) / ||X||, with ||K||^2 from the Grams
 * ata[0..nmodes-1] (row-major upper triangles read, see
 * ptiMatrixSolveNormals) and <X,K> from mats[nmodes-1] and the last MTTKRP
 * output mats[nmodes].  FP64 sums.  The reference reads the triangle of the
 * Grams that ssyrk leaves unwritten (ktensor.c:128-143); this reads the
 * written one.  ata[nmodes] is overwritten with the Hadamard product. */
double KruskalTensorFit(ptiSparseTensor const * const ptien, ptiValue const * const lambda,
                        ptiMatrix ** mats, ptiMatrix ** ata);
double KruskalTensorFrobeniusNormSquared(ptiIndex const nmodes, ptiValue const * const lambda,
                                         ptiMatrix ** ata);
double SparseKruskalTensorInnerProduct(ptiIndex const nmodes, ptiValue const * const lambda,
                                       ptiMatrix ** mats);
double KruskalTensorFitHiCOO(ptiSparseTensorHiCOO const * const hitsr, ptiValue const * const lambda,
                             ptiRankMatrix ** mats, ptiRankMatrix ** ata);
double KruskalTensorFrobeniusNormSquaredRank(ptiIndex const nmodes, ptiValue const * const lambda,
                                             ptiRankMatrix ** ata);
double SparseKruskalTensorInnerProductRank(ptiIndex const nmodes, ptiValue const * const lambda,
                                           ptiRankMatrix ** mats);
int    ptiNewRankKruskalTensor(ptiRankKruskalTensor *ktsr, ptiIndex nmodes, const ptiIndex ndims[],
                               ptiElementIndex rank);
void   ptiFreeRankKruskalTensor(ptiRankKruskalTensor *ktsr);
int    ptiDumpRankKruskalTensor(ptiRankKruskalTensor *ktsr, FILE *fp);

#ifdef __cplusplus
}
#endif

#endif /* HIPARTI_H */
