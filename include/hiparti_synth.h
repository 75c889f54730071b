/*
 * Seeded synthetic sparse tensors (libHiParTI.so).  No reference counterpart:
 * the reference only loads files (src/sptensor/load.c).  The generator is
 * counter-based, so its output depends on the arguments and the seed only,
 * never on the thread count or the machine.
 *
 * Index draws per mode: uniform on [0, ndims[m]) or Zipf(s) with
 * P(k) proportional to (k+1)^-s over the same range, optionally relabelled by
 * a seeded random permutation of the mode.  Duplicate coordinates are removed
 * by drawing further rounds; the nnz unique coordinates with the lowest draw
 * numbers are kept and returned in lexicographic order (mode 0 slowest).
 * Values are then drawn in that order: U(0,1) (open interval) or
 * lognormal(0,1).
 */
#ifndef HIPARTI_SYNTH_H
#define HIPARTI_SYNTH_H

#include "HiParTI.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { PTI_SYNTH_UNIFORM = 0, PTI_SYNTH_ZIPF = 1 };
enum { PTI_SYNTH_VAL_UNIFORM = 0, PTI_SYNTH_VAL_LOGNORMAL = 1 };

/* dist[m] in {PTI_SYNTH_UNIFORM, PTI_SYNTH_ZIPF}; zipf_s[m] used for Zipf
 * modes; permute[m] != 0 relabels mode m.  nthreads <= 0 uses all cores.
 * Fails with PTIERR_VALUE_ERROR when nnz exceeds the number of cells. */
int ptiSynthSparseTensor(ptiSparseTensor *tsr, ptiIndex nmodes, const ptiIndex ndims[],
                         ptiNnzIndex nnz, const int dist[], const double zipf_s[],
                         const int permute[], int value_dist, uint64_t seed, int nthreads);

/* Fill a matrix (nrows x ncols, padding columns left untouched) with U(0,1)
 * values drawn from (seed, row, column). */
int ptiSynthUniformMatrix(ptiMatrix *mtx, uint64_t seed);

#ifdef __cplusplus
}
#endif

#endif /* HIPARTI_SYNTH_H */
