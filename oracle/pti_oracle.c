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
