/* CPU restatement of the pair rows over lattice bins -- TEST INFRASTRUCTURE ONLY (the pair path's CPU baseline and a
 * checker).  s_r = sum_j S(ZA_r, ZB_j) w_j for tables whose columns are offset rows (kind 0), circular rows (1) or rank
 * tables (3), the layout of the reference's ProductKernel.tables (blml_kernels.py:133-200), read as its row kernels
 * read them (blml_local.py:86-96 _col_index, 203-282 _rows_numba; blml_local_gpu.py:42-105): the column factors
 * multiply in column order in the table's precision (REAL = float when f32), the row's sum is FP64 over the sources
 * in order, skipping zero weights.  OpenMP over rows.  Pinned by tests/test_oracle_pairs.py against the numpy
 * restatement and the dense Gram. */
#include <stdint.h>

static inline long col_index(int kd, long za, long zb, long base, long J, long mk, long nb, long nT) {
    long i;
    if (kd == 0) {
        i = base + za - zb + J;
    } else if (kd == 3) {
        i = base + za * mk + zb;
    } else {
        long m = za - zb;
        if (m < 0) m += nb;
        else if (m >= nb) m -= nb;
        if (m > nb - m) m = nb - m;
        i = base + m;
    }
    return i < 0 ? 0 : (i >= nT ? nT - 1 : i);
}

#define ROWS_BODY(REAL)                                                                                          \
    _Pragma("omp parallel for schedule(dynamic, 16) num_threads(nthreads)")                                      \
    for (long r = 0; r < R; ++r) {                                                                               \
        const int32_t *za = ZA + r * d;                                                                          \
        double acc = 0.0;                                                                                        \
        for (long j = 0; j < M; ++j) {                                                                           \
            const double wj = w[j];                                                                              \
            if (wj == 0.0) continue;                                                                             \
            const int32_t *zb = ZB + j * d;                                                                      \
            REAL v = (REAL)1;                                                                                    \
            for (int k = 0; k < d; ++k)                                                                          \
                v *= T[col_index(kind[k], za[k], zb[k], base[k], J[k], mk[k], nbins[k], nT)];                     \
            acc += (double)v * wj;                                                                               \
        }                                                                                                        \
        s[r] = acc;                                                                                              \
    }

/* Returns 0, or 1 when a column kind is not 0, 1 or 3 (groups are not restated here). */
int pair_rows_cpu(int d, const int32_t *kind, const int64_t *base, const int64_t *J, const int64_t *mk,
                  const int64_t *nbins, int64_t nT, const void *table, int f32, const int32_t *ZA, const int32_t *ZB,
                  const double *w, int64_t R, int64_t M, double *s, int nthreads) {
    for (int k = 0; k < d; ++k)
        if (kind[k] != 0 && kind[k] != 1 && kind[k] != 3) return 1;
    if (f32) {
        const float *T = (const float *)table;
        ROWS_BODY(float)
    } else {
        const double *T = (const double *)table;
        ROWS_BODY(double)
    }
    return 0;
}
