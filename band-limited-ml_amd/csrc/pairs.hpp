// Pair accumulation over lattice bins: rows s_r = sum_j S(ZA_r, ZB_j) w_j and the explicit block S(ZA, ZB), with
// the kernel read from the flat value tables of the reference's ProductKernel.tables (blml_kernels.py:133-200) as
// its row kernels read them (blml_local.py:86-96 _col_index, 230-253; blml_local_gpu.py:42-105, 307-361).
//
// rows: one thread per row r (its coordinates in registers); the sources are split into slices (grid.y) and each
// slice is walked in chunks staged in LDS (coordinates and weights: every thread reads the same source, an LDS
// broadcast).  A table small enough sits in LDS too.  Each (row, slice) partial is an FP64 sum in source order;
// a second kernel adds the slices in order, so the result does not depend on scheduling.  Factors multiply in
// column order, then the groups (the reference's rows multiply in unit order: the same product up to rounding).
#pragma once
#include "common.hpp"
#include "blml_hip.h"

#define BLML_PAIR_NT 256
#define BLML_PAIR_CH 256         // sources per LDS chunk
#define BLML_RADIAL_Q 4.0        // radial rows: grid points per unit of lattice distance (blml_trivial.py:218)

// The table index of a column factor (kinds 0, 1, 3), clamped into [0, nT) (a coordinate outside the table's span
// is a stale table: the binding checks spans before the launch).
__device__ __forceinline__ long pair_col_index(int kd, long za, long zb, long base, long J, long mk, long nb, long nT) {
    long i;
    if (kd == 0) {
        i = base + za - zb + J;
    } else if (kd == 3) {
        i = base + za * mk + zb;
    } else {
        long m = za - zb;                   // |za - zb| <= nb: fold with branches
        if (m < 0) m += nb;
        else if (m >= nb) m -= nb;
        if (m > nb - m) m = nb - m;
        i = base + m;
    }
    return i < 0 ? 0 : (i >= nT ? nT - 1 : i);
}

// S(za, zb) = vn * vd: vd the product over the columns in depmask (groups with any column there), vn over the
// others; zb: a pointer to d coordinates.
template <typename R_>
__device__ __forceinline__ void pair_value(const blml_pairtab_t &t, const R_ *T, const R_ *RT, const int *za,
                                           const int *zb, long depmask, R_ &vn, R_ &vdep) {
    R_ v = (R_)1, vd = (R_)1;
    long r2[BLML_PAIR_MAXG];
    bool tor[BLML_PAIR_MAXG], gdep[BLML_PAIR_MAXG];
#pragma unroll
    for (int g = 0; g < BLML_PAIR_MAXG; ++g) { r2[g] = 0; tor[g] = false; gdep[g] = false; }
#pragma unroll
    for (int k = 0; k < BLML_PAIR_MAXD; ++k) {
        if (k < t.d) {
            const int kd = t.kind[k];
            const long a = za[k], b = zb[k], off = a - b;
            const bool dep = (depmask >> k) & 1;
            if (kd == 2 || kd == 4) {
                const int gk = t.grp[k];
#pragma unroll
                for (int g = 0; g < BLML_PAIR_MAXG; ++g) {
                    if (g == gk) {
                        if (kd == 2) {
                            r2[g] += off * off;
                        } else {             // torus: row-major circular offset over the group's columns
                            const long nb = t.nbins[k];
                            r2[g] = r2[g] * nb + (off < 0 ? off + nb : off);
                            tor[g] = true;
                        }
                        gdep[g] = gdep[g] || dep;
                    }
                }
            } else {
                const R_ tv = T[pair_col_index(kd, a, b, t.base[k], t.J[k], t.mk[k], t.nbins[k], t.nT)];
                if (dep) vd *= tv;
                else v *= tv;
            }
        }
    }
#pragma unroll
    for (int g = 0; g < BLML_PAIR_MAXG; ++g) {
        if (g < t.G) {
            R_ tv;
            if (tor[g]) {
                const long i = r2[g] < t.LR ? r2[g] : t.LR - 1;
                tv = RT[(long)g * t.LR + i];
            } else {
                const double x = sqrt((double)r2[g]) * BLML_RADIAL_Q;
                long i0 = (long)x;
                const double f = x - (double)i0;
                if (i0 > t.LR - 2) i0 = t.LR - 2;
                tv = (R_)((double)RT[(long)g * t.LR + i0] * (1.0 - f) + (double)RT[(long)g * t.LR + i0 + 1] * f);
            }
            if (gdep[g]) vd *= tv;
            else v *= tv;
        }
    }
    vn = v;
    vdep = vd;
}

// part[slice][r] = sum over the slice's sources j (w_j != 0) of S(ZA_r, ZB_j) w_j.  LDS: [T if TL][ZB chunk][w chunk].
template <typename R_, bool TL>
__global__ void __launch_bounds__(BLML_PAIR_NT) pair_rows_kernel(blml_pairtab_t t, const int *ZA, const int *ZB,
                                                                 const double *w, int R, int M, int slice,
                                                                 double *part) {
    unsigned char *lds = g_dyn_lds;
    const long tbytes = TL ? ((t.nT * (long)sizeof(R_) + 15) / 16) * 16 : 0;
    R_ *lT = reinterpret_cast<R_ *>(lds);
    int *lz = reinterpret_cast<int *>(lds + tbytes);
    double *lw = reinterpret_cast<double *>(lds + tbytes + ((BLML_PAIR_CH * BLML_PAIR_MAXD * 4 + 15) / 16) * 16);
    const R_ *T = reinterpret_cast<const R_ *>(t.T);
    if (TL) {
        for (long i = threadIdx.x; i < t.nT; i += BLML_PAIR_NT) lT[i] = T[i];
        T = lT;
    }
    const R_ *RT = reinterpret_cast<const R_ *>(t.RT);
    const int d = t.d;
    const int r = blockIdx.x * BLML_PAIR_NT + threadIdx.x;
    const bool has = r < R;
    int za[BLML_PAIR_MAXD];
#pragma unroll
    for (int k = 0; k < BLML_PAIR_MAXD; ++k) za[k] = (has && k < d) ? ZA[(long)r * d + k] : 0;
    const long j0 = (long)blockIdx.y * slice;
    const long j1 = j0 + slice < M ? j0 + slice : M;
    double acc = 0.0;
    for (long c0 = j0; c0 < j1; c0 += BLML_PAIR_CH) {
        const int n = (int)(j1 - c0 < BLML_PAIR_CH ? j1 - c0 : BLML_PAIR_CH);
        __syncthreads();
        for (int i = threadIdx.x; i < n * d; i += BLML_PAIR_NT) lz[i] = ZB[c0 * d + i];
        for (int i = threadIdx.x; i < n; i += BLML_PAIR_NT) lw[i] = w[c0 + i];
        __syncthreads();
        for (int jj = 0; jj < n; ++jj) {
            const double wj = lw[jj];
            if (wj == 0.0) continue;
            R_ vn, vd;
            pair_value<R_>(t, T, RT, za, lz + jj * d, 0, vn, vd);
            acc += (double)vn * wj;
        }
    }
    if (has) part[(long)blockIdx.y * R + r] = acc;
}

// s[r] = sum over slices, in slice order
__global__ void pair_rows_sum_kernel(const double *part, int R, int nslice, double *s) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    double a = 0.0;
    for (int y = 0; y < nslice; ++y) a += part[(long)y * R + r];
    s[r] = a;
}

// out[r * M + j] = S(ZA_r, ZB_j); with depmask, out_m[r * M + j] = the product over the columns not in depmask.
template <typename R_>
__global__ void __launch_bounds__(BLML_PAIR_NT) pair_cols_kernel(blml_pairtab_t t, const int *ZA, const int *ZB,
                                                                 long R, long M, long depmask, R_ *out, R_ *out_m) {
    const R_ *T = reinterpret_cast<const R_ *>(t.T);
    const R_ *RT = reinterpret_cast<const R_ *>(t.RT);
    const int d = t.d;
    const long n = R * M;
    for (long idx = (long)blockIdx.x * BLML_PAIR_NT + threadIdx.x; idx < n; idx += (long)gridDim.x * BLML_PAIR_NT) {
        const long r = idx / M, j = idx - r * M;
        int za[BLML_PAIR_MAXD], zb[BLML_PAIR_MAXD];
#pragma unroll
        for (int k = 0; k < BLML_PAIR_MAXD; ++k) {
            za[k] = k < d ? ZA[r * d + k] : 0;
            zb[k] = k < d ? ZB[j * d + k] : 0;
        }
        R_ vn, vd;
        pair_value<R_>(t, T, RT, za, zb, depmask, vn, vd);
        if (depmask != 0) out_m[idx] = vn;
        out[idx] = vn * vd;
    }
}
