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

// A group's factor from its row of LR values: a torus table read at the row-major circular offset r2, or a row on
// the r grid interpolated linearly at sqrt(r2) * RADIAL_Q (indexes clamped into the row).
template <typename R_>
__device__ __forceinline__ R_ pair_group_row(const R_ *row, long LR, long r2, bool tor) {
    if (tor) return row[r2 < LR ? r2 : LR - 1];
    const double x = sqrt((double)r2) * BLML_RADIAL_Q;
    long i0 = (long)x;
    const double f = x - (double)i0;
    if (i0 > LR - 2) i0 = LR - 2;
    return (R_)((double)row[i0] * (1.0 - f) + (double)row[i0 + 1] * f);
}

// S(za, zb) = vn * vd: vd the product over the columns in depmask (groups with any column there), vn over the
// others; zb: a pointer to d coordinates.  NG: the groups the instantiation holds (t.G <= NG); NG = 0 for tables
// without groups (every column of kind 0, 1 or 3, as the binding checks) compiles the group code out.
template <typename R_, int NG = BLML_PAIR_MAXG>
__device__ __forceinline__ void pair_value(const blml_pairtab_t &t, const R_ *T, const R_ *RT, const int *za,
                                           const int *zb, long depmask, R_ &vn, R_ &vdep) {
    constexpr int NGA = NG > 0 ? NG : 1;
    R_ v = (R_)1, vd = (R_)1;
    long r2[NGA];
    bool tor[NGA], gdep[NGA];
#pragma unroll
    for (int g = 0; g < NGA; ++g) { r2[g] = 0; tor[g] = false; gdep[g] = false; }
#pragma unroll
    for (int k = 0; k < BLML_PAIR_MAXD; ++k) {
        if (k < t.d) {
            const int kd = t.kind[k];
            const long a = za[k], b = zb[k], off = a - b;
            const bool dep = (depmask >> k) & 1;
            if (NG > 0 && (kd == 2 || kd == 4)) {
                const int gk = t.grp[k];
#pragma unroll
                for (int g = 0; g < NGA; ++g) {
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
    for (int g = 0; g < NGA; ++g) {
        if (NG > 0 && g < t.G) {
            const R_ tv = pair_group_row<R_>(RT + (long)g * t.LR, t.LR, r2[g], tor[g]);
            if (gdep[g]) vd *= tv;
            else v *= tv;
        }
    }
    vn = v;
    vdep = vd;
}

// The table index of a column factor (kinds 0, 1, 3) in 32-bit arithmetic (the binding keeps every table below 2^31
// entries and every coordinate below 2^30), brought into [0, nT) by one unsigned minimum: a coordinate outside the
// table's span (a stale table, which the binding rejects before the launch) reads a table entry, never past it.
__device__ __forceinline__ int pair_col_index32(int kd, int za, int zb, int baseJ, int base, int mk, int nb,
                                                unsigned nT1) {
    int i;
    if (kd == 0) {
        i = baseJ + (za - zb);
    } else if (kd == 3) {
        i = base + za * mk + zb;
    } else {
        int m = za - zb;
        if (m < 0) m += nb;
        else if (m >= nb) m -= nb;
        if (m > nb - m) m = nb - m;
        i = base + m;
    }
    return (int)min((unsigned)i, nT1);
}

// pair_rows_kernel for tables without radial or torus groups (every column of kind 0, 1 or 3), U sources at a time:
// per column one uniform branch on its kind, then the U sources' coordinates, indexes and table reads issue
// together (independent LDS reads in flight instead of one dependent chain per source).  Each pair's product runs
// in column order and the FP64 partial adds the sources in order, skipping zero weights: the same sums as
// pair_rows_kernel<R_, TL, 0>, to the bit.
template <typename R_, bool TL, int U>
__global__ void __launch_bounds__(BLML_PAIR_NT) pair_rows_u_kernel(blml_pairtab_t t, const int *ZA, const int *ZB,
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
    const int d = t.d;
    const unsigned nT1 = (unsigned)(t.nT - 1);
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
        for (int jj = 0; jj < n; jj += U) {
            R_ v[U];
            double wv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                v[u] = (R_)1;
                wv[u] = jj + u < n ? lw[jj + u] : 0.0;
            }
#pragma unroll
            for (int k = 0; k < BLML_PAIR_MAXD; ++k) {
                if (k < d) {
                    const int kd = t.kind[k];
                    const int base = (int)t.base[k], baseJ = (int)(t.base[k] + t.J[k]);
                    const int mk = (int)t.mk[k], nb = (int)t.nbins[k];
                    int idx[U];
#pragma unroll
                    for (int u = 0; u < U; ++u)
                        idx[u] = jj + u < n ? pair_col_index32(kd, za[k], lz[(jj + u) * d + k], baseJ, base, mk, nb,
                                                               nT1)
                                            : 0;
#pragma unroll
                    for (int u = 0; u < U; ++u) v[u] *= T[idx[u]];
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (wv[u] != 0.0) acc += (double)v[u] * wv[u];
        }
    }
    if (has) part[(long)blockIdx.y * R + r] = acc;
}

// The rows of a table in LDS whose D columns are all offset rows (kind 0: sinc / slits lattices, the commonest
// table): each thread folds its row's coordinates and the columns' bases into one constant per column, so a pair
// costs per column one subtraction, one LDS address, one table read and one product -- no per-column kind branch,
// no clamp (an index outside the table, which only a stale table gives and the binding rejects, reads other LDS of
// this workgroup or, past its allocation, zero: never memory outside it).  U sources at a time; the same products
// (column order) and the same ordered FP64 sums as pair_rows_kernel, to the bit (with finite tables).
template <typename R_, int D, int U>
__global__ void __launch_bounds__(BLML_PAIR_NT) pair_rows_off_kernel(blml_pairtab_t t, const int *ZA, const int *ZB,
                                                                     const double *w, int R, int M, int slice,
                                                                     double *part) {
    unsigned char *lds = g_dyn_lds;
    const long tbytes = ((t.nT * (long)sizeof(R_) + 15) / 16) * 16;
    R_ *lT = reinterpret_cast<R_ *>(lds);
    int *lz = reinterpret_cast<int *>(lds + tbytes);
    double *lw = reinterpret_cast<double *>(lds + tbytes + ((BLML_PAIR_CH * BLML_PAIR_MAXD * 4 + 15) / 16) * 16);
    const R_ *T = reinterpret_cast<const R_ *>(t.T);
    for (long i = threadIdx.x; i < t.nT; i += BLML_PAIR_NT) lT[i] = T[i];
    const int r = blockIdx.x * BLML_PAIR_NT + threadIdx.x;
    const bool has = r < R;
    // byte offsets: the sources' coordinates staged as zb sizeof(R_), so a pair's table entry sits zb sizeof(R_)
    // bytes before the row's entry point (base + J + za) sizeof(R_) of each column
    constexpr int ES = (int)sizeof(R_);
    const unsigned char *pk[D];       // the row's entry point into each column's table (LDS)
#pragma unroll
    for (int k = 0; k < D; ++k)
        pk[k] = reinterpret_cast<const unsigned char *>(lT) +
                ((int)(t.base[k] + t.J[k]) + (has ? ZA[(long)r * D + k] : 0)) * ES;
    const long j0 = (long)blockIdx.y * slice;
    const long j1 = j0 + slice < M ? j0 + slice : M;
    double acc = 0.0;
    for (long c0 = j0; c0 < j1; c0 += BLML_PAIR_CH) {
        const int n = (int)(j1 - c0 < BLML_PAIR_CH ? j1 - c0 : BLML_PAIR_CH);
        __syncthreads();
        for (int i = threadIdx.x; i < n * D; i += BLML_PAIR_NT) lz[i] = ZB[c0 * D + i] * ES;
        for (int i = threadIdx.x; i < n; i += BLML_PAIR_NT) lw[i] = w[c0 + i];
        // past the chunk's end (up to U - 1 sources): the chunk's first source again, so that every table read of
        // the last group of U stays inside the table (their weights are taken as zero and skipped below)
        for (int i = n * D + threadIdx.x; i < (n + U) * D && i < BLML_PAIR_CH * BLML_PAIR_MAXD; i += BLML_PAIR_NT)
            lz[i] = ZB[c0 * D + i % D] * ES;
        __syncthreads();
        for (int jj = 0; jj < n; jj += U) {
            R_ v[U];
            double wv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                v[u] = (R_)1;
                wv[u] = jj + u < n ? lw[jj + u] : 0.0;
            }
#pragma unroll
            for (int k = 0; k < D; ++k) {
#pragma unroll
                for (int u = 0; u < U; ++u)
                    v[u] *= *reinterpret_cast<const R_ *>(pk[k] - lz[(jj + u) * D + k]);
            }
            // a zero weight adds a zero product, which leaves the sum as it is (the tables are finite): the same
            // sum as skipping it
#pragma unroll
            for (int u = 0; u < U; ++u) acc += (double)v[u] * wv[u];
        }
    }
    if (has) part[(long)blockIdx.y * R + r] = acc;
}

// part[slice][r] = sum over the slice's sources j (w_j != 0) of S(ZA_r, ZB_j) w_j.  LDS: [T if TL][ZB chunk][w chunk].
template <typename R_, bool TL, int NG>
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
            pair_value<R_, NG>(t, T, RT, za, lz + jj * d, 0, vn, vd);
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

// ------------------------------------------------------------------------------------------------ fused rows
// Several densities in one pass, with derivatives (the reference's rows_multi_kernel, blml_local_gpu.py:107-180, and
// rows_kernel's grad branch, :94-103): density m has weights Wt[j * Mm + m] and a unit mask (bit u set: unit u's
// factor enters its product, the others count as 1).  Accumulator entry e is density em[e]'s sum of derivative
// entry eq[e]: its derivative value times the product of the density's other units, zero when the density's mask
// leaves the entry's unit out.  Every per-entry value is resolved on the host, so the kernel indexes its arrays with
// compile-time indexes only.

struct pair_multi_dev {
    int Mm, nE, U;
    unsigned umask[BLML_PAIR_MAXM];
    unsigned eumask[BLML_PAIR_MAXW];
    int em[BLML_PAIR_MAXW], eq[BLML_PAIR_MAXW], eunit[BLML_PAIR_MAXW], ecol[BLML_PAIR_MAXW], egrp[BLML_PAIR_MAXW];
    int erank[BLML_PAIR_MAXW];
    int dim_unit[BLML_PAIR_MAXD], grp_unit[BLML_PAIR_MAXG];
    long edbase[BLML_PAIR_MAXW], eJ[BLML_PAIR_MAXW], emk[BLML_PAIR_MAXW];
    long ndT;
    const void *dT, *dR;
};

// part[(slice * (Mm + nE) + m) * R + r] = the slice's sum of density m's S(ZA_r, ZB_j) Wt[j, m];
// part[(slice * (Mm + nE) + Mm + e) * R + r] = its sum for accumulator entry e.  Staging as pair_rows_kernel.
// NE: the accumulator entries the instantiation holds (nE <= NE; 4 keeps the common small passes light on registers);
// NG: the groups it holds (t.G <= NG; 0 compiles the group code out), as pair_rows_kernel.
template <typename R_, bool TL, int NE, int NG>
__global__ void __launch_bounds__(BLML_PAIR_NT) pair_rows_multi_kernel(blml_pairtab_t t, pair_multi_dev gd, const int *ZA,
                                                                       const int *ZB, const double *Wt, int R, int M,
                                                                       int slice, double *part) {
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
    const R_ *dT = reinterpret_cast<const R_ *>(gd.dT);
    const R_ *dR = reinterpret_cast<const R_ *>(gd.dR);
    const int d = t.d, Mm = gd.Mm;
    const int r = blockIdx.x * BLML_PAIR_NT + threadIdx.x;
    const bool has = r < R;
    int za[BLML_PAIR_MAXD];
#pragma unroll
    for (int k = 0; k < BLML_PAIR_MAXD; ++k) za[k] = (has && k < d) ? ZA[(long)r * d + k] : 0;
    const long j0 = (long)blockIdx.y * slice;
    const long j1 = j0 + slice < M ? j0 + slice : M;
    double acc[BLML_PAIR_MAXM], dacc[NE];
#pragma unroll
    for (int m = 0; m < BLML_PAIR_MAXM; ++m) acc[m] = 0.0;
#pragma unroll
    for (int e = 0; e < NE; ++e) dacc[e] = 0.0;
    for (long c0 = j0; c0 < j1; c0 += BLML_PAIR_CH) {
        const int n = (int)(j1 - c0 < BLML_PAIR_CH ? j1 - c0 : BLML_PAIR_CH);
        __syncthreads();
        for (int i = threadIdx.x; i < n * d; i += BLML_PAIR_NT) lz[i] = ZB[c0 * d + i];
        for (int i = threadIdx.x; i < n * Mm; i += BLML_PAIR_NT) lw[i] = Wt[c0 * Mm + i];
        __syncthreads();
        for (int jj = 0; jj < n; ++jj) {
            const double *wj = lw + jj * Mm;
            bool any = false;
#pragma unroll
            for (int m = 0; m < BLML_PAIR_MAXM; ++m)
                if (m < Mm && wj[m] != 0.0) any = true;
            if (!any) continue;
            const int *zb = lz + jj * d;
            constexpr int NGA = NG > 0 ? NG : 1;
            R_ fac[BLML_PAIR_MAXD];
            long r2[NGA];
            bool tor[NGA];
#pragma unroll
            for (int u = 0; u < BLML_PAIR_MAXD; ++u) fac[u] = (R_)1;
#pragma unroll
            for (int g = 0; g < NGA; ++g) { r2[g] = 0; tor[g] = false; }
#pragma unroll
            for (int k = 0; k < BLML_PAIR_MAXD; ++k) {
                if (k < d) {
                    const int kd = t.kind[k];
                    const long a = za[k], b = zb[k], off = a - b;
                    if (NG > 0 && (kd == 2 || kd == 4)) {
                        const int gk = t.grp[k];
#pragma unroll
                        for (int g = 0; g < NGA; ++g) {
                            if (g == gk) {
                                if (kd == 2) {
                                    r2[g] += off * off;
                                } else {
                                    const long nb = t.nbins[k];
                                    r2[g] = r2[g] * nb + (off < 0 ? off + nb : off);
                                    tor[g] = true;
                                }
                            }
                        }
                    } else {
                        const R_ tv = T[pair_col_index(kd, a, b, t.base[k], t.J[k], t.mk[k], t.nbins[k], t.nT)];
                        const int uk = gd.dim_unit[k];
#pragma unroll
                        for (int u = 0; u < BLML_PAIR_MAXD; ++u)
                            if (u == uk) fac[u] = tv;
                    }
                }
            }
#pragma unroll
            for (int g = 0; g < NGA; ++g) {
                if (NG > 0 && g < t.G) {
                    const R_ tv = pair_group_row<R_>(RT + (long)g * t.LR, t.LR, r2[g], tor[g]);
                    const int ug = gd.grp_unit[g];
#pragma unroll
                    for (int u = 0; u < BLML_PAIR_MAXD; ++u)
                        if (u == ug) fac[u] = tv;
                }
            }
#pragma unroll
            for (int m = 0; m < BLML_PAIR_MAXM; ++m) {
                if (m < Mm && wj[m] != 0.0) {
                    const unsigned um = gd.umask[m];
                    R_ v = (R_)1;
#pragma unroll
                    for (int u = 0; u < BLML_PAIR_MAXD; ++u)
                        if (u < gd.U && ((um >> u) & 1u)) v *= fac[u];
                    acc[m] += (double)v * wj[m];
                }
            }
#pragma unroll
            for (int e = 0; e < NE; ++e) {
                if (e < gd.nE) {
                    const double we = wj[gd.em[e]];
                    const unsigned um = gd.eumask[e];
                    const int uq = gd.eunit[e];
                    if (we != 0.0 && ((um >> uq) & 1u)) {
                        const int g = gd.egrp[e], kc = gd.ecol[e];
                        long a = 0;
#pragma unroll
                        for (int k = 0; k < BLML_PAIR_MAXD; ++k)
                            if (k == kc) a = za[k];
                        const long b = kc >= 0 ? zb[kc] : 0;
                        R_ dv;
                        if (NG > 0 && g >= 0) {
                            long rg = 0;
                            bool tg = false;
#pragma unroll
                            for (int gg = 0; gg < NGA; ++gg)
                                if (gg == g) { rg = r2[gg]; tg = tor[gg]; }
                            dv = pair_group_row<R_>(dR + (long)gd.eq[e] * t.LR, t.LR, rg, tg);
                            if (!tg && kc >= 0) dv *= (R_)(a - b);
                        } else {
                            long i = gd.erank[e] ? gd.edbase[e] + a * gd.emk[e] + b : gd.edbase[e] + a - b + gd.eJ[e];
                            i = i < 0 ? 0 : (i >= gd.ndT ? gd.ndT - 1 : i);
                            dv = dT[i];
                        }
                        R_ ex = (R_)1;
#pragma unroll
                        for (int u = 0; u < BLML_PAIR_MAXD; ++u)
                            if (u < gd.U && u != uq && ((um >> u) & 1u)) ex *= fac[u];
                        dacc[e] += (double)(dv * ex) * we;
                    }
                }
            }
        }
    }
    if (has) {
        const long at = (long)blockIdx.y * (Mm + gd.nE);
#pragma unroll
        for (int m = 0; m < BLML_PAIR_MAXM; ++m)
            if (m < Mm) part[(at + m) * R + r] = acc[m];
#pragma unroll
        for (int e = 0; e < NE; ++e)
            if (e < gd.nE) part[(at + Mm + e) * R + r] = dacc[e];
    }
}

// s[r * Mm + m], ds[r * nE + e]: the slices' partials added in slice order
__global__ void pair_rows_multi_sum_kernel(const double *part, int R, int nslice, int Mm, int nE, double *s,
                                           double *ds) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    const long stride = (long)(Mm + nE) * R;
    for (int m = 0; m < Mm + nE; ++m) {
        double a = 0.0;
        for (int y = 0; y < nslice; ++y) a += part[y * stride + (long)m * R + r];
        if (m < Mm) s[(long)r * Mm + m] = a;
        else ds[(long)r * nE + (m - Mm)] = a;
    }
}

// out[r * M + j] = S(ZA_r, ZB_j); with depmask, out_m[r * M + j] = the product over the columns not in depmask.
template <typename R_, int NG>
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
        pair_value<R_, NG>(t, T, RT, za, zb, depmask, vn, vd);
        if (depmask != 0) out_m[idx] = vn;
        out[idx] = vn * vd;
    }
}
