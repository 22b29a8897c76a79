// Mat-vec of a separable mode (every factor a LINE factor on a short axis) by dense line operators applied to
// LDS tiles that hold whole groups of axes, so that one pass over HBM covers several axes.
//
// A LINE factor of length m is a symmetric m x m operator M: Toeplitz M[i][j] = data[|i - j|] for a padded axis
// (the zero-padded FFT convolution of the reference, blml_binned.py:156-165, is exactly this product), circulant
// M[i][j] = c[(i - j) mod m], c[k] = sum_q spec_q cos(2 pi F_q k / m), for a circular one.  Along an axis of a
// tile, Y = M X over blocks of lines runs on the matrix cores: M's columns sit in the A operand (registers, loaded
// once per axis), each K-step reads one value per lane of the lines' block from LDS (the B operand), and the
// accumulators are written back in place.  FP32: v_mfma_f32_32x32x2f32 (32 rows x 32 lines, K = 2); FP64:
// v_mfma_f64_16x16x4f64 (16 x 16, K = 4).
//
// Passes: the lattice (axes 0..d-1, then the nc components) is cut into groups of consecutive axes [lo, hi).  A
// pass's tile is OB outer indices x R = prod m[lo:hi] rows x w contiguous values of the inner extent
// rs = nc * prod m[hi:], i.e. element (ob, r, col) of tile (o0, c0) is at (o0 + ob) R rs + r rs + c0 + col.
// The innermost pass reads contiguous blocks; outer passes read rows of w contiguous values (w a power of two).
#pragma once
#include "common.hpp"

#define BLML_SEP_MAXPASS 5

struct SepPass {
    int lo, hi;          // axes of the pass
    int OB, R, w;        // tile: OB outer indices x R rows x w values
    int pad_;
    long rs;             // inner extent (values) = row stride
    long nout;           // outer indices
    long ntiles;         // ceil(nout / OB) * ceil(rs / w)
    long nchunks;        // ceil(rs / w)
};

struct SepPlan {
    int npass;
    int tile;            // LDS values of the largest tile (with the bank padding)
    int pq;              // LDS values of the largest pass's operator tables
    int pad_;
    SepPass pass[BLML_SEP_MAXPASS];
};

// a factor's table: M zero-padded to P x P, P = the MFMA block (BLK) or twice it
__host__ __device__ inline int sep_table_dim(int m, int blk) { return m <= blk ? blk : 2 * blk; }
// LDS address of tile value e: one pad value every 32, so lines of 32 consecutive values and columns of a
// power-of-two stride spread over the banks
__device__ __forceinline__ int sep_lds(int e) { return e + (e >> 5); }

template <typename T>
struct SepMfma;
template <>
struct SepMfma<float> {
    static constexpr int BLK = 32, KS = 2, NACC = 16, MMAX = 64;
    typedef float acc_t __attribute__((ext_vector_type(16)));
    __device__ static acc_t step(float a, float b, acc_t c) { return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0); }
    // accumulator register r of lane l holds row (r & 3) + 8 (r >> 2) + 4 (l >> 5), line l & 31
    __device__ static int row(int r, int l) { return (r & 3) + 8 * (r >> 2) + 4 * (l >> 5); }
};
template <>
struct SepMfma<double> {
    static constexpr int BLK = 16, KS = 4, NACC = 4, MMAX = 32;
    typedef double acc_t __attribute__((ext_vector_type(4)));
    __device__ static acc_t step(double a, double b, acc_t c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    // accumulator register r of lane l holds row 4 r + (l >> 4), line l & 15
    __device__ static int row(int r, int l) { return 4 * r + (l >> 4); }
};

// Y = M X along every line of one tile axis (length m <= MB * BLK, tile stride st), in place.  M: the factor's
// table (global, zero-padded to MB BLK x MB BLK).  Blocks of BLK lines are dealt to the waves; a lane's A operand for K-step k and row
// block I is M[I BLK + (lane % BLK)][k KS + lane / BLK], its B operand the tile value of line
// (block BLK + lane % BLK) at position k KS + lane / BLK.
template <typename T, int MB>
__device__ void sep_axis_mfma(int toff, int nvals, int m, int st, const T *Mt) {
    using F = SepMfma<T>;
    constexpr int KMAX = MB * F::BLK / F::KS;
    m = __builtin_amdgcn_readfirstlane(m);
    st = __builtin_amdgcn_readfirstlane(st);
    nvals = __builtin_amdgcn_readfirstlane(nvals);
    T *tile = reinterpret_cast<T *>(g_dyn_lds) + toff;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane % F::BLK, lk = lane / F::BLK;
    T a[MB][KMAX];
    constexpr int P = MB * F::BLK;       // the table is M zero-padded to P x P: no bounds tests
#pragma unroll
    for (int I = 0; I < MB; ++I)
#pragma unroll
        for (int k = 0; k < KMAX; ++k) a[I][k] = Mt[(I * F::BLK + li) * P + k * F::KS + lk];
    const int nlines = nvals / m;
    const int nblk = (nlines + F::BLK - 1) / F::BLK;
    const int span = m * st;
    const float inv_st = 1.0f / (float)st;
    for (int blk = wave; blk < nblk; blk += BLML_NWAVE) {
        const int line = blk * F::BLK + li;
        const bool valid = line < nlines;
        int lo;
        const int hi = fdivmod(valid ? line : 0, st, inv_st, lo);
        const int base = hi * span + lo;
        // the block's B operands, all loads first (clamped positions past m meet zero columns of A; lines past the
        // end read line 0 and are never stored)
        T xv[KMAX];
#pragma unroll
        for (int k = 0; k < KMAX; ++k) xv[k] = tile[sep_lds(base + min(k * F::KS + lk, m - 1) * st)];
        typename F::acc_t acc[MB];
#pragma unroll
        for (int I = 0; I < MB; ++I)
#pragma unroll
            for (int r = 0; r < F::NACC; ++r) acc[I][r] = (T)0;
#pragma unroll
        for (int k = 0; k < KMAX; ++k)
#pragma unroll
            for (int I = 0; I < MB; ++I) acc[I] = F::step(a[I][k], xv[k], acc[I]);
#pragma unroll
        for (int I = 0; I < MB; ++I)
#pragma unroll
            for (int r = 0; r < F::NACC; ++r) {
                const int row = I * F::BLK + F::row(r, lane);
                if (valid && row < m) tile[sep_lds(base + row * st)] = acc[I][r];
            }
    }
}

template <typename T>
__device__ __noinline__ void sep_axis(int toff, int nvals, int m, int st, const T *Mt) {
    if (m <= SepMfma<T>::BLK) sep_axis_mfma<T, 1>(toff, nvals, m, st, Mt);
    else if (m <= 2 * SepMfma<T>::BLK) sep_axis_mfma<T, 2>(toff, nvals, m, st, Mt);
}

// Moves one tile between global memory and LDS, one value per thread at a time (UB values per thread in flight).
// load: tile = (T)(src - srcb), zero where the tile has no value; store: dst = (To)(tile + dstb) on the valid
// values only.  Both compute in FP64.
template <typename T, typename G, bool LOAD>
__device__ __noinline__ void sep_move(const SepPass &ps, long gbase, int obv, int wv, const G *src, const G *srcb,
                                      G *dst, const G *dstb) {
    T *tile = reinterpret_cast<T *>(g_dyn_lds);
    const int tsize = ps.OB * ps.R * ps.w;
    const int RW = ps.R * ps.w;
    const float inv_w = 1.0f / (float)ps.w, inv_rw = 1.0f / (float)RW;
    constexpr int UB = 8;
    for (int e0 = threadIdx.x; e0 < tsize; e0 += UB * BLML_NT) {
        T v[UB];
        long g[UB];
#pragma unroll
        for (int u = 0; u < UB; ++u) {
            const int e = e0 + u * BLML_NT;
            g[u] = -1;
            if (e < tsize) {
                int col, r;
                const int q = fdivmod(e, ps.w, inv_w, col);
                const int ob = fdivmod(e, RW, inv_rw, r);
                r = q - ob * ps.R;
                if (ob < obv && col < wv) g[u] = gbase + ((long)ob * ps.R + r) * ps.rs + col;
            }
        }
        if (LOAD) {
#pragma unroll
            for (int u = 0; u < UB; ++u) {
                v[u] = (T)0;
                if (g[u] >= 0) {
                    double x = (double)src[g[u]];
                    if (srcb) x -= (double)srcb[g[u]];
                    v[u] = (T)x;
                }
            }
#pragma unroll
            for (int u = 0; u < UB; ++u) {
                const int e = e0 + u * BLML_NT;
                if (e < tsize) tile[sep_lds(e)] = v[u];
            }
        } else {
#pragma unroll
            for (int u = 0; u < UB; ++u)
                if (g[u] >= 0) v[u] = tile[sep_lds(e0 + u * BLML_NT)];
#pragma unroll
            for (int u = 0; u < UB; ++u) {
                if (g[u] >= 0) {
                    double x = (double)v[u];
                    if (dstb) x += (double)dstb[g[u]];
                    dst[g[u]] = (G)x;
                }
            }
        }
    }
}

// One pass of a separable mode: dst = (prod over the pass's axes of M_a) src.  The first pass may subtract a base
// from its input, the last may add a base to its output (the incremental form U = Ubase + S (E - Ebase)); values
// are converted Ti -> T on load and T -> To on store (the base arithmetic in FP64).  Tiles are dealt to the team's
// members round robin.
template <typename T, typename Ti, typename To>
__device__ void sep_pass(const SepPass &ps, int d, int nc, const int *shape, const int *fac_of, const T *tabs,
                         const long *tab_off, const Ti *src, const Ti *srcb, To *dst, const To *dstb, int rank,
                         int nranks) {
    (void)nc;
    (void)d;
    const int tsize = ps.OB * ps.R * ps.w;
    for (long t = rank; t < ps.ntiles; t += nranks) {
        const long to = t / ps.nchunks, tc = t - to * ps.nchunks;
        const long o0 = to * ps.OB, c0 = tc * ps.w;
        const int obv = (int)((ps.nout - o0) < ps.OB ? (ps.nout - o0) : ps.OB);
        const int wv = (int)((ps.rs - c0) < ps.w ? (ps.rs - c0) : ps.w);
        const long gbase = o0 * ps.R * ps.rs + c0;
        sep_move<T, Ti, true>(ps, gbase, obv, wv, src, srcb, nullptr, nullptr);
        __syncthreads();
        // the pass's axes, innermost first: axis a has tile stride w * prod m[a+1:hi]
        int st = ps.w;
        for (int a = ps.hi - 1; a >= ps.lo; --a) {
            if (fac_of[a] >= 0 && shape[a] > 1) {
                sep_axis<T>(0, tsize, shape[a], st, tabs + tab_off[fac_of[a]]);
                __syncthreads();
            }
            st *= shape[a];
        }
        sep_move<T, To, false>(ps, gbase, obv, wv, nullptr, nullptr, dst, dstb);
        __syncthreads();
    }
}
