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
    int tile;            // LDS values of the largest tile (with the bank padding and 64 dump slots)
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

// Y = M X along every line of one tile axis (length m <= MB * BLK, tile stride st), in place.  The factor's
// table (M zero-padded to P = MB BLK square) is staged in LDS at aoff with row stride P + 1 (bank spread).  For
// MB = 1 pairs of BLK-line blocks are dealt to the waves and run as two independent accumulator chains (MB = 2
// has two row blocks, so two chains, already); a lane's A operand for
// K-step k and row block I is M[I BLK + (lane % BLK)][k KS + lane / BLK], its B operand the tile value of line
// (block BLK + lane % BLK) at position k KS + lane / BLK.
template <typename T, int MB>
__device__ __forceinline__ void sep_axis_mfma(int toff, int nvals, int m, int st, int aoff) {
    using F = SepMfma<T>;
    constexpr int KMAX = MB * F::BLK / F::KS;
    constexpr int P = MB * F::BLK;
    constexpr int NH = MB == 1 ? 2 : 1;     // line blocks per wave step: independent accumulator chains
    m = __builtin_amdgcn_readfirstlane(m);
    st = __builtin_amdgcn_readfirstlane(st);
    nvals = __builtin_amdgcn_readfirstlane(nvals);
    T *tile = reinterpret_cast<T *>(g_dyn_lds) + toff;
    const T *At = reinterpret_cast<const T *>(g_dyn_lds) + aoff;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane % F::BLK, lk = lane / F::BLK;
    T a[MB][KMAX];
#pragma unroll
    for (int I = 0; I < MB; ++I)
#pragma unroll
        for (int k = 0; k < KMAX; ++k) a[I][k] = At[(I * F::BLK + li) * (P + 1) + k * F::KS + lk];
    const int nlines = nvals / m;
    const int nblk = (nlines + F::BLK - 1) / F::BLK;
    const int span = m * st;
    const float inv_st = 1.0f / (float)st;
    const int dump = sep_lds(nvals) + 1 + lane;
    for (int b0 = NH * wave; b0 < nblk; b0 += NH * BLML_NWAVE) {
        int base[NH];
        bool valid[NH];
#pragma unroll
        for (int h = 0; h < NH; ++h) {
            const int line = (b0 + h) * F::BLK + li;
            valid[h] = line < nlines;             // lines past the end read line 0 and are never stored
            int lo;
            const int hi = fdivmod(valid[h] ? line : 0, st, inv_st, lo);
            base[h] = hi * span + lo;
        }
        // B operands of both blocks, all loads first (clamped positions past m meet zero columns of A)
        T xv[NH][KMAX];
#pragma unroll
        for (int h = 0; h < NH; ++h)
#pragma unroll
            for (int k = 0; k < KMAX; ++k) xv[h][k] = tile[sep_lds(base[h] + min(k * F::KS + lk, m - 1) * st)];
        typename F::acc_t acc[NH][MB];
#pragma unroll
        for (int h = 0; h < NH; ++h)
#pragma unroll
            for (int I = 0; I < MB; ++I)
#pragma unroll
                for (int r = 0; r < F::NACC; ++r) acc[h][I][r] = (T)0;
#pragma unroll
        for (int k = 0; k < KMAX; ++k)
#pragma unroll
            for (int I = 0; I < MB; ++I)
#pragma unroll
                for (int h = 0; h < NH; ++h) acc[h][I] = F::step(a[I][k], xv[h][k], acc[h][I]);
        // stores of rows past m or lines past the end go to this lane's dump slot (no branches)
#pragma unroll
        for (int h = 0; h < NH; ++h)
#pragma unroll
            for (int I = 0; I < MB; ++I)
#pragma unroll
                for (int r = 0; r < F::NACC; ++r) {
                    const int row = I * F::BLK + F::row(r, lane);
                    tile[(valid[h] && row < m) ? sep_lds(base[h] + row * st) : dump] = acc[h][I][r];
                }
    }
}

template <typename T>
__device__ __forceinline__ void sep_axis(int toff, int nvals, int m, int st, int aoff) {
    if (m <= SepMfma<T>::BLK) sep_axis_mfma<T, 1>(toff, nvals, m, st, aoff);
    else if (m <= 2 * SepMfma<T>::BLK) sep_axis_mfma<T, 2>(toff, nvals, m, st, aoff);
}

// LDS values a staged table of an axis of length m takes (P rows of P + 1)
template <typename T>
__host__ __device__ inline int sep_stage_len(int m) {
    const int P = sep_table_dim(m, SepMfma<T>::BLK);
    return P * (P + 1);
}

// Moves one tile between global memory and LDS (UB accesses per thread in flight).  The tile's valid values are
// one contiguous run of obv R w values when w == rs, else obv R rows of wv values at stride rs (w then a power of
// two, so tile index e = row w + col splits by shifts).  Offsets within the tile's region are 32-bit (every fit's
// values stay below 2^31).  Loads are branch-free: an invalid slot reads the region's first value and keeps 0.
// load: tile = (T)(src - srcb); store: dst = (To)(tile + dstb) on the valid values.  Both compute in FP64.
template <typename T, typename G, bool LOAD>
__device__ __forceinline__ void sep_move(const SepPass &ps, long gbase, int obv, int wv, const G *src, const G *srcb,
                                         G *dst, const G *dstb) {
    T *tile = reinterpret_cast<T *>(g_dyn_lds);
    const int tsize = ps.OB * ps.R * ps.w;
    const bool contig = ps.w == ps.rs;
    const int nrow = obv * ps.R;
    const int nval = contig ? nrow * ps.w : 0;
    const int lw = 31 - __clz(ps.w), wm = ps.w - 1;
    const int rs = (int)ps.rs;
    const G *sp = LOAD ? src + gbase : nullptr;
    const G *sbp = (LOAD && srcb) ? srcb + gbase : nullptr;
    G *dp = LOAD ? nullptr : dst + gbase;
    const G *dbp = (!LOAD && dstb) ? dstb + gbase : nullptr;
    constexpr int UB = sizeof(G) == 4 ? 32 : 16;
    for (int e0 = threadIdx.x; e0 < tsize; e0 += UB * BLML_NT) {
        int off[UB];
        bool ok[UB];
#pragma unroll
        for (int u = 0; u < UB; ++u) {
            const int e = e0 + u * BLML_NT;
            if (contig) {
                ok[u] = e < nval;
                off[u] = ok[u] ? e : 0;
            } else {
                const int row = e >> lw, col = e & wm;
                ok[u] = row < nrow && col < wv;
                off[u] = ok[u] ? row * rs + col : 0;
            }
        }
        if (LOAD) {
            double v[UB];
#pragma unroll
            for (int u = 0; u < UB; ++u) v[u] = (double)sp[off[u]];
            if (sbp) {
#pragma unroll
                for (int u = 0; u < UB; ++u) v[u] -= (double)sbp[off[u]];
            }
#pragma unroll
            for (int u = 0; u < UB; ++u) {
                const int e = e0 + u * BLML_NT;
                if (e < tsize) tile[sep_lds(e)] = ok[u] ? (T)v[u] : (T)0;
            }
        } else {
            double v[UB];
#pragma unroll
            for (int u = 0; u < UB; ++u) v[u] = (double)tile[sep_lds(ok[u] ? e0 + u * BLML_NT : 0)];
            if (dbp) {
#pragma unroll
                for (int u = 0; u < UB; ++u) v[u] += (double)dbp[off[u]];
            }
#pragma unroll
            for (int u = 0; u < UB; ++u)
                if (ok[u]) dp[off[u]] = (G)v[u];
        }
    }
}

// One pass of a separable mode: dst = (prod over the pass's axes of M_a) src.  The first pass may subtract a base
// from its input, the last may add a base to its output (the incremental form U = Ubase + S (E - Ebase)); values
// are converted Ti -> T on load and T -> To on store (the base arithmetic in FP64).  Tiles are dealt to the team's
// members round robin.
template <typename T, typename Ti, typename To>
__device__ __forceinline__ void sep_pass(const SepPass &ps, int d, int nc, const int *shape, const int *fac_of, const T *tabs,
                         const long *tab_off, const Ti *src, const Ti *srcb, To *dst, const To *dstb, int rank,
                         int nranks) {
    (void)nc;
    (void)d;
    const int tsize = ps.OB * ps.R * ps.w;
    // the operator tables of the pass's axes, staged once in LDS after the tile and its dump slots (one batched
    // loop over all of them: the loads of a batch are in flight together)
    int aoff[BLML_MAXD], tdim[BLML_MAXD], tcum[BLML_MAXD + 1];
    const T *tsrc[BLML_MAXD];
    int nst = 0;
    {
        int o = sep_lds(tsize) + 1 + 64;
        tcum[0] = 0;
        for (int a = ps.lo; a < ps.hi; ++a) {
            aoff[a] = -1;
            if (fac_of[a] < 0 || shape[a] < 2) continue;
            const int P = sep_table_dim(shape[a], SepMfma<T>::BLK);
            aoff[a] = o;
            tdim[nst] = P;
            tsrc[nst] = tabs + tab_off[fac_of[a]];
            tcum[nst + 1] = tcum[nst] + P * P;
            ++nst;
            o += sep_stage_len<T>(shape[a]);
        }
        T *lds = reinterpret_cast<T *>(g_dyn_lds);
        constexpr int SB = 8;
        for (int j0 = threadIdx.x; j0 < tcum[nst]; j0 += SB * BLML_NT) {
            T v[SB];
            int dsto[SB];
#pragma unroll
            for (int u = 0; u < SB; ++u) {
                const int j = j0 + u * BLML_NT;
                dsto[u] = -1;
                v[u] = (T)0;
                if (j < tcum[nst]) {
                    int q = 0, ao = sep_lds(tsize) + 1 + 64;
                    while (j >= tcum[q + 1]) {
                        ao += tdim[q] * (tdim[q] + 1);
                        ++q;
                    }
                    const int i = j - tcum[q], P = tdim[q];
                    v[u] = tsrc[q][i];
                    dsto[u] = ao + (i / P) * (P + 1) + i % P;
                }
            }
#pragma unroll
            for (int u = 0; u < SB; ++u)
                if (dsto[u] >= 0) lds[dsto[u]] = v[u];
        }
        __syncthreads();
    }
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
                sep_axis<T>(0, tsize, shape[a], st, aoff[a]);
                __syncthreads();
            }
            st *= shape[a];
        }
        sep_move<T, To, false>(ps, gbase, obv, wv, nullptr, nullptr, dst, dstb);
        __syncthreads();
    }
}
