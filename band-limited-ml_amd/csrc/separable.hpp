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
#include <type_traits>

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
        __builtin_amdgcn_sched_barrier(0);        // all B operands in flight before the first product
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

// The same product for lines of exactly BLK values whose blocks pair up evenly (nlines a multiple of 2 BLK) and
// whose stride st is 1 or a multiple of 32.  Then position j of a line starting at tile index base sits at LDS
// sep_lds(base) + j stp: with st a multiple of 32, adding j st leaves base's low five bits alone, so the bank
// padding grows by j st / 32 (stp = st + st / 32); with st = 1 a line is 32-aligned (base = line BLK, a multiple
// of 32 for BLK = 32; for BLK = 16 the line stays inside one 32-value run), so stp = 1.  Every row is below m and
// every line exists: no clamps and no dump slots.
#ifndef BLML_SEP_FULL_NH
#define BLML_SEP_FULL_NH 2
#endif
template <typename T>
__device__ __forceinline__ void sep_axis_mfma_full(int nvals, int st, int aoff) {
    using F = SepMfma<T>;
    constexpr int BLK = F::BLK, KMAX = BLK / F::KS, NH = BLML_SEP_FULL_NH;
    st = __builtin_amdgcn_readfirstlane(st);
    nvals = __builtin_amdgcn_readfirstlane(nvals);
    T *tile = reinterpret_cast<T *>(g_dyn_lds);
    const T *At = reinterpret_cast<const T *>(g_dyn_lds) + aoff;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane % BLK, lk = lane / BLK;
    T a[KMAX];
#pragma unroll
    for (int k = 0; k < KMAX; ++k) a[k] = At[li * (BLK + 1) + k * F::KS + lk];
    const int nblk = nvals / (BLK * BLK);
    const int span = BLK * st;
    const int stp = st == 1 ? 1 : st + (st >> 5);
    const float inv_st = 1.0f / (float)st;
    const int lane_row = F::row(0, lane) * stp;     // row(r, lane) = row(r, 0) + row(0, lane): a uniform part per r
    for (int b0 = NH * wave; b0 < nblk; b0 += NH * BLML_NWAVE) {
        int lb[NH];
#pragma unroll
        for (int h = 0; h < NH; ++h) {
            int lo;
            const int hi = fdivmod((b0 + h) * BLK + li, st, inv_st, lo);
            lb[h] = sep_lds(hi * span + lo);
        }
        T xv[NH][KMAX];
#pragma unroll
        for (int h = 0; h < NH; ++h)
#pragma unroll
            for (int k = 0; k < KMAX; ++k) xv[h][k] = tile[lb[h] + (k * F::KS + lk) * stp];
        // every B operand is read before the first product: left to itself the scheduler sinks each LDS read to
        // just before its product, which then waits out a full LDS round trip per product
        __builtin_amdgcn_sched_barrier(0);
        typename F::acc_t acc[NH];
#pragma unroll
        for (int h = 0; h < NH; ++h)
#pragma unroll
            for (int r = 0; r < F::NACC; ++r) acc[h][r] = (T)0;
#pragma unroll
        for (int k = 0; k < KMAX; ++k)
#pragma unroll
            for (int h = 0; h < NH; ++h) acc[h] = F::step(a[k], xv[h][k], acc[h]);
#pragma unroll
        for (int h = 0; h < NH; ++h)
#pragma unroll
            for (int r = 0; r < F::NACC; ++r) tile[lb[h] + lane_row + F::row(r, 0) * stp] = acc[h][r];
    }
}

// The full-block product of sep_axis_mfma_full (FP32, lines of 32 values) for a centrosymmetric operator
// (M[31-i][31-j] = M[i][j]: every LINE operator is, being symmetric Toeplitz or symmetric circulant) at half the
// matrix-core work.  With s_j = x_j + x_{31-j} and d_j = x_j - x_{31-j} (j < 16),
//     y_i = (C s)_i + (D d)_i,   y_{31-i} = (C s)_i - (D d)_i   (i < 16),
//     C_ij = (M_ij + M_{i,31-j}) / 2,   D_ij = (M_ij - M_{i,31-j}) / 2,
// two 16 x 16 products per line instead of one 32 x 32.  The staged table (32 rows of 33, sep_centro_table) holds
// [C | D] in its first 16 rows.  On v_mfma_f32_16x16x4f32 a block is 16 lines; K-step k's A operand in lane l is
// C[l % 16][4 k + l / 16] (D likewise), its B operand s (d) at position 4 k + l / 16 of the block's line l % 16, and
// accumulator register r of lane l holds row 4 (l / 16) + r of line l % 16.  A wave takes NH blocks at a time:
// 16 NH lines, block h holding lines (h & 1) + 2 c + 32 (h >> 1) of the group (c = l % 16), so that the 16 lines
// of a block sit on different LDS banks along the contiguous axis.
template <int NH>
__device__ __forceinline__ void sep_axis_centro(int nvals, int st, int aoff) {
    static_assert(NH % 2 == 0, "blocks pair up on odd and even lines");
    typedef float acc4_t __attribute__((ext_vector_type(4)));
    st = __builtin_amdgcn_readfirstlane(st);
    nvals = __builtin_amdgcn_readfirstlane(nvals);
    float *tile = reinterpret_cast<float *>(g_dyn_lds);
    const float *At = reinterpret_cast<const float *>(g_dyn_lds) + aoff;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int c = lane & 15, q = lane >> 4;
    float ca[4], da[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        ca[k] = At[c * 33 + 4 * k + q];
        da[k] = At[c * 33 + 16 + 4 * k + q];
    }
    const int ngrp = nvals / (32 * 16 * NH);
    const int span = 32 * st;
    const int stp = st == 1 ? 1 : st + (st >> 5);
    const float inv_st = 1.0f / (float)st;
    for (int g = wave; g < ngrp; g += BLML_NWAVE) {
        int lb[NH];
#pragma unroll
        for (int h = 0; h < NH; ++h) {
            int lo;
            const int hi = fdivmod(g * 16 * NH + (h & 1) + 2 * c + 32 * (h >> 1), st, inv_st, lo);
            lb[h] = sep_lds(hi * span + lo);
        }
        float xa[NH][4], xb[NH][4];
#pragma unroll
        for (int h = 0; h < NH; ++h)
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                xa[h][k] = tile[lb[h] + (4 * k + q) * stp];
                xb[h][k] = tile[lb[h] + (31 - 4 * k - q) * stp];
            }
        __builtin_amdgcn_sched_barrier(0);        // every operand read before the first product
        acc4_t u[NH], v[NH];
#pragma unroll
        for (int h = 0; h < NH; ++h)
#pragma unroll
            for (int r = 0; r < 4; ++r) u[h][r] = v[h][r] = 0.0f;
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int h = 0; h < NH; ++h) {
                u[h] = __builtin_amdgcn_mfma_f32_16x16x4f32(ca[k], xa[h][k] + xb[h][k], u[h], 0, 0, 0);
                v[h] = __builtin_amdgcn_mfma_f32_16x16x4f32(da[k], xa[h][k] - xb[h][k], v[h], 0, 0, 0);
            }
#pragma unroll
        for (int h = 0; h < NH; ++h)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = 4 * q + r;
                tile[lb[h] + i * stp] = u[h][r] + v[h][r];
                tile[lb[h] + (31 - i) * stp] = u[h][r] - v[h][r];
            }
    }
}
#ifndef BLML_SEP_CENTRO_NH
#define BLML_SEP_CENTRO_NH 4
#endif

template <typename T>
__device__ __forceinline__ void sep_axis(int toff, int nvals, int m, int st, int aoff) {
    constexpr int BLK = SepMfma<T>::BLK;
    if (toff == 0 && m == BLK && nvals % (BLML_SEP_FULL_NH * BLK * BLK) == 0 && (st == 1 || st % 32 == 0))
        sep_axis_mfma_full<T>(nvals, st, aoff);
    else if (m <= BLK) sep_axis_mfma<T, 1>(toff, nvals, m, st, aoff);
    else if (m <= 2 * BLK) sep_axis_mfma<T, 2>(toff, nvals, m, st, aoff);
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

// A whole tile moved with neither a base nor a type change, UB values per thread in flight.  Tile value
// e = e0 + t + NT u (e0 a multiple of UB NT) sits at LDS sep_lds(e0 + t) + (NT + NT / 32) u, since NT u is a
// multiple of 32.  Its global offset is linear in u: contiguous tiles (w == rs) at gbase + e; row tiles with w
// dividing NT at gbase + row rs + col with col = t mod w and the row growing by NT / w per u; row tiles with w a
// multiple of UB NT at gbase + row rs + (e0 mod w) + t + NT u (the chunk stays inside one row).  The caller checks
// that the tile is whole and that tsize is a multiple of UB NT, so every access is one the general move makes.
// global-memory views of plain pointers (global loads and stores count on vmcnt only, so the LDS waits of the
// tile's products do not wait for them)
template <typename T>
using sep_gptr = __attribute__((address_space(1))) T *;
template <typename T>
using sep_gcptr = const __attribute__((address_space(1))) T *;

#ifndef BLML_SEP_FAST_UB
#define BLML_SEP_FAST_UB 16
#endif
__device__ __forceinline__ bool sep_fast_move_ok(const SepPass &ps, int obv, int wv) {
    const long tsize = (long)ps.OB * ps.R * ps.w;
    return obv == ps.OB && wv == ps.w && tsize % (BLML_SEP_FAST_UB * BLML_NT) == 0 &&
           (ps.w == ps.rs || ps.w <= BLML_NT || ps.w % (BLML_SEP_FAST_UB * BLML_NT) == 0);
}

template <typename T, bool LOAD>
__device__ __forceinline__ void sep_move_fast(const SepPass &ps, long gbase, const T *src, T *dst) {
    constexpr int UB = BLML_SEP_FAST_UB;
    T *tile = reinterpret_cast<T *>(g_dyn_lds);
    const int tsize = ps.OB * ps.R * ps.w;
    const int t = threadIdx.x;
    const int lw = 31 - __clz(ps.w);
    const int mode = ps.w == ps.rs ? 0 : (ps.w <= BLML_NT ? 1 : 2);
    const long ustep = mode == 1 ? (long)(BLML_NT >> lw) * ps.rs : (long)BLML_NT;
    constexpr int lstep = BLML_NT + BLML_NT / 32;
    for (int e0 = 0; e0 < tsize; e0 += UB * BLML_NT) {
        long g0;
        if (mode == 0) g0 = gbase + e0 + t;
        else if (mode == 1) g0 = gbase + (long)((e0 + t) >> lw) * ps.rs + (t & (ps.w - 1));
        else g0 = gbase + (long)(e0 >> lw) * ps.rs + (e0 & (ps.w - 1)) + t;
        const int l0 = sep_lds(e0 + t);
        T v[UB];
        if (LOAD) {
            sep_gcptr<T> p = (sep_gcptr<T>)(src + g0);
#pragma unroll
            for (int u = 0; u < UB; ++u) v[u] = p[u * ustep];
#pragma unroll
            for (int u = 0; u < UB; ++u) tile[l0 + lstep * u] = v[u];
        } else {
            sep_gptr<T> p = (sep_gptr<T>)(dst + g0);
#pragma unroll
            for (int u = 0; u < UB; ++u) v[u] = tile[l0 + lstep * u];
#pragma unroll
            for (int u = 0; u < UB; ++u) p[u * ustep] = v[u];
        }
    }
}

// A pass whose tiles are all whole, of NV NT values, with neither bases nor type changes, software-pipelined: the
// next tile's values are loaded into registers (NV per thread) while the matrix cores work on the current one in
// LDS, and a tile's results leave through registers in chunks (the stores drain while the loop goes on).  Global
// memory moves in 16-byte accesses (VW = 16 / sizeof(T) values), so a thread has NV / VW loads and as many stores
// in flight per tile: well inside the 63 vector-memory operations a wave can count, so the next tile's loads never
// wait for this tile's stores to retire.  Tile value e = VW (t + NT i) + j (j < VW) sits at LDS
// sep_lds(VW t) + j + (VW NT + VW NT / 32) i (a run of VW values never crosses a 32-value pad boundary) and at
// global offset start + j + i step:
//   w == rs (contiguous):             start = gbase + VW t,                              step = VW NT
//   w divides NT, VW divides w (rows): start = gbase + (VW t / w) rs + (VW t mod w),      step = (VW NT / w) rs
// The caller guarantees src and dst 16-byte aligned and every tile's first value at an offset that VW divides.
// stage() runs once, after the first tile's loads are issued: it may write LDS outside the tile (operator tables).
// CENTRO (FP32): the staged tables are sep_centro_table's [C | D] and the axes take sep_axis_centro.
template <typename T, int NV, bool CENTRO, typename Stage>
__device__ __forceinline__ void sep_pass_pipe(const SepPass &ps, const int *shape, const int *fac_of, const int *aoff,
                                              const T *src, T *dst, int rank, int nranks, const Stage &stage) {
    static_assert(!CENTRO || sizeof(T) == 4, "the centrosymmetric product runs on the FP32 matrix cores");
    constexpr int VW = 16 / (int)sizeof(T);
    constexpr int NQ = NV / VW;
    typedef T vec_t __attribute__((ext_vector_type(VW)));
    T *tile = reinterpret_cast<T *>(g_dyn_lds);
    const int t = threadIdx.x;
    // the pass's axes: per axis (indexed by constants only, so the arrays stay in registers) whether it has an
    // operator, its tile stride and its staged-table offset; nothing inside the tile loop reads memory other than
    // the tiles (a read there would wait for the prefetched loads)
    bool act[BLML_MAXD];
    int ast[BLML_MAXD], aao[BLML_MAXD];
    {
        const int lo = __builtin_amdgcn_readfirstlane(ps.lo), hi = __builtin_amdgcn_readfirstlane(ps.hi);
        int st = __builtin_amdgcn_readfirstlane(ps.w);
#pragma unroll
        for (int a = BLML_MAXD - 1; a >= 0; --a) {
            const bool in = a >= lo && a < hi;
            act[a] = in && fac_of[a] >= 0 && shape[a] > 1;
            ast[a] = __builtin_amdgcn_readfirstlane(st);
            aao[a] = __builtin_amdgcn_readfirstlane(act[a] ? aoff[a] : 0);
            if (in) st *= shape[a];
        }
    }
    const long rs = ((long)__builtin_amdgcn_readfirstlane((int)(ps.rs >> 32)) << 32) |
                    (unsigned)__builtin_amdgcn_readfirstlane((int)ps.rs);
    const int w = __builtin_amdgcn_readfirstlane(ps.w);
    const long R = __builtin_amdgcn_readfirstlane(ps.R), OB = __builtin_amdgcn_readfirstlane(ps.OB);
    const long nchunks = ps.nchunks, ntiles = ps.ntiles;
    const bool contig = w == rs;
    const int lw = 31 - __clz(w);
    const long qstep = contig ? (long)BLML_NT : (long)((VW * BLML_NT) >> lw) * (rs / VW);   // in VW-value units
    const long toff = contig ? (long)VW * t : (long)((VW * t) >> lw) * rs + ((VW * t) & (w - 1));
    const int l0 = sep_lds(VW * t);
    constexpr int lstep = VW * BLML_NT + VW * BLML_NT / 32;
    const int tsize = NV * BLML_NT;
    vec_t v[NQ];
    long tn = rank;
    if (tn < ntiles) {
        const long to = tn / nchunks, tc = tn - to * nchunks;
        sep_gcptr<vec_t> p = (sep_gcptr<vec_t>)(src + (to * OB * R * rs + tc * w + toff));
#pragma unroll
        for (int i = 0; i < NQ; ++i) {
            v[i] = *p;
            p += qstep;
        }
    }
    stage();            // tables staged while the first tile's loads are in flight (the loop's barrier covers both)
    for (; tn < ntiles; tn += nranks) {
        const long to = tn / nchunks, tc = tn - to * nchunks;
        const long gbase = to * OB * R * rs + tc * w;
#pragma unroll
        for (int i = 0; i < NQ; ++i)
#pragma unroll
            for (int j = 0; j < VW; ++j) tile[l0 + lstep * i + j] = v[i][j];
        __syncthreads();
        const long tnext = tn + nranks;
        if (tnext < ntiles) {
            const long to2 = tnext / nchunks, tc2 = tnext - to2 * nchunks;
            sep_gcptr<vec_t> p = (sep_gcptr<vec_t>)(src + (to2 * OB * R * rs + tc2 * w + toff));
#pragma unroll
            for (int i = 0; i < NQ; ++i) {
                v[i] = *p;
                p += qstep;
            }
        }
#pragma unroll
        for (int a = BLML_MAXD - 1; a >= 0; --a) {          // innermost axis first
            if (act[a]) {
                if constexpr (CENTRO)
                    sep_axis_centro<BLML_SEP_CENTRO_NH>(tsize, ast[a], aao[a]);
                else
                    sep_axis_mfma_full<T>(tsize, ast[a], aao[a]);
                __syncthreads();
            }
        }
        sep_gptr<vec_t> q = (sep_gptr<vec_t>)(dst + (gbase + toff));
        constexpr int CH = 4;
#pragma unroll
        for (int i0 = 0; i0 < NQ; i0 += CH) {
            vec_t o[CH];
#pragma unroll
            for (int i = 0; i < CH; ++i)
#pragma unroll
                for (int j = 0; j < VW; ++j) o[i][j] = tile[l0 + lstep * (i0 + i) + j];
#pragma unroll
            for (int i = 0; i < CH; ++i) {
                *q = o[i];
                q += qstep;
            }
        }
        __syncthreads();
    }
}

// whether every tile of the pass is whole, of NV NT values, contiguous or in rows of w dividing NT (and divisible
// by the 16-byte access's value count), and every axis of the pass takes the full-block product (m == BLK, stride
// 1 or a multiple of 32)
template <typename T>
__device__ __forceinline__ bool sep_pipe_ok(const SepPass &ps, int nv, const int *shape, const int *fac_of) {
    constexpr int BLK = SepMfma<T>::BLK;
    const long tsize = (long)ps.OB * ps.R * ps.w;
    constexpr int VW = 16 / (int)sizeof(T);
    if (!(tsize == (long)nv * BLML_NT && ps.nout % ps.OB == 0 && ps.rs % ps.w == 0 &&
          (ps.w == ps.rs || (ps.w <= BLML_NT && BLML_NT % ps.w == 0 && ps.w % VW == 0))))
        return false;
    if (tsize % (BLML_SEP_FULL_NH * BLK * BLK)) return false;
    long st = ps.w;
    for (int a = ps.hi - 1; a >= ps.lo; --a) {
        if (fac_of[a] >= 0 && shape[a] > 1 && !(shape[a] == BLK && (st == 1 || st % 32 == 0))) return false;
        st *= shape[a];
    }
    return true;
}

// Stages the operator tables of a pass's axes in LDS after the tile and its dump slots (one batched loop over all
// of them: the loads of a batch are in flight together) and returns each axis's table offset in aoff (-1: none).
template <typename T>
__device__ __forceinline__ void sep_stage_tables(const SepPass &ps, const int *shape, const int *fac_of, const T *tabs,
                                                 const long *tab_off, int *aoff) {
    const int tsize = ps.OB * ps.R * ps.w;
    int tdim[BLML_MAXD], tcum[BLML_MAXD + 1];
    const T *tsrc[BLML_MAXD];
    int nst = 0;
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

struct SepNoStage {
    __device__ __forceinline__ void operator()() const {}
};

// Stages the operator tables of every axis with aoff[a] >= 0 (each BLK x BLK, the pipelined path's only table
// size) in LDS at aoff[a], rows of BLK + 1: every table's loads are issued before the first LDS write, so the whole
// staging costs about one memory latency.  The arrays are the caller's and indexed by constants only.
template <typename T>
struct SepStageAll {
    const int *fac_of;
    const T *tabs;
    const long *tab_off;
    const int *aoff;
    __device__ __forceinline__ void operator()() const {
        constexpr int BLK = SepMfma<T>::BLK, PP = BLK * BLK, NS = (PP + BLML_NT - 1) / BLML_NT;
        T *lds = reinterpret_cast<T *>(g_dyn_lds);
        T tv[BLML_MAXD][NS];
#pragma unroll
        for (int a = 0; a < BLML_MAXD; ++a) {
            if (aoff[a] >= 0) {
                const T *ts = tabs + tab_off[fac_of[a]];
#pragma unroll
                for (int s = 0; s < NS; ++s) {
                    const int i = threadIdx.x + s * BLML_NT;
                    tv[a][s] = i < PP ? ts[i] : (T)0;
                }
            }
        }
#pragma unroll
        for (int a = 0; a < BLML_MAXD; ++a) {
            if (aoff[a] >= 0) {
#pragma unroll
                for (int s = 0; s < NS; ++s) {
                    const int i = threadIdx.x + s * BLML_NT;
                    if (i < PP) lds[aoff[a] + (i / BLK) * (BLK + 1) + i % BLK] = tv[a][s];
                }
            }
        }
    }
};

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
    int aoff[BLML_MAXD];
    sep_stage_tables<T>(ps, shape, fac_of, tabs, tab_off, aoff);
    if (std::is_same<Ti, T>::value && std::is_same<To, T>::value && !srcb && !dstb) {
        constexpr int NV = sizeof(T) == 4 ? 64 : 32;          // 128 KB tiles
        const bool al16 = ((reinterpret_cast<unsigned long>(src) | reinterpret_cast<unsigned long>(dst)) & 15) == 0;
        if (al16 && sep_pipe_ok<T>(ps, NV, shape, fac_of)) {
            sep_pass_pipe<T, NV, false>(ps, shape, fac_of, aoff, reinterpret_cast<const T *>(src), reinterpret_cast<T *>(dst),
                                 rank, nranks, SepNoStage{});
            return;
        }
    }
    for (long t = rank; t < ps.ntiles; t += nranks) {
        const long to = t / ps.nchunks, tc = t - to * ps.nchunks;
        const long o0 = to * ps.OB, c0 = tc * ps.w;
        const int obv = (int)((ps.nout - o0) < ps.OB ? (ps.nout - o0) : ps.OB);
        const int wv = (int)((ps.rs - c0) < ps.w ? (ps.rs - c0) : ps.w);
        const long gbase = o0 * ps.R * ps.rs + c0;
        const bool fast = sep_fast_move_ok(ps, obv, wv);
        if (std::is_same<Ti, T>::value && !srcb && fast)
            sep_move_fast<T, true>(ps, gbase, reinterpret_cast<const T *>(src), nullptr);
        else
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
        if (std::is_same<To, T>::value && !dstb && fast)
            sep_move_fast<T, false>(ps, gbase, nullptr, reinterpret_cast<T *>(dst));
        else
            sep_move<T, To, false>(ps, gbase, obv, wv, nullptr, nullptr, dst, dstb);
        __syncthreads();
    }
}
