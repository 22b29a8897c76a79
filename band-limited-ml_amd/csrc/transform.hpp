// The Gram mat-vec S_k e of one mode, done by one workgroup on one fit (the reference's
// irfftn(rfftn(e, s=pad) * Kf, s=pad)[:m], blml_binned.py:156-165).
//
// A mode's kernel is a product of units (blml_trivial.py:330-335); each unit is an operator along its own lattice
// axes, so S_k is the composition of the units' operators.  Every operator is applied line by line: the lines (all
// bins along one axis with the other indices fixed) are staged in LDS in rounds of several lines, transformed, and
// written back.
//   padded axis   : linear convolution with an even row, as a zero-padded power-of-two FFT (Stockham autosort,
//                   radices 16/8/4/2 with the radix-R butterflies in registers; natural order in and out), the
//                   spectrum multiplied in between;
//   circular axis : the periodic kernel's spectrum is a harmonic indicator, so the DFT is evaluated only at the
//                   band's frequencies (a band-restricted DFT) and inverted from them.
// Multi-axis units (ball, pball) transform axis by axis into a scratch buffer, multiply by their dense spectrum,
// and transform back.  Data are complex (interleaved); a real vector enters with a zero imaginary part (the
// kernel is real and even, so S keeps real vectors real).
#pragma once
#include "common.hpp"
#include "blml_hip.h"
#include "separable.hpp"

enum {
    OP_CONV_PAD = 0,   // m -> m: zero-pad to L, FFT, * spec, IFFT, first m
    OP_CONV_CIRC = 1,  // m -> m: DFT at nF frequencies, * spec, inverse from them
    OP_FFT_FWD = 2,    // m -> L: zero-padded FFT, natural order out
    OP_FFT_INV = 3,    // L -> m: inverse FFT (unscaled), first m
    OP_DFT_FWD = 4,    // m -> nF
    OP_DFT_INV = 5,    // nF -> m (unscaled)
    OP_FFT_FULL = 6,   // L -> L forward (spectra of padded tables at plan creation)
    OP_TOEP = 7,       // m -> m: the padded op's linear convolution summed directly (y_i = sum_j t_|i-j| x_j)
};

#define BLML_DFT_TEAM 16   // threads per line of the band-restricted DFT passes
#define BLML_TOEP_R 8      // outputs per thread and inputs per step of the direct convolution (OP_TOEP)
#ifndef BLML_TOEP_MAX_DEFAULT
#define BLML_TOEP_MAX_DEFAULT 128   // FP32 plans: the longest padded line summed directly
#endif

// OP_TOEP's padded line: m rounded up to whole steps of BLML_TOEP_R
__host__ __device__ inline int toep_len(int m) { return (m + BLML_TOEP_R - 1) / BLML_TOEP_R * BLML_TOEP_R; }

// LDS layout of an FFT line: element i at i ^ ((i >> 4) & 15), a permutation within each block of 16.  A radix-16
// stage writes with stride 16 (one LDS bank for the whole wavefront in plain order); the swizzle spreads those
// writes over 16 offsets, while consecutive reads stay within a block (conflict-free), at no extra LDS.
__host__ __device__ __forceinline__ int pidx(int i) { return i ^ ((i >> 4) & 15); }

struct PassShape {
    int team;   // threads per line
    int lines;  // lines per round
    int buf;    // complex LDS elements per line
    int roots;  // complex LDS elements of the shared table: roots of unity (circular ops) or the two-level
                // twiddle table lo[64], hi[L / 64] (FFT ops)
};

__host__ __device__ inline bool is_fft_op(int op) {
    return op == OP_CONV_PAD || op == OP_FFT_FWD || op == OP_FFT_INV || op == OP_FFT_FULL;
}

// Threads and LDS per line of an op, for a pass over nlines lines.  The host uses the same rule to size the
// dynamic LDS.
//   FFT ops: L / 16 threads per line (each holds 16 values: one radix-16 butterfly, or 16 / R radix-R ones), so up
//            to BLML_NT / (L / 16) lines per round (L <= 16 BLML_NT); a line shorter than 16 is one thread's.
//   DFT ops: 16 threads per line, up to BLML_NT / 16 lines per round.
//   OP_TOEP: the rounds of OP_CONV_PAD at the same L (its lines per round), each line toep_len(m) values at an odd
//            stride (b64 reads of consecutive lines spread over the banks); the shared table is the even row over
//            offsets -(P - 1) .. P - 1 (P = toep_len(m)), 2P - 1 values of T in P complex elements.  Both fit in
//            OP_CONV_PAD's LDS when m > BLML_TOEP_R (then P + 1 <= 2m - 1 <= L).
__host__ __device__ inline PassShape pass_shape(int op, int m, int L, int nF, long nlines) {
    PassShape s{1, 1, 0, 0};
    int cap;
    if (op == OP_TOEP) {
        int t = L / 16;
        if (t < 1) t = 1;
        s.team = t;
        cap = BLML_NT / t;
        s.buf = toep_len(m) | 1;
        s.roots = toep_len(m);
    } else if (is_fft_op(op)) {
        int t = L / 16;
        if (t < 1) t = 1;
        s.team = t;
        cap = BLML_NT / t;
        s.buf = L;
        s.roots = L <= 1024 ? (L > 1 ? L : 2) : 64 + L / 64;
    } else {
        s.team = BLML_DFT_TEAM;
        cap = BLML_NT / BLML_DFT_TEAM;
        s.buf = m + nF;
        s.roots = m;
    }
    s.lines = nlines < cap ? (int)(nlines > 0 ? nlines : 1) : cap;
    return s;
}

__host__ __device__ inline int pass_lds_elems(int op, int m, int L, int nF, long nlines) {
    PassShape s = pass_shape(op, m, L, nF, nlines);
    return s.roots + s.lines * s.buf;
}

// cos / sin of 2 pi k / 16
__device__ __forceinline__ double c16cos(int k) {
    const double c1 = 0.92387953251128674, c2 = 0.70710678118654752, c3 = 0.38268343236508977;
    switch (k & 15) {
        case 0: return 1.0;   case 1: return c1;   case 2: return c2;   case 3: return c3;
        case 4: return 0.0;   case 5: return -c3;  case 6: return -c2;  case 7: return -c1;
        case 8: return -1.0;  case 9: return -c1;  case 10: return -c2; case 11: return -c3;
        case 12: return 0.0;  case 13: return c3;  case 14: return c2;  default: return c1;
    }
}
__device__ __forceinline__ double c16sin(int k) { return c16cos(k - 4); }

__host__ __device__ constexpr int rev_bits(int r, int bits) {
    int o = 0;
    for (int i = 0; i < bits; ++i) o |= ((r >> i) & 1) << (bits - 1 - i);
    return o;
}
template <int R>
struct Log2 { static constexpr int v = R <= 1 ? 0 : 1 + Log2<R / 2>::v; };
template <>
struct Log2<1> { static constexpr int v = 0; };

// In-register R-point DFT (R = 2..16), natural order in and out: y_k = sum_r v_r exp(dir 2 pi i r k / R).
template <typename T, int R>
__device__ __forceinline__ void dft_reg(cpx<T> *v, int dir) {
    constexpr int LB = Log2<R>::v;
    cpx<T> u[R];
#pragma unroll
    for (int r = 0; r < R; ++r) u[rev_bits(r, LB)] = v[r];
#pragma unroll
    for (int h = 1; h < R; h <<= 1) {
#pragma unroll
        for (int j = 0; j < R / 2; ++j) {
            const int pos = j & (h - 1);
            const int i0 = ((j - pos) << 1) + pos, i1 = i0 + h;
            const int k = pos * (16 / (2 * h));
            const cpx<T> wv = {(T)c16cos(k), (T)(dir * c16sin(k))};
            const cpx<T> t = cmul(wv, u[i1]);
            u[i1] = csub(u[i0], t);
            u[i0] = cadd(u[i0], t);
        }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = u[r];
}

// exp(dir 2 pi i k / L) from the twiddle table in LDS: for L <= 1024 the table itself, tl[k] = exp(-2 pi i k / L);
// beyond, two levels lo[j] = exp(-2 pi i j / L) (j < 64), hi[i] = exp(-2 pi i 64 i / L) (i < L / 64), so
// exp(-2 pi i k / L) = hi[k >> 6] lo[k & 63].
template <typename T>
__device__ __forceinline__ cpx<T> twiddle(const cpx<T> *tl, int L, int k, int dir) {
    k &= L - 1;
    cpx<T> w;
    if (L <= 1024) w = tl[k];
    else w = cmul(tl[k & 63], tl[64 + (k >> 6)]);
    return dir < 0 ? w : cpx<T>{w.x, -w.y};
}

// Fills the table of length-L twiddles from the full-circle table tw[k] = exp(-2 pi i k / Lmax).
template <typename T>
__device__ __forceinline__ void load_twiddles(cpx<T> *tl, int L, const cpx<T> *tw, int Lmax) {
    const int st = Lmax / L;
    if (L <= 1024) {
        for (int j = threadIdx.x; j < L; j += BLML_NT) tl[j] = tw[(long)j * st];
        return;
    }
    for (int j = threadIdx.x; j < 64 + L / 64; j += BLML_NT) {
        if (j < 64) tl[j] = tw[(long)j * st];
        else tl[j] = tw[(long)(j - 64) * 64 * st];
    }
}

// One Stockham stage of radix R on a line x[0..L) in LDS (natural order in, natural order out).  The line's
// butterflies j = j0 + b * nthr, b < NB = 16 / R, belong to this thread (nthr = L / 16 threads per line).
// mul: multiply the inputs by spec (real) as they are read (the first inverse stage of a convolution).
template <typename T, int R, int NB>
__device__ __noinline__ void stockham_stage(int xoff, int L, int Ns, int j0, int nthr, int toff, int dir,
                                            const T *mul) {
    cpx<T> *x = dyn_lds<T>() + xoff;             // the line (LDS)
    const cpx<T> *tl = dyn_lds<T>() + toff;      // its twiddle table (LDS)
    cpx<T> v[R * NB];
    const int stride = L / R;
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int j = j0 + b * nthr;
        const int k = j & (Ns - 1);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            cpx<T> a = x[pidx(j + r * stride)];
            if (mul) {
                const T s = mul[j + r * stride];
                a = {a.x * s, a.y * s};
            }
            if (r > 0 && Ns > 1) a = cmul(a, twiddle(tl, L, k * r * (L / (Ns * R)), dir));
            v[b * R + r] = a;
        }
        dft_reg<T, R>(v + b * R, dir);
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int j = j0 + b * nthr;
        const int k = j & (Ns - 1);
        const int base = (j - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) x[pidx(base + r * Ns)] = v[b * R + r];
    }
    __syncthreads();
}

// Whole FFT of a line of length L (<= 16 BLML_NT) in LDS, all threads participating (lines of a round share the
// schedule).  dir = -1 forward, +1 inverse (unscaled).  mul (optional): spectrum applied on the first stage's reads.
// sched: the order of the radix stages -- 0: radix 16 while 16 divides what is left, the remainder (8, 4 or 2)
// last; 1: the remainder first, then radix 16 (the same stages, another register / LDS pattern; the plan's
// autotuner picks by timing).
template <typename T>
__device__ void fft_line(int xoff, int L, int j0, int team, int toff, int dir, const T *mul, int sched = 0) {
    if (L < 16) {                       // one thread per line, one stage of radix L
        switch (L) {
            case 2: stockham_stage<T, 2, 1>(xoff, L, 1, 0, 1, toff, dir, mul); break;
            case 4: stockham_stage<T, 4, 1>(xoff, L, 1, 0, 1, toff, dir, mul); break;
            case 8: stockham_stage<T, 8, 1>(xoff, L, 1, 0, 1, toff, dir, mul); break;
            default: __syncthreads(); __syncthreads(); break;   // L = 1: identity (same barriers as a stage)
        }
        return;
    }
    int Ns = 1;
    bool first = true;
    // the remainder radix of L = 16^k r (r = 1, 2, 4 or 8)
    int r0 = L;
    while (r0 >= 16) r0 >>= 4;
    while (Ns < L) {
        const int rem = L / Ns;
        const T *m_ = first ? mul : nullptr;
        if (sched == 1 && first && r0 > 1) {
            if (r0 == 8) stockham_stage<T, 8, 2>(xoff, L, Ns, j0, team, toff, dir, m_);
            else if (r0 == 4) stockham_stage<T, 4, 4>(xoff, L, Ns, j0, team, toff, dir, m_);
            else stockham_stage<T, 2, 8>(xoff, L, Ns, j0, team, toff, dir, m_);
            Ns *= r0;
        } else if (rem >= 16) stockham_stage<T, 16, 1>(xoff, L, Ns, j0, team, toff, dir, m_), Ns *= 16;
        else if (rem == 8) stockham_stage<T, 8, 2>(xoff, L, Ns, j0, team, toff, dir, m_), Ns *= 8;
        else if (rem == 4) stockham_stage<T, 4, 4>(xoff, L, Ns, j0, team, toff, dir, m_), Ns *= 4;
        else stockham_stage<T, 2, 8>(xoff, L, Ns, j0, team, toff, dir, m_), Ns *= 2;
        first = false;
    }
}

template <typename T>
struct Pass {
    int op;
    int m, L, nF;
    int n_in, n_out;      // line length in / out
    int outer, inner;     // lines = outer * inner; element (o, i, r) at (o * n * inner + i * inner + r)
    const int *F;         // circular: frequencies (0 <= F < m)
    const T *spec;        // CONV ops: spectrum values (real)
    const double *row;    // OP_TOEP: the even row t_0 .. t_{m-1}
    int rank, nranks;     // the team member running this call: rounds rank, rank + nranks, ... of the lines
    bool pairs;           // CONV ops on real data: two lines per transform, as real and imaginary part (exact: the
                          // operator is real, so S (x + i y) = S x + i S y)
    int sched;            // FFT ops: the radix-stage order of fft_line
};

// Moves a round of lines between global memory and LDS (line l of the round at lbuf[l * stride_l + off + i]).
// load: positions [0, width) of each LDS line, from global positions < nvalid (zero beyond; zero lines past the
// end); store: positions [0, width) to global.  lofs[l]: the global offset of the round's line l (or -1 past the
// end).  Contiguous lines (inner == 1) are walked element-fastest, strided lines line-fastest, so consecutive
// threads touch consecutive addresses either way.
template <typename T>
__device__ __forceinline__ void round_io(bool load, int inner, int nl, int nvalid, int width, int off, cpx<T> *lbuf,
                                         int stride_l, const int *lofs, const int *lofs2, const cpx<T> *src,
                                         cpx<T> *dst, bool padded = false) {
    constexpr int UB = 8;                 // elements per thread in flight: the loads of a batch issue together
    const int total = nl * width;
    const int dv = inner == 1 ? width : nl;
    const float inv = 1.0f / (float)dv;
    for (int b0 = threadIdx.x; b0 < total; b0 += UB * BLML_NT) {
        cpx<T> v[UB];
        int cell[UB];
        long g[UB], g2[UB];
#pragma unroll
        for (int u = 0; u < UB; ++u) {
            const int idx = b0 + u * BLML_NT;
            cell[u] = -1;
            g[u] = -1;
            g2[u] = -1;
            v[u] = {(T)0, (T)0};
            if (idx < total) {
                int l, i, rr;
                if (inner == 1) { l = fdivmod(idx, dv, inv, rr); i = rr; }
                else { i = fdivmod(idx, dv, inv, rr); l = rr; }
                const int lo = lofs[l];
                cell[u] = l * stride_l + off + (padded ? pidx(i) : i);
                if (lo >= 0 && (!load || i < nvalid)) {
                    g[u] = lo + (long)i * inner;
                    if (lofs2 && lofs2[l] >= 0) g2[u] = (long)lofs2[l] + (long)i * inner;
                }
            }
        }
        if (load) {
            if (lofs2) {
#pragma unroll
                for (int u = 0; u < UB; ++u) {
                    if (g[u] >= 0) v[u].x = src[g[u]].x;
                    if (g2[u] >= 0) v[u].y = src[g2[u]].x;
                }
            } else {
#pragma unroll
                for (int u = 0; u < UB; ++u)
                    if (g[u] >= 0) v[u] = src[g[u]];
            }
#pragma unroll
            for (int u = 0; u < UB; ++u)
                if (cell[u] >= 0) lbuf[cell[u]] = v[u];
        } else {
#pragma unroll
            for (int u = 0; u < UB; ++u)
                if (g[u] >= 0) v[u] = lbuf[cell[u]];
            if (lofs2) {
#pragma unroll
                for (int u = 0; u < UB; ++u) {
                    if (g[u] >= 0) dst[g[u]] = {v[u].x, (T)0};
                    if (g2[u] >= 0) dst[g2[u]] = {v[u].y, (T)0};
                }
            } else {
#pragma unroll
                for (int u = 0; u < UB; ++u)
                    if (g[u] >= 0) dst[g[u]] = v[u];
            }
        }
    }
}

// OP_TOEP on one round of nl lines staged in LDS (line l at lines + l * stride, P = toep_len(m) values, zero past
// m): y_i = sum_j t_{i-j} x_j for i < m, written straight to dst at lofs[l] + i * inner (with lofs2: the real and
// imaginary parts as two real lines, like round_io's store).  t2[k] = t_{k-(P-1)}, the even row over offsets
// -(P - 1) .. P - 1 (zero where |offset| >= m).  A thread takes BLML_TOEP_R consecutive outputs of one line and
// steps through the inputs BLML_TOEP_R at a time: 2R - 1 table values and R inputs per R^2 products.  Consecutive
// threads take consecutive lines at the same outputs, so the table reads are the same address across a wavefront.
// The sums run in FP64 (the products of two FP32 values are exact there, and FP64 FMAs issue at the rate of
// unpacked FP32 ones), so an FP32 plan's result is its FP32 inputs' convolution rounded once.  Every output of the
// round is computed from LDS before this returns, so dst may be the round's own source.
template <typename T>
__device__ __noinline__ void toep_round(const cpx<T> *lines, int stride, const T *t2, int m, int nl, int inner,
                                        const int *lofs, const int *lofs2, cpx<T> *dst) {
    constexpr int R = BLML_TOEP_R;
    const int P = toep_len(m);
    const int nb = P / R;
    for (int it = threadIdx.x; it < nl * nb; it += BLML_NT) {
        const int b = it / nl, l = it - b * nl;
        const int i0 = b * R;
        const cpx<T> *x = lines + l * stride;
        double ar[R], ai[R];
#pragma unroll
        for (int r = 0; r < R; ++r) ar[r] = ai[r] = 0.0;
        for (int j0 = 0; j0 < P; j0 += R) {
            const T *tp = t2 + (i0 - j0) + (P - 1) - (R - 1);    // tv[s] = t_{i0 - j0 - (R - 1) + s}
            double tv[2 * R - 1];
#pragma unroll
            for (int s = 0; s < 2 * R - 1; ++s) tv[s] = (double)tp[s];
            double xr[R], xi[R];
#pragma unroll
            for (int u = 0; u < R; ++u) {
                const cpx<T> v = x[j0 + u];
                xr[u] = (double)v.x;
                xi[u] = (double)v.y;
            }
#pragma unroll
            for (int u = 0; u < R; ++u)
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const double c = tv[r - u + R - 1];             // t_{(i0 + r) - (j0 + u)}
                    ar[r] = fma(c, xr[u], ar[r]);
                    ai[r] = fma(c, xi[u], ai[r]);
                }
        }
        const int lo = lofs[l];
        if (lo >= 0) {
            const int lo2 = lofs2 ? lofs2[l] : -1;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int i = i0 + r;
                if (i < m) {
                    if (lofs2) {
                        dst[lo + (long)i * inner] = {(T)ar[r], (T)0};
                        if (lo2 >= 0) dst[lo2 + (long)i * inner] = {(T)ai[r], (T)0};
                    } else {
                        dst[lo + (long)i * inner] = {(T)ar[r], (T)ai[r]};
                    }
                }
            }
        }
    }
}

// Applies one pass to every line.  src and dst may alias when n_in == n_out (a round's lines are staged in LDS
// before any is written back, and lines are disjoint).  All threads execute the same sequence of barriers.
template <typename T>
__device__ __noinline__ void run_pass(const Pass<T> &p, const cpx<T> *src, cpx<T> *dst, const cpx<T> *tw, int Lmax) {
    cpx<T> *lds = dyn_lds<T>();
    const long nlines0 = (long)p.outer * p.inner;
    const long nlines = p.pairs ? (nlines0 + 1) / 2 : nlines0;    // transforms (line pairs when pairing)
    const PassShape sh = pass_shape(p.op, p.m, p.L, p.nF, nlines);
    const int team = sh.team, nl = sh.lines;
    const int tm = threadIdx.x / team, tid = threadIdx.x - tm * team;
    const bool has_line = tm < nl;
    cpx<T> *roots = lds;
    cpx<T> *lines = lds + sh.roots;
    cpx<T> *buf = lines + (has_line ? tm : 0) * sh.buf;
    if (p.op == OP_TOEP) {
        T *t2 = reinterpret_cast<T *>(roots);
        const int P = toep_len(p.m);
        for (int k = threadIdx.x; k < 2 * P - 1; k += BLML_NT) {
            const int o = k < P - 1 ? P - 1 - k : k - (P - 1);
            t2[k] = o < p.m ? (T)p.row[o] : (T)0;
        }
    } else if (is_fft_op(p.op)) {
        load_twiddles<T>(roots, p.L, tw, Lmax);
    } else {
        for (int j = threadIdx.x; j < p.m; j += BLML_NT) {
            double sn, cs;
            sincospi(-2.0 * (double)j / (double)p.m, &sn, &cs);
            roots[j] = {(T)cs, (T)sn};
        }
    }
    // line offsets into the scratch (int: the plan keeps every intermediate below 2^31 values)
    __shared__ int lofs_in[BLML_NT], lofs_out[BLML_NT], lofs_in2[BLML_NT], lofs_out2[BLML_NT];
    const int *li2 = p.pairs ? lofs_in2 : nullptr;
    const int *lo2 = p.pairs ? lofs_out2 : nullptr;
    const float inv_inner = 1.0f / (float)p.inner;
    for (long base = (long)p.rank * nl; base < nlines; base += (long)p.nranks * nl) {
        for (int l = threadIdx.x; l < nl; l += BLML_NT) {
            const long t = base + l;
            const long la = p.pairs ? 2 * t : t, lb = p.pairs ? 2 * t + 1 : -1;
            lofs_in[l] = lofs_out[l] = lofs_in2[l] = lofs_out2[l] = -1;
            if (t < nlines) {
                int r;
                int o = fdivmod((int)la, p.inner, inv_inner, r);
                lofs_in[l] = (int)((long)o * p.n_in * p.inner + r);
                lofs_out[l] = (int)((long)o * p.n_out * p.inner + r);
                if (lb >= 0 && lb < nlines0) {
                    o = fdivmod((int)lb, p.inner, inv_inner, r);
                    lofs_in2[l] = (int)((long)o * p.n_in * p.inner + r);
                    lofs_out2[l] = (int)((long)o * p.n_out * p.inner + r);
                }
            }
        }
        __syncthreads();
        if (p.op == OP_TOEP) {
            const int P = toep_len(p.m);
            round_io<T>(true, p.inner, nl, p.m, P, 0, lines, sh.buf, lofs_in, li2, src, nullptr);
            __syncthreads();
            toep_round<T>(lines, sh.buf, reinterpret_cast<const T *>(roots), p.m, nl, p.inner, lofs_out, lo2, dst);
            __syncthreads();
        } else if (is_fft_op(p.op)) {
            const int nin = (p.op == OP_FFT_FULL || p.op == OP_FFT_INV) ? p.L : p.m;
            round_io<T>(true, p.inner, nl, nin, p.L, 0, lines, sh.buf, lofs_in, li2, src, nullptr, true);
            __syncthreads();
            const int j0 = has_line ? tid : 0;
            const int boff = (int)(buf - lds), roff = 0;
            if (p.op == OP_FFT_INV) {
                fft_line<T>(boff, p.L, j0, team, roff, +1, nullptr, p.sched);
            } else {
                fft_line<T>(boff, p.L, j0, team, roff, -1, nullptr, p.sched);
                if (p.op == OP_CONV_PAD) fft_line<T>(boff, p.L, j0, team, roff, +1, p.spec, p.sched);
            }
            const int nout = (p.op == OP_FFT_FWD || p.op == OP_FFT_FULL) ? p.L : p.m;
            round_io<T>(false, p.inner, nl, nout, nout, 0, lines, sh.buf, lofs_out, lo2, nullptr, dst, true);
            __syncthreads();
        } else {
            // band-restricted DFT on line buffers [0, m) (values) and [m, m + nF) (band)
            cpx<T> *xl = buf, *Xf = buf + p.m;
            if (p.op == OP_DFT_INV)
                round_io<T>(true, p.inner, nl, p.nF, p.nF, p.m, lines, sh.buf, lofs_in, nullptr, src, nullptr);
            else
                round_io<T>(true, p.inner, nl, p.m, p.m, 0, lines, sh.buf, lofs_in, li2, src, nullptr);
            __syncthreads();
            if (p.op != OP_DFT_INV) {
                for (int q = tid; q < p.nF; q += team) {
                    const int fq = p.F[q];
                    int idx = 0;
                    cpx<T> acc = {(T)0, (T)0};
                    for (int j = 0; j < p.m; ++j) {
                        acc = cadd(acc, cmul(xl[j], roots[idx]));
                        idx += fq;
                        if (idx >= p.m) idx -= p.m;
                    }
                    if (p.op == OP_CONV_CIRC) {
                        const T s = p.spec[q];
                        acc = {acc.x * s, acc.y * s};
                    }
                    if (has_line) Xf[q] = acc;
                }
                __syncthreads();
            }
            if (p.op == OP_DFT_FWD) {
                round_io<T>(false, p.inner, nl, p.nF, p.nF, p.m, lines, sh.buf, lofs_out, nullptr, nullptr, dst);
            } else {
                // inverse from the band: out[i] = sum_q Xf[q] exp(+2 pi i F_q i / m), into [0, m)
                const float inv_m = 1.0f / (float)p.m;
                for (int i = tid; i < p.m; i += team) {
                    cpx<T> acc = {(T)0, (T)0};
                    for (int q = 0; q < p.nF; ++q) {
                        int idx;
                        fdivmod(p.F[q] * i, p.m, inv_m, idx);
                        acc = cadd(acc, cmulc(Xf[q], roots[idx]));
                    }
                    if (has_line) xl[i] = acc;
                }
                __syncthreads();
                round_io<T>(false, p.inner, nl, p.m, p.m, 0, lines, sh.buf, lofs_out, lo2, nullptr, dst);
            }
            __syncthreads();
        }
    }
}

// Everything a workgroup needs to apply the modes' operators of one fit.
template <typename T>
struct MatvecCtx {
    int d;
    int shape[BLML_MAXD];
    const blml_factor_t *facs;   // the plan's factor table
    const T *spec;             // the plan's spectra
    const int *freqs;          // the plan's frequency lists
    const cpx<T> *tw;          // full-circle twiddles of the longest FFT
    int Lmax;
    cpx<T> *X, *Y;             // scratch (scratch_len complex each)
    int fft_sched;             // the radix-stage order of every FFT pass (fft_line's sched)
    Team tm;                   // who runs the mat-vec (passes split their lines over the team)
    const SepPlan *sep;        // separable modes: dense line operators on LDS tiles (null: transform passes)
    bool sep_last_sync;        // team barrier after a separable mat-vec's last pass (false: the kernel ends there)
    const T *sep_tab;          // per factor: P, Q, c_mid, middle row
    const long *sep_off;
    const double *fdata = nullptr;   // the plan's factor data (OP_TOEP reads a padded factor's row there)
    int toep_max = 0;          // padded line factors with BLML_TOEP_R < m <= toep_max take OP_TOEP (0: none)
};

// U_k = S_k E_k of a separable mode by tile passes (see separable.hpp); the intermediate lives in c.X.
template <typename T, typename Ti, typename To>
__device__ __forceinline__ void matvec_sep_body(const MatvecCtx<T> &c, const int fac_begin, const int fac_count,
                                                const Ti *E, To *U, int nc, const Ti *Ebase, const To *Ubase) {
    int fac_of[BLML_MAXD];
    for (int a = 0; a < BLML_MAXD; ++a) fac_of[a] = -1;
    for (int f = 0; f < fac_count; ++f) fac_of[c.facs[fac_begin + f].axes[0]] = fac_begin + f;
    T *Z = reinterpret_cast<T *>(c.X);
    const int np = c.sep->npass;
    for (int q = 0; q < np; ++q) {
        const SepPass ps = c.sep->pass[q];
        const bool first = q == 0, last = q == np - 1;
        if (first && last)
            sep_pass<T, Ti, To>(ps, c.d, nc, c.shape, fac_of, c.sep_tab, c.sep_off, E, Ebase, U, Ubase, c.tm.rank,
                                c.tm.nranks);
        else if (first)
            sep_pass<T, Ti, T>(ps, c.d, nc, c.shape, fac_of, c.sep_tab, c.sep_off, E, Ebase, Z, (const T *)nullptr,
                               c.tm.rank, c.tm.nranks);
        else if (last)
            sep_pass<T, T, To>(ps, c.d, nc, c.shape, fac_of, c.sep_tab, c.sep_off, Z, (const T *)nullptr, U, Ubase,
                               c.tm.rank, c.tm.nranks);
        else
            sep_pass<T, T, T>(ps, c.d, nc, c.shape, fac_of, c.sep_tab, c.sep_off, Z, (const T *)nullptr, Z,
                              (const T *)nullptr, c.tm.rank, c.tm.nranks);
        if (!last || c.sep_last_sync) c.tm.sync();
    }
}

// The call boundary of the separable path inside the solver kernels (its tile loops are inlined, so no registers are
// saved per tile or axis; the stand-alone mat-vec kernel inlines the body itself and saves none at all).
template <typename T, typename Ti, typename To>
__device__ __noinline__ void matvec_sep(const MatvecCtx<T> &c, const int fac_begin, const int fac_count, const Ti *E,
                                        To *U, int nc, const Ti *Ebase, const To *Ubase) {
    matvec_sep_body<T, Ti, To>(c, fac_begin, fac_count, E, U, nc, Ebase, Ubase);
}

template <typename T>
__device__ void mul_group_spec(const MatvecCtx<T> &c, const blml_factor_t &fa, cpx<T> *buf, const int *ext,
                               const T *spec) {
    int total = 1;
    for (int j = 0; j < c.d; ++j) total *= ext[j];
    const bool all = fa.naxes == c.d;      // the group spans every axis: the spectrum index is the element index
    float inv[BLML_MAXD];
    for (int j = 0; j < c.d; ++j) inv[j] = 1.0f / (float)ext[j];
    for (long idx_ = c.tm.first(); idx_ < total; idx_ += c.tm.stride()) {
        const int idx = (int)idx_;
        int g = idx;
        if (!all) {
            int rem = idx;
            int coord[BLML_MAXD];
            for (int j = c.d - 1; j >= 0; --j) {
                int r;
                rem = fdivmod(rem, ext[j], inv[j], r);
                coord[j] = r;
            }
            g = 0;
            for (int a = 0; a < fa.naxes; ++a) g = g * fa.tlen[a] + coord[fa.axes[a]];
        }
        const T s = spec[g];
        buf[idx] = {buf[idx].x * s, buf[idx].y * s};
    }
    c.tm.sync();
}

template <typename T>
__device__ inline void line_geom(const int *ext, int d, int axis, int &outer, int &inner) {
    outer = 1;
    inner = 1;
    for (int j = 0; j < axis; ++j) outer *= ext[j];
    for (int j = axis + 1; j < d; ++j) inner *= ext[j];
}

// Applies one factor to the complex lattice array in c.X (in place; group factors use c.Y as well).
template <typename T>
__device__ void apply_factor(const MatvecCtx<T> &c, const blml_factor_t &fa, bool real) {
    if (fa.kind == BLML_FAC_LINE) {
        const int a = fa.axes[0];
        Pass<T> p;
        p.rank = c.tm.rank;
        p.nranks = c.tm.nranks;
        p.sched = c.fft_sched;
        p.pairs = real;
        p.m = c.shape[a];
        p.n_in = p.n_out = p.m;
        line_geom<T>(c.shape, c.d, a, p.outer, p.inner);
        p.spec = c.spec + fa.spec_off;
        if (fa.circ[0]) {
            p.op = OP_CONV_CIRC;
            p.nF = fa.tlen[0];
            p.F = c.freqs + fa.freq_off[0];
            p.L = 0;
        } else {
            p.op = p.m > BLML_TOEP_R && p.m <= c.toep_max ? OP_TOEP : OP_CONV_PAD;
            p.L = fa.tlen[0];
            p.nF = 0;
            p.F = nullptr;
            p.row = c.fdata + fa.data_off;
        }
        run_pass<T>(p, c.X, c.X, c.tw, c.Lmax);
        c.tm.sync();
        return;
    }
    int ext[BLML_MAXD];
    for (int j = 0; j < c.d; ++j) ext[j] = c.shape[j];
    cpx<T> *cur = c.X, *oth = c.Y;
    for (int ai = fa.naxes - 1; ai >= 0; --ai) {          // forward, last axis first
        const int a = fa.axes[ai];
        Pass<T> p;
        p.rank = c.tm.rank;
        p.nranks = c.tm.nranks;
        p.sched = c.fft_sched;
        p.m = c.shape[a];
        p.n_in = ext[a];
        line_geom<T>(ext, c.d, a, p.outer, p.inner);
        if (fa.circ[ai]) {
            p.op = OP_DFT_FWD; p.nF = fa.tlen[ai]; p.F = c.freqs + fa.freq_off[ai]; p.L = 0;
            p.n_out = p.nF;
        } else {
            p.op = OP_FFT_FWD; p.L = fa.tlen[ai]; p.nF = 0; p.F = nullptr;
            p.n_out = p.L;
        }
        p.spec = nullptr;
        p.pairs = false;
        run_pass<T>(p, cur, oth, c.tw, c.Lmax);
        c.tm.sync();
        ext[a] = p.n_out;
        cpx<T> *t = cur; cur = oth; oth = t;
    }
    mul_group_spec<T>(c, fa, cur, ext, c.spec + fa.spec_off);
    for (int ai = 0; ai < fa.naxes; ++ai) {               // inverse, first axis first
        const int a = fa.axes[ai];
        Pass<T> p;
        p.rank = c.tm.rank;
        p.nranks = c.tm.nranks;
        p.sched = c.fft_sched;
        p.m = c.shape[a];
        p.n_in = ext[a];
        p.n_out = p.m;
        line_geom<T>(ext, c.d, a, p.outer, p.inner);
        if (fa.circ[ai]) {
            p.op = OP_DFT_INV; p.nF = fa.tlen[ai]; p.F = c.freqs + fa.freq_off[ai]; p.L = 0;
        } else {
            p.op = OP_FFT_INV; p.L = fa.tlen[ai]; p.nF = 0; p.F = nullptr;
        }
        p.spec = nullptr;
        p.pairs = false;
        run_pass<T>(p, cur, oth, c.tw, c.Lmax);
        c.tm.sync();
        ext[a] = p.m;
        cpx<T> *t = cur; cur = oth; oth = t;
    }
    // 2 * naxes passes: the result is back in c.X
}

// U_k = S_k E_k for one mode.  E, U: B * nc values (nc = 1 real, 2 interleaved complex) of any real type; the
// transforms run in T.  With bases: U = Ubase + S_k (E - Ebase) -- the incremental form, whose rounding error in T is
// relative to the increment, not to U (E and Ebase / U and Ubase may not alias each other's outputs).
template <typename T, typename Ti, typename To>
__device__ void matvec_mode(const MatvecCtx<T> &c, const int fac_begin, const int fac_count, const Ti *E, To *U,
                            int B, int nc, const Ti *Ebase = nullptr, const To *Ubase = nullptr) {
    if (c.sep) {
        matvec_sep<T, Ti, To>(c, fac_begin, fac_count, E, U, nc, Ebase, Ubase);
        return;
    }
    constexpr int UB = 4;                 // bins per thread in flight
    const long st = c.tm.stride();
    for (long b0 = c.tm.first(); b0 < B; b0 += UB * st) {
        double re[UB], im[UB];
#pragma unroll
        for (int u = 0; u < UB; ++u) {
            const long b = b0 + u * st;
            re[u] = 0.0; im[u] = 0.0;
            if (b < B) {
                re[u] = (double)E[nc * b];
                if (nc == 2) im[u] = (double)E[2 * b + 1];
                if (Ebase) {
                    re[u] -= (double)Ebase[nc * b];
                    if (nc == 2) im[u] -= (double)Ebase[2 * b + 1];
                }
            }
        }
#pragma unroll
        for (int u = 0; u < UB; ++u) {
            const long b = b0 + u * st;
            if (b < B) c.X[b] = {(T)re[u], (T)im[u]};
        }
    }
    c.tm.sync();
    for (int f = 0; f < fac_count; ++f) apply_factor<T>(c, c.facs[fac_begin + f], nc == 1);
    for (long b0 = c.tm.first(); b0 < B; b0 += UB * st) {
        double re[UB], im[UB];
#pragma unroll
        for (int u = 0; u < UB; ++u) {
            const long b = b0 + u * st;
            re[u] = 0.0; im[u] = 0.0;
            if (b < B) {
                const cpx<T> x = c.X[b];
                re[u] = (double)x.x;
                im[u] = (double)x.y;
                if (Ubase) {
                    re[u] += (double)Ubase[nc * b];
                    if (nc == 2) im[u] += (double)Ubase[2 * b + 1];
                }
            }
        }
#pragma unroll
        for (int u = 0; u < UB; ++u) {
            const long b = b0 + u * st;
            if (b < B) {
                U[nc * b] = (To)re[u];
                if (nc == 2) U[2 * b + 1] = (To)im[u];
            }
        }
    }
    c.tm.sync();
}
