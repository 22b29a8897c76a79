// The Gram mat-vec S_k e of one mode, done by one workgroup on one fit (the reference's
// irfftn(rfftn(e, s=pad) * Kf, s=pad)[:m], blml_binned.py:156-165).
//
// A mode's kernel is a product of units (blml_trivial.py:330-335); each unit is an operator along its own lattice
// axes, so S_k is the composition of the units' operators.  Every operator is applied line by line: a line (all
// bins along one axis with the other indices fixed) is staged in LDS, transformed, and written back.
//   padded axis   : linear convolution with an even row, as a zero-padded power-of-two FFT (in-place radix-2,
//                   DIT forward from bit-reversed input, DIF inverse to bit-reversed output), spectrum in between;
//   circular axis : the periodic kernel's spectrum is a harmonic indicator, so the DFT is evaluated only at the
//                   band's frequencies (a band-restricted DFT) and inverted from them.
// Multi-axis units (ball, pball) transform axis by axis into a scratch buffer, multiply by their dense spectrum,
// and transform back.  Data are complex (interleaved); a real vector enters with a zero imaginary part (the
// kernel is real and even, so S keeps real vectors real).
#pragma once
#include "common.hpp"
#include "blml_hip.h"

enum {
    OP_CONV_PAD = 0,   // m -> m: zero-pad to L, FFT, * spec, IFFT, first m
    OP_CONV_CIRC = 1,  // m -> m: DFT at nF frequencies, * spec, inverse from them
    OP_FFT_FWD = 2,    // m -> L: zero-padded FFT, natural order out
    OP_FFT_INV = 3,    // L -> m: inverse FFT (unscaled), first m
    OP_DFT_FWD = 4,    // m -> nF
    OP_DFT_INV = 5,    // nF -> m (unscaled)
    OP_FFT_FULL = 6,   // L -> L forward (spectra of padded tables at plan creation)
};

struct PassShape {
    int team;   // threads per line
    int buf;    // complex LDS elements per team
    int roots;  // complex LDS elements of the shared root table (circular ops)
};

// Threads per line and LDS per line of an op.  The host uses the same rule to size the dynamic LDS.
__host__ __device__ inline PassShape pass_shape(int op, int m, int L, int nF) {
    PassShape s{1, 0, 0};
    if (op == OP_CONV_PAD || op == OP_FFT_FWD || op == OP_FFT_INV || op == OP_FFT_FULL) {
        int t = L / 2;
        if (t < 1) t = 1;
        if (t > BLML_NT) t = BLML_NT;
        s.team = t;
        s.buf = L;
    } else {
        s.team = (m + nF <= 512) ? 64 : BLML_NT;
        s.buf = (op == OP_CONV_CIRC) ? (m + nF) : (op == OP_DFT_FWD ? m : nF);
        s.roots = m;
    }
    return s;
}

__host__ __device__ inline int pass_lds_elems(int op, int m, int L, int nF) {
    PassShape s = pass_shape(op, m, L, nF);
    return s.roots + (BLML_NT / s.team) * s.buf;
}

__device__ __forceinline__ int bitrev(int i, int logL) {
    return logL == 0 ? 0 : (int)(__brev((unsigned)i) >> (32 - logL));
}

// Forward FFT in place on x[0..L): bit-reversed input, natural output.  tw[k] = exp(-2 pi i k / Lmax).
template <typename T>
__device__ void fft_dit_fwd(cpx<T> *x, int L, const cpx<T> *tw, int Lmax, int tid, int team) {
    for (int h = 1; h < L; h <<= 1) {
        const int step = Lmax / (2 * h);
        for (int j = tid; j < L / 2; j += team) {
            const int pos = j & (h - 1);
            const int i0 = ((j - pos) << 1) + pos, i1 = i0 + h;
            const cpx<T> w = tw[pos * step];
            const cpx<T> a = x[i0], t = cmul(w, x[i1]);
            x[i0] = cadd(a, t);
            x[i1] = csub(a, t);
        }
        __syncthreads();
    }
}

// Inverse FFT (unscaled) in place: natural input, bit-reversed output.
template <typename T>
__device__ void fft_dif_inv(cpx<T> *x, int L, const cpx<T> *tw, int Lmax, int tid, int team) {
    for (int h = L / 2; h >= 1; h >>= 1) {
        const int step = Lmax / (2 * h);
        for (int j = tid; j < L / 2; j += team) {
            const int pos = j & (h - 1);
            const int i0 = ((j - pos) << 1) + pos, i1 = i0 + h;
            const cpx<T> w = tw[pos * step];
            const cpx<T> a = x[i0], b = x[i1];
            x[i0] = cadd(a, b);
            x[i1] = cmulc(csub(a, b), w);   // * conj(w): the inverse transform's twiddle
        }
        __syncthreads();
    }
}

template <typename T>
struct Pass {
    int op;
    int m, L, logL, nF;
    int n_in, n_out;      // line length in / out
    int outer, inner;     // lines = outer * inner; element (o, i, r) at (o * n * inner + i * inner + r)
    const int *F;         // circular: frequencies (0 <= F < m)
    const T *spec;        // CONV ops: spectrum values (real)
};

// Applies one pass to every line.  src and dst may alias when n_in == n_out (each line is staged in LDS before it
// is written back, and lines are disjoint).  All threads execute the same sequence of barriers.
template <typename T>
__device__ void run_pass(const Pass<T> &p, const cpx<T> *src, cpx<T> *dst, cpx<T> *lds, const cpx<T> *tw,
                         int Lmax) {
    const PassShape sh = pass_shape(p.op, p.m, p.L, p.nF);
    const int team = sh.team, nteams = BLML_NT / team;
    const int tm = threadIdx.x / team, tid = threadIdx.x % team;
    cpx<T> *roots = lds;
    cpx<T> *buf = lds + sh.roots + tm * sh.buf;
    const bool circ = sh.roots > 0;
    if (circ) {
        for (int j = threadIdx.x; j < p.m; j += BLML_NT) {
            double sn, cs;
            sincospi(-2.0 * (double)j / (double)p.m, &sn, &cs);
            roots[j] = {(T)cs, (T)sn};
        }
        __syncthreads();
    }
    const long nlines = (long)p.outer * p.inner;
    for (long base = 0; base < nlines; base += nteams) {
        const long li = base + tm;
        const bool act = li < nlines;
        long o = 0, r = 0;
        if (act) {
            o = li / p.inner;
            r = li - o * p.inner;
        }
        const cpx<T> *sl = src + (size_t)o * p.n_in * p.inner + r;
        cpx<T> *dl = dst + (size_t)o * p.n_out * p.inner + r;
        const long st = p.inner;
        switch (p.op) {
            case OP_CONV_PAD:
            case OP_FFT_FWD:
            case OP_FFT_FULL: {
                const int nin = (p.op == OP_FFT_FULL) ? p.L : p.m;
                for (int i = tid; i < p.L; i += team) buf[i] = {(T)0, (T)0};
                __syncthreads();
                if (act)
                    for (int i = tid; i < nin; i += team) buf[bitrev(i, p.logL)] = sl[i * st];
                __syncthreads();
                fft_dit_fwd(buf, p.L, tw, Lmax, tid, team);
                if (p.op == OP_CONV_PAD) {
                    for (int q = tid; q < p.L; q += team) {
                        const T s = p.spec[q];
                        buf[q] = {buf[q].x * s, buf[q].y * s};
                    }
                    __syncthreads();
                    fft_dif_inv(buf, p.L, tw, Lmax, tid, team);
                    if (act)
                        for (int i = tid; i < p.m; i += team) dl[i * st] = buf[bitrev(i, p.logL)];
                } else {
                    if (act)
                        for (int q = tid; q < p.L; q += team) dl[q * st] = buf[q];
                }
                __syncthreads();
                break;
            }
            case OP_FFT_INV: {
                if (act)
                    for (int q = tid; q < p.L; q += team) buf[q] = sl[q * st];
                __syncthreads();
                fft_dif_inv(buf, p.L, tw, Lmax, tid, team);
                if (act)
                    for (int i = tid; i < p.m; i += team) dl[i * st] = buf[bitrev(i, p.logL)];
                __syncthreads();
                break;
            }
            case OP_CONV_CIRC:
            case OP_DFT_FWD: {
                cpx<T> *xl = buf;
                cpx<T> *Xf = buf + p.m;
                if (act)
                    for (int i = tid; i < p.m; i += team) xl[i] = sl[i * st];
                __syncthreads();
                for (int q = tid; q < p.nF; q += team) {
                    const int f = p.F[q];
                    int idx = 0;
                    cpx<T> acc = {(T)0, (T)0};
                    for (int j = 0; j < p.m; ++j) {
                        acc = cadd(acc, cmul(xl[j], roots[idx]));
                        idx += f;
                        if (idx >= p.m) idx -= p.m;
                    }
                    if (p.op == OP_CONV_CIRC) {
                        const T s = p.spec[q];
                        Xf[q] = {acc.x * s, acc.y * s};
                    } else if (act) {
                        dl[q * st] = acc;
                    }
                }
                __syncthreads();
                if (p.op == OP_CONV_CIRC) {
                    for (int i = tid; i < p.m; i += team) {
                        cpx<T> acc = {(T)0, (T)0};
                        for (int q = 0; q < p.nF; ++q) {
                            const int idx = (int)(((long)p.F[q] * i) % p.m);
                            acc = cadd(acc, cmulc(Xf[q], roots[idx]));
                        }
                        if (act) dl[i * st] = acc;
                    }
                    __syncthreads();
                }
                break;
            }
            case OP_DFT_INV: {
                if (act)
                    for (int q = tid; q < p.nF; q += team) buf[q] = sl[q * st];
                __syncthreads();
                for (int i = tid; i < p.m; i += team) {
                    cpx<T> acc = {(T)0, (T)0};
                    for (int q = 0; q < p.nF; ++q) {
                        const int idx = (int)(((long)p.F[q] * i) % p.m);
                        acc = cadd(acc, cmulc(buf[q], roots[idx]));
                    }
                    if (act) dl[i * st] = acc;
                }
                __syncthreads();
                break;
            }
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
    const cpx<T> *tw;          // twiddles of the longest FFT
    int Lmax;
    cpx<T> *X, *Y;             // scratch (scratch_len complex each)
    cpx<T> *lds;
};

template <typename T>
__device__ void mul_group_spec(const MatvecCtx<T> &c, const blml_factor_t &fa, cpx<T> *buf, const int *ext,
                               const T *spec) {
    long total = 1;
    for (int j = 0; j < c.d; ++j) total *= ext[j];
    for (long idx = threadIdx.x; idx < total; idx += BLML_NT) {
        long rem = idx;
        long g = 0;
        int coord[BLML_MAXD];
        for (int j = c.d - 1; j >= 0; --j) {
            coord[j] = (int)(rem % ext[j]);
            rem /= ext[j];
        }
        for (int a = 0; a < fa.naxes; ++a) g = g * fa.tlen[a] + coord[fa.axes[a]];
        const T s = spec[g];
        buf[idx] = {buf[idx].x * s, buf[idx].y * s};
    }
    __syncthreads();
}

template <typename T>
__device__ inline void line_geom(const int *ext, int d, int axis, int &outer, int &inner) {
    outer = 1;
    inner = 1;
    for (int j = 0; j < axis; ++j) outer *= ext[j];
    for (int j = axis + 1; j < d; ++j) inner *= ext[j];
}

__device__ inline int ilog2(int L) {
    int l = 0;
    while ((1 << l) < L) ++l;
    return l;
}

// Applies one factor to the complex lattice array in c.X (in place; group factors use c.Y as well).
template <typename T>
__device__ void apply_factor(const MatvecCtx<T> &c, const blml_factor_t &fa) {
    if (fa.kind == BLML_FAC_LINE) {
        const int a = fa.axes[0];
        Pass<T> p;
        p.m = c.shape[a];
        p.n_in = p.n_out = p.m;
        line_geom<T>(c.shape, c.d, a, p.outer, p.inner);
        p.spec = c.spec + fa.spec_off;
        if (fa.circ[0]) {
            p.op = OP_CONV_CIRC;
            p.nF = fa.tlen[0];
            p.F = c.freqs + fa.freq_off[0];
            p.L = 0;
            p.logL = 0;
        } else {
            p.op = OP_CONV_PAD;
            p.L = fa.tlen[0];
            p.logL = ilog2(p.L);
            p.nF = 0;
            p.F = nullptr;
        }
        run_pass<T>(p, c.X, c.X, c.lds, c.tw, c.Lmax);
        return;
    }
    int ext[BLML_MAXD];
    for (int j = 0; j < c.d; ++j) ext[j] = c.shape[j];
    cpx<T> *cur = c.X, *oth = c.Y;
    for (int ai = fa.naxes - 1; ai >= 0; --ai) {          // forward, last axis first
        const int a = fa.axes[ai];
        Pass<T> p;
        p.m = c.shape[a];
        p.n_in = ext[a];
        line_geom<T>(ext, c.d, a, p.outer, p.inner);
        if (fa.circ[ai]) {
            p.op = OP_DFT_FWD; p.nF = fa.tlen[ai]; p.F = c.freqs + fa.freq_off[ai]; p.L = 0; p.logL = 0;
            p.n_out = p.nF;
        } else {
            p.op = OP_FFT_FWD; p.L = fa.tlen[ai]; p.logL = ilog2(p.L); p.nF = 0; p.F = nullptr;
            p.n_out = p.L;
        }
        p.spec = nullptr;
        run_pass<T>(p, cur, oth, c.lds, c.tw, c.Lmax);
        ext[a] = p.n_out;
        cpx<T> *t = cur; cur = oth; oth = t;
    }
    mul_group_spec<T>(c, fa, cur, ext, c.spec + fa.spec_off);
    for (int ai = 0; ai < fa.naxes; ++ai) {               // inverse, first axis first
        const int a = fa.axes[ai];
        Pass<T> p;
        p.m = c.shape[a];
        p.n_in = ext[a];
        p.n_out = p.m;
        line_geom<T>(ext, c.d, a, p.outer, p.inner);
        if (fa.circ[ai]) {
            p.op = OP_DFT_INV; p.nF = fa.tlen[ai]; p.F = c.freqs + fa.freq_off[ai]; p.L = 0; p.logL = 0;
        } else {
            p.op = OP_FFT_INV; p.L = fa.tlen[ai]; p.logL = ilog2(p.L); p.nF = 0; p.F = nullptr;
        }
        p.spec = nullptr;
        run_pass<T>(p, cur, oth, c.lds, c.tw, c.Lmax);
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
    for (int b = threadIdx.x; b < B; b += BLML_NT) {
        double re = (double)E[nc * b], im = nc == 2 ? (double)E[2 * b + 1] : 0.0;
        if (Ebase) {
            re -= (double)Ebase[nc * b];
            if (nc == 2) im -= (double)Ebase[2 * b + 1];
        }
        c.X[b] = {(T)re, (T)im};
    }
    __syncthreads();
    for (int f = 0; f < fac_count; ++f) apply_factor<T>(c, c.facs[fac_begin + f]);
    for (int b = threadIdx.x; b < B; b += BLML_NT) {
        double re = (double)c.X[b].x, im = (double)c.X[b].y;
        if (Ubase) {
            re += (double)Ubase[nc * b];
            if (nc == 2) im += (double)Ubase[2 * b + 1];
        }
        U[nc * b] = (To)re;
        if (nc == 2) U[2 * b + 1] = (To)im;
    }
    __syncthreads();
}
