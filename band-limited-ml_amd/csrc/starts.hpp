// The start search of a complex counts fit (blml_trivial.py:1605-1676, called by blml_gpu.py:470-486): candidate
// band-limited fields g with |g| matching sqrt(counts) on the bin box, by hybrid input-output (HIO) and error
// reduction (ER) on the FFT grid of the box (power-of-two sides, at least twice the box per axis).
//
// Every run is one workgroup of a persistent launch (runs dealt round robin) and runs in FP64 on its own two grids
// in HBM, P (the spectrum) and X (the field):
//   start   random run r < n_rand:   P = band * exp(2 pi i U_r),  U_r the host's uniform draws;
//           symmetry-breaking run:   P = band * fftn(A) * exp(i kick_r)   (A: sqrt(counts) on the box);
//   step    X = ifftn(P); on the box X <- A s X / |X| with s = sqrt(sum_box |X|^2 / sum counts) (X = A s where
//           X = 0); Pn = fftn(X); P <- Pn on the band, P - beta Pn off it during the HIO steps, 0 off it during the
//           ER steps;
//   end     P <- P band; X = ifftn(P); score = sum_occ w log max(|X|^2 / sum_box |X|^2, 1e-300) (FP64, fixed order);
//           the field at the occupied cells is kept for the caller.
// The FFTs are the engine's Stockham line passes (run_pass, OP_FFT_FULL forward / OP_FFT_INV inverse, in place),
// one pass per axis; ifftn's 1/ncell is applied in the elementwise step that follows.
#pragma once
#include "transform.hpp"

struct HioArgs {
    int d;
    int shape[BLML_MAXD];       // grid sides (powers of two)
    int box[BLML_MAXD];         // the data box: cells [0, box_j) per axis
    long ncell;                 // prod shape
    int n_hio, n_er;            // steps of a random run; a symmetry-breaking run takes n_er ER steps only
    int n_rand, n_kick;         // runs: n_rand random-phase runs, then n_kick symmetry-breaking runs
    double beta;
    double ab2;                 // sum of the counts on the box (sum A^2)
    long nocc;
    const unsigned char *band;  // (ncell) the band indicator
    const double *amp;          // (ncell) A: sqrt(counts) on the box, 0 elsewhere
    const long *occ;            // (nocc) grid cells of the occupied bins
    const double *wocc;         // (nocc) their counts
    const double *phase;        // (n_rand, ncell) uniform draws of the random runs
    const double *kick;         // (n_kick, ncell) phase kick (sign * delta * direction) of the symmetry-breaking runs
    cpx<double> *P;             // (runs, ncell) out: the final band-limited spectra
    cpx<double> *X;             // (runs, ncell) work: the fields
    cpx<double> *gocc;          // (runs, nocc) out: the final fields at the occupied cells
    double *score;              // (runs) out
    const cpx<double> *tw;      // twiddles exp(-2 pi i k / Lmax)
    int Lmax;
};

// whether cell c of the grid lies in the box (row-major cells, last axis fastest; sides are powers of two)
__device__ __forceinline__ bool hio_in_box(const HioArgs &a, long c) {
    bool in = true;
    for (int j = a.d - 1; j >= 0; --j) {
        const int m = a.shape[j];
        in = in && (int)(c & (m - 1)) < a.box[j];
        c >>= 31 - __clz(m);
    }
    return in;
}

// fftn (dir -1, unscaled) or the unscaled inverse (dir +1) of the grid X, in place, axis by axis
__device__ void hio_fft(const HioArgs &a, cpx<double> *X, int dir) {
    for (int ax = 0; ax < a.d; ++ax) {
        Pass<double> p;
        p.op = dir < 0 ? OP_FFT_FULL : OP_FFT_INV;
        p.L = p.m = p.n_in = p.n_out = a.shape[ax];
        p.nF = 0;
        p.F = nullptr;
        p.spec = nullptr;
        p.rank = 0;
        p.nranks = 1;
        p.pairs = false;
        p.sched = 0;
        line_geom<double>(a.shape, a.d, ax, p.outer, p.inner);
        run_pass<double>(p, X, X, a.tw, a.Lmax);
    }
}

__device__ void hio_run(const HioArgs &a, int run, double *red) {
    const long M = a.ncell;
    cpx<double> *P = a.P + (long)run * M;
    cpx<double> *X = a.X + (long)run * M;
    const double inv_m = 1.0 / (double)M;
    const bool rnd = run < a.n_rand;
    // the start
    if (rnd) {
        const double *U = a.phase + (long)run * M;
        for (long c = threadIdx.x; c < M; c += BLML_NT) {
            double sn, cs;
            sincospi(2.0 * U[c], &sn, &cs);
            P[c] = a.band[c] ? cpx<double>{cs, sn} : cpx<double>{0.0, 0.0};
        }
    } else {
        for (long c = threadIdx.x; c < M; c += BLML_NT) X[c] = {a.amp[c], 0.0};
        __syncthreads();
        hio_fft(a, X, -1);
        const double *K = a.kick + (long)(run - a.n_rand) * M;
        for (long c = threadIdx.x; c < M; c += BLML_NT) {
            double sn, cs;
            sincos(K[c], &sn, &cs);
            P[c] = a.band[c] ? cmul(X[c], cpx<double>{cs, sn}) : cpx<double>{0.0, 0.0};
        }
    }
    __syncthreads();
    const int nh = rnd ? a.n_hio : 0;
    const int nit = nh + a.n_er;
    for (int it = 0; it < nit; ++it) {
        for (long c = threadIdx.x; c < M; c += BLML_NT) X[c] = P[c];
        __syncthreads();
        hio_fft(a, X, +1);
        // the data constraint on the box, after ifftn's 1/M
        double s2 = 0.0;
        for (long c = threadIdx.x; c < M; c += BLML_NT) {
            const cpx<double> v = {X[c].x * inv_m, X[c].y * inv_m};
            X[c] = v;
            if (hio_in_box(a, c)) s2 += v.x * v.x + v.y * v.y;
        }
        const double s = sqrt(block_sum1(s2, red) / a.ab2);
        for (long c = threadIdx.x; c < M; c += BLML_NT) {
            if (hio_in_box(a, c)) {
                const cpx<double> v = X[c];
                const double r = sqrt(v.x * v.x + v.y * v.y);
                const double av = a.amp[c] * s;
                X[c] = r > 0.0 ? cpx<double>{av * v.x / r, av * v.y / r} : cpx<double>{av, 0.0};
            }
        }
        __syncthreads();
        hio_fft(a, X, -1);
        const bool hio = it < nh;
        for (long c = threadIdx.x; c < M; c += BLML_NT) {
            const cpx<double> pn = X[c];
            if (a.band[c]) P[c] = pn;
            else P[c] = hio ? cpx<double>{P[c].x - a.beta * pn.x, P[c].y - a.beta * pn.y} : cpx<double>{0.0, 0.0};
        }
        __syncthreads();
    }
    // band-limit, the field, the score
    for (long c = threadIdx.x; c < M; c += BLML_NT) {
        if (!a.band[c]) P[c] = {0.0, 0.0};
        X[c] = a.band[c] ? P[c] : cpx<double>{0.0, 0.0};
    }
    __syncthreads();
    hio_fft(a, X, +1);
    double fb = 0.0;
    for (long c = threadIdx.x; c < M; c += BLML_NT) {
        const cpx<double> v = {X[c].x * inv_m, X[c].y * inv_m};
        X[c] = v;
        if (hio_in_box(a, c)) fb += v.x * v.x + v.y * v.y;
    }
    fb = block_sum1(fb, red);
    double sc = 0.0;
    cpx<double> *g = a.gocc + (long)run * a.nocc;
    for (long i = threadIdx.x; i < a.nocc; i += BLML_NT) {
        const cpx<double> v = X[a.occ[i]];
        g[i] = v;
        sc += a.wocc[i] * log(fmax((v.x * v.x + v.y * v.y) / fb, 1e-300));
    }
    sc = block_sum1(sc, red);
    if (threadIdx.x == 0) a.score[run] = sc;
    __syncthreads();
}

// The field of a start at the original bins when the search grid is coarser than the lattice (q_j > 1 on some
// axis): g_b = sum over the band's cells n of c_n exp(2 pi i sum_j dz_bj k_nj / (m_j q_j)), c = P / ncell, with the
// phase reduced exactly in integers (dz: the bin's lattice offset from the box corner, k: the cell's signed
// frequency index).
struct BandEvalArgs {
    int d;
    int mq[BLML_MAXD];          // m_j q_j
    long nb, nband;
    const int *dz;              // (nb, d)
    const int *kf;              // (nband, d) signed frequency indices
    const cpx<double> *c;       // (nband) coefficients
    cpx<double> *g;             // (nb) out
};

__global__ void __launch_bounds__(256) band_eval_kernel(BandEvalArgs a) {
    const long b = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= a.nb) return;
    int z[BLML_MAXD];
    for (int j = 0; j < a.d; ++j) z[j] = a.dz[b * a.d + j];
    cpx<double> acc = {0.0, 0.0};
    for (long n = 0; n < a.nband; ++n) {
        double ph = 0.0;
        for (int j = 0; j < a.d; ++j) {
            const long m = a.mq[j];
            long r = ((long)z[j] * (long)a.kf[n * a.d + j]) % m;
            if (r < 0) r += m;
            ph += (double)r / (double)m;
        }
        ph -= floor(ph);
        double sn, cs;
        sincospi(2.0 * ph, &sn, &cs);
        acc = cadd(acc, cmul(a.c[n], cpx<double>{cs, sn}));
    }
    a.g[b] = acc;
}
