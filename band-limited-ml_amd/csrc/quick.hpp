// BLMLQuick on counts on a fixed lattice, real rank 1 (blml_gpu.py:249-429, the counts path of
// _BLMLQuickCUDA.fit; CPU twin blml_trivial.py:3516-3598 with _solve_weighted_cd): block-Jacobi coordinate descent on
//   Phi(e) = (1/2 e'Se - n sum_b w_b log e_b) / n,   (S e)_b e_b = n w_b at the optimum,
// S the kernel's Gram matrix over the occupied bins (times the signs, D S D, when they are fixed).
//
// A block I of bins moves each e_b to its own one-coordinate minimiser (the positive root of
// sbb e^2 + (Se_b - sbb e_b) e - n w_b = 0, written without cancellation), then Se += S[:, I] dlt.  S[:, I] dlt
// is one pair-rows accumulation (all bins as rows, the block as sources): the B x |I| block of the reference's
// cols kernel is never stored.  With the safeguard on, the step is the largest t in {1, 1/2, ..., 2^-30} that does
// not raise Phi along dlt (blml_gpu.py:352-366), chosen on the device from FP64 block sums.
#pragma once
#include "common.hpp"

#define BLML_QUICK_NT 256
#define BLML_QUICK_NSTEP 31

// one thread per block position j: gathers the block's coordinates and forms the step
__global__ void __launch_bounds__(BLML_QUICK_NT) quick_block_kernel(const int *Z, int d, const int *I, int m,
                                                                    const double *e, const double *Se,
                                                                    const double *w, const double *sg, double n,
                                                                    double sbb, int *Zb, double *eold, double *dlt,
                                                                    double *estar, double *src) {
    const int j = blockIdx.x * BLML_QUICK_NT + threadIdx.x;
    if (j >= m) return;
    const long b = I[j];
    for (int k = 0; k < d; ++k) Zb[(long)j * d + k] = Z[b * d + k];
    const double eo = e[b];
    const double bt = Se[b] - sbb * eo;
    const double cc = n * w[b];
    const double disc = sqrt(bt * bt + 4.0 * sbb * cc);
    const double es = bt >= 0.0 ? 2.0 * cc / (bt + disc) : (-bt + disc) / (2.0 * sbb);
    const double dl = es - eo;
    eold[j] = eo;
    dlt[j] = dl;
    estar[j] = es;
    src[j] = sg ? dl * sg[b] : dl;
}

// the safeguard's step (one workgroup): t = the largest 2^-k, k < 31, with
// t lin + t^2/2 quad - sum_j n w_j log(1 + t dlt_j / eold_j) <= 0, lin = dlt . Se_I, quad = dlt . (S dlt)_I; else 0.
// v: S[:, I] dlt before the signs of the rows.  FP64 sums in a fixed order.
__global__ void __launch_bounds__(BLML_QUICK_NT) quick_step_kernel(const int *I, int m, const double *Se,
                                                                   const double *v, const double *w,
                                                                   const double *sg, const double *eold,
                                                                   const double *dlt, double n, double *t_out) {
    __shared__ double red[BLML_QUICK_NSTEP + 2][BLML_QUICK_NT / 64];
    double acc[BLML_QUICK_NSTEP + 2];
#pragma unroll
    for (int k = 0; k < BLML_QUICK_NSTEP + 2; ++k) acc[k] = 0.0;
    for (int j = threadIdx.x; j < m; j += BLML_QUICK_NT) {
        const long b = I[j];
        const double dl = dlt[j], eo = eold[j], cc = n * w[b];
        const double vb = sg ? v[b] * sg[b] : v[b];
        acc[0] += dl * Se[b];
        acc[1] += dl * vb;
        double t = 1.0;
#pragma unroll
        for (int k = 0; k < BLML_QUICK_NSTEP; ++k) {
            acc[2 + k] += cc * log1p(t * dl / eo);
            t *= 0.5;
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < BLML_QUICK_NSTEP + 2; ++k) {
        double x = acc[k];
        for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
        if (lane == 0) red[k][wave] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s[BLML_QUICK_NSTEP + 2];
        for (int k = 0; k < BLML_QUICK_NSTEP + 2; ++k) {
            double x = 0.0;
            for (int q = 0; q < BLML_QUICK_NT / 64; ++q) x += red[k][q];
            s[k] = x;
        }
        double t = 1.0, pick = 0.0;
        for (int k = 0; k < BLML_QUICK_NSTEP; ++k) {
            if (t * s[0] + 0.5 * t * t * s[1] - s[2 + k] <= 0.0) {   // false for NaN (a step past e = 0)
                pick = t;
                break;
            }
            t *= 0.5;
        }
        *t_out = pick;
    }
}

// Se += t (signs) v over every bin; e[I_j] = eold_j + t dlt_j with the safeguard (t_dev), else the minimisers
__global__ void __launch_bounds__(BLML_QUICK_NT) quick_apply_kernel(long B, const int *I, int m, const double *v,
                                                                    const double *sg, const double *eold,
                                                                    const double *dlt, const double *estar,
                                                                    const double *t_dev, double *e, double *Se) {
    const double t = t_dev ? *t_dev : 1.0;
    const long i = (long)blockIdx.x * BLML_QUICK_NT + threadIdx.x;
    if (i < B) Se[i] += t_dev ? t * (sg ? v[i] * sg[i] : v[i]) : (sg ? v[i] * sg[i] : v[i]);
    if (i < m) e[I[i]] = t_dev ? eold[i] + t * dlt[i] : estar[i];
}

// per-workgroup partials of (e . Se, sum_b w_b log max(e_b, 1e-300), max rho, non-finite e count) over bins;
// rho_b = |Se_b e_b / (n w_b) - 1| on bins with e_b > 0 and w_b > 0 (blml_gpu.py:268-271)
__global__ void __launch_bounds__(BLML_QUICK_NT) quick_stats_kernel(long B, const double *e, const double *Se,
                                                                    const double *w, double n, double *part) {
    __shared__ double red[4][BLML_QUICK_NT / 64];
    double a = 0.0, l = 0.0, r = 0.0, bad = 0.0;
    for (long i = (long)blockIdx.x * BLML_QUICK_NT + threadIdx.x; i < B; i += (long)gridDim.x * BLML_QUICK_NT) {
        const double ei = e[i], si = Se[i], wi = w[i];
        a += ei * si;
        l += wi * log(fmax(ei, 1e-300));
        if (ei > 0.0 && wi > 0.0) r = fmax(r, fabs(si * ei / (n * wi) - 1.0));
        if (!isfinite(ei)) bad += 1.0;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_down(a, o, 64);
        l += __shfl_down(l, o, 64);
        r = fmax(r, __shfl_down(r, o, 64));
        bad += __shfl_down(bad, o, 64);
    }
    if (lane == 0) { red[0][wave] = a; red[1][wave] = l; red[2][wave] = r; red[3][wave] = bad; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
        for (int q = 0; q < BLML_QUICK_NT / 64; ++q) {
            s0 += red[0][q]; s1 += red[1][q]; s2 = fmax(s2, red[2][q]); s3 += red[3][q];
        }
        part[4 * blockIdx.x + 0] = s0;
        part[4 * blockIdx.x + 1] = s1;
        part[4 * blockIdx.x + 2] = s2;
        part[4 * blockIdx.x + 3] = s3;
    }
}

// out = (e . Se, sum w log e, max rho, non-finite count): the partials added in workgroup order
__global__ void quick_stats_sum_kernel(const double *part, int np, double *out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    for (int q = 0; q < np; ++q) {
        s0 += part[4 * q]; s1 += part[4 * q + 1]; s2 = fmax(s2, part[4 * q + 2]); s3 += part[4 * q + 3];
    }
    out[0] = s0; out[1] = s1; out[2] = s2; out[3] = s3;
}

// src = e (times the signs); after the rows pass, Se = v (times the signs)
__global__ void quick_signed_kernel(long B, const double *x, const double *sg, double *y) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < B) y[i] = sg ? x[i] * sg[i] : x[i];
}
