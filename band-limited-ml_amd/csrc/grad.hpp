// The reductions of BLMLBinned.param_grad and log_density (blml_binned.py:543-563, 685-693) over one fit's
// coefficients E, S e = U (K modes) and derivative mat-vecs dU_j = (dS_{k_j} / dlog fc_u) E_{k_j}, for counts m on
// the lattice:
//   F_b = sum_k |U_kb|^2, Z = Re sum conj(E) U, M = sum m, wn_b = m_b / max(F_b + eps Z, 1e-300) (m_b > 0, else 0),
//   out = [Z, M, sum_b wn_b F_b, sum_{m_b > 0} m_b log max(F_b / Z + eps, 1e-300),
//          then per entry j: sum_b wn_b Re(conj(U_{k_j} b) dU_jb), Re sum_b conj(E_{k_j} b) dU_jb]
// FP64 throughout; every sum is block partials in index order then an ordered sum of the blocks, so the result does
// not depend on scheduling.  Z comes first (a pass of its own): wn needs it.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#define BLML_GRAD_NT 256
#define BLML_GRAD_MAXE 16     // derivative entries per call

struct GradArgs {
    int K, nE, nc;
    long B;
    int emode[BLML_GRAD_MAXE];
    double eps;
};

template <typename T>
__device__ __forceinline__ double gd_re(const T *X, long i) { return (double)X[i]; }

__device__ __forceinline__ double gd_block_sum(double v, double *red) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double a = 0.0;
    for (int w = 0; w < BLML_GRAD_NT / 64; ++w) a += red[w];
    return a;
}

// pass 1: partials of Z = Re sum conj(E) U and M = sum m
template <typename T>
__global__ void __launch_bounds__(BLML_GRAD_NT) grad_pass1_kernel(GradArgs g, const T *E, const T *U,
                                                                  const double *m, double *part) {
    __shared__ double red[BLML_GRAD_NT / 64];
    const long V = (long)g.K * g.B * g.nc;
    double z = 0.0, mm = 0.0;
    for (long i = (long)blockIdx.x * BLML_GRAD_NT + threadIdx.x; i < V; i += (long)gridDim.x * BLML_GRAD_NT)
        z += gd_re(E, i) * gd_re(U, i);
    for (long b = (long)blockIdx.x * BLML_GRAD_NT + threadIdx.x; b < g.B; b += (long)gridDim.x * BLML_GRAD_NT)
        mm += m[b];
    z = gd_block_sum(z, red);
    mm = gd_block_sum(mm, red);
    if (threadIdx.x == 0) {
        part[blockIdx.x * 2] = z;
        part[blockIdx.x * 2 + 1] = mm;
    }
}

// ordered sum of nb partial rows of width w: out[i] = sum_b part[b * w + i]
__global__ void grad_sum_kernel(const double *part, int nb, int w, double *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= w) return;
    double a = 0.0;
    for (int b = 0; b < nb; ++b) a += part[(long)b * w + i];
    out[i] = a;
}

// pass 2: partials of [sum wn F, sum m log(F / Z + eps), (a_j, c_j) per entry]; zm = [Z, M] from pass 1
template <typename T>
__global__ void __launch_bounds__(BLML_GRAD_NT) grad_pass2_kernel(GradArgs g, const T *E, const T *U, const T *dU,
                                                                  const double *m, const double *zm, double *part) {
    __shared__ double red[BLML_GRAD_NT / 64];
    const double Z = zm[0];
    const int W = 2 + 2 * g.nE;
    double acc[2 + 2 * BLML_GRAD_MAXE];
    for (int i = 0; i < W; ++i) acc[i] = 0.0;
    const long B = g.B;
    const int nc = g.nc;
    for (long b = (long)blockIdx.x * BLML_GRAD_NT + threadIdx.x; b < B; b += (long)gridDim.x * BLML_GRAD_NT) {
        double F = 0.0;
        for (int k = 0; k < g.K; ++k)
            for (int c = 0; c < nc; ++c) {
                const double u = gd_re(U, ((long)k * B + b) * nc + c);
                F += u * u;
            }
        const double mb = m[b];
        const double wn = mb > 0 ? mb / fmax(F + g.eps * Z, 1e-300) : 0.0;
        acc[0] += wn * F;
        if (mb > 0) acc[1] += mb * log(fmax(F / Z + g.eps, 1e-300));
        for (int j = 0; j < g.nE; ++j) {
            const int k = g.emode[j];
            double a = 0.0, c2 = 0.0;
            for (int c = 0; c < nc; ++c) {
                const double d = gd_re(dU, ((long)j * B + b) * nc + c);
                a += gd_re(U, ((long)k * B + b) * nc + c) * d;
                c2 += gd_re(E, ((long)k * B + b) * nc + c) * d;
            }
            acc[2 + 2 * j] += wn * a;
            acc[3 + 2 * j] += c2;
        }
    }
    for (int i = 0; i < W; ++i) {
        const double v = gd_block_sum(acc[i], red);
        if (threadIdx.x == 0) part[(long)blockIdx.x * W + i] = v;
    }
}

// ------------------------------------------------------------------------------- the refit response (implicit_grad)
// Elementwise terms of the Hessian-vector product and the mixed derivative of the fit's objective at its stationary
// point (blml_binned.py:583-628), FP64 over (K, B, nc) vectors laid out mode-major like the fit's outputs.  occ: w > 0
// (the training counts); F = sum_k |U_k|^2, Fs = F on occupied bins, 1 elsewhere.
#define BLML_IFT_INNER 0   // out = occ ? (w / Fs) dU - (w dF / Fs^2) U : 0, dF = 2 Re sum_k conj(U_k) dU_k
#define BLML_IFT_REAL 1    // out = dU + (occ ? n w / e^2 : 0) D                    (real rank 1: X = E)
#define BLML_IFT_MIXED 2   // out = -wd U_j + [j == k] ws dU_k, out2 = ws U_k; ws = occ ? w / Fs : 0,
                           // wd = occ ? w dF_k / Fs^2 : 0, dF_k = 2 Re conj(U_k) dU_k (dU: mode k's B nc values)
#define BLML_IFT_DPDE 3    // out = (m / max(F + eps Q, 1e-300) / M) U             (the score's gradient sources)
#define BLML_IFT_MASK 4    // out = occ ? X : 0

struct IftArgs {
    int op, K, nc, k;
    long B;
    double n, eps, Q, M;
};

__global__ void __launch_bounds__(BLML_GRAD_NT) ift_elem_kernel(IftArgs a, const double *U, const double *dU,
                                                                const double *X, const double *w, const double *m,
                                                                double *out, double *out2) {
    const long b = (long)blockIdx.x * BLML_GRAD_NT + threadIdx.x;
    if (b >= a.B) return;
    const int K = a.K, nc = a.nc;
    const long B = a.B;
    auto at = [&](int k, int c) { return ((long)k * B + b) * nc + c; };
    const double wb = w[b];
    const bool occ = wb > 0;
    if (a.op == BLML_IFT_MASK) {
        for (int k = 0; k < K; ++k)
            for (int c = 0; c < nc; ++c) out[at(k, c)] = occ ? X[at(k, c)] : 0.0;
        return;
    }
    if (a.op == BLML_IFT_REAL) {
        const double e = X[at(0, 0)];
        const double q = occ ? a.n * wb / (e * e) : 0.0;
        for (int k = 0; k < K; ++k)
            for (int c = 0; c < nc; ++c) out[at(k, c)] = dU[at(k, c)] + q * m[at(k, c)];
        return;
    }
    double F = 0.0;
    for (int k = 0; k < K; ++k)
        for (int c = 0; c < nc; ++c) F += U[at(k, c)] * U[at(k, c)];
    if (a.op == BLML_IFT_DPDE) {
        const double wn = m[b] / fmax(F + a.eps * a.Q, 1e-300) / a.M;
        for (int k = 0; k < K; ++k)
            for (int c = 0; c < nc; ++c) out[at(k, c)] = wn * U[at(k, c)];
        return;
    }
    const double Fs = occ ? F : 1.0;
    if (a.op == BLML_IFT_INNER) {
        double dF = 0.0;
        for (int k = 0; k < K; ++k)
            for (int c = 0; c < nc; ++c) dF += U[at(k, c)] * dU[at(k, c)];
        dF *= 2.0;
        const double p = wb / Fs, q = wb * dF / (Fs * Fs);
        for (int k = 0; k < K; ++k)
            for (int c = 0; c < nc; ++c) out[at(k, c)] = occ ? p * dU[at(k, c)] - q * U[at(k, c)] : 0.0;
        return;
    }
    // BLML_IFT_MIXED
    double dF = 0.0;
    for (int c = 0; c < nc; ++c) dF += U[at(a.k, c)] * dU[b * nc + c];
    dF *= 2.0;
    const double ws = occ ? wb / Fs : 0.0, wd = occ ? wb * dF / (Fs * Fs) : 0.0;
    for (int j = 0; j < K; ++j)
        for (int c = 0; c < nc; ++c)
            out[at(j, c)] = -wd * U[at(j, c)] + (j == a.k ? ws * dU[b * nc + c] : 0.0);
    for (int c = 0; c < nc; ++c) out2[b * nc + c] = ws * U[at(a.k, c)];
}
