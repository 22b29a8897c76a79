// BLMLQuick's complex / rank-K fits on counts on a fixed lattice (blml_gpu.py:565-744 _psd_solve, with
// blml_trivial.py:2590-2651 lbfgs_psd for solver='lbfgs'): the coefficients E (B bins x K modes, real or complex)
// maximise L = sum_b w_b log F_b - n log Q, F_b = sum_k |(S_k e_k)_b|^2, Q = sum_k e_k^H S_k e_k.
//
// State on the device: E, SE = S E as (B, K, nc) FP64 values (nc = 2: interleaved complex, 1: real), the layout of
// the reference's (B, K) arrays.  S_k is one pair table for every mode (a shared kernel) or one per mode.
//
// A sweep visits blocks I of bins (Gauss-Seidel between blocks, Jacobi inside one): each row moves to its own
// stationary point R_bk + s_k e_bk = n w_b e_bk / |e_b|^2 (R = S e - s e, s_k = S_k(z, z)), in closed form for a
// shared kernel and by the secular equation's bisection for per-mode kernels; then SE += S[:, I] d, one pair-rows
// pass whose sources carry every mode's column of d.  With the safeguard the step is the first t in 2^-k, k < 31,
// that does not lower L along d (every bin's F at every t, FP64 block sums, chosen on the device).
#pragma once
#include "common.hpp"
#include "pairs.hpp"

#define BLML_PSD_NT 256
#define BLML_PSD_NSTEP 31
#define BLML_PSD_NACC (BLML_PSD_NSTEP + 5)   // sum w log F(t) per step, then L0, a0, Re<E, V>, Re<d, SE_I>, Re<d, V_I>
#ifndef BLML_VEC_MAXV
#define BLML_VEC_MAXV 8
#endif

struct PsdArgs {
    int K, nc;
    int shared;
    double n;
    double s[BLML_QUICK_MAXK];
};

// The block's coordinates for one pair table.
__global__ void psd_gather_kernel(const int *Z, int d, const int *I, int m, int *Zb) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m) return;
    for (int c = 0; c < d; ++c) Zb[(long)j * d + c] = Z[(long)I[j] * d + c];
}

// One thread per block position j: the row's new value and d = new - old.  EI, dl: (m, K, nc); Wm: d by mode,
// (K, m, nc), the sources of per-mode pair passes.
__global__ void __launch_bounds__(BLML_PSD_NT) psd_block_kernel(PsdArgs a, const int *I, int m, const double *E,
                                                                const double *SE, const double *w, double *EI,
                                                                double *dl, double *Wm) {
    const int j = blockIdx.x * BLML_PSD_NT + threadIdx.x;
    if (j >= m) return;
    const int K = a.K, nc = a.nc, V = K * nc;
    const long b = I[j];
    double R[2 * BLML_QUICK_MAXK], e[2 * BLML_QUICK_MAXK], nv[2 * BLML_QUICK_MAXK];
    for (int i = 0; i < V; ++i) {
        e[i] = E[b * V + i];
        R[i] = SE[b * V + i] - a.s[i / nc] * e[i];
    }
    const double cc = a.n * w[b];
    if (a.shared) {
        double ar = 0.0, en = 0.0;
        for (int i = 0; i < V; ++i) {
            ar += R[i] * R[i];
            en += e[i] * e[i];
        }
        ar = sqrt(ar);
        en = sqrt(en);
        const double s0 = a.s[0];
        const double t = 2.0 * cc / (ar + sqrt(ar * ar + 4.0 * s0 * cc));
        for (int i = 0; i < V; ++i) nv[i] = t * (ar > 0.0 ? R[i] / ar : e[i] / fmax(en, 1e-30));
    } else {
        double a2[BLML_QUICK_MAXK];
        double smax = a.s[0];
        for (int k = 0; k < K; ++k) {
            double q = 0.0;
            for (int c = 0; c < nc; ++c) q += R[k * nc + c] * R[k * nc + c];
            a2[k] = q;
            smax = fmax(smax, a.s[k]);
        }
        auto phi = [&](double x) {
            double acc = 0.0;
            for (int k = 0; k < K; ++k) {
                const double den = smax + x - a.s[k];
                acc += a2[k] / (den * den);
            }
            return (smax + x) * acc;
        };
        double lo = 0.0, hi = 1.0;
        for (int it = 0; it < 200 && phi(hi) >= cc; ++it) hi *= 2.0;
        for (int it = 0; it < 60; ++it) {
            const double mid = 0.5 * (lo + hi);
            if (phi(mid) > cc) lo = mid;
            else hi = mid;
        }
        const double lam = smax + 0.5 * (lo + hi);
        for (int i = 0; i < V; ++i) nv[i] = R[i] / (lam - a.s[i / nc]);
    }
    for (int i = 0; i < V; ++i) {
        const double di = nv[i] - e[i];
        EI[(long)j * V + i] = e[i];
        dl[(long)j * V + i] = di;
        Wm[((long)(i / nc) * m + j) * nc + (i % nc)] = di;
    }
}

__device__ __forceinline__ double psd_wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// The safeguard's sums, per workgroup: over every bin, sum_b w_b log max(F_b(t), 1e-300) with
// F_b(t) = sum |SE_b + t V_b|^2 for the 31 steps, the same at t = 0 (L0's sum), Re<E, SE>, Re<E, V>; over the
// block, Re<d, SE_I> and Re<d, V_I>.  part[wg * BLML_PSD_NACC + q].
__global__ void __launch_bounds__(BLML_PSD_NT) psd_step_part_kernel(PsdArgs a, long B, const int *I, int m,
                                                                    const double *E, const double *SE,
                                                                    const double *Vv, const double *w,
                                                                    const double *dl, double *part) {
    __shared__ double red[BLML_PSD_NACC][BLML_PSD_NT / 64];
    const int V = a.K * a.nc;
    double acc[BLML_PSD_NACC];
#pragma unroll
    for (int q = 0; q < BLML_PSD_NACC; ++q) acc[q] = 0.0;
    const long stride = (long)gridDim.x * BLML_PSD_NT;
    for (long b = (long)blockIdx.x * BLML_PSD_NT + threadIdx.x; b < B; b += stride) {
        double F0 = 0.0;
        for (int i = 0; i < V; ++i) {
            const double s = SE[b * V + i], e = E[b * V + i], v = Vv[b * V + i];
            F0 += s * s;
            acc[BLML_PSD_NSTEP + 1] += e * s;
            acc[BLML_PSD_NSTEP + 2] += e * v;
        }
        const double wb = w[b];
        acc[BLML_PSD_NSTEP] += wb * log(fmax(F0, 1e-300));
        double t = 1.0;
#pragma unroll
        for (int q = 0; q < BLML_PSD_NSTEP; ++q) {
            double F = 0.0;
            for (int i = 0; i < V; ++i) {
                const double x = SE[b * V + i] + t * Vv[b * V + i];
                F += x * x;
            }
            acc[q] += wb * log(fmax(F, 1e-300));
            t *= 0.5;
        }
    }
    for (long j = (long)blockIdx.x * BLML_PSD_NT + threadIdx.x; j < m; j += stride) {
        const long b = I[j];
        for (int i = 0; i < V; ++i) {
            const double di = dl[j * V + i];
            acc[BLML_PSD_NSTEP + 3] += di * SE[b * V + i];
            acc[BLML_PSD_NSTEP + 4] += di * Vv[b * V + i];
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < BLML_PSD_NACC; ++q) {
        const double x = psd_wave_sum(acc[q]);
        if (lane == 0) red[q][wave] = x;
    }
    __syncthreads();
    if (threadIdx.x < BLML_PSD_NACC) {
        double x = 0.0;
        for (int q = 0; q < BLML_PSD_NT / 64; ++q) x += red[threadIdx.x][q];
        part[(long)blockIdx.x * BLML_PSD_NACC + threadIdx.x] = x;
    }
}

// The step: the partials added in workgroup order; Ls(t) = S(t) - n log max(a0 + a1 t + a2 t^2, 1e-300) against
// L0 = S(0) - n log max(a0, 1e-300) (a1 = Re<d, SE_I> + Re<E, V>, a2 = Re<d, V_I>); the first t with Ls >= L0,
// else 0 (blml_gpu.py:668-677).
__global__ void psd_step_pick_kernel(const double *part, int nb, double n, double *t_out) {
    __shared__ double s[BLML_PSD_NACC];
    if (threadIdx.x < BLML_PSD_NACC) {
        double x = 0.0;
        for (int g = 0; g < nb; ++g) x += part[(long)g * BLML_PSD_NACC + threadIdx.x];
        s[threadIdx.x] = x;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    const double a0 = s[BLML_PSD_NSTEP + 1];
    const double a1 = s[BLML_PSD_NSTEP + 3] + s[BLML_PSD_NSTEP + 2];
    const double a2 = s[BLML_PSD_NSTEP + 4];
    const double L0 = s[BLML_PSD_NSTEP] - n * log(fmax(a0, 1e-300));
    double t = 1.0, pick = 0.0;
    for (int q = 0; q < BLML_PSD_NSTEP; ++q) {
        const double Ls = s[q] - n * log(fmax(a0 + a1 * t + a2 * t * t, 1e-300));
        if (Ls >= L0) {          // false for NaN
            pick = t;
            break;
        }
        t *= 0.5;
    }
    *t_out = pick;
}

// SE += t V over every value; E[I_j] = EI_j + t d_j (t = 1 without the safeguard: t_dev null)
__global__ void __launch_bounds__(BLML_PSD_NT) psd_apply_kernel(long nval, int V, const int *I, int m,
                                                                const double *Vv, const double *EI,
                                                                const double *dl, const double *t_dev, double *E,
                                                                double *SE) {
    const double t = t_dev ? *t_dev : 1.0;
    const long i = (long)blockIdx.x * BLML_PSD_NT + threadIdx.x;
    if (i < nval) SE[i] += t * Vv[i];
    if (i < (long)m * V) {
        const long j = i / V;
        const int q = (int)(i - j * V);
        E[(long)I[j] * V + q] = EI[i] + t * dl[i];
    }
}

// Per-workgroup partials of (Re<E, SE>, sum w log max(F, 1e-300), max rho, non-finite E count) at (sc E, sc SE):
// rho_b = |SE_b |e_b|^2 / (n w_b) - e_b| / |e_b| (blml_gpu.py:616-620).  Added by quick_stats_sum_kernel.
__global__ void __launch_bounds__(BLML_PSD_NT) psd_stats_kernel(long B, int V, const double *E, const double *SE,
                                                                const double *w, double n, double sc, double *part) {
    __shared__ double red[4][BLML_PSD_NT / 64];
    double tr = 0.0, wl = 0.0, rho = 0.0, bad = 0.0;
    for (long b = (long)blockIdx.x * BLML_PSD_NT + threadIdx.x; b < B; b += (long)gridDim.x * BLML_PSD_NT) {
        double F = 0.0, en2 = 0.0;
        for (int i = 0; i < V; ++i) {
            const double e = sc * E[b * V + i], s = sc * SE[b * V + i];
            F += s * s;
            en2 += e * e;
            tr += e * s;
            if (!isfinite(E[b * V + i])) bad += 1.0;
        }
        const double wb = w[b];
        wl += wb * log(fmax(F, 1e-300));
        const double q = en2 / (n * wb);
        double dev = 0.0;
        for (int i = 0; i < V; ++i) {
            const double x = sc * SE[b * V + i] * q - sc * E[b * V + i];
            dev += x * x;
        }
        rho = fmax(rho, sqrt(dev / fmax(en2, 1e-300)));
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    tr = psd_wave_sum(tr);
    wl = psd_wave_sum(wl);
    bad = psd_wave_sum(bad);
    for (int off = 32; off > 0; off >>= 1) rho = fmax(rho, __shfl_xor(rho, off, 64));
    if (lane == 0) { red[0][wave] = tr; red[1][wave] = wl; red[2][wave] = rho; red[3][wave] = bad; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
        for (int q = 0; q < BLML_PSD_NT / 64; ++q) {
            s0 += red[0][q]; s1 += red[1][q]; s2 = fmax(s2, red[2][q]); s3 += red[3][q];
        }
        part[4 * blockIdx.x + 0] = s0;
        part[4 * blockIdx.x + 1] = s1;
        part[4 * blockIdx.x + 2] = s2;
        part[4 * blockIdx.x + 3] = s3;
    }
}

// out = (coef (w_b / max(F_b, 1e-300))) SE_b, in the (B, K, nc) layout or by mode (K, B, nc); coef_dev (a device
// scalar, null: coef) -- the L-BFGS gradient's sources and the canonical form (Q / n) (w / F) S e.
__global__ void __launch_bounds__(BLML_PSD_NT) psd_wscale_kernel(long B, int K, int nc, const double *SE,
                                                                 const double *w, double coef, int by_mode,
                                                                 double *out) {
    const long b = (long)blockIdx.x * BLML_PSD_NT + threadIdx.x;
    if (b >= B) return;
    const int V = K * nc;
    double F = 0.0;
    for (int i = 0; i < V; ++i) F += SE[b * V + i] * SE[b * V + i];
    const double f = coef * (w[b] / fmax(F, 1e-300));
    for (int i = 0; i < V; ++i) {
        const double x = f * SE[b * V + i];
        if (by_mode) out[((long)(i / nc) * B + b) * nc + (i % nc)] = x;
        else out[b * V + i] = x;
    }
}

// (B, K, nc) -> (K, B, nc): each mode's column contiguous, the sources of its own pair pass
__global__ void psd_by_mode_kernel(long B, int K, int nc, const double *X, double *out) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int V = K * nc;
    if (i >= B * V) return;
    const long b = i / V;
    const int q = (int)(i - b * V);
    out[((long)(q / nc) * B + b) * nc + (q % nc)] = X[i];
}

// G = 2 (Vv - (n / Q) SE), Q = stats[0] (Re<E, SE>)
__global__ void psd_grad_kernel(long nval, const double *Vv, const double *SE, const double *stats, double n,
                                double *G) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nval) return;
    const double Q = stats[0];
    G[i] = 2.0 * (Vv[i] - (n / Q) * SE[i]);
}

// ------------------------------------------------------------------------------- pair rows with several columns
// part[(slice * nw + i) * R + r] = the slice's sum over sources j of S(ZA_r, ZB_j) Wt[j * nw + i], i < nw <= NW:
// pair_rows_kernel with NW weight columns staged per chunk (a source whose columns are all zero is skipped).
template <typename R_, bool TL, int NG, int NW>
__global__ void __launch_bounds__(BLML_PAIR_NT) pair_rows_nw_kernel(blml_pairtab_t t, const int *ZA, const int *ZB,
                                                                    const double *Wt, int nw, int R, int M,
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
    const int d = t.d;
    const int r = blockIdx.x * BLML_PAIR_NT + threadIdx.x;
    const bool has = r < R;
    int za[BLML_PAIR_MAXD];
#pragma unroll
    for (int k = 0; k < BLML_PAIR_MAXD; ++k) za[k] = (has && k < d) ? ZA[(long)r * d + k] : 0;
    const long j0 = (long)blockIdx.y * slice;
    const long j1 = j0 + slice < M ? j0 + slice : M;
    double acc[NW];
#pragma unroll
    for (int i = 0; i < NW; ++i) acc[i] = 0.0;
    for (long c0 = j0; c0 < j1; c0 += BLML_PAIR_CH) {
        const int n = (int)(j1 - c0 < BLML_PAIR_CH ? j1 - c0 : BLML_PAIR_CH);
        __syncthreads();
        for (int i = threadIdx.x; i < n * d; i += BLML_PAIR_NT) lz[i] = ZB[c0 * d + i];
        for (int i = threadIdx.x; i < n * NW; i += BLML_PAIR_NT) {
            const int jj = i / NW, q = i - jj * NW;
            lw[i] = q < nw ? Wt[(c0 + jj) * nw + q] : 0.0;
        }
        __syncthreads();
        for (int jj = 0; jj < n; ++jj) {
            bool any = false;
#pragma unroll
            for (int i = 0; i < NW; ++i) any = any || lw[jj * NW + i] != 0.0;
            if (!any) continue;
            R_ vn, vd;
            pair_value<R_, NG>(t, T, RT, za, lz + jj * d, 0, vn, vd);
            const double v = (double)vn;
#pragma unroll
            for (int i = 0; i < NW; ++i) acc[i] += v * lw[jj * NW + i];
        }
    }
    if (has) {
#pragma unroll
        for (int i = 0; i < NW; ++i)
            if (i < nw) part[((long)blockIdx.y * nw + i) * R + r] = acc[i];
    }
}

// out[r * ostride + ooff + i] = sum over slices (in order) of part[(slice * nw + i) * R + r]
__global__ void pair_rows_nw_sum_kernel(const double *part, int R, int nsl, int nw, int ostride, int ooff,
                                        double *out) {
    const long x = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= (long)R * nw) return;
    const int r = (int)(x / nw), i = (int)(x - (long)r * nw);
    double a = 0.0;
    for (int y = 0; y < nsl; ++y) a += part[((long)y * nw + i) * R + r];
    out[(long)r * ostride + ooff + i] = a;
}

// ----------------------------------------------------------------------------------- FP64 vector operations
// The Anderson mixing's and L-BFGS's vector work on the device: Gram matrices of combinations of vectors, linear
// combinations, and dots kept on the device for the two-loop recursion.  Every sum is block partials in index
// order then an ordered sum of the blocks.
struct VecArgs {
    int nv, ny;
    const double *X[BLML_VEC_MAXV];
    double A[BLML_VEC_MAXV][BLML_VEC_MAXV];   // Y_i = sum_j A[i][j] X_j
    double c[BLML_VEC_MAXV];
};

// part[wg * ny * ny + i * ny + j] (i <= j): the workgroup's sum of Y_i Y_j
__global__ void __launch_bounds__(BLML_PSD_NT) vec_gram_part_kernel(VecArgs a, long len, double *part) {
    __shared__ double red[BLML_PSD_NT / 64];
    const int ny = a.ny, nv = a.nv;
    double acc[BLML_VEC_MAXV * (BLML_VEC_MAXV + 1) / 2];
    const int np = ny * (ny + 1) / 2;
    for (int q = 0; q < np; ++q) acc[q] = 0.0;
    for (long x = (long)blockIdx.x * BLML_PSD_NT + threadIdx.x; x < len; x += (long)gridDim.x * BLML_PSD_NT) {
        double xv[BLML_VEC_MAXV], y[BLML_VEC_MAXV];
        for (int j = 0; j < nv; ++j) xv[j] = a.X[j][x];
        for (int i = 0; i < ny; ++i) {
            double s = 0.0;
            for (int j = 0; j < nv; ++j)
                if (a.A[i][j] != 0.0) s += a.A[i][j] * xv[j];
            y[i] = s;
        }
        int q = 0;
        for (int i = 0; i < ny; ++i)
            for (int j = i; j < ny; ++j) acc[q++] += y[i] * y[j];
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int q = 0;
    for (int i = 0; i < ny; ++i)
        for (int j = i; j < ny; ++j, ++q) {
            const double v = psd_wave_sum(acc[q]);
            __syncthreads();
            if (lane == 0) red[wave] = v;
            __syncthreads();
            if (threadIdx.x == 0) {
                double s = 0.0;
                for (int w = 0; w < BLML_PSD_NT / 64; ++w) s += red[w];
                part[(long)blockIdx.x * ny * ny + i * ny + j] = s;
            }
        }
}

// out[i * ny + j] = out[j * ny + i] = sum over workgroups of part
__global__ void vec_gram_sum_kernel(const double *part, int nb, int ny, double *out) {
    const int x = threadIdx.x;
    if (x >= ny * ny) return;
    const int i = x / ny, j = x % ny;
    if (j < i) return;
    double s = 0.0;
    for (int g = 0; g < nb; ++g) s += part[(long)g * ny * ny + x];
    out[i * ny + j] = s;
    out[j * ny + i] = s;
}

// mode 0: out = sum_j c_j X_j (in order); mode 1 (Anderson): out = X_{nv-1} - sum_{j < nv-1} c_j (X_{j+1} - X_j).
// out may be one of the X (every input is read before the element is written).
__global__ void __launch_bounds__(BLML_PSD_NT) vec_lincomb_kernel(VecArgs a, int mode, long len, double *out) {
    const long x = (long)blockIdx.x * BLML_PSD_NT + threadIdx.x;
    if (x >= len) return;
    double v;
    if (mode == 0) {
        v = a.c[0] * a.X[0][x];
        for (int j = 1; j < a.nv; ++j) v += a.c[j] * a.X[j][x];
    } else {
        v = a.X[a.nv - 1][x];
        for (int j = 0; j + 1 < a.nv; ++j) v -= a.c[j] * (a.X[j + 1][x] - a.X[j][x]);
    }
    out[x] = v;
}

// part[wg] = the workgroup's sum of X Y
__global__ void __launch_bounds__(BLML_PSD_NT) vec_dot_part_kernel(const double *X, const double *Y, long len,
                                                                   double *part) {
    __shared__ double red[BLML_PSD_NT / 64];
    double acc = 0.0;
    for (long x = (long)blockIdx.x * BLML_PSD_NT + threadIdx.x; x < len; x += (long)gridDim.x * BLML_PSD_NT)
        acc += X[x] * Y[x];
    acc = psd_wave_sum(acc);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) red[wave] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < BLML_PSD_NT / 64; ++w) s += red[w];
        part[blockIdx.x] = s;
    }
}

// *out = rho (sum of part) (mode 0), or *out = a_in[0] - rho (sum of part) (mode 1)
__global__ void vec_dot_final_kernel(const double *part, int nb, double rho, int mode, const double *a_in,
                                     double *out) {
    if (threadIdx.x != 0) return;
    double s = 0.0;
    for (int g = 0; g < nb; ++g) s += part[g];
    *out = mode == 0 ? rho * s : *a_in - rho * s;
}

// y += sgn c_dev[0] x
__global__ void vec_axpy_dev_kernel(long len, const double *c_dev, double sgn, const double *X, double *Y) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= len) return;
    Y[i] += (sgn * *c_dev) * X[i];
}
