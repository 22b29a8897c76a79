// The continuous kernel at arbitrary points (BLMLBinned.sqrt_density / predict / predict_grid,
// blml_binned.py:452-499): s_q = sum_j K(x_q - z_j) w_j for query points x_q and lattice bins z_j, K the product of
// a mode's units evaluated from their closed forms (blml_trivial.py: sinc :368-370, slits :442-444, dirichlet
// :509-515, ball :588-591 with the profile :135-143, periodic ball :656-665).  One thread per query point, the
// bins staged in LDS in chunks (every lane reads the same bin: a broadcast), FP64 throughout; sources split into
// slices over grid.y, each (point, slice) partial an FP64 sum in bin order, the slices added in order by a second
// kernel, so the result does not depend on scheduling.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#include "blml_hip.h"

#define BLML_POINT_NT 256

struct PointTab {
    int32_t d, U;
    blml_pointunit_t u[BLML_POINT_MAXU];
};

__device__ __forceinline__ double pt_sinc(double x) { return x == 0.0 ? 1.0 : sinpi(x) / (M_PI * x); }

// FT of the ball |nu| <= 1/2 in D dimensions at radius s: (1/2)^(D/2) J_{D/2}(pi s) / s^(D/2), its volume at 0;
// half-integer orders through the spherical Bessel functions (J_{n+1/2}(x) = sqrt(2 x / pi) j_n(x)), whose
// small-argument series avoid the cancellation of their closed forms
__device__ inline double pt_ball_profile(double s, int D) {
    const double x = M_PI * s;
    switch (D) {
        case 1:
            return pt_sinc(s);
        case 2:
            return s == 0.0 ? M_PI / 4.0 : 0.5 * j1(x) / s;
        case 3: {
            if (s == 0.0) return M_PI / 6.0;
            const double x2 = x * x;
            const double sj1 = x < 1e-2 ? x / 3.0 - x * x2 / 30.0 + x * x2 * x2 / 840.0
                                        : (sin(x) / x - cos(x)) / x;
            return 0.5 * sj1 / s;
        }
        case 4: {
            if (s == 0.0) return M_PI * M_PI / 32.0;
            const double J2 = x < 1e-3 ? x * x / 8.0 - x * x * x * x / 96.0 : jn(2, x);
            return 0.25 * J2 / (s * s);
        }
        default: {   // 5
            if (s == 0.0) return M_PI * M_PI / 60.0;
            const double x2 = x * x;
            const double sj2 = x < 5e-2 ? x2 / 15.0 - x2 * x2 / 210.0 + x2 * x2 * x2 / 7560.0
                                        : (3.0 / x2 - 1.0) * sin(x) / x - 3.0 * cos(x) / x2;
            return 0.25 * sj2 / (s * s);
        }
    }
}

// one unit at the coordinate differences t (per column)
__device__ inline double pt_unit(const blml_pointunit_t &u, const double *t, const int32_t *harm) {
    switch (u.kind) {
        case 0: {                                   // sinc
            const double f = u.fc[0];
            return f * pt_sinc(f * t[u.col[0]]);
        }
        case 1: {                                   // slits: fc sinc(fc t) - fi sinc(fi t)
            const double f = u.fc[0], tt = t[u.col[0]];
            return f * pt_sinc(f * tt) - u.fi * pt_sinc(u.fi * tt);
        }
        case 2: {                                   // dirichlet (with fill): the folded distance
            const double P = u.period[0];
            double tt = fmod(fabs(t[u.col[0]]), P);
            tt = fmin(tt, P - tt);
            if (tt < 1e-10) return u.scale * (double)(u.M - u.Mi);
            const double r = tt / P;
            return u.scale * (sinpi(u.M * r) - sinpi(u.Mi * r)) / sinpi(r);
        }
        case 3: {                                   // ball over its columns
            double s2 = 0.0, pf = 1.0;
            for (int i = 0; i < u.ncol; ++i) {
                const double v = u.fc[i] * t[u.col[i]];
                s2 += v * v;
                pf *= u.fc[i];
            }
            return pf * pt_ball_profile(sqrt(s2), u.ncol);
        }
        default: {                                  // periodic ball: the harmonic sum
            double acc = 0.0;
            const int32_t *h = harm + u.hoff;
            for (int q = 0; q < u.nharm; ++q) {
                double ph = 0.0;
                for (int i = 0; i < u.ncol; ++i) ph += (double)h[q * u.ncol + i] * t[u.col[i]] / u.period[i];
                acc += cospi(2.0 * ph);
            }
            return u.scale * acc;
        }
    }
}

template <int NW>
__global__ void __launch_bounds__(BLML_POINT_NT) point_rows_kernel(PointTab tab, const int32_t *harm, const double *X,
                                                                   const double *Z, const double *W, int Nq, int M,
                                                                   int slice, double *part) {
    __shared__ double sz[BLML_POINT_NT * BLML_MAXD];
    __shared__ double sw[BLML_POINT_NT * NW];
    const int d = tab.d;
    const int q = blockIdx.x * BLML_POINT_NT + threadIdx.x;
    const bool act = q < Nq;
    double x[BLML_MAXD];
#pragma unroll
    for (int c = 0; c < BLML_MAXD; ++c) x[c] = (act && c < d) ? X[(long)q * d + c] : 0.0;
    double acc[NW];
#pragma unroll
    for (int i = 0; i < NW; ++i) acc[i] = 0.0;
    const int j0 = blockIdx.y * slice, j1 = min(M, j0 + slice);
    for (int c0 = j0; c0 < j1; c0 += BLML_POINT_NT) {
        const int n = min(BLML_POINT_NT, j1 - c0);
        __syncthreads();
        for (int i = threadIdx.x; i < n * d; i += BLML_POINT_NT) sz[i] = Z[(long)c0 * d + i];
        for (int i = threadIdx.x; i < n * NW; i += BLML_POINT_NT) sw[i] = W[(long)c0 * NW + i];
        __syncthreads();
        for (int jj = 0; jj < n; ++jj) {
            bool any = false;
#pragma unroll
            for (int i = 0; i < NW; ++i) any = any || sw[jj * NW + i] != 0.0;
            if (!any) continue;
            double t[BLML_MAXD];
#pragma unroll
            for (int c = 0; c < BLML_MAXD; ++c) t[c] = c < d ? x[c] - sz[jj * d + c] : 0.0;
            double K = 1.0;
            for (int u = 0; u < tab.U; ++u) K *= pt_unit(tab.u[u], t, harm);
#pragma unroll
            for (int i = 0; i < NW; ++i) acc[i] += K * sw[jj * NW + i];
        }
    }
    if (act) {
#pragma unroll
        for (int i = 0; i < NW; ++i) part[((long)blockIdx.y * Nq + q) * NW + i] = acc[i];
    }
}

// s[q, i] = sum over slices (in order) of part[slice, q, i]
__global__ void point_rows_sum_kernel(const double *part, int Nq, int nsl, int nw, double *s) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)Nq * nw) return;
    double a = 0.0;
    for (int y = 0; y < nsl; ++y) a += part[(long)y * Nq * nw + i];
    s[i] = a;
}
