// The warm-start network's kernels: 3x3 "same" convolutions on NCHW images with BF16 operands on the matrix cores
// (v_mfma_f32_32x32x8bf16_1k, FP32 accumulation), their weight gradients, GELU, the masked squared loss and Adam.
//
// conv:  Z[n, co, p] = sum_{ci, t} W[co, ci, t] X[n, ci, p + delta_t] + b[co]  (t = 3 dy + dx, delta_t = (dy-1, dx-1),
//        zero padding), and optionally A = bf16(GELU(Z)).  A workgroup (4 waves) takes two image rows (W = 64:
//        128 pixels, 32 per wave) of one image; every wave computes a 32 (output channels) x 32 (pixels) tile over
//        K = Cin * 9 in steps of 8.  The input rows with their halo and the weights are staged in LDS.  The data
//        gradient is the same convolution with W'[ci, co, t] = W[co, ci, 8 - t] applied to dZ.
// dW:    dW[co, ci, t] = sum_{n, p} dZ[n, co, p] X[n, ci, p + delta_t]: one workgroup per image (pixels are the K
//        dimension), partial sums per image, then an ordered reduction over the images.
#pragma once
#include "common.hpp"

#define BLML_NET_NT 256
#define BLML_NET_W 64            // image width the conv tiling is written for
#define BLML_NET_RT 2            // image rows per workgroup
#define BLML_NET_PW (BLML_NET_W + 2)
#define BLML_NET_PATCH ((BLML_NET_RT + 2) * BLML_NET_PW)

typedef unsigned short bf16_t;
typedef short net_v4s __attribute__((ext_vector_type(4)));
typedef float net_v16f __attribute__((ext_vector_type(16)));

__device__ __forceinline__ float bf2f(bf16_t h) { return __uint_as_float((unsigned)h << 16); }
// round to nearest even
__device__ __forceinline__ bf16_t f2bf(float f) {
    unsigned u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (bf16_t)(u >> 16);     // inf / nan: truncate
    u += 0x7fffu + ((u >> 16) & 1u);
    return (bf16_t)(u >> 16);
}
__device__ __forceinline__ float net_gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float net_gelu_d(float x) {
    return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}

// accumulator register r of lane l: row (r & 3) + 8 (r >> 2) + 4 (l >> 5), column l & 31
__device__ __forceinline__ int net_acc_row(int r, int l) { return (r & 3) + 8 * (r >> 2) + 4 * (l >> 5); }

// LDS: Xs [Cin][RT + 2][W + 2] bf16 (+ one zero), Ws [32][Kpad] bf16, ks [Kpad] int (patch offset of k, -1: pad)
__global__ void __launch_bounds__(BLML_NET_NT) net_conv_kernel(const bf16_t *X, const bf16_t *Wk, const float *bias,
                                                               float *Z, bf16_t *A, int Cin, int Cout, int H,
                                                               int gelu) {
    const int W = BLML_NET_W;
    const int K = Cin * 9, Kpad = (K + 7) / 8 * 8;
    const int n = blockIdx.y, y0 = blockIdx.x * BLML_NET_RT;
    bf16_t *Xs = reinterpret_cast<bf16_t *>(g_dyn_lds);
    const int xs_len = Cin * BLML_NET_PATCH;
    bf16_t *Ws = Xs + ((xs_len + 1 + 7) / 8) * 8;
    int *ks = reinterpret_cast<int *>(Ws + 32 * Kpad);
    for (int i = threadIdx.x; i < xs_len; i += BLML_NET_NT) {
        const int ci = i / BLML_NET_PATCH, rem = i - ci * BLML_NET_PATCH;
        const int rr = rem / BLML_NET_PW, cc = rem - rr * BLML_NET_PW;
        const int row = y0 - 1 + rr, col = cc - 1;
        bf16_t v = 0;
        if (row >= 0 && row < H && col >= 0 && col < W) v = X[(((long)n * Cin + ci) * H + row) * W + col];
        Xs[i] = v;
    }
    for (int i = threadIdx.x; i < 32 * Kpad; i += BLML_NET_NT) {
        const int co = i / Kpad, k = i - co * Kpad;
        Ws[i] = (co < Cout && k < K) ? Wk[(long)co * K + k] : (bf16_t)0;
    }
    for (int k = threadIdx.x; k < Kpad; k += BLML_NET_NT) {
        int o = -1;
        if (k < K) {
            const int ci = k / 9, t = k - ci * 9;
            o = ci * BLML_NET_PATCH + (t / 3) * BLML_NET_PW + (t % 3);
        }
        ks[k] = o;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int rr = wave >> 1, c0 = (wave & 1) * 32;
    const int pixbase = rr * BLML_NET_PW + c0 + (lane & 31);
    const int half = lane >> 5;
    net_v16f acc;
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
    for (int kk = 0; kk < Kpad; kk += 8) {
        const int kb = kk + 4 * half;
        net_v4s a, b;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            a[e] = (short)Ws[(lane & 31) * Kpad + kb + e];
            const int o = ks[kb + e];
            b[e] = (short)(o >= 0 ? Xs[pixbase + o] : (bf16_t)0);
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
    }
    const int y = y0 + rr, x = c0 + (lane & 31);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int co = net_acc_row(r, lane);
        if (co < Cout) {
            const float z = acc[r] + (bias ? bias[co] : 0.0f);
            const long o = (((long)n * Cout + co) * H + y) * W + x;
            Z[o] = z;
            if (gelu) A[o] = f2bf(net_gelu(z));
        }
    }
}

// part[n][co][j] = sum_p dZ[n, co, p] X[n, ci, p + delta_t], j = ci * 9 + t < Cin * 9.  One workgroup per image;
// LDS: Xs (as above), Ds [32][128] bf16 (dZ of the two rows, channels past Cout zero).
__global__ void __launch_bounds__(BLML_NET_NT) net_conv_dw_kernel(const bf16_t *X, const bf16_t *dZ, float *part,
                                                                  int Cin, int Cout, int H) {
    const int W = BLML_NET_W, PX = BLML_NET_RT * BLML_NET_W;
    const int K = Cin * 9, ntile = (K + 31) / 32;
    const int n = blockIdx.x;
    bf16_t *Xs = reinterpret_cast<bf16_t *>(g_dyn_lds);
    const int xs_len = Cin * BLML_NET_PATCH;
    bf16_t *Ds = Xs + ((xs_len + 7) / 8) * 8;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, half = lane >> 5;
    constexpr int MT = 3;                 // tiles per wave: K = 288 is 9 tiles of 32 over 4 waves
    net_v16f acc[MT];
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[m][r] = 0.0f;
    // this lane's column of each of its tiles: (ci, t) -> patch offset (-1: past K)
    int colo[MT];
#pragma unroll
    for (int m = 0; m < MT; ++m) {
        const int j = (wave + 4 * m) * 32 + (lane & 31);
        colo[m] = -1;
        if (wave + 4 * m < ntile && j < K) {
            const int ci = j / 9, t = j - ci * 9;
            colo[m] = ci * BLML_NET_PATCH + (t / 3) * BLML_NET_PW + (t % 3);
        }
    }
    for (int y0 = 0; y0 < H; y0 += BLML_NET_RT) {
        __syncthreads();
        for (int i = threadIdx.x; i < xs_len; i += BLML_NET_NT) {
            const int ci = i / BLML_NET_PATCH, rem = i - ci * BLML_NET_PATCH;
            const int rr = rem / BLML_NET_PW, cc = rem - rr * BLML_NET_PW;
            const int row = y0 - 1 + rr, col = cc - 1;
            bf16_t v = 0;
            if (row >= 0 && row < H && col >= 0 && col < W) v = X[(((long)n * Cin + ci) * H + row) * W + col];
            Xs[i] = v;
        }
        for (int i = threadIdx.x; i < 32 * PX; i += BLML_NET_NT) {
            const int co = i / PX, p = i - co * PX;
            Ds[i] = co < Cout ? dZ[(((long)n * Cout + co) * H + y0) * W + p] : (bf16_t)0;
        }
        __syncthreads();
        for (int kk = 0; kk < PX; kk += 8) {
            const int pb = kk + 4 * half;
            net_v4s a;
            int pofs[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                a[e] = (short)Ds[(lane & 31) * PX + pb + e];
                const int p = pb + e, prr = p / BLML_NET_W, pcc = p - prr * BLML_NET_W;
                pofs[e] = prr * BLML_NET_PW + pcc;
            }
#pragma unroll
            for (int m = 0; m < MT; ++m) {
                if (wave + 4 * m < ntile) {
                    net_v4s b;
#pragma unroll
                    for (int e = 0; e < 4; ++e) b[e] = (short)(colo[m] >= 0 ? Xs[pofs[e] + colo[m]] : (bf16_t)0);
                    acc[m] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc[m], 0, 0, 0);
                }
            }
        }
    }
#pragma unroll
    for (int m = 0; m < MT; ++m) {
        const int j = (wave + 4 * m) * 32 + (lane & 31);
        if (wave + 4 * m >= ntile || j >= K) continue;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int co = net_acc_row(r, lane);
            if (co < Cout) part[((long)n * Cout + co) * K + j] = acc[m][r];
        }
    }
}

// out[i] = sum over n of part[n][i] (ordered), i < len
__global__ void net_sum_parts_kernel(const float *part, int N, int len, float *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= len) return;
    float a = 0.0f;
    for (int n = 0; n < N; ++n) a += part[(long)n * len + i];
    out[i] = a;
}

// db_part[n][co] = sum_p dZ[n, co, p] (bf16 operand values, FP32 sum); one workgroup per (n, co)
__global__ void __launch_bounds__(BLML_NET_NT) net_bias_grad_kernel(const bf16_t *dZ, int Cout, int HW, float *part) {
    __shared__ float red[BLML_NWAVE];
    const long base = (long)blockIdx.x * HW;
    float a = 0.0f;
    for (int p = threadIdx.x; p < HW; p += BLML_NET_NT) a += bf2f(dZ[base + p]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
        float s = 0.0f;
        for (int w = 0; w < BLML_NET_NT / 64; ++w) s += red[w];
        part[blockIdx.x] = s;
    }
}

// dZ = bf16(dA * GELU'(Z))
__global__ void net_gelu_bwd_kernel(const float *dA, const float *Z, bf16_t *dZ, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dZ[i] = f2bf(dA[i] * net_gelu_d(Z[i]));
}

// masked squared loss: dZ = bf16(2 M (Z - Y) / msum); loss_part[block] = sum M (Z - Y)^2 / msum
__global__ void __launch_bounds__(BLML_NET_NT) net_mse_kernel(const float *Z, const float *Y, const float *Mk,
                                                              bf16_t *dZ, float *loss_part, long n, float inv_msum) {
    __shared__ float red[BLML_NWAVE];
    const long i = (long)blockIdx.x * BLML_NET_NT + threadIdx.x;
    float l = 0.0f;
    if (i < n) {
        const float r = Z[i] - Y[i], m = Mk[i];
        dZ[i] = f2bf(2.0f * m * r * inv_msum);
        l = m * r * r * inv_msum;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = l;
    __syncthreads();
    if (threadIdx.x == 0) {
        float s = 0.0f;
        for (int w = 0; w < BLML_NET_NT / 64; ++w) s += red[w];
        loss_part[blockIdx.x] = s;
    }
}

// W'[ci, co, t] = W[co, ci, 8 - t] (the data gradient's weights), bf16
__global__ void net_flip_kernel(const bf16_t *Wk, bf16_t *Wf, int Cin, int Cout) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Cin * Cout * 9) return;
    const int ci = i / (Cout * 9), rem = i - ci * Cout * 9, co = rem / 9, t = rem - co * 9;
    Wf[i] = Wk[((long)co * Cin + ci) * 9 + (8 - t)];
}

// Adam on FP32 master weights, with the BF16 copy the convolutions read (bf may be null: biases stay FP32)
__global__ void net_adam_kernel(float *p, const float *g, float *m, float *v, bf16_t *bf, long n, float lr, float b1,
                                float b2, float eps, float c1, float c2) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    p[i] = pi;
    if (bf) bf[i] = f2bf(pi);
}

// bf16 <- f32
__global__ void net_to_bf16_kernel(const float *x, bf16_t *y, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = f2bf(x[i]);
}
