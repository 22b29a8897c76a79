// Shared device helpers of the binned-fit engine: complex type, workgroup reductions in FP64.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define BLML_NT 512                 // threads per workgroup (8 wavefronts of 64)
#define BLML_NWAVE (BLML_NT / 64)
#define BLML_NRED 24                // simultaneous FP64 sums one reduction can carry

template <typename T>
struct cpx {
    T x, y;
};

// The launch's dynamic LDS.  Functions that are not inlined reach LDS only through this declaration (and offsets
// into it), never through pointer arguments: a pointer argument is generic, which turns every LDS access into a
// FLAT instruction, while a pointer derived from a __shared__ object in the same function compiles to ds_ ops.
extern __shared__ __align__(16) unsigned char g_dyn_lds[];
template <typename T>
__device__ __forceinline__ cpx<T> *dyn_lds() { return reinterpret_cast<cpx<T> *>(g_dyn_lds); }

template <typename T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ cpx<T> cmulc(cpx<T> a, cpx<T> b) {  // a * conj(b)
    return {a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
template <typename T>
__device__ __forceinline__ cpx<T> cadd(cpx<T> a, cpx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> csub(cpx<T> a, cpx<T> b) { return {a.x - b.x, a.y - b.y}; }

__device__ __forceinline__ double wave_sum(double a) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
    return a;
}
__device__ __forceinline__ double wave_max(double a) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) a = fmax(a, __shfl_xor(a, off, 64));
    return a;
}

// Sums N per-thread FP64 partials over the workgroup.  Every thread ends with the same totals, summed in the
// same fixed order (lanes by butterfly, then waves 0..3), so results are deterministic run to run and identical
// across threads (the solver's scalar control flow depends on them).  red: LDS, >= BLML_NWAVE * N doubles.
template <int N>
__device__ __forceinline__ void block_sum(double (&v)[N], double *red) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = wave_sum(v[i]);
    __syncthreads();
    if (lane == 0) {
#pragma unroll
        for (int i = 0; i < N; ++i) red[wid * N + i] = v[i];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < N; ++i) {
        double a = red[i];
#pragma unroll
        for (int w = 1; w < BLML_NWAVE; ++w) a += red[w * N + i];
        v[i] = a;
    }
}

__device__ __forceinline__ double block_max(double v, double *red) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    v = wave_max(v);
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double a = red[0];
#pragma unroll
    for (int w = 1; w < BLML_NWAVE; ++w) a = fmax(a, red[w]);
    return a;
}

__device__ __forceinline__ double block_sum1(double v, double *red) {
    double a[1] = {v};
    block_sum<1>(a, red);
    return a[0];
}
