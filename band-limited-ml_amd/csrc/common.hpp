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

// two / four consecutive FP32 values read as one 8- / 16-byte access (the address aligned to the access's size)
struct alignas(8) f32x2 {
    float v[2];
};
struct alignas(16) f32x4 {
    float v[4];
};
template <int W>
struct f32xw;
template <>
struct f32xw<2> { typedef f32x2 type; };
template <>
struct f32xw<4> { typedef f32x4 type; };

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

// q = a / d, r = a % d for 0 <= a and 1 <= d, from a float reciprocal with one correction step when a < 2^24
// (exact there; integer division is a long instruction sequence on the GPU), by integer division beyond.
__device__ __forceinline__ int fdivmod(int a, int d, float inv_d, int &r) {
    if (a >= (1 << 24)) {
        const int q = a / d;
        r = a - q * d;
        return q;
    }
    int q = (int)((float)a * inv_d);
    r = a - q * d;
    if (r < 0) { --q; r += d; }
    else if (r >= d) { ++q; r -= d; }
    return q;
}

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

__device__ __forceinline__ double block_max1(double v, double *red) { return block_max(v, red); }

// ------------------------------------------------------------------------------------------------ teams
// A team is the set of workgroups that runs one fit: one workgroup (rank 0 of 1, the batched path), or every
// workgroup of a cooperative launch (a fit too large for one CU).  Element loops stride over the team; sync()
// orders every team member's global writes before any member's later reads; sum / max reduce over the team in
// FP64 and give every thread of every member the same totals (partials combined in rank order), so the solver's
// scalar control flow stays identical across the team.

#define BLML_BAR_SPIN (1L << 26)    // polls (each with a short sleep) before a grid barrier declares a timeout

// bar[0]: arrivals, bar[1]: generation, bar[2]: set when a wait timed out (the results carry it to the host).
// Every thread first waits until its own memory operations are complete (its stores have reached the L2); after
// the workgroup barrier one thread publishes them with a release fence at agent scope (the L2's dirty lines are
// written back: the L2s of the XCDs are not coherent by themselves), arrives, waits, and after the wait takes one
// acquire fence at agent scope (this XCD's L2 and the CU's caches are invalidated), so plain global stores made
// before the barrier by any workgroup are visible to loads made after it by any other.  The barrier words are
// only touched by agent-scope atomics, which are coherent without fences: the polls are relaxed loads, so a
// waiting workgroup does not invalidate its L2 once per poll while the others are still writing back.  The
// generation is read to completion before the arrival is counted (else the last arrival could overtake it and
// the wait would never see a change), and the reset of the count completes before the generation advances.
// One fence pair per workgroup, not per wave: each is a cache-wide operation.  The wait is bounded: the kernel
// always ends, and a timeout is reported instead of hanging.
__device__ inline void grid_barrier(unsigned *bar, int nranks) {
    __builtin_amdgcn_s_waitcnt(0);      // vmcnt = expcnt = lgkmcnt = 0
    __syncthreads();
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const unsigned gen = __hip_atomic_load(&bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __builtin_amdgcn_s_waitcnt(0);
        const unsigned arrived = __hip_atomic_fetch_add(&bar[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (arrived == (unsigned)nranks - 1) {
            __hip_atomic_store(&bar[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __builtin_amdgcn_s_waitcnt(0);
            __hip_atomic_fetch_add(&bar[1], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            long spins = 0;
            // after any timeout every later wait is skipped, so a failed team still finishes promptly
            while (__hip_atomic_load(&bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen &&
                   __hip_atomic_load(&bar[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u) {
                __builtin_amdgcn_s_sleep(2);
                if (++spins > BLML_BAR_SPIN) {
                    __hip_atomic_fetch_or(&bar[2], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
}

struct Team {
    int rank, nranks;
    unsigned *bar;       // grid teams: the barrier words
    double *part;        // grid teams: 2 * nranks * BLML_NRED partial sums (two buffers, alternating)
    int parity;

    __device__ long first() const { return (long)rank * BLML_NT + threadIdx.x; }
    __device__ long stride() const { return (long)nranks * BLML_NT; }
    __device__ bool lead() const { return rank == 0 && threadIdx.x == 0; }
    __device__ void sync() const {
        if (nranks > 1) grid_barrier(bar, nranks);
        else __syncthreads();
    }
    // combines one value per member (already reduced over the workgroup) into red[i] for i < N, in rank order;
    // op 0: sum, 1: max
    template <int N>
    __device__ void combine(const double (&v)[N], double *red, int op) {
        double *buf = part + (long)parity * nranks * BLML_NRED;
        parity ^= 1;
        if (threadIdx.x == 0)
            for (int i = 0; i < N; ++i) buf[(long)rank * BLML_NRED + i] = v[i];
        grid_barrier(bar, nranks);
        const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
        for (int i = wid; i < N; i += BLML_NWAVE) {
            double a = op ? -INFINITY : 0.0;
            for (int r = lane; r < nranks; r += 64) {
                const double x = __hip_atomic_load(&buf[(long)r * BLML_NRED + i], __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT);
                a = op ? fmax(a, x) : a + x;
            }
            a = op ? wave_max(a) : wave_sum(a);
            if (lane == 0) red[BLML_NWAVE * BLML_NRED - BLML_NRED + i] = a;
        }
        __syncthreads();
    }
    template <int N>
    __device__ void sum(double (&v)[N], double *red) {
        block_sum<N>(v, red);
        if (nranks == 1) return;
        combine<N>(v, red, 0);
#pragma unroll
        for (int i = 0; i < N; ++i) v[i] = red[BLML_NWAVE * BLML_NRED - BLML_NRED + i];
        __syncthreads();
    }
    __device__ double sum1(double v, double *red) {
        double a[1] = {v};
        sum<1>(a, red);
        return a[0];
    }
    __device__ double max(double v, double *red) {
        double a[1] = {block_max(v, red)};
        if (nranks == 1) return a[0];
        combine<1>(a, red, 1);
        a[0] = red[BLML_NWAVE * BLML_NRED - BLML_NRED];
        __syncthreads();
        return a[0];
    }
};

__device__ inline Team block_team() { return Team{0, 1, nullptr, nullptr, 0}; }
