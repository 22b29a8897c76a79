// A CPU stand-in for the parts of HIP that the engine's device headers use, so that the fit's device code
// (common.hpp, transform.hpp, solver.hpp, fit.hpp) compiles and runs on the host under AddressSanitizer: every
// global (device), LDS and private access of one workgroup is then checked against its allocation.  Test
// infrastructure only; the library never includes it.
//
// One workgroup = blockDim.x fibers on the calling thread, run round robin (a stack switch of their own: glibc's
// swapcontext makes a system call per switch, and AddressSanitizer's interceptor clears the whole target stack).  __syncthreads() parks the
// fiber until every fiber of the block has arrived; a wavefront shuffle goes through an exchange slot per thread and
// two barriers of its 64-lane wave.  A fiber that ends while others wait at a barrier, or a barrier that no fiber can
// pass, is reported as a deadlock (a divergent barrier on the device).  A grid of several workgroups (a cooperative
// launch) runs the same way, every workgroup with its own barrier, its own LDS objects and its own dynamic LDS (each
// a heap block of the exact size, so AddressSanitizer bounds it); the build rewrites the device headers' LDS
// declarations into calls of lds_obj() / dyn_lds_base() (replay.py).  A grid barrier's polling thread yields at
// each poll (s_sleep).  The matrix-core instructions are emulated on a wave's 64 lanes with the lane layouts the
// separable path documents.
#pragma once
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <functional>
#include <vector>

#if defined(__has_feature)
#if __has_feature(address_sanitizer)
#define CPUHIP_ASAN 1
#endif
#endif
#if defined(__SANITIZE_ADDRESS__)
#define CPUHIP_ASAN 1
#endif
#ifdef CPUHIP_ASAN
#include <sanitizer/asan_interface.h>
#include <sanitizer/common_interface_defs.h>
#endif

#define __device__
#define __host__
#define __global__
#define __forceinline__ inline __attribute__((always_inline))
#define __noinline__ __attribute__((noinline))
#define __launch_bounds__(...)
#define __align__(n) __attribute__((aligned(n)))

#define __HIP_MEMORY_SCOPE_AGENT 4   // clang's __hip_atomic_* builtins (host code too) take its scope numbers
#define __builtin_amdgcn_readfirstlane(x) (x)
#define __builtin_amdgcn_s_waitcnt(x) ((void)0)
#define __builtin_amdgcn_s_sleep(x) cpuhip::yield()
#define __builtin_amdgcn_fence(o, s) ((void)0)
#define __builtin_amdgcn_sched_barrier(x) ((void)0)
#define __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, x, y, z) cpuhip::mfma_f32_32x32x2(a, b, c)
#define __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, x, y, z) cpuhip::mfma_f64_16x16x4(a, b, c)
#define __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, x, y, z) cpuhip::mfma_f32_16x16x4(a, b, c)

struct cpuhip_dim3 {
    unsigned x = 0, y = 0, z = 0;
};
extern cpuhip_dim3 threadIdx, blockIdx, blockDim, gridDim;

namespace cpuhip {

// x86-64 System V stack switch: saves the callee-saved registers, MXCSR and the x87 control word on the current
// stack, stores the stack pointer in *from and resumes the stack `to` (a new fiber's stack is primed to "return"
// into cpuhip_trampoline, which calls r12(rbx)).
extern "C" void cpuhip_switch(void **from, void *to);
extern "C" void cpuhip_trampoline();

struct Fiber {
    void *sp = nullptr;
    char *stack = nullptr;
    size_t stack_size = 0;
    int tid = 0;
    int state = 0;        // 0 runnable, 1 at the block barrier, 2 at its wave barrier, 3 done
};

struct Block {
    int n = 0;
    unsigned id = 0;
    std::vector<Fiber> f;
    void *main_sp = nullptr;
    int cur = -1;
    int arrived = 0;
    int done = 0;
    std::vector<int> wave_arrived;
    std::vector<double> xch, xa, xb;
    std::function<void()> body;
    long long clock = 0;
    long long barriers = 0;
    void *main_fake = nullptr;
    const void *main_bottom = nullptr;
    size_t main_size = 0;
    unsigned char *dyn = nullptr;            // this workgroup's dynamic LDS (exactly the launch's size)
    std::vector<void *> objs;                // its LDS objects, by slot
    std::vector<size_t> obj_bytes;
};

extern Block *g_blk;
extern long long g_clock;

inline unsigned char *dyn_lds_base() { return g_blk->dyn; }

// the workgroup's LDS object `slot` of n values of T (one heap block each, filled with a pattern: LDS is not
// cleared at launch)
template <typename T>
inline T *lds_obj(int slot, size_t n = 1) {
    Block &b = *g_blk;
    if ((int)b.objs.size() <= slot) {
        b.objs.resize(slot + 1, nullptr);
        b.obj_bytes.resize(slot + 1, 0);
    }
    if (!b.objs[slot]) {
        b.obj_bytes[slot] = sizeof(T) * n;
        b.objs[slot] = malloc(sizeof(T) * n);
        memset(b.objs[slot], 0xA5, sizeof(T) * n);
    } else if (b.obj_bytes[slot] != sizeof(T) * n) {
        fprintf(stderr, "cpuhip: LDS object %d asked for with two sizes\n", slot);
        abort();
    }
    return (T *)b.objs[slot];
}

[[noreturn]] inline void die(const char *msg) {
    fprintf(stderr, "cpuhip: %s (thread %d)\n", msg, g_blk ? g_blk->cur : -1);
    fflush(stderr);
    abort();
}

// fiber -> scheduler
inline void yield() {
    Block &b = *g_blk;
    Fiber &me = b.f[b.cur];
#ifdef CPUHIP_ASAN
    void *fake = nullptr;
    __sanitizer_start_switch_fiber(&fake, b.main_bottom, b.main_size);
    cpuhip_switch(&me.sp, b.main_sp);
    __sanitizer_finish_switch_fiber(fake, nullptr, nullptr);
#else
    cpuhip_switch(&me.sp, b.main_sp);
#endif
}

inline void syncthreads() {
    Block &b = *g_blk;
    ++b.barriers;
    if (++b.arrived == b.n) {
        b.arrived = 0;
        for (auto &x : b.f)
            if (x.state == 1) x.state = 0;
        return;
    }
    b.f[b.cur].state = 1;
    yield();
}

inline void wave_sync() {
    Block &b = *g_blk;
    const int w = b.cur / 64;
    const int wn = std::min(64, b.n - w * 64);
    if (++b.wave_arrived[w] == wn) {
        b.wave_arrived[w] = 0;
        for (int t = w * 64; t < w * 64 + wn; ++t)
            if (b.f[t].state == 2) b.f[t].state = 0;
        return;
    }
    b.f[b.cur].state = 2;
    yield();
}

inline double shfl_xor(double v, int mask, int width) {
    Block &b = *g_blk;
    const int t = b.cur;
    if (width != 64) die("__shfl_xor: width 64 only");
    b.xch[t] = v;
    wave_sync();
    const int src = (t & ~63) | ((t & 63) ^ mask);
    if (src >= b.n) die("__shfl_xor: lane outside the block");
    const double r = b.xch[src];
    wave_sync();
    return r;
}

// v_mfma_f64_16x16x4f64 on one wave: A[i][k] in lane i + 16 k, B[k][j] in lane j + 16 k; accumulator register r of
// lane l holds D[4 r + (l >> 4)][l & 15]
typedef double cpuhip_d4 __attribute__((ext_vector_type(4)));
inline cpuhip_d4 mfma_f64_16x16x4(double a, double b, cpuhip_d4 c) {
    Block &B = *g_blk;
    const int t = B.cur, w0 = t & ~63, l = t & 63;
    B.xa[t] = a;
    B.xb[t] = b;
    wave_sync();
    cpuhip_d4 d = c;
    for (int r = 0; r < 4; ++r) {
        const int i = 4 * r + (l >> 4), j = l & 15;
        double acc = d[r];
        for (int k = 0; k < 4; ++k) acc += B.xa[w0 + i + 16 * k] * B.xb[w0 + j + 16 * k];
        d[r] = acc;
    }
    wave_sync();
    return d;
}
// v_mfma_f32_32x32x2f32 on one wave: A[i][k] in lane i + 32 k, B[k][j] in lane j + 32 k; accumulator register r of
// lane l holds D[(r & 3) + 8 (r >> 2) + 4 (l >> 5)][l & 31]
typedef float cpuhip_f16 __attribute__((ext_vector_type(16)));
inline cpuhip_f16 mfma_f32_32x32x2(float a, float b, cpuhip_f16 c) {
    Block &B = *g_blk;
    const int t = B.cur, w0 = t & ~63, l = t & 63;
    B.xa[t] = a;
    B.xb[t] = b;
    wave_sync();
    cpuhip_f16 d = c;
    for (int r = 0; r < 16; ++r) {
        const int i = (r & 3) + 8 * (r >> 2) + 4 * (l >> 5), j = l & 31;
        float acc = d[r];
        for (int k = 0; k < 2; ++k) acc += (float)B.xa[w0 + i + 32 * k] * (float)B.xb[w0 + j + 32 * k];
        d[r] = acc;
    }
    wave_sync();
    return d;
}

// v_mfma_f32_16x16x4f32 on one wave: A[i][k] in lane i + 16 k, B[k][j] in lane j + 16 k; accumulator register r of
// lane l holds D[4 (l >> 4) + r][l & 15]
typedef float cpuhip_f4 __attribute__((ext_vector_type(4)));
inline cpuhip_f4 mfma_f32_16x16x4(float a, float b, cpuhip_f4 c) {
    Block &B = *g_blk;
    const int t = B.cur, w0 = t & ~63, l = t & 63;
    B.xa[t] = a;
    B.xb[t] = b;
    wave_sync();
    cpuhip_f4 d = c;
    for (int r = 0; r < 4; ++r) {
        const int i = 4 * (l >> 4) + r, j = l & 15;
        float acc = d[r];
        for (int k = 0; k < 4; ++k) acc += (float)B.xa[w0 + i + 16 * k] * (float)B.xb[w0 + j + 16 * k];
        d[r] = acc;
    }
    wave_sync();
    return d;
}

void fiber_entry(Block *b);

// Runs body() as a grid of nblocks workgroups of nthreads fibers, each workgroup with `lds` bytes of dynamic LDS;
// returns the workgroup barriers crossed (per workgroup, summed over the grid).
inline long long run_grid(int nblocks, int nthreads, size_t lds, std::function<void()> body,
                          size_t stack_bytes = 1 << 18) {
    std::vector<Block> blocks(nblocks);
    Block *prev = g_blk;
    gridDim.x = nblocks;
    blockDim.x = nthreads;
    for (int bi = 0; bi < nblocks; ++bi) {
        Block &b = blocks[bi];
        b.n = nthreads;
        b.id = bi;
        b.f.resize(nthreads);
        b.wave_arrived.assign((nthreads + 63) / 64, 0);
        b.xch.assign(nthreads, 0.0);
        b.xa.assign(nthreads, 0.0);
        b.xb.assign(nthreads, 0.0);
        b.body = body;
        b.dyn = (unsigned char *)malloc(lds ? lds : 1);
        memset(b.dyn, 0xA5, lds ? lds : 1);
        for (int t = 0; t < nthreads; ++t) {
            Fiber &fb = b.f[t];
            fb.tid = t;
            fb.stack_size = stack_bytes;
            fb.stack = (char *)malloc(stack_bytes);
            if (!fb.stack) die("out of memory for fiber stacks");
            // the primed frame, from the top: return address (the trampoline), rbp, rbx = the block, r12 = the
            // entry, r13, r14, r15, MXCSR and x87 control word; the trampoline's call sees a 16-byte aligned stack
            uintptr_t top = ((uintptr_t)fb.stack + stack_bytes) & ~(uintptr_t)15;
            uint64_t *q = (uint64_t *)top;
            *--q = (uint64_t)(uintptr_t)cpuhip_trampoline;
            *--q = 0;                                   // rbp
            *--q = (uint64_t)(uintptr_t)&b;             // rbx
            *--q = (uint64_t)(uintptr_t)fiber_entry;    // r12
            *--q = 0;                                   // r13
            *--q = 0;                                   // r14
            *--q = 0;                                   // r15
            *--q = (uint64_t)0x1F80u | ((uint64_t)0x037Fu << 32);
            fb.sp = q;
        }
    }
    int blocks_done = 0;
    while (blocks_done < nblocks) {
        bool moved = false;
        for (int bi = 0; bi < nblocks; ++bi) {
            Block &b = blocks[bi];
            if (b.done == b.n) continue;
            g_blk = &b;
            blockIdx.x = b.id;
            for (int t = 0; t < b.n; ++t) {
                Fiber &fb = b.f[t];
                if (fb.state != 0) continue;
                moved = true;
                b.cur = t;
                threadIdx.x = (unsigned)t;
#ifdef CPUHIP_ASAN
                __sanitizer_start_switch_fiber(&b.main_fake, fb.stack, fb.stack_size);
                cpuhip_switch(&b.main_sp, fb.sp);
                __sanitizer_finish_switch_fiber(b.main_fake, &b.main_bottom, &b.main_size);
#else
                cpuhip_switch(&b.main_sp, fb.sp);
#endif
                if (fb.state == 3 && ++b.done == b.n) ++blocks_done;
            }
        }
        if (!moved && blocks_done < nblocks) die("deadlock: no fiber can pass its barrier (a divergent barrier)");
    }
    long long nb = 0;
    for (Block &b : blocks) {
        if (b.arrived != 0) die("a barrier some threads reached and others skipped before the end");
        for (auto &fb : b.f) free(fb.stack);
        for (void *o : b.objs) free(o);
        free(b.dyn);
        nb += b.barriers / b.n;
    }
    g_blk = prev;
    return nb;
}

inline void fiber_entry(Block *b) {
#ifdef CPUHIP_ASAN
    __sanitizer_finish_switch_fiber(nullptr, &b->main_bottom, &b->main_size);
#endif
    b->body();
    Fiber &me = b->f[b->cur];
    me.state = 3;
#ifdef CPUHIP_ASAN
    __sanitizer_start_switch_fiber(nullptr, b->main_bottom, b->main_size);
#endif
    cpuhip_switch(&me.sp, b->main_sp);      // never resumed
    __builtin_trap();
}

}  // namespace cpuhip

#ifdef CPUHIP_DEFINE_SWITCH
asm(R"(
    .text
    .globl cpuhip_switch
    .type cpuhip_switch, @function
cpuhip_switch:
    pushq %rbp
    pushq %rbx
    pushq %r12
    pushq %r13
    pushq %r14
    pushq %r15
    subq $8, %rsp
    stmxcsr (%rsp)
    fnstcw 4(%rsp)
    movq %rsp, (%rdi)
    movq %rsi, %rsp
    ldmxcsr (%rsp)
    fldcw 4(%rsp)
    addq $8, %rsp
    popq %r15
    popq %r14
    popq %r13
    popq %r12
    popq %rbx
    popq %rbp
    ret
    .size cpuhip_switch, .-cpuhip_switch
    .globl cpuhip_trampoline
    .type cpuhip_trampoline, @function
cpuhip_trampoline:
    movq %rbx, %rdi
    callq *%r12
    ud2
    .size cpuhip_trampoline, .-cpuhip_trampoline
)");
#endif

inline void __syncthreads() { cpuhip::syncthreads(); }
inline double __shfl_xor(double v, int mask, int width) { return cpuhip::shfl_xor(v, mask, width); }
inline long long wall_clock64() { return ++cpuhip::g_clock; }
inline void sincospi(double x, double *s, double *c) {
    *s = sin(M_PI * x);
    *c = cos(M_PI * x);
}
inline double cospi(double x) { return cos(M_PI * x); }
inline int atomicAdd(int *p, int v) {
    const int o = *p;
    *p = o + v;
    return o;
}
inline int __clz(int x) { return x == 0 ? 32 : __builtin_clz((unsigned)x); }
inline int min(int a, int b) { return a < b ? a : b; }
inline int max(int a, int b) { return a > b ? a : b; }
