// libblml_hip.so: kernels and C-ABI of the binned-fit engine (declarations and contracts in include/blml_hip.h).
//
// Kernels
//   spec_kernel      one workgroup per factor: the spectrum of every kernel unit (plan creation; replaces
//                    _ModeKernel._table / _ffts, blml_binned.py:79-90)
//   matvec_kernel    one workgroup per fit: U = [S_k E_k]_k (_Sall, blml_binned.py:156-168)
//   fit_kernel       persistent: each workgroup takes fits from a queue (longest first) and runs the whole fit
//                    (blml_binned.py:383-445) with the solvers of solver.hpp
//   cert_kernel      one workgroup per fit: power iteration for certificate() (blml_binned.py:501-518)
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <limits>
#include <mutex>
#include <cstdlib>
#include <string>
#include <vector>

#include "blml_hip.h"
#include "net.hpp"
#include "pairs.hpp"
#include "points.hpp"
#include "grad.hpp"
#include "quick.hpp"
#include "quick_psd.hpp"
#include "fit.hpp"
#include "empirical.hpp"
#include "starts.hpp"
#include "plan_layout.hpp"

namespace {

#define HIPCHECK(x)                                                                              \
    do {                                                                                         \
        hipError_t e_ = (x);                                                                     \
        if (e_ != hipSuccess) return fail(BLML_ERUNTIME, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Device scratch of one call, allocated on the call's stream and given back on it however the call returns (an
// error return included); release() is the checked release of the normal path.
struct StreamScratch {
    void *p = nullptr;
    hipStream_t st;
    explicit StreamScratch(hipStream_t s) : st(s) {}
    StreamScratch(const StreamScratch &) = delete;
    StreamScratch &operator=(const StreamScratch &) = delete;
    ~StreamScratch() {
        if (p) (void)hipFreeAsync(p, st);
    }
    hipError_t alloc(size_t bytes) { return hipMallocAsync(&p, bytes, st); }
    hipError_t release() {
        const hipError_t e = p ? hipFreeAsync(p, st) : hipSuccess;
        p = nullptr;
        return e;
    }
};


// ------------------------------------------------------------------------------------------------ kernels
template <typename T>
__global__ void twiddle_kernel(cpx<T> *tw, int Lmax) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < Lmax) {
        double sn, cs;
        sincospi(-2.0 * (double)i / (double)Lmax, &sn, &cs);
        tw[i] = {(T)cs, (T)sn};
    }
}

template <typename T>
__global__ void __launch_bounds__(BLML_NT) spec_kernel(DevPlan<T> p, int nfacs, const int *fac_m, cpx<T> *tmp,
                                                       const long *tmp_off) {
    if ((int)blockIdx.x >= nfacs) return;
    spec_factor<T>(p, blockIdx.x, fac_m, tmp, tmp_off);
}

// one workgroup per fit
template <typename T>
__global__ void __launch_bounds__(BLML_NT) matvec_kernel(DevPlan<T> p, const T *E, T *U) {
    if ((int)blockIdx.x >= p.nfits) return;
    run_matvec<T>(p, blockIdx.x, block_team(), E, U, nullptr);
}

// a plan of one fit, spread over every workgroup of a cooperative launch
template <typename T>
__global__ void __launch_bounds__(BLML_NT) matvec_grid_kernel(DevPlan<T> p, unsigned *bar, double *part, const T *E,
                                                             T *U) {
    run_matvec<T>(p, 0, Team{(int)blockIdx.x, (int)gridDim.x, bar, part, 0}, E, U, p.sep);
}

// One pass of one mode of the stand-alone mat-vec of a separable spread plan whose every pass is a pipelined
// whole-tile pass (the host checks it): only that path is compiled in, so the kernel's registers serve it alone
// and nothing is saved to scratch.  The passes are separate launches on one stream (the stream orders them; no
// grid barrier, no cooperative launch).  FP32 plans take the centrosymmetric split of the line operators.
template <typename T>
__global__ void __launch_bounds__(BLML_NT) matvec_pipe_kernel(DevPlan<T> p, int k, int q, const T *E, T *U) {
    if constexpr (sizeof(T) == 4) {
        if (p.sep_tab_c) {
            run_matvec_pipe_pass<T, true>(p, (int)blockIdx.x, (int)gridDim.x, k, q, E, U);
            return;
        }
    }
    run_matvec_pipe_pass<T, false>(p, (int)blockIdx.x, (int)gridDim.x, k, q, E, U);
}

// The batched path: a persistent grid whose workgroups take fits from an atomic queue, longest first; one
// workgroup runs one whole fit.
template <typename T>
__global__ void __launch_bounds__(BLML_NT) fit_kernel(DevPlan<T> p, const int *order, int *queue,
                                                      const double *counts, const double *draws,
                                                      const double *init, const double *signs, T *e_out, T *se_out,
                                                      double *density, blml_result_t *results, double *merit_hist,
                                                      double *rho_hist, double tick_hz) {
    __shared__ int s_fit;
    for (;;) {
        if (threadIdx.x == 0) s_fit = atomicAdd(queue, 1);
        __syncthreads();
        // the fit's index is uniform: readfirstlane lets the compiler load its table row with scalar loads and
        // keep everything derived from it in scalar registers (an LDS or vector-loaded value counts as divergent)
        const int qi = __builtin_amdgcn_readfirstlane(s_fit);
        __syncthreads();
        if (qi >= p.nfits) break;
        const int fi = __builtin_amdgcn_readfirstlane(order[qi]);
        run_fit<T>(p, fi, block_team(), counts, draws, init, signs, e_out, se_out, density, results, merit_hist,
                   rho_hist, tick_hz);
    }
}

// A plan of one large fit: every workgroup of a cooperative launch works on it (grid barriers between the
// solver's data-dependent steps, FP64 team reductions).
template <typename T>
__global__ void __launch_bounds__(BLML_NT) fit_grid_kernel(DevPlan<T> p, unsigned *bar, double *part,
                                                           const double *counts, const double *draws,
                                                           const double *init, const double *signs, T *e_out,
                                                           T *se_out, double *density, blml_result_t *results,
                                                           double *merit_hist, double *rho_hist, double tick_hz) {
    run_fit<T>(p, 0, Team{(int)blockIdx.x, (int)gridDim.x, bar, part, 0}, counts, draws, init, signs, e_out,
               se_out, density, results, merit_hist, rho_hist, tick_hz);
}

// BLMLEmpirical over a batched plan: the persistent queue of fit_kernel, each fit by run_empirical
template <typename T>
__global__ void __launch_bounds__(BLML_NT) emp_kernel(DevPlan<T> p, const int *order, int *queue,
                                                      blml_emp_params_t ep, const double *counts,
                                                      const double *draws, const double *init, const double *aux,
                                                      T *e_out, T *se_out, double *density, blml_result_t *results,
                                                      double *j_hist, double *emp_out, double tick_hz) {
    __shared__ int s_fit;
    for (;;) {
        if (threadIdx.x == 0) s_fit = atomicAdd(queue, 1);
        __syncthreads();
        const int qi = __builtin_amdgcn_readfirstlane(s_fit);
        __syncthreads();
        if (qi >= p.nfits) break;
        const int fi = __builtin_amdgcn_readfirstlane(order[qi]);
        run_empirical<T>(p, fi, block_team(), ep, counts, draws, init, aux, e_out, se_out, density, results, j_hist,
                         emp_out, tick_hz);
    }
}

template <typename Te>
__global__ void __launch_bounds__(BLML_NT) cert_kernel(DevPlan<double> p, const Te *e_hat, const double *counts,
                                                       int max_iter, double tol, double *lam_out, int *iters) {
    if ((int)blockIdx.x >= p.nfits) return;
    run_cert<Te>(p, blockIdx.x, block_team(), e_hat, counts, max_iter, tol, lam_out, iters);
}

template <typename Te>
__global__ void __launch_bounds__(BLML_NT) cert_grid_kernel(DevPlan<double> p, unsigned *bar, double *part,
                                                            const Te *e_hat, const double *counts, int max_iter,
                                                            double tol, double *lam_out, int *iters) {
    run_cert<Te>(p, 0, Team{(int)blockIdx.x, (int)gridDim.x, bar, part, 0}, e_hat, counts, max_iter, tol, lam_out,
                 iters);
}

__global__ void __launch_bounds__(BLML_NT) finish64_kernel(DevPlan<double> p, float *se_out,
                                                          const double *counts, double *density,
                                                          blml_result_t *results) {
    if ((int)blockIdx.x >= p.nfits) return;
    run_finish64(p, blockIdx.x, block_team(), se_out, counts, density, results);
}

__global__ void __launch_bounds__(BLML_NT) finish64_grid_kernel(DevPlan<double> p, unsigned *bar, double *part,
                                                               float *se_out, const double *counts, double *density,
                                                               blml_result_t *results) {
    run_finish64(p, 0, Team{(int)blockIdx.x, (int)gridDim.x, bar, part, 0}, se_out, counts, density, results);
}

// raises `st` to the static LDS (bytes of __shared__ arrays) of kernel `kern`
int static_lds(const void *kern, int &st) {
    hipFuncAttributes a;
    HIPCHECK(hipFuncGetAttributes(&a, kern));
    st = std::max(st, (int)a.sharedSizeBytes);
    return BLML_OK;
}

// A spread plan whose modes are all LINE factors on axes of at most SepMfma<T>::MMAX points takes the separable
// path: the factors' dense tables on the device and two pass plans sized for the LDS the mat-vec and fit kernels
// leave free (sep_plan_host).
template <typename T>
int build_sep(Plan &pl, const double *fdata, const int32_t *freqs, int max_lds) {
    int st_mv = 0, st_fit = 0, st_pipe = 0;
    int rc = static_lds((const void *)matvec_grid_kernel<T>, st_mv);
    if (!rc) rc = static_lds((const void *)fit_grid_kernel<T>, st_fit);
    if (!rc) rc = static_lds((const void *)matvec_pipe_kernel<T>, st_pipe);
    if (rc) return rc;
    SepHost<T> sh;
    if (!sep_plan_host<T>(pl, fdata, freqs, max_lds, st_mv, st_fit, st_pipe, sh)) return BLML_OK;
    HIPCHECK(hipMalloc(&pl.d_sep, sizeof(SepPlan) * 2));
    HIPCHECK(hipMalloc(&pl.d_sep_tab, sizeof(T) * std::max<size_t>(1, sh.tab.size())));
    HIPCHECK(hipMalloc(&pl.d_sep_off, sizeof(long) * sh.off.size()));
    HIPCHECK(hipMemcpy(pl.d_sep, sh.sp, sizeof(SepPlan) * 2, hipMemcpyHostToDevice));
    if (!sh.tab.empty())
        HIPCHECK(hipMemcpy(pl.d_sep_tab, sh.tab.data(), sizeof(T) * sh.tab.size(), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(pl.d_sep_off, sh.off.data(), sizeof(long) * sh.off.size(), hipMemcpyHostToDevice));
    return BLML_OK;
}

// the dynamic LDS of the batched launches: the line buffers, or with resident intermediates their offset + size
static int lds_mv_launch(const Plan &pl) { return pl.res_mv ? pl.res_off + pl.res_mv : pl.lds_bytes; }
static int lds_fit_launch(const Plan &pl) { return pl.res_fit ? pl.res_off + pl.res_fit : pl.lds_bytes; }

template <typename T>
int create_t(Plan &pl, const double *fdata, int64_t n_fdata, const int32_t *freqs, int64_t n_freqs) {
    std::vector<int> fac_m;
    std::vector<long> tmp_off;
    long tmp_total = 0;
    int64_t spec_total = 0;
    int rc = layout(pl, fdata, n_fdata, freqs, n_freqs, fac_m, tmp_off, tmp_total, spec_total);
    if (rc) return rc;
    // a launch's LDS = its kernel's static arrays + the dynamic line buffers; both must fit one workgroup's LDS
    int dev = 0, max_lds = 0, st = 0, st64 = 0;
    HIPCHECK(hipGetDevice(&dev));
    HIPCHECK(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    rc = static_lds((const void *)fit_kernel<T>, st);
    if (!rc) rc = static_lds((const void *)matvec_kernel<T>, st);
    if (!rc && sizeof(T) == 8) rc = static_lds((const void *)cert_kernel<double>, st);
    if (!rc && sizeof(T) == 4) rc = static_lds((const void *)cert_kernel<float>, st64);
    if (!rc) rc = static_lds((const void *)spec_kernel<T>, st);
    if (!rc && sizeof(T) == 4) rc = static_lds((const void *)finish64_kernel, st64);
    if (!rc && sizeof(T) == 4) rc = static_lds((const void *)spec_kernel<double>, st64);
    if (rc) return rc;
    if (pl.lds_bytes + st > max_lds)
        return fail(BLML_EVALUE, "a transform line needs " + std::to_string(pl.lds_bytes) + " bytes of LDS (limit " +
                                     std::to_string(max_lds - st) + "): lattice axis too long");
    if (sizeof(T) == 4 && pl.lds64_bytes + st64 > max_lds)
        return fail(BLML_EVALUE, "the FP64 finishing pass of a line needs " + std::to_string(pl.lds64_bytes) +
                                     " bytes of LDS (limit " + std::to_string(max_lds - st64) +
                                     "): lattice axis too long");
    HIPCHECK(hipMalloc(&pl.d_fits, sizeof(blml_fit_t) * std::max(1, pl.nfits)));
    HIPCHECK(hipMalloc(&pl.d_facs, sizeof(blml_factor_t) * std::max(1, pl.nfacs)));
    HIPCHECK(hipMalloc(&pl.d_fdata, sizeof(double) * std::max<int64_t>(1, n_fdata)));
    HIPCHECK(hipMalloc(&pl.d_freqs, sizeof(int) * std::max<int64_t>(1, n_freqs)));
    HIPCHECK(hipMalloc(&pl.d_spec, sizeof(T) * std::max<int64_t>(1, spec_total)));
    HIPCHECK(hipMalloc(&pl.d_tw, sizeof(cpx<T>) * pl.Lmax));
    HIPCHECK(hipMalloc(&pl.d_ws, std::max<int64_t>(256, pl.ws_elems)));
    HIPCHECK(hipMalloc(&pl.d_order, sizeof(int) * std::max(1, pl.nfits)));
    HIPCHECK(hipMalloc(&pl.d_queue, sizeof(int)));
    HIPCHECK(hipMemcpy(pl.d_fits, pl.hfits.data(), sizeof(blml_fit_t) * pl.nfits, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(pl.d_facs, pl.hfacs.data(), sizeof(blml_factor_t) * pl.nfacs, hipMemcpyHostToDevice));
    if (n_fdata) HIPCHECK(hipMemcpy(pl.d_fdata, fdata, sizeof(double) * n_fdata, hipMemcpyHostToDevice));
    if (n_freqs) HIPCHECK(hipMemcpy(pl.d_freqs, freqs, sizeof(int) * n_freqs, hipMemcpyHostToDevice));
    // longest fits first (the persistent queue's order)
    std::vector<int> order(pl.nfits);
    for (int i = 0; i < pl.nfits; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        const blml_fit_t &fa = pl.hfits[a], &fb = pl.hfits[b];
        return (double)fa.B * fa.K * (1 + fa.complex_g) > (double)fb.B * fb.K * (1 + fb.complex_g);
    });
    HIPCHECK(hipMemcpy(pl.d_order, order.data(), sizeof(int) * pl.nfits, hipMemcpyHostToDevice));
    twiddle_kernel<T><<<(pl.Lmax + 255) / 256, 256>>>((cpx<T> *)pl.d_tw, pl.Lmax);
    HIPCHECK(hipGetLastError());
    int *d_fac_m = nullptr;
    long *d_tmp_off = nullptr;
    cpx<T> *d_tmp = nullptr;
    HIPCHECK(hipMalloc(&d_fac_m, sizeof(int) * std::max<size_t>(1, fac_m.size())));
    HIPCHECK(hipMalloc(&d_tmp_off, sizeof(long) * std::max(1, pl.nfacs)));
    HIPCHECK(hipMalloc(&d_tmp, sizeof(cpx<T>) * std::max<long>(1, tmp_total)));
    HIPCHECK(hipMemcpy(d_fac_m, fac_m.data(), sizeof(int) * fac_m.size(), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(d_tmp_off, tmp_off.data(), sizeof(long) * pl.nfacs, hipMemcpyHostToDevice));
    if (pl.nfacs > 0) {
        spec_kernel<T><<<pl.nfacs, BLML_NT, pl.lds_bytes>>>(dev_plan<T>(pl), pl.nfacs, d_fac_m, d_tmp, d_tmp_off);
        HIPCHECK(hipGetLastError());
    }
    if (sizeof(T) == 4) {
        HIPCHECK(hipMalloc(&pl.d_spec64, sizeof(double) * std::max<int64_t>(1, spec_total)));
        HIPCHECK(hipMalloc(&pl.d_tw64, sizeof(cpx<double>) * pl.Lmax));
        twiddle_kernel<double><<<(pl.Lmax + 255) / 256, 256>>>((cpx<double> *)pl.d_tw64, pl.Lmax);
        HIPCHECK(hipGetLastError());
        cpx<double> *d_tmp64 = nullptr;
        HIPCHECK(hipMalloc(&d_tmp64, sizeof(cpx<double>) * std::max<long>(1, tmp_total)));
        if (pl.nfacs > 0) {
            spec_kernel<double><<<pl.nfacs, BLML_NT, pl.lds64_bytes>>>(dev_plan64(pl), pl.nfacs, d_fac_m, d_tmp64,
                                                                       d_tmp_off);
            HIPCHECK(hipGetLastError());
        }
        HIPCHECK(hipDeviceSynchronize());
        HIPCHECK(hipFree(d_tmp64));
    }
    HIPCHECK(hipDeviceSynchronize());
    HIPCHECK(hipFree(d_fac_m));
    HIPCHECK(hipFree(d_tmp_off));
    HIPCHECK(hipFree(d_tmp));
    // FP32 plans sum the padded line factors of up to BLML_TOEP_MAX_DEFAULT bins directly (OP_TOEP) instead of by
    // FFT; BLML_TOEP_MAX (environment) sets another bound (0: every padded line by FFT)
    if (sizeof(T) == 4) {
        const char *env = getenv("BLML_TOEP_MAX");
        pl.toep_max = env ? atoi(env) : BLML_TOEP_MAX_DEFAULT;
    }
    // LDS-resident intermediates of the batched kernels: after the line buffers, as much as the workgroup's LDS
    // leaves; the launch reserves the largest need among the fits that fit (the others stay in the workspace).
    // BLML_LDS_RESIDENT (environment) names the kernels that use it: 'm' matvec_kernel, 'f' fit_kernel (default
    // both; "0" neither).
    {
        const char *env = getenv("BLML_LDS_RESIDENT");
        const std::string use = env ? env : "mf";
        int st_mvk = 0, st_fitk = 0;
        rc = static_lds((const void *)matvec_kernel<T>, st_mvk);
        if (!rc) rc = static_lds((const void *)fit_kernel<T>, st_fitk);
        if (rc) return rc;
        pl.res_off = (pl.lds_bytes + 15) / 16 * 16;
        const long cap_mv = (long)max_lds - st_mvk - pl.res_off, cap_fit = (long)max_lds - st_fitk - pl.res_off;
        long need_mv = 0, need_fit = 0;
        for (const blml_fit_t &f : pl.hfits) {
            const long need = res_need(f, pl.hfacs.data(), (int)sizeof(T));
            if (need <= cap_mv) need_mv = std::max(need_mv, need);
            if (need <= cap_fit) need_fit = std::max(need_fit, need);
        }
        pl.res_need_mv = (int)need_mv;
        pl.res_need_fit = (int)need_fit;
        pl.res_mv = use.find('m') != std::string::npos ? (int)need_mv : 0;
        pl.res_fit = use.find('f') != std::string::npos ? (int)need_fit : 0;
    }
    // persistent grid: as many workgroups as fit on the device at once (capped by the number of fits)
    int ncu = 0, occ = 0;
    HIPCHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fit_kernel<T>, BLML_NT, lds_fit_launch(pl)));
    int khz = 0;
    HIPCHECK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    if (khz > 0) pl.tick_hz = 1e3 * khz;
    if (occ < 1) occ = 1;
    pl.grid = std::max(1, std::min(pl.nfits, ncu * occ));
    // a plan of one large fit: every workgroup that fits on the device at once works on it
    const long V0 = vmax_of(pl.hfits[0]);
    pl.spread = pl.nfits == 1 && (pl.params.spread > 0 || (pl.params.spread == 0 && V0 >= kSpreadMin));
    if (pl.spread) {
        int coop_ok = 0;
        HIPCHECK(hipDeviceGetAttribute(&coop_ok, hipDeviceAttributeCooperativeLaunch, dev));
        if (!coop_ok) return fail(BLML_ERUNTIME, "the device does not support cooperative launches");
        int o1 = 0, o2 = 0, o3 = 0, o4 = 0;
        HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&o1, fit_grid_kernel<T>, BLML_NT, pl.lds_bytes));
        HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&o2, matvec_grid_kernel<T>, BLML_NT, pl.lds_bytes));
        if (sizeof(T) == 8)
            HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&o3, cert_grid_kernel<double>, BLML_NT,
                                                                  pl.lds_bytes));
        else
            HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&o3, cert_grid_kernel<float>, BLML_NT,
                                                                  pl.lds64_bytes));
        o4 = o1;
        if (sizeof(T) == 4)
            HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&o4, finish64_grid_kernel, BLML_NT,
                                                                  pl.lds64_bytes));
        const int oc = std::min(std::min(o1, o2), std::min(o3, o4));
        if (oc < 1) return fail(BLML_ERUNTIME, "no workgroup of the spread kernels fits a CU");
        pl.coop = ncu * oc;
        HIPCHECK(hipMalloc(&pl.d_bar, 4 * sizeof(unsigned)));
        HIPCHECK(hipMalloc(&pl.d_part, sizeof(double) * 2 * (size_t)pl.coop * BLML_NRED));
        HIPCHECK(hipMemset(pl.d_bar, 0, 4 * sizeof(unsigned)));
        pl.lds_mv = pl.lds_fit = pl.lds_bytes;
        rc = build_sep<T>(pl, fdata, freqs, max_lds);
        if (rc) return rc;
        if (pl.sep) {   // the spread launches use the separable path; their co-resident grid follows its LDS
            int a = 0, b = 0;
            HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, fit_grid_kernel<T>, BLML_NT, pl.lds_fit));
            HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, matvec_grid_kernel<T>, BLML_NT, pl.lds_mv));
            const int oc2 = std::min(std::min(a, b), oc);
            if (pl.pipe_mv) {
                int c = 0;
                HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&c, matvec_pipe_kernel<T>, BLML_NT, pl.lds_pipe));
                pl.pipe_mv = c >= std::max(oc2, 1);
            }
            if (oc2 < 1) return fail(BLML_ERUNTIME, "no workgroup of the separable spread kernels fits a CU");
            if (oc2 < oc) {
                pl.coop = ncu * oc2;
                HIPCHECK(hipFree(pl.d_part));
                HIPCHECK(hipMalloc(&pl.d_part, sizeof(double) * 2 * (size_t)pl.coop * BLML_NRED));
            }
        }
    }
    return BLML_OK;
}

// a cooperative launch of a spread plan's kernel (the barrier words start cleared)
int launch_coop(Plan &pl, const void *kern, void **args, int lds, hipStream_t st) {
    HIPCHECK(hipMemsetAsync(pl.d_bar, 0, 4 * sizeof(unsigned), st));
    HIPCHECK(hipLaunchCooperativeKernel(kern, dim3(pl.coop), dim3(BLML_NT), args, (unsigned)lds, st));
    return BLML_OK;
}

void free_plan(Plan *pl) {
    if (!pl) return;
    void *ptrs[] = {pl->d_fits, pl->d_facs, pl->d_fdata, pl->d_freqs, pl->d_spec, pl->d_tw,
                    pl->d_spec64, pl->d_tw64, pl->d_ws, pl->d_order, pl->d_queue, pl->d_bar, pl->d_part,
                    pl->d_sep, pl->d_sep_tab, pl->d_sep_off};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    delete pl;
}

}  // namespace

template <typename T>
int empirical_launch(Plan &pl, const blml_emp_params_t &ep, const double *counts, const double *draws,
                            const double *init, const double *aux, void *e_out, void *se_out, double *density,
                            blml_result_t *results, double *j_hist, double *emp_out, hipStream_t st) {
    const void *kern = (const void *)emp_kernel<T>;
    int dev = 0, max_lds = 0, st_lds = 0, ncu = 0, occ = 0;
    HIPCHECK(hipGetDevice(&dev));
    HIPCHECK(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    int rc = static_lds(kern, st_lds);
    if (rc) return rc;
    if (pl.lds_bytes + st_lds > max_lds)
        return fail(BLML_EVALUE, "a transform line of this plan does not fit the LDS beside the empirical solver");
    HIPCHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, BLML_NT, pl.lds_bytes));
    const int grid = std::max(1, std::min(pl.nfits, ncu * std::max(occ, 1)));
    DevPlan<T> dp = dev_plan<T>(pl);
    double tick = pl.tick_hz;
    void *args[] = {&dp, &pl.d_order, &pl.d_queue, (void *)&ep, &counts, &draws, &init, &aux, &e_out, &se_out,
                    &density, &results, &j_hist, &emp_out, &tick};
    HIPCHECK(hipMemsetAsync(pl.d_queue, 0, sizeof(int), st));
    HIPCHECK(hipLaunchKernel(kern, dim3(grid), dim3(BLML_NT), args, (size_t)pl.lds_bytes, st));
    return BLML_OK;
}

// a deterministic, denormal-free test vector for the autotuner's timed mat-vecs
template <typename T>
__global__ void tune_fill_kernel(T *E, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        E[i] = (T)(0.5 + 0.25 * sin(0.37 * (double)(i % 9973)));
}

// The start search: every run one workgroup, dealt round robin (runs are few: tens).
__global__ void __launch_bounds__(BLML_NT) hio_kernel(HioArgs a) {
    __shared__ double red[BLML_NWAVE];
    const int runs = a.n_rand + a.n_kick;
    for (int r = blockIdx.x; r < runs; r += gridDim.x) hio_run(a, r, red);
}

// ------------------------------------------------------------------------------------------------ C-ABI
extern "C" {

const char *blml_last_error(void) { return g_err.c_str(); }

int blml_plan_create(const blml_fit_t *fits, int32_t nfits, const blml_factor_t *facs, int32_t nfacs,
                     const double *fdata, int64_t n_fdata, const int32_t *freqs, int64_t n_freqs, int32_t dtype,
                     const blml_params_t *params, void **plan) {
    if (!plan || !params || (nfits > 0 && !fits) || (nfacs > 0 && !facs)) return fail(BLML_EVALUE, "null argument");
    if (nfits < 1) return fail(BLML_EVALUE, "a plan needs at least one fit");
    if (dtype != BLML_F32 && dtype != BLML_F64) return fail(BLML_EVALUE, "dtype: BLML_F32 or BLML_F64");
    Plan *pl = new Plan();
    pl->dtype = dtype;
    pl->nfits = nfits;
    pl->nfacs = nfacs;
    pl->params = *params;
    pl->hfits.assign(fits, fits + nfits);
    pl->hfacs.assign(facs, facs + nfacs);
    pl->esize = dtype == BLML_F32 ? 4 : 8;
    const int rc = dtype == BLML_F32 ? create_t<float>(*pl, fdata, n_fdata, freqs, n_freqs)
                                     : create_t<double>(*pl, fdata, n_fdata, freqs, n_freqs);
    if (rc) {
        free_plan(pl);
        return rc;
    }
    *plan = pl;
    return BLML_OK;
}

int blml_plan_destroy(void *plan) {
    free_plan((Plan *)plan);
    return BLML_OK;
}

int blml_plan_info(void *plan, int64_t *out_elems, int64_t *ws_bytes, int32_t *lds_bytes, int32_t *grid) {
    Plan *pl = (Plan *)plan;
    if (!pl) return fail(BLML_EVALUE, "null plan");
    if (out_elems) *out_elems = pl->out_elems;
    if (ws_bytes) *ws_bytes = pl->ws_elems;
    if (lds_bytes) *lds_bytes = pl->lds_bytes;
    if (grid) *grid = pl->spread ? -pl->coop : pl->grid;   // negative: one fit spread over that many workgroups
    return BLML_OK;
}

// The plan's autotuner: times the stand-alone mat-vec (the same passes the fit kernels run) under every candidate
// configuration and keeps the fastest.  Candidates: the FFT passes' radix-stage order (fft_line's sched 0 / 1) and,
// for batched plans, whether a fit's transform intermediates stay in LDS (the launch then reserves it).  A spread
// plan on the separable path has no FFT passes: its one configuration is timed.  Every candidate computes the same
// operator (the stage order changes the rounding of FP32 transforms only), so the choice is a speed choice.
int blml_plan_autotune(void *plan, int32_t reps, double *times, int32_t *choice, void *stream) {
    Plan *pl = (Plan *)plan;
    if (!pl || !times || !choice) return fail(BLML_EVALUE, "null argument");
    if (reps < 1) return fail(BLML_EVALUE, "reps must be >= 1");
    hipStream_t st = (hipStream_t)stream;
    const size_t bytes = std::max<size_t>(16, (size_t)pl->out_elems * pl->esize);
    StreamScratch sE(st), sU(st);
    HIPCHECK(sE.alloc(bytes));
    HIPCHECK(sU.alloc(bytes));
    if (pl->dtype == BLML_F32)
        tune_fill_kernel<float><<<256, 256, 0, st>>>((float *)sE.p, (long)pl->out_elems);
    else
        tune_fill_kernel<double><<<256, 256, 0, st>>>((double *)sE.p, (long)pl->out_elems);
    HIPCHECK(hipGetLastError());
    hipEvent_t e0, e1;
    HIPCHECK(hipEventCreate(&e0));
    HIPCHECK(hipEventCreate(&e1));
    struct Cand { int sched; bool res; };
    const bool batched = !pl->spread;
    const bool fft = !pl->sep;
    Cand cands[BLML_TUNE_MAX];
    int nc = 0;
    for (int sched = 0; sched < (fft ? 2 : 1); ++sched)
        for (int res = 1; res >= 0; --res)
            if (res == 0 || (batched && (pl->res_need_mv > 0 || pl->res_need_fit > 0))) cands[nc++] = {sched, res == 1};
    for (int i = 0; i < BLML_TUNE_MAX; ++i) times[i] = NAN;
    int best = 0;
    int rc = BLML_OK;
    for (int c = 0; c < nc && rc == BLML_OK; ++c) {
        pl->fft_sched = cands[c].sched;
        pl->res_mv = cands[c].res ? pl->res_need_mv : 0;
        pl->res_fit = cands[c].res ? pl->res_need_fit : 0;
        rc = blml_matvec(pl, sE.p, sU.p, st);                       // warm-up
        if (rc == BLML_OK && hipEventRecord(e0, st) != hipSuccess) rc = fail(BLML_ERUNTIME, "hipEventRecord");
        for (int r = 0; r < reps && rc == BLML_OK; ++r) rc = blml_matvec(pl, sE.p, sU.p, st);
        if (rc == BLML_OK && hipEventRecord(e1, st) != hipSuccess) rc = fail(BLML_ERUNTIME, "hipEventRecord");
        if (rc == BLML_OK && hipEventSynchronize(e1) != hipSuccess) rc = fail(BLML_ERUNTIME, "hipEventSynchronize");
        float ms = 0.0f;
        if (rc == BLML_OK && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = fail(BLML_ERUNTIME, "hipEventElapsedTime");
        if (rc == BLML_OK) {
            times[c] = (double)ms * 1e-3 / reps;
            if (times[c] < times[best]) best = c;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc == BLML_OK && nc > 0) {
        pl->fft_sched = cands[best].sched;
        pl->res_mv = cands[best].res ? pl->res_need_mv : 0;
        pl->res_fit = cands[best].res ? pl->res_need_fit : 0;
        *choice = cands[best].sched * 2 + (cands[best].res ? 0 : 1);
        if (batched) {      // the persistent fit grid follows the fit launch's LDS
            int dev = 0, ncu = 0, occ = 0;
            HIPCHECK(hipGetDevice(&dev));
            HIPCHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
            const void *fk = pl->dtype == BLML_F32 ? (const void *)fit_kernel<float> : (const void *)fit_kernel<double>;
            HIPCHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fk, BLML_NT, lds_fit_launch(*pl)));
            pl->grid = std::max(1, std::min(pl->nfits, ncu * std::max(occ, 1)));
        }
        // the times in candidate-code order: code = sched * 2 + (resident ? 0 : 1)
        double byc[BLML_TUNE_MAX];
        for (int i = 0; i < BLML_TUNE_MAX; ++i) byc[i] = NAN;
        for (int c = 0; c < nc; ++c) byc[cands[c].sched * 2 + (cands[c].res ? 0 : 1)] = times[c];
        for (int i = 0; i < BLML_TUNE_MAX; ++i) times[i] = byc[i];
    }
    if (rc) return rc;
    HIPCHECK(sE.release());
    HIPCHECK(sU.release());
    return BLML_OK;
}

int blml_plan_matvec_kernel(void *plan, int32_t *which) {
    Plan *pl = (Plan *)plan;
    if (!pl || !which) return fail(BLML_EVALUE, "null argument");
    *which = pl->spread ? (pl->pipe_mv ? 2 : 1) : 0;
    return BLML_OK;
}

int blml_matvec(void *plan, const void *E, void *U, void *stream) {
    Plan *pl = (Plan *)plan;
    if (!pl || !E || !U) return fail(BLML_EVALUE, "null argument");
    hipStream_t st = (hipStream_t)stream;
    if (pl->spread) {
        // the pipelined kernel moves 16-byte vectors: every mode's slice of E and U must start 16-byte aligned
        const blml_fit_t &f = pl->hfits[0];
        const long mb = (long)f.B * (f.complex_g ? 2 : 1) * pl->esize;
        const bool al16 = (((unsigned long)E | (unsigned long)U) & 15) == 0 && (f.out_off * pl->esize) % 16 == 0 &&
                          (f.K == 1 || mb % 16 == 0);
        const bool pipe = pl->pipe_mv && al16;
        if (pipe) {           // every pass of every mode a launch of its own, in order on the stream
            for (int k = 0; k < f.K; ++k)
                for (int q = 0; q < pl->sep_npass; ++q) {
                    if (pl->dtype == BLML_F32)
                        matvec_pipe_kernel<float><<<pl->coop, BLML_NT, pl->lds_pipe, st>>>(
                            dev_plan<float>(*pl), k, q, (const float *)E, (float *)U);
                    else
                        matvec_pipe_kernel<double><<<pl->coop, BLML_NT, pl->lds_pipe, st>>>(
                            dev_plan<double>(*pl), k, q, (const double *)E, (double *)U);
                    HIPCHECK(hipGetLastError());
                }
            return BLML_OK;
        }
        if (pl->dtype == BLML_F32) {
            DevPlan<float> dp = dev_plan<float>(*pl);
            void *args[] = {&dp, &pl->d_bar, &pl->d_part, &E, &U};
            return launch_coop(*pl, (const void *)matvec_grid_kernel<float>, args, pl->lds_mv, st);
        }
        DevPlan<double> dp = dev_plan<double>(*pl);
        void *args[] = {&dp, &pl->d_bar, &pl->d_part, &E, &U};
        return launch_coop(*pl, (const void *)matvec_grid_kernel<double>, args, pl->lds_mv, st);
    }
    if (pl->dtype == BLML_F32) {
        DevPlan<float> dp = dev_plan<float>(*pl);
        dp.res_off = pl->res_off;
        dp.res_cap = pl->res_mv;
        matvec_kernel<float><<<pl->nfits, BLML_NT, lds_mv_launch(*pl), st>>>(dp, (const float *)E, (float *)U);
    } else {
        DevPlan<double> dp = dev_plan<double>(*pl);
        dp.res_off = pl->res_off;
        dp.res_cap = pl->res_mv;
        matvec_kernel<double><<<pl->nfits, BLML_NT, lds_mv_launch(*pl), st>>>(dp, (const double *)E, (double *)U);
    }
    HIPCHECK(hipGetLastError());
    return BLML_OK;
}

int blml_fit(void *plan, const double *counts, const double *draws, const double *init, const double *signs,
             void *e_out, void *se_out, double *density, blml_result_t *results, double *merit_hist, double *rho_hist,
             void *stream) {
    Plan *pl = (Plan *)plan;
    if (!pl || !counts || !draws || !e_out || !se_out || !density || !results) return fail(BLML_EVALUE, "null argument");
    for (const blml_fit_t &f : pl->hfits) {
        if (f.init_K != 0 && !init) return fail(BLML_EVALUE, "a warm-started fit needs init");
        if (f.has_signs && !signs) return fail(BLML_EVALUE, "a fit with signs needs signs");
    }
    hipStream_t st = (hipStream_t)stream;
    if (pl->spread) {
        double tick = pl->tick_hz;
        int rc;
        if (pl->dtype == BLML_F32) {
            DevPlan<float> dp = dev_plan<float>(*pl);
            void *args[] = {&dp, &pl->d_bar, &pl->d_part, &counts, &draws, &init, &signs, &e_out, &se_out,
                            &density, &results, &merit_hist, &rho_hist, &tick};
            rc = launch_coop(*pl, (const void *)fit_grid_kernel<float>, args, pl->lds_fit, st);
            if (rc) return rc;
            DevPlan<double> d64 = dev_plan64(*pl);
            void *a64[] = {&d64, &pl->d_bar, &pl->d_part, &se_out, &counts, &density, &results};
            return launch_coop(*pl, (const void *)finish64_grid_kernel, a64, pl->lds64_bytes, st);
        }
        DevPlan<double> dp = dev_plan<double>(*pl);
        void *args[] = {&dp, &pl->d_bar, &pl->d_part, &counts, &draws, &init, &signs, &e_out, &se_out,
                        &density, &results, &merit_hist, &rho_hist, &tick};
        return launch_coop(*pl, (const void *)fit_grid_kernel<double>, args, pl->lds_fit, st);
    }
    HIPCHECK(hipMemsetAsync(pl->d_queue, 0, sizeof(int), st));
    if (pl->dtype == BLML_F32) {
        DevPlan<float> dp = dev_plan<float>(*pl);
        dp.res_off = pl->res_off;
        dp.res_cap = pl->res_fit;
        fit_kernel<float><<<pl->grid, BLML_NT, lds_fit_launch(*pl), st>>>(
            dp, pl->d_order, pl->d_queue, counts, draws, init, signs, (float *)e_out,
            (float *)se_out, density, results, merit_hist, rho_hist, pl->tick_hz);
        HIPCHECK(hipGetLastError());
        finish64_kernel<<<pl->nfits, BLML_NT, pl->lds64_bytes, st>>>(dev_plan64(*pl), (float *)se_out, counts,
                                                                     density, results);
    } else {
        DevPlan<double> dp = dev_plan<double>(*pl);
        dp.res_off = pl->res_off;
        dp.res_cap = pl->res_fit;
        fit_kernel<double><<<pl->grid, BLML_NT, lds_fit_launch(*pl), st>>>(
            dp, pl->d_order, pl->d_queue, counts, draws, init, signs, (double *)e_out,
            (double *)se_out, density, results, merit_hist, rho_hist, pl->tick_hz);
    }
    HIPCHECK(hipGetLastError());
    return BLML_OK;
}

int blml_empirical(void *plan, const blml_emp_params_t *ep, const double *counts, const double *draws,
                   const double *init, const double *aux, void *e_out, void *se_out, double *density,
                   blml_result_t *results, double *j_hist, double *emp_out, void *stream) {
    Plan *pl = (Plan *)plan;
    if (!pl || !ep || !counts || !draws || !aux || !e_out || !se_out || !density || !results || !emp_out)
        return fail(BLML_EVALUE, "null argument");
    if (pl->spread) return fail(BLML_EVALUE, "BLMLEmpirical runs on batched plans (create the plan with spread = -1)");
    if (ep->reg < 0 || ep->reg > 2) return fail(BLML_EVALUE, "reg: 'ridge' | 'l2' | 'l1'");
    if (ep->normalise < 0 || ep->normalise > 1) return fail(BLML_EVALUE, "normalise: 'window' | 'plane'");
    if (!(ep->lam >= 0) || !(ep->tol >= 0) || !(ep->outer_tol >= 0))
        return fail(BLML_EVALUE, "lam, tol and outer_tol must be >= 0");
    if (ep->max_iter < 1 || ep->max_outer < 0) return fail(BLML_EVALUE, "max_iter >= 1 and max_outer >= 0");
    if (ep->anderson < 0 || ep->anderson > pl->params.anderson)
        return fail(BLML_EVALUE, "anderson: 0.." + std::to_string(pl->params.anderson) +
                                     " (the plan's anderson sizes the state pool)");
    for (const blml_fit_t &f : pl->hfits) {
        if ((f.complex_g || f.K > 1) && !ep->retrieve)
            return fail(BLML_EVALUE, "BLMLEmpirical: rank K > 1 or complex_g need retrieve=True (the hidden phase)");
        if (f.init_K != 0 && (!init || (ep->retrieve ? f.init_K != f.K : f.init_K != 1)))
            return fail(BLML_EVALUE, "a warm-started empirical fit needs init with every mode (retrieve) or mode 0");
        if (f.has_signs) return fail(BLML_EVALUE, "BLMLEmpirical takes no signs");
    }
    hipStream_t st = (hipStream_t)stream;
    if (pl->dtype == BLML_F32)
        return empirical_launch<float>(*pl, *ep, counts, draws, init, aux, e_out, se_out, density, results, j_hist,
                                       emp_out, st);
    return empirical_launch<double>(*pl, *ep, counts, draws, init, aux, e_out, se_out, density, results, j_hist,
                                    emp_out, st);
}

int blml_certificate(void *plan, const void *e_hat, const double *counts, int32_t max_iter, double tol,
                     double *lam, int32_t *iters, void *stream) {
    Plan *pl = (Plan *)plan;
    if (!pl || !e_hat || !counts || !lam || !iters) return fail(BLML_EVALUE, "null argument");
    if (max_iter < 1) return fail(BLML_EVALUE, "max_iter must be >= 1");
    if (!(tol >= 0)) return fail(BLML_EVALUE, "tol must be >= 0");
    hipStream_t st = (hipStream_t)stream;
    // FP64 transforms for every plan: an FP32 plan's FP64 view (its finishing pass's spectra, twiddles and scratch)
    const bool f32 = pl->dtype == BLML_F32;
    DevPlan<double> dp = f32 ? dev_plan64(*pl) : dev_plan<double>(*pl);
    const int lds = f32 ? pl->lds64_bytes : pl->lds_bytes;
    if (pl->spread) {
        void *args[] = {&dp, &pl->d_bar, &pl->d_part, &e_hat, &counts, &max_iter, &tol, &lam, &iters};
        return launch_coop(*pl, f32 ? (const void *)cert_grid_kernel<float> : (const void *)cert_grid_kernel<double>,
                           args, lds, st);
    }
    if (f32)
        cert_kernel<float><<<pl->nfits, BLML_NT, lds, st>>>(dp, (const float *)e_hat, counts, max_iter, tol, lam,
                                                             iters);
    else
        cert_kernel<double><<<pl->nfits, BLML_NT, lds, st>>>(dp, (const double *)e_hat, counts, max_iter, tol, lam,
                                                              iters);
    HIPCHECK(hipGetLastError());
    return BLML_OK;
}

// ------------------------------------------------------------------------------------------------ pair path
static int pair_check(const blml_pairtab_t *t) {
    if (!t) return fail(BLML_EVALUE, "null table");
    if (t->d < 1 || t->d > BLML_PAIR_MAXD)
        return fail(BLML_EVALUE, "pair tables: 1.." + std::to_string(BLML_PAIR_MAXD) + " columns on this engine");
    if (t->G < 0 || t->G > BLML_PAIR_MAXG)
        return fail(BLML_EVALUE, "pair tables: at most " + std::to_string(BLML_PAIR_MAXG) + " radial groups");
    if (!t->T || t->nT < 1) return fail(BLML_EVALUE, "pair tables: empty T");
    if (t->nT >= (1L << 31) - 1) return fail(BLML_EVALUE, "pair tables: at most 2^31 - 2 values in T");
    if (t->G > 0 && (!t->RT || t->LR < 2)) return fail(BLML_EVALUE, "pair tables: radial rows missing");
    for (int k = 0; k < t->d; ++k) {
        const int kd = t->kind[k];
        if (kd < 0 || kd > 4) return fail(BLML_EVALUE, "pair tables: unknown column kind");
        if ((kd == 2 || kd == 4) && (t->grp[k] < 0 || t->grp[k] >= t->G))
            return fail(BLML_EVALUE, "pair tables: a radial column without its group");
        if ((kd == 1 || kd == 4) && t->nbins[k] < 1) return fail(BLML_EVALUE, "pair tables: circular bins missing");
    }
    return BLML_OK;
}

// the rows launch's shape: rb workgroups of rows, nsl slices of `slice` sources (a multiple of the LDS chunk)
static void pair_rows_grid(long R, long M, long &rb, long &nsl, long &slice) {
    rb = (R + BLML_PAIR_NT - 1) / BLML_PAIR_NT;
    nsl = std::max<long>(1, std::min<long>((2048 + rb - 1) / rb, (M + BLML_PAIR_CH - 1) / BLML_PAIR_CH));
    slice = ((M + nsl - 1) / nsl + BLML_PAIR_CH - 1) / BLML_PAIR_CH * BLML_PAIR_CH;
    nsl = (M + slice - 1) / slice;
}

// s = rows(ZA, ZB, w) for R, M > 0 with a partials buffer of nsl * R values (pair_rows_grid)
static int pair_rows_launch(const blml_pairtab_t &t, const int32_t *ZA, const int32_t *ZB, const double *w, long R,
                            long M, double *part, double *s, hipStream_t st) {
    long rb, nsl, slice;
    pair_rows_grid(R, M, rb, nsl, slice);
    const size_t es = t.f32 ? 4 : 8;
    const bool tl = t.nT * (long)es <= 32768;
    const long tbytes = tl ? ((t.nT * (long)es + 15) / 16) * 16 : 0;
    const int lds = (int)(tbytes + ((BLML_PAIR_CH * BLML_PAIR_MAXD * 4 + 15) / 16) * 16 + BLML_PAIR_CH * 8);
    const dim3 grid((unsigned)rb, (unsigned)nsl);
#define BLML_ROWS_LAUNCH(R_, TL, NG) \
    pair_rows_kernel<R_, TL, NG><<<grid, BLML_PAIR_NT, lds, st>>>(t, ZA, ZB, w, (int)R, (int)M, (int)slice, part)
#define BLML_ROWS_BY_G(R_, TL) \
    do { \
        if (t.G == 0) pair_rows_u_kernel<R_, TL, 4><<<grid, BLML_PAIR_NT, lds, st>>>(t, ZA, ZB, w, (int)R, (int)M, \
                                                                                   (int)slice, part); \
        else if (t.G == 1) BLML_ROWS_LAUNCH(R_, TL, 1); \
        else BLML_ROWS_LAUNCH(R_, TL, BLML_PAIR_MAXG); \
    } while (0)
    // tables in LDS whose columns are all offset rows: pair_rows_off_kernel, by column count
    bool all_off = tl && t.G == 0 && t.d <= 5;
    for (int k = 0; k < t.d && all_off; ++k) all_off = t.kind[k] == 0;
    if (all_off) {
#define BLML_OFF_LAUNCH(R_, D) \
        pair_rows_off_kernel<R_, D, 4><<<grid, BLML_PAIR_NT, lds, st>>>(t, ZA, ZB, w, (int)R, (int)M, (int)slice, part)
#define BLML_OFF_BY_D(R_) \
        do { \
            switch (t.d) { \
                case 1: BLML_OFF_LAUNCH(R_, 1); break; \
                case 2: BLML_OFF_LAUNCH(R_, 2); break; \
                case 3: BLML_OFF_LAUNCH(R_, 3); break; \
                case 4: BLML_OFF_LAUNCH(R_, 4); break; \
                default: BLML_OFF_LAUNCH(R_, 5); break; \
            } \
        } while (0)
        if (t.f32) BLML_OFF_BY_D(float);
        else BLML_OFF_BY_D(double);
#undef BLML_OFF_BY_D
#undef BLML_OFF_LAUNCH
        HIPCHECK(hipGetLastError());
        pair_rows_sum_kernel<<<(unsigned)((R + 255) / 256), 256, 0, st>>>(part, (int)R, (int)nsl, s);
        HIPCHECK(hipGetLastError());
        return BLML_OK;
    }
    // pair_check: a column of kind 2 or 4 names a group < G, so G = 0 means no such column
    if (t.f32) {
        if (tl) BLML_ROWS_BY_G(float, true);
        else BLML_ROWS_BY_G(float, false);
    } else {
        if (tl) BLML_ROWS_BY_G(double, true);
        else BLML_ROWS_BY_G(double, false);
    }
#undef BLML_ROWS_BY_G
#undef BLML_ROWS_LAUNCH
    HIPCHECK(hipGetLastError());
    pair_rows_sum_kernel<<<(unsigned)((R + 255) / 256), 256, 0, st>>>(part, (int)R, (int)nsl, s);
    HIPCHECK(hipGetLastError());
    return BLML_OK;
}

int blml_pair_rows(const blml_pairtab_t *tab, const int32_t *ZA, const int32_t *ZB, const double *w, int64_t R,
                   int64_t M, double *s, void *stream) {
    int rc = pair_check(tab);
    if (rc) return rc;
    if (R < 0 || M < 0 || R >= (1L << 31) || M >= (1L << 31)) return fail(BLML_EVALUE, "R, M: 0 .. 2^31 - 1");
    if ((R > 0 && (!ZA || !s)) || (M > 0 && (!ZB || !w))) return fail(BLML_EVALUE, "null argument");
    hipStream_t st = (hipStream_t)stream;
    if (R == 0) return BLML_OK;
    if (M == 0) {
        HIPCHECK(hipMemsetAsync(s, 0, sizeof(double) * R, st));
        return BLML_OK;
    }
    long rb, nsl, slice;
    pair_rows_grid(R, M, rb, nsl, slice);
    StreamScratch sc(st);
    HIPCHECK(sc.alloc(sizeof(double) * nsl * R));
    rc = pair_rows_launch(*tab, ZA, ZB, w, R, M, (double *)sc.p, s, st);
    if (rc) return rc;
    HIPCHECK(sc.release());
    return BLML_OK;
}

// ---- blml_pair_rows_noalloc: blml_pair_rows' launch with its partial sums in a fixed array of the library's own.
// The partials of one call, nsl * R doubles (pair_rows_grid), live in `scratch`: one array per GPU (the code
// object's copy on each device), held until the process ends.  Host code reaches that memory only through the
// address hipGetSymbolAddress gives on each device, kept per device in g_scratch_addr (the array's name on the host
// is a stand-in, not a device address).  The rows kernel writes every (row, slice) partial before the sum kernel
// reads it, so nothing needs clearing between calls; every call on one GPU uses the same array, so calls that
// overlap on one GPU race.
static constexpr long kNoAllocDoubles = BLML_PAIR_NOALLOC_MAX;
static __device__ __align__(256) double scratch[kNoAllocDoubles];

static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_addr[kMaxDevices] = {};

// the current device's address of `scratch`, looked up on the first call on that device (with it current)
static int scratch_addr(double **out) {
    int dev = 0;
    HIPCHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(BLML_ERUNTIME, "device index beyond the library's table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_addr[dev]) {
        void *p = nullptr;
        HIPCHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(BLML_ERUNTIME, "hipGetSymbolAddress(scratch) gave no address");
        g_scratch_addr[dev] = p;
    }
    *out = (double *)g_scratch_addr[dev];
    return BLML_OK;
}

static __global__ void fill_zero_kernel(double *s, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) s[i] = 0.0;
}

int blml_pair_rows_noalloc(const blml_pairtab_t *tab, const int32_t *ZA, const int32_t *ZB, const double *w,
                           int64_t R, int64_t M, double *s, void *stream) {
    int rc = pair_check(tab);
    if (rc) return rc;
    if (R < 0 || M < 0 || R >= (1L << 31) || M >= (1L << 31)) return fail(BLML_EVALUE, "R, M: 0 .. 2^31 - 1");
    if ((R > 0 && (!ZA || !s)) || (M > 0 && (!ZB || !w))) return fail(BLML_EVALUE, "null argument");
    hipStream_t st = (hipStream_t)stream;
    if (R == 0) return BLML_OK;
    if (M == 0) {      // a fill kernel, not a memset node: the caller may be capturing a graph
        fill_zero_kernel<<<(unsigned)((R + 255) / 256), 256, 0, st>>>(s, R);
        HIPCHECK(hipGetLastError());
        return BLML_OK;
    }
    long rb, nsl, slice;
    pair_rows_grid(R, M, rb, nsl, slice);
    if (nsl * R > kNoAllocDoubles)
        return fail(BLML_EVALUE, "blml_pair_rows_noalloc: " + std::to_string(R) + " rows need " +
                                     std::to_string(nsl * R) + " partial sums, more than its fixed " +
                                     std::to_string(kNoAllocDoubles) + "; use blml_pair_rows");
    double *part = nullptr;
    rc = scratch_addr(&part);
    if (rc) return rc;
    return pair_rows_launch(*tab, ZA, ZB, w, R, M, part, s, st);
}

// rows for Mm densities with derivative entries: Wt (M x Mm, device), umask (Mm x U, host; NULL: every unit in
// every density), s (R x Mm), ds (R x Mm x nW)
static int pair_multi_run(const blml_pairtab_t *tab, const blml_pairgrad_t *grad, const int32_t *ZA,
                          const int32_t *ZB, const double *Wt, const int32_t *umask, int Mm, int64_t R, int64_t M,
                          double *s, double *ds, hipStream_t st) {
    int rc = pair_check(tab);
    if (rc) return rc;
    if (!grad) return fail(BLML_EVALUE, "null unit / derivative entries");
    const blml_pairgrad_t &g = *grad;
    if (Mm < 1 || Mm > BLML_PAIR_MAXM)
        return fail(BLML_EVALUE, "densities: 1.." + std::to_string(BLML_PAIR_MAXM) + " per pass on this engine");
    if (g.nW < 0 || (long)Mm * g.nW > BLML_PAIR_MAXW)
        return fail(BLML_EVALUE, "derivative entries: densities x entries at most " +
                                     std::to_string(BLML_PAIR_MAXW) + " on this engine");
    if (g.U < 1 || g.U > BLML_PAIR_MAXD) return fail(BLML_EVALUE, "derivative rows: 1.." +
                                                     std::to_string(BLML_PAIR_MAXD) + " kernel units");
    pair_multi_dev gd{};
    gd.Mm = Mm;
    gd.nE = Mm * g.nW;
    gd.U = g.U;
    for (int m = 0; m < Mm; ++m) {
        unsigned um = 0;
        for (int u = 0; u < g.U; ++u)
            if (!umask || umask[(long)m * g.U + u]) um |= 1u << u;
        gd.umask[m] = um;
    }
    for (int k = 0; k < BLML_PAIR_MAXD; ++k) gd.dim_unit[k] = -1;
    for (int k = 0; k < tab->d; ++k) {
        const int kd = tab->kind[k];
        if (kd == 2 || kd == 4) continue;
        if (g.dim_unit[k] < 0 || g.dim_unit[k] >= g.U) return fail(BLML_EVALUE, "derivative rows: a column's unit");
        gd.dim_unit[k] = g.dim_unit[k];
    }
    for (int c = 0; c < tab->G; ++c) {
        if (g.grp_unit[c] < 0 || g.grp_unit[c] >= g.U) return fail(BLML_EVALUE, "derivative rows: a group's unit");
        gd.grp_unit[c] = g.grp_unit[c];
    }
    bool cols = false, grps = false;
    for (int q = 0; q < g.nW; ++q) {
        const int gq = g.wgrp[q], kc = g.wcol[q];
        if (g.wunit[q] < 0 || g.wunit[q] >= g.U) return fail(BLML_EVALUE, "derivative entry: unit out of range");
        if (gq >= 0) {
            if (gq >= tab->G || kc < -1 || (kc >= 0 && (kc >= tab->d || tab->grp[kc] != gq)))
                return fail(BLML_EVALUE, "derivative entry: a group's entry outside its group");
            grps = true;
        } else {
            if (kc < 0 || kc >= tab->d || tab->kind[kc] == 2 || tab->kind[kc] == 4)
                return fail(BLML_EVALUE, "derivative entry: a column entry needs a column of kind 0, 1 or 3");
            if (g.dbase[q] < 0 || g.dbase[q] >= g.ndT) return fail(BLML_EVALUE, "derivative entry: dbase outside dT");
            cols = true;
        }
        for (int m = 0; m < Mm; ++m) {
            const int e = m * g.nW + q;
            gd.em[e] = m;
            gd.eq[e] = q;
            gd.eumask[e] = gd.umask[m];
            gd.eunit[e] = g.wunit[q];
            gd.egrp[e] = gq;
            gd.ecol[e] = kc;
            if (gq < 0) {
                gd.erank[e] = tab->kind[kc] == 3;
                gd.eJ[e] = tab->J[kc];
                gd.emk[e] = tab->mk[kc];
                gd.edbase[e] = g.dbase[q];
            }
        }
    }
    if (cols && !g.dT) return fail(BLML_EVALUE, "derivative rows: dT missing");
    if (grps && (!g.dR || tab->LR < 2)) return fail(BLML_EVALUE, "derivative rows: dR missing");
    gd.ndT = g.ndT;
    gd.dT = g.dT;
    gd.dR = g.dR;
    if (R < 0 || M < 0 || R >= (1L << 31) || M >= (1L << 31)) return fail(BLML_EVALUE, "R, M: 0 .. 2^31 - 1");
    if ((R > 0 && (!ZA || !s || (gd.nE > 0 && !ds))) || (M > 0 && (!ZB || !Wt))) return fail(BLML_EVALUE, "null argument");
    if (R == 0) return BLML_OK;
    if (M == 0) {
        HIPCHECK(hipMemsetAsync(s, 0, sizeof(double) * R * Mm, st));
        if (gd.nE > 0) HIPCHECK(hipMemsetAsync(ds, 0, sizeof(double) * R * gd.nE, st));
        return BLML_OK;
    }
    long rb, nsl, slice;
    pair_rows_grid(R, M, rb, nsl, slice);
    StreamScratch sc(st);
    HIPCHECK(sc.alloc(sizeof(double) * nsl * (Mm + gd.nE) * R));
    double *part = (double *)sc.p;
    const blml_pairtab_t t = *tab;
    const size_t es = t.f32 ? 4 : 8;
    const bool tl = t.nT * (long)es <= 32768;
    const long tbytes = tl ? ((t.nT * (long)es + 15) / 16) * 16 : 0;
    const int lds = (int)(tbytes + ((BLML_PAIR_CH * BLML_PAIR_MAXD * 4 + 15) / 16) * 16 +
                          BLML_PAIR_CH * BLML_PAIR_MAXM * 8);
    const dim3 grid((unsigned)rb, (unsigned)nsl);
#define BLML_MULTI_LAUNCH_G(R_, TL, NE, NG) \
    pair_rows_multi_kernel<R_, TL, NE, NG><<<grid, BLML_PAIR_NT, lds, st>>>(t, gd, ZA, ZB, Wt, (int)R, (int)M, (int)slice, part)
#define BLML_MULTI_LAUNCH(R_, TL, NE) \
    do { \
        if (t.G == 0) BLML_MULTI_LAUNCH_G(R_, TL, NE, 0); \
        else if (t.G == 1) BLML_MULTI_LAUNCH_G(R_, TL, NE, 1); \
        else BLML_MULTI_LAUNCH_G(R_, TL, NE, BLML_PAIR_MAXG); \
    } while (0)
    const bool small = gd.nE <= 4;     // a group entry (egrp >= 0) needs t.G >= 1, checked above
    if (t.f32) {
        if (tl) { if (small) BLML_MULTI_LAUNCH(float, true, 4); else BLML_MULTI_LAUNCH(float, true, BLML_PAIR_MAXW); }
        else { if (small) BLML_MULTI_LAUNCH(float, false, 4); else BLML_MULTI_LAUNCH(float, false, BLML_PAIR_MAXW); }
    } else {
        if (tl) { if (small) BLML_MULTI_LAUNCH(double, true, 4); else BLML_MULTI_LAUNCH(double, true, BLML_PAIR_MAXW); }
        else { if (small) BLML_MULTI_LAUNCH(double, false, 4); else BLML_MULTI_LAUNCH(double, false, BLML_PAIR_MAXW); }
    }
#undef BLML_MULTI_LAUNCH
#undef BLML_MULTI_LAUNCH_G
    HIPCHECK(hipGetLastError());
    pair_rows_multi_sum_kernel<<<(unsigned)((R + 255) / 256), 256, 0, st>>>(part, (int)R, (int)nsl, Mm, gd.nE, s, ds);
    HIPCHECK(hipGetLastError());
    HIPCHECK(sc.release());
    return BLML_OK;
}

int blml_pair_rows_grad(const blml_pairtab_t *tab, const blml_pairgrad_t *grad, const int32_t *ZA, const int32_t *ZB,
                        const double *w, int64_t R, int64_t M, double *s, double *ds, void *stream) {
    if (grad && grad->nW < 1) return fail(BLML_EVALUE, "derivative rows: no derivative entries");
    return pair_multi_run(tab, grad, ZA, ZB, w, nullptr, 1, R, M, s, ds, (hipStream_t)stream);
}

int blml_pair_rows_multi(const blml_pairtab_t *tab, const blml_pairgrad_t *grad, const int32_t *ZA, const int32_t *ZB,
                         const double *Wt, const int32_t *umask, int32_t Mm, int64_t R, int64_t M, double *s, double *ds,
                         void *stream) {
    if (!umask) return fail(BLML_EVALUE, "null unit mask");
    return pair_multi_run(tab, grad, ZA, ZB, Wt, umask, Mm, R, M, s, ds, (hipStream_t)stream);
}

int blml_pair_cols(const blml_pairtab_t *tab, const int32_t *ZA, const int32_t *ZB, int64_t R, int64_t M,
                   int64_t depmask, void *out, void *out_m, void *stream) {
    int rc = pair_check(tab);
    if (rc) return rc;
    if (R < 0 || M < 0) return fail(BLML_EVALUE, "R, M must be >= 0");
    if (R == 0 || M == 0) return BLML_OK;
    if (!ZA || !ZB || !out || (depmask != 0 && !out_m)) return fail(BLML_EVALUE, "null argument");
    hipStream_t st = (hipStream_t)stream;
    const long n = R * M;
    const unsigned blocks = (unsigned)std::min<long>((n + BLML_PAIR_NT - 1) / BLML_PAIR_NT, 16384);
    const blml_pairtab_t t = *tab;
#define BLML_COLS_LAUNCH(R_, NG) \
    pair_cols_kernel<R_, NG><<<blocks, BLML_PAIR_NT, 0, st>>>(t, ZA, ZB, R, M, depmask, (R_ *)out, \
                                                             (R_ *)(out_m ? out_m : out))
#define BLML_COLS_BY_G(R_) \
    do { \
        if (t.G == 0) BLML_COLS_LAUNCH(R_, 0); \
        else if (t.G == 1) BLML_COLS_LAUNCH(R_, 1); \
        else BLML_COLS_LAUNCH(R_, BLML_PAIR_MAXG); \
    } while (0)
    if (tab->f32) BLML_COLS_BY_G(float);
    else BLML_COLS_BY_G(double);
#undef BLML_COLS_BY_G
#undef BLML_COLS_LAUNCH
    HIPCHECK(hipGetLastError());
    return BLML_OK;
}

// ------------------------------------------------------------------------------------------------ points
int blml_point_rows(const blml_pointunit_t *units, int32_t U, int32_t d, const int32_t *harm, const double *X,
                    const double *Z, const double *W, int32_t nw, int64_t Nq, int64_t M, double *S, void *stream) {
    if (!units || U < 1 || U > BLML_POINT_MAXU) return fail(BLML_EVALUE, "point rows: 1..8 kernel units");
    if (d < 1 || d > BLML_MAXD) return fail(BLML_EVALUE, "point rows: 1..5 columns");
    if (nw != 1 && nw != 2) return fail(BLML_EVALUE, "point rows: nw 1 (real) or 2 (complex)");
    if (Nq < 0 || M < 0 || Nq >= (1L << 31) || M >= (1L << 31)) return fail(BLML_EVALUE, "Nq, M: 0 .. 2^31 - 1");
    if ((Nq > 0 && (!X || !S)) || (M > 0 && (!Z || !W))) return fail(BLML_EVALUE, "null argument");
    PointTab t;
    t.d = d;
    t.U = U;
    for (int u = 0; u < U; ++u) {
        const blml_pointunit_t &a = units[u];
        if (a.kind < 0 || a.kind > 4) return fail(BLML_EVALUE, "point rows: unknown unit kind");
        const int nc = a.kind == 3 || a.kind == 4 ? a.ncol : 1;
        if (nc < 1 || nc > d || (a.kind == 3 && nc > 5)) return fail(BLML_EVALUE, "point rows: unit columns");
        for (int i = 0; i < nc; ++i)
            if (a.col[i] < 0 || a.col[i] >= d) return fail(BLML_EVALUE, "point rows: a unit's column outside 0..d-1");
        if (a.kind == 2 && !(a.period[0] > 0)) return fail(BLML_EVALUE, "point rows: dirichlet period");
        if (a.kind == 4 && (a.nharm < 0 || (a.nharm > 0 && (!harm || a.hoff < 0))))
            return fail(BLML_EVALUE, "point rows: periodic ball harmonics");
        t.u[u] = a;
        t.u[u].ncol = nc;
    }
    hipStream_t st = (hipStream_t)stream;
    if (Nq == 0) return BLML_OK;
    if (M == 0) {
        fill_zero_kernel<<<(unsigned)((Nq * nw + 255) / 256), 256, 0, st>>>(S, Nq * nw);
        HIPCHECK(hipGetLastError());
        return BLML_OK;
    }
    const long rb = (Nq + BLML_POINT_NT - 1) / BLML_POINT_NT;
    long nsl = std::max<long>(1, std::min<long>((2048 + rb - 1) / rb, (M + BLML_POINT_NT - 1) / BLML_POINT_NT));
    const long slice = ((M + nsl - 1) / nsl + BLML_POINT_NT - 1) / BLML_POINT_NT * BLML_POINT_NT;
    nsl = (M + slice - 1) / slice;
    StreamScratch sc(st);
    HIPCHECK(sc.alloc(sizeof(double) * nsl * Nq * nw));
    double *part = (double *)sc.p;
    const dim3 grid((unsigned)rb, (unsigned)nsl);
    if (nw == 1)
        point_rows_kernel<1><<<grid, BLML_POINT_NT, 0, st>>>(t, harm, X, Z, W, (int)Nq, (int)M, (int)slice, part);
    else
        point_rows_kernel<2><<<grid, BLML_POINT_NT, 0, st>>>(t, harm, X, Z, W, (int)Nq, (int)M, (int)slice, part);
    HIPCHECK(hipGetLastError());
    point_rows_sum_kernel<<<(unsigned)((Nq * nw + 255) / 256), 256, 0, st>>>(part, (int)Nq, (int)nsl, nw, S);
    HIPCHECK(hipGetLastError());
    HIPCHECK(sc.release());
    return BLML_OK;
}

// ------------------------------------------------------------------------------------------------ gradients
}  // extern "C"
template <typename T>
static int grad_dots_t(const GradArgs &g, const T *E, const T *U, const T *dU, const double *m, double *out,
                       hipStream_t st) {
    const long V = (long)g.K * g.B * g.nc;
    const int nb1 = (int)std::min<long>(1024, (V + BLML_GRAD_NT - 1) / BLML_GRAD_NT);
    const int nb2 = (int)std::min<long>(1024, (g.B + BLML_GRAD_NT - 1) / BLML_GRAD_NT);
    const int W = 2 + 2 * g.nE;
    StreamScratch sc(st);
    HIPCHECK(sc.alloc(sizeof(double) * (2L * nb1 + (long)W * nb2)));
    double *p1 = (double *)sc.p, *p2 = p1 + 2L * nb1;
    grad_pass1_kernel<T><<<nb1, BLML_GRAD_NT, 0, st>>>(g, E, U, m, p1);
    HIPCHECK(hipGetLastError());
    grad_sum_kernel<<<1, 64, 0, st>>>(p1, nb1, 2, out);
    HIPCHECK(hipGetLastError());
    grad_pass2_kernel<T><<<nb2, BLML_GRAD_NT, 0, st>>>(g, E, U, dU, m, out, p2);
    HIPCHECK(hipGetLastError());
    grad_sum_kernel<<<(W + 63) / 64, 64, 0, st>>>(p2, nb2, W, out + 2);
    HIPCHECK(hipGetLastError());
    HIPCHECK(sc.release());
    return BLML_OK;
}
extern "C" {

int blml_ift_elem(int32_t op, int32_t K, int32_t nc, int64_t B, int32_t k, double n, double eps, double Q, double M,
                  const double *U, const double *dU, const double *X, const double *w, const double *m, double *out,
                  double *out2, void *stream) {
    if (op < 0 || op > 4) return fail(BLML_EVALUE, "ift: unknown op");
    if (K < 1 || K > BLML_MAXK || (nc != 1 && nc != 2) || B < 1 || !w || !out)
        return fail(BLML_EVALUE, "ift: K, nc, B, w, out");
    if ((op == BLML_IFT_INNER || op == BLML_IFT_MIXED || op == BLML_IFT_DPDE) && !U) return fail(BLML_EVALUE, "ift: U");
    if ((op == BLML_IFT_INNER || op == BLML_IFT_MIXED || op == BLML_IFT_REAL) && !dU) return fail(BLML_EVALUE, "ift: dU");
    if ((op == BLML_IFT_REAL || op == BLML_IFT_MASK) && !X) return fail(BLML_EVALUE, "ift: X");
    if ((op == BLML_IFT_REAL || op == BLML_IFT_DPDE) && !m) return fail(BLML_EVALUE, "ift: m");
    if (op == BLML_IFT_MIXED && (k < 0 || k >= K || !out2)) return fail(BLML_EVALUE, "ift: mode k, out2");
    if (op == BLML_IFT_REAL && K != 1) return fail(BLML_EVALUE, "ift: the real rank-1 product has K = 1");
    IftArgs a{op, K, nc, k, B, n, eps, Q, M};
    hipStream_t st = (hipStream_t)stream;
    ift_elem_kernel<<<(unsigned)((B + BLML_GRAD_NT - 1) / BLML_GRAD_NT), BLML_GRAD_NT, 0, st>>>(a, U, dU, X, w, m, out,
                                                                                               out2);
    HIPCHECK(hipGetLastError());
    return BLML_OK;
}

int blml_grad_dots(int32_t dtype, int32_t K, int32_t nc, int64_t B, const void *E, const void *U, const void *dU,
                   int32_t nE, const int32_t *emode, const double *m, double eps, double *out, void *stream) {
    if (dtype != BLML_F32 && dtype != BLML_F64) return fail(BLML_EVALUE, "dtype: BLML_F32 or BLML_F64");
    if (K < 1 || K > BLML_MAXK || (nc != 1 && nc != 2) || B < 1) return fail(BLML_EVALUE, "grad dots: K, nc, B");
    if (nE < 0 || nE > BLML_GRAD_MAXE) return fail(BLML_EVALUE, "grad dots: 0..16 derivative entries");
    if (!E || !U || !m || !out || (nE > 0 && (!dU || !emode))) return fail(BLML_EVALUE, "null argument");
    if (!(eps >= 0)) return fail(BLML_EVALUE, "grad dots: eps >= 0");
    GradArgs g;
    g.K = K;
    g.nE = nE;
    g.nc = nc;
    g.B = B;
    g.eps = eps;
    for (int j = 0; j < BLML_GRAD_MAXE; ++j) g.emode[j] = 0;
    for (int j = 0; j < nE; ++j) {
        if (emode[j] < 0 || emode[j] >= K) return fail(BLML_EVALUE, "grad dots: an entry's mode outside 0..K-1");
        g.emode[j] = emode[j];
    }
    hipStream_t st = (hipStream_t)stream;
    return dtype == BLML_F32 ? grad_dots_t<float>(g, (const float *)E, (const float *)U, (const float *)dU, m, out, st)
                             : grad_dots_t<double>(g, (const double *)E, (const double *)U, (const double *)dU, m, out,
                                                   st);
}

// ------------------------------------------------------------------------------------------------ BLMLQuick counts
int blml_quick_sweep(const blml_pairtab_t *tab, const int32_t *Z, int64_t B, const double *w, const double *sg,
                     const int32_t *perm, int64_t Bs, double n, double sbb, int32_t safeguard, double *e, double *Se,
                     void *stream) {
    int rc = pair_check(tab);
    if (rc) return rc;
    if (B < 1 || B >= (1L << 31)) return fail(BLML_EVALUE, "B: 1 .. 2^31 - 1 bins");
    if (Bs < 1 || Bs > B) return fail(BLML_EVALUE, "block size: 1 .. B");
    if (!Z || !w || !perm || !e || !Se) return fail(BLML_EVALUE, "null argument");
    if (!(n > 0) || !(sbb > 0)) return fail(BLML_EVALUE, "n and S_bb must be positive");
    hipStream_t st = (hipStream_t)stream;
    const int d = tab->d;
    long rb, nsl, slice, nsl_r = 0;
    pair_rows_grid(B, Bs, rb, nsl, slice);
    if (B % Bs) pair_rows_grid(B, B % Bs, rb, nsl_r, slice);     // the last, shorter block
    nsl = std::max(nsl, nsl_r);
    // one allocation: partials (nsl B), v (B), eold, dlt, estar, src (Bs each), t, block coordinates (Bs d int32)
    const size_t nd = (size_t)nsl * B + B + 4 * (size_t)Bs + 1;
    StreamScratch sc(st);
    HIPCHECK(sc.alloc(sizeof(double) * nd + sizeof(int32_t) * (size_t)Bs * d));
    double *part = (double *)sc.p, *v = part + (size_t)nsl * B, *eold = v + B, *dlt = eold + Bs, *estar = dlt + Bs,
           *src = estar + Bs, *tdev = src + Bs;
    int32_t *Zb = (int32_t *)(tdev + 1);
    for (long s0 = 0; s0 < B; s0 += Bs) {
        const long m = std::min<long>(Bs, B - s0);
        const int32_t *I = perm + s0;
        quick_block_kernel<<<(unsigned)((m + BLML_QUICK_NT - 1) / BLML_QUICK_NT), BLML_QUICK_NT, 0, st>>>(
            Z, d, I, (int)m, e, Se, w, sg, n, sbb, Zb, eold, dlt, estar, src);
        HIPCHECK(hipGetLastError());
        rc = pair_rows_launch(*tab, Z, Zb, src, B, m, part, v, st);
        if (rc) return rc;
        if (safeguard) {
            quick_step_kernel<<<1, BLML_QUICK_NT, 0, st>>>(I, (int)m, Se, v, w, sg, eold, dlt, n, tdev);
            HIPCHECK(hipGetLastError());
        }
        quick_apply_kernel<<<(unsigned)((B + BLML_QUICK_NT - 1) / BLML_QUICK_NT), BLML_QUICK_NT, 0, st>>>(
            B, I, (int)m, v, sg, eold, dlt, estar, safeguard ? tdev : nullptr, e, Se);
        HIPCHECK(hipGetLastError());
    }
    HIPCHECK(sc.release());
    return BLML_OK;
}

int blml_quick_exact(const blml_pairtab_t *tab, const int32_t *Z, int64_t B, const double *e, const double *sg,
                     double *Se, void *stream) {
    int rc = pair_check(tab);
    if (rc) return rc;
    if (B < 1 || B >= (1L << 31)) return fail(BLML_EVALUE, "B: 1 .. 2^31 - 1 bins");
    if (!Z || !e || !Se) return fail(BLML_EVALUE, "null argument");
    hipStream_t st = (hipStream_t)stream;
    long rb, nsl, slice;
    pair_rows_grid(B, B, rb, nsl, slice);
    StreamScratch sc(st);
    HIPCHECK(sc.alloc(sizeof(double) * ((size_t)nsl * B + 2 * (size_t)B)));
    double *part = (double *)sc.p, *src = part + (size_t)nsl * B, *v = src + B;
    const unsigned g = (unsigned)((B + 255) / 256);
    quick_signed_kernel<<<g, 256, 0, st>>>(B, e, sg, src);
    HIPCHECK(hipGetLastError());
    rc = pair_rows_launch(*tab, Z, Z, src, B, B, part, v, st);
    if (rc) return rc;
    quick_signed_kernel<<<g, 256, 0, st>>>(B, v, sg, Se);
    HIPCHECK(hipGetLastError());
    HIPCHECK(sc.release());
    return BLML_OK;
}

int blml_quick_stats(const double *e, const double *Se, const double *w, int64_t B, double n, double *out,
                     void *stream) {
    if (B < 1) return fail(BLML_EVALUE, "B must be >= 1");
    if (!e || !Se || !w || !out) return fail(BLML_EVALUE, "null argument");
    hipStream_t st = (hipStream_t)stream;
    const int np = (int)std::min<long>((B + BLML_QUICK_NT - 1) / BLML_QUICK_NT, 1024);
    StreamScratch sc(st);
    HIPCHECK(sc.alloc(sizeof(double) * 4 * np));
    double *part = (double *)sc.p;
    quick_stats_kernel<<<np, BLML_QUICK_NT, 0, st>>>(B, e, Se, w, n, part);
    HIPCHECK(hipGetLastError());
    quick_stats_sum_kernel<<<1, 64, 0, st>>>(part, np, out);
    HIPCHECK(hipGetLastError());
    HIPCHECK(sc.release());
    return BLML_OK;
}

// ------------------------------------------------------------------------------- complex / rank-K counts fits
// out[r * ostride + ooff + i] = sum_j S(ZA_r, ZB_j) Wt[j * nw + i] for i < nw <= 16; part: nsl * nw * R doubles
// (pair_rows_grid(R, M))
static int psd_rows_launch(const blml_pairtab_t &t, const int32_t *ZA, const int32_t *ZB, const double *Wt, int nw,
                           long R, long M, double *part, double *out, int ostride, int ooff, hipStream_t st) {
    long rb, nsl, slice;
    pair_rows_grid(R, M, rb, nsl, slice);
    const size_t es = t.f32 ? 4 : 8;
    const bool tl = t.nT * (long)es <= 32768;
    const int NWl = nw <= 2 ? 2 : nw <= 4 ? 4 : nw <= 8 ? 8 : 16;
    const long tbytes = tl ? ((t.nT * (long)es + 15) / 16) * 16 : 0;
    const int lds = (int)(tbytes + ((BLML_PAIR_CH * BLML_PAIR_MAXD * 4 + 15) / 16) * 16 + BLML_PAIR_CH * 8 * NWl);
    const dim3 grid((unsigned)rb, (unsigned)nsl);
#define BLML_NW_LAUNCH(R_, TL, NG, NW) \
    pair_rows_nw_kernel<R_, TL, NG, NW><<<grid, BLML_PAIR_NT, lds, st>>>(t, ZA, ZB, Wt, nw, (int)R, (int)M, \
                                                                         (int)slice, part)
#define BLML_NW_BY_W(R_, TL, NG) \
    do { \
        if (NWl == 2) BLML_NW_LAUNCH(R_, TL, NG, 2); \
        else if (NWl == 4) BLML_NW_LAUNCH(R_, TL, NG, 4); \
        else if (NWl == 8) BLML_NW_LAUNCH(R_, TL, NG, 8); \
        else BLML_NW_LAUNCH(R_, TL, NG, 16); \
    } while (0)
#define BLML_NW_BY_G(R_, TL) \
    do { \
        if (t.G == 0) BLML_NW_BY_W(R_, TL, 0); \
        else BLML_NW_BY_W(R_, TL, BLML_PAIR_MAXG); \
    } while (0)
    if (t.f32) {
        if (tl) BLML_NW_BY_G(float, true);
        else BLML_NW_BY_G(float, false);
    } else {
        if (tl) BLML_NW_BY_G(double, true);
        else BLML_NW_BY_G(double, false);
    }
#undef BLML_NW_BY_G
#undef BLML_NW_BY_W
#undef BLML_NW_LAUNCH
    HIPCHECK(hipGetLastError());
    const long tot = R * nw;
    pair_rows_nw_sum_kernel<<<(unsigned)((tot + 255) / 256), 256, 0, st>>>(part, (int)R, (int)nsl, nw, ostride, ooff,
                                                                         out);
    HIPCHECK(hipGetLastError());
    return BLML_OK;
}

static int psd_check(const blml_pairtab_t *tabs, int32_t ntab, const int32_t *const *Z, int64_t B, int32_t K,
                     int32_t nc) {
    if (!tabs || !Z) return fail(BLML_EVALUE, "null argument");
    if (K < 1 || K > BLML_QUICK_MAXK) return fail(BLML_EVALUE, "K: 1 .. " + std::to_string(BLML_QUICK_MAXK));
    if (nc != 1 && nc != 2) return fail(BLML_EVALUE, "nc: 1 (real) or 2 (complex)");
    if (ntab != 1 && ntab != K) return fail(BLML_EVALUE, "tables: one shared, or one per mode");
    if (B < 1 || B >= (1L << 30)) return fail(BLML_EVALUE, "B: 1 .. 2^30 - 1 bins");
    for (int q = 0; q < ntab; ++q) {
        const int rc = pair_check(tabs + q);
        if (rc) return rc;
        if (tabs[q].d != tabs[0].d) return fail(BLML_EVALUE, "tables: the same number of columns");
        if (!Z[q]) return fail(BLML_EVALUE, "null bins");
    }
    return BLML_OK;
}

// SE (ostride K nc) = S src: src (B, K, nc) for a shared table (one pass of K nc columns); per-mode tables read
// byk, the sources by mode (K, B, nc), one pass of nc columns each
static int psd_apply_all(const blml_pairtab_t *tabs, int32_t ntab, const int32_t *const *Z, long B, int K, int nc,
                         const double *src, const double *byk, double *part, double *SE, hipStream_t st) {
    if (ntab == 1) return psd_rows_launch(tabs[0], Z[0], Z[0], src, K * nc, B, B, part, SE, K * nc, 0, st);
    for (int k = 0; k < K; ++k) {
        const int rc = psd_rows_launch(tabs[k], Z[k], Z[k], byk + (size_t)k * B * nc, nc, B, B, part, SE, K * nc,
                                       k * nc, st);
        if (rc) return rc;
    }
    return BLML_OK;
}

static size_t psd_part_len(long R, long M, int nw) {
    long rb, nsl, slice;
    pair_rows_grid(R, M, rb, nsl, slice);
    return (size_t)nsl * nw * R;
}

int blml_quick_psd_exact(const blml_pairtab_t *tabs, int32_t ntab, const int32_t *const *Z, int64_t B, int32_t K,
                         int32_t nc, const double *E, double *SE, void *stream) {
    int rc = psd_check(tabs, ntab, Z, B, K, nc);
    if (rc) return rc;
    if (!E || !SE) return fail(BLML_EVALUE, "null argument");
    hipStream_t st = (hipStream_t)stream;
    const size_t V = (size_t)B * K * nc;
    const size_t np = psd_part_len(B, B, ntab == 1 ? K * nc : nc);
    StreamScratch sc(st);
    HIPCHECK(sc.alloc(sizeof(double) * (np + (ntab == 1 ? 0 : V))));
    double *part = (double *)sc.p, *byk = part + np;
    if (ntab > 1) {
        psd_by_mode_kernel<<<(unsigned)((V + 255) / 256), 256, 0, st>>>(B, K, nc, E, byk);
        HIPCHECK(hipGetLastError());
    }
    rc = psd_apply_all(tabs, ntab, Z, B, K, nc, E, byk, part, SE, st);
    if (rc) return rc;
    HIPCHECK(sc.release());
    return BLML_OK;
}

int blml_quick_psd_sweep(const blml_pairtab_t *tabs, int32_t ntab, const int32_t *const *Z, int64_t B, int32_t K,
                         int32_t nc, const double *w, const int32_t *perm, int64_t Bs, double n, const double *s,
                         int32_t safeguard, double *E, double *SE, void *stream) {
    int rc = psd_check(tabs, ntab, Z, B, K, nc);
    if (rc) return rc;
    if (!w || !perm || !s || !E || !SE) return fail(BLML_EVALUE, "null argument");
    if (Bs < 1 || Bs > B) return fail(BLML_EVALUE, "block size: 1 .. B");
    if (!(n > 0)) return fail(BLML_EVALUE, "n must be positive");
    PsdArgs a;
    a.K = K;
    a.nc = nc;
    a.shared = ntab == 1;
    a.n = n;
    for (int k = 0; k < BLML_QUICK_MAXK; ++k) a.s[k] = k < K ? s[k] : 0.0;
    for (int k = 0; k < K; ++k)
        if (!(a.s[k] > 0)) return fail(BLML_EVALUE, "S_k(z, z) must be positive");
    hipStream_t st = (hipStream_t)stream;
    const int d = tabs[0].d, Vn = K * nc;
    const int nw = ntab == 1 ? Vn : nc;
    size_t np = psd_part_len(B, Bs, nw);
    if (B % Bs) np = std::max(np, psd_part_len(B, B % Bs, nw));
    const int nbs = (int)std::min<long>((B + BLML_PSD_NT - 1) / BLML_PSD_NT, 512);
    // partials, V (B K nc), EI, d, d by mode (Bs K nc each), step partials, t; then the block's bins per table
    const size_t nd = np + (size_t)B * Vn + 3 * (size_t)Bs * Vn + (size_t)nbs * BLML_PSD_NACC + 1;
    StreamScratch sc(st);
    HIPCHECK(sc.alloc(sizeof(double) * nd + sizeof(int32_t) * (size_t)ntab * Bs * d));
    double *part = (double *)sc.p, *Vv = part + np, *EI = Vv + (size_t)B * Vn, *dl = EI + (size_t)Bs * Vn,
           *Wm = dl + (size_t)Bs * Vn, *spart = Wm + (size_t)Bs * Vn, *tdev = spart + (size_t)nbs * BLML_PSD_NACC;
    int32_t *Zb = (int32_t *)(tdev + 1);
    for (long s0 = 0; s0 < B; s0 += Bs) {
        const long m = std::min<long>(Bs, B - s0);
        const int32_t *I = perm + s0;
        for (int q = 0; q < ntab; ++q) {
            psd_gather_kernel<<<(unsigned)((m + 255) / 256), 256, 0, st>>>(Z[q], d, I, (int)m, Zb + (size_t)q * Bs * d);
            HIPCHECK(hipGetLastError());
        }
        psd_block_kernel<<<(unsigned)((m + BLML_PSD_NT - 1) / BLML_PSD_NT), BLML_PSD_NT, 0, st>>>(a, I, (int)m, E, SE,
                                                                                              w, EI, dl, Wm);
        HIPCHECK(hipGetLastError());
        if (ntab == 1) {
            rc = psd_rows_launch(tabs[0], Z[0], Zb, dl, Vn, B, m, part, Vv, Vn, 0, st);
            if (rc) return rc;
        } else {
            for (int k = 0; k < K; ++k) {
                rc = psd_rows_launch(tabs[k], Z[k], Zb + (size_t)k * Bs * d, Wm + (size_t)k * m * nc, nc, B, m, part,
                                     Vv, Vn, k * nc, st);
                if (rc) return rc;
            }
        }
        if (safeguard) {
            psd_step_part_kernel<<<nbs, BLML_PSD_NT, 0, st>>>(a, B, I, (int)m, E, SE, Vv, w, dl, spart);
            HIPCHECK(hipGetLastError());
            psd_step_pick_kernel<<<1, 64, 0, st>>>(spart, nbs, n, tdev);
            HIPCHECK(hipGetLastError());
        }
        const long nval = (long)B * Vn;
        psd_apply_kernel<<<(unsigned)((nval + BLML_PSD_NT - 1) / BLML_PSD_NT), BLML_PSD_NT, 0, st>>>(
            nval, Vn, I, (int)m, Vv, EI, dl, safeguard ? tdev : nullptr, E, SE);
        HIPCHECK(hipGetLastError());
    }
    HIPCHECK(sc.release());
    return BLML_OK;
}

int blml_quick_psd_stats(const double *E, const double *SE, const double *w, int64_t B, int32_t K, int32_t nc,
                         double n, double scale, double *out, void *stream) {
    if (B < 1 || K < 1 || K > BLML_QUICK_MAXK || (nc != 1 && nc != 2)) return fail(BLML_EVALUE, "B, K, nc");
    if (!E || !SE || !w || !out) return fail(BLML_EVALUE, "null argument");
    hipStream_t st = (hipStream_t)stream;
    const int np = (int)std::min<long>((B + BLML_PSD_NT - 1) / BLML_PSD_NT, 1024);
    StreamScratch sc(st);
    HIPCHECK(sc.alloc(sizeof(double) * 4 * np));
    double *part = (double *)sc.p;
    psd_stats_kernel<<<np, BLML_PSD_NT, 0, st>>>(B, K * nc, E, SE, w, n, scale, part);
    HIPCHECK(hipGetLastError());
    quick_stats_sum_kernel<<<1, 64, 0, st>>>(part, np, out);
    HIPCHECK(hipGetLastError());
    HIPCHECK(sc.release());
    return BLML_OK;
}

int blml_quick_psd_wscale(const double *SE, const double *w, int64_t B, int32_t K, int32_t nc, double coef,
                          double *out, void *stream) {
    if (B < 1 || K < 1 || K > BLML_QUICK_MAXK || (nc != 1 && nc != 2)) return fail(BLML_EVALUE, "B, K, nc");
    if (!SE || !w || !out) return fail(BLML_EVALUE, "null argument");
    hipStream_t st = (hipStream_t)stream;
    psd_wscale_kernel<<<(unsigned)((B + BLML_PSD_NT - 1) / BLML_PSD_NT), BLML_PSD_NT, 0, st>>>(B, K, nc, SE, w, coef,
                                                                                            0, out);
    HIPCHECK(hipGetLastError());
    return BLML_OK;
}

int blml_quick_psd_grad(const blml_pairtab_t *tabs, int32_t ntab, const int32_t *const *Z, int64_t B, int32_t K,
                        int32_t nc, const double *w, double n, const double *E, const double *SE, double *G,
                        void *stream) {
    int rc = psd_check(tabs, ntab, Z, B, K, nc);
    if (rc) return rc;
    if (!w || !E || !SE || !G) return fail(BLML_EVALUE, "null argument");
    hipStream_t st = (hipStream_t)stream;
    const size_t V = (size_t)B * K * nc;
    const size_t np = psd_part_len(B, B, ntab == 1 ? K * nc : nc);
    const int ns = (int)std::min<long>((B + BLML_PSD_NT - 1) / BLML_PSD_NT, 1024);
    StreamScratch sc(st);
    // partials, sources, S sources, stats partials, stats
    HIPCHECK(sc.alloc(sizeof(double) * (np + 2 * V + 4 * (size_t)ns + 4)));
    double *part = (double *)sc.p, *src = part + np, *Vv = src + V, *spart = Vv + V, *stats = spart + 4 * (size_t)ns;
    psd_stats_kernel<<<ns, BLML_PSD_NT, 0, st>>>(B, K * nc, E, SE, w, n, 1.0, spart);
    HIPCHECK(hipGetLastError());
    quick_stats_sum_kernel<<<1, 64, 0, st>>>(spart, ns, stats);
    HIPCHECK(hipGetLastError());
    psd_wscale_kernel<<<(unsigned)((B + BLML_PSD_NT - 1) / BLML_PSD_NT), BLML_PSD_NT, 0, st>>>(B, K, nc, SE, w, 1.0,
                                                                                            ntab > 1, src);
    HIPCHECK(hipGetLastError());
    rc = psd_apply_all(tabs, ntab, Z, B, K, nc, src, src, part, Vv, st);
    if (rc) return rc;
    psd_grad_kernel<<<(unsigned)((V + 255) / 256), 256, 0, st>>>((long)V, Vv, SE, stats, n, G);
    HIPCHECK(hipGetLastError());
    HIPCHECK(sc.release());
    return BLML_OK;
}

static int vec_args(int32_t nv, const double *const *X, VecArgs &a) {
    if (nv < 1 || nv > BLML_VEC_MAXV || !X) return fail(BLML_EVALUE, "vectors: 1 .. " + std::to_string(BLML_VEC_MAXV));
    a.nv = nv;
    a.ny = 0;
    for (int j = 0; j < BLML_VEC_MAXV; ++j) {
        a.X[j] = j < nv ? X[j] : nullptr;
        a.c[j] = 0.0;
        for (int i = 0; i < BLML_VEC_MAXV; ++i) a.A[i][j] = 0.0;
    }
    for (int j = 0; j < nv; ++j)
        if (!X[j]) return fail(BLML_EVALUE, "null vector");
    return BLML_OK;
}

int blml_vec_gram(int32_t nv, const double *const *X, int32_t ny, const double *A, int64_t len, double *out,
                  void *stream) {
    VecArgs a;
    int rc = vec_args(nv, X, a);
    if (rc) return rc;
    if (ny < 1 || ny > BLML_VEC_MAXV || !A || !out || len < 1) return fail(BLML_EVALUE, "ny, A, out, len");
    a.ny = ny;
    for (int i = 0; i < ny; ++i)
        for (int j = 0; j < nv; ++j) a.A[i][j] = A[i * nv + j];
    hipStream_t st = (hipStream_t)stream;
    const int nb = (int)std::min<long>((len + BLML_PSD_NT - 1) / BLML_PSD_NT, 256);
    StreamScratch sc(st);
    HIPCHECK(sc.alloc(sizeof(double) * (size_t)nb * ny * ny));
    vec_gram_part_kernel<<<nb, BLML_PSD_NT, 0, st>>>(a, len, (double *)sc.p);
    HIPCHECK(hipGetLastError());
    vec_gram_sum_kernel<<<1, 64, 0, st>>>((const double *)sc.p, nb, ny, out);
    HIPCHECK(hipGetLastError());
    HIPCHECK(sc.release());
    return BLML_OK;
}

int blml_vec_lincomb(int32_t nv, const double *const *X, const double *c, int32_t mode, int64_t len, double *out,
                     void *stream) {
    VecArgs a;
    int rc = vec_args(nv, X, a);
    if (rc) return rc;
    if (!c || !out || len < 1 || (mode != 0 && mode != 1)) return fail(BLML_EVALUE, "c, out, len, mode");
    for (int j = 0; j < nv; ++j) a.c[j] = c[j];
    hipStream_t st = (hipStream_t)stream;
    vec_lincomb_kernel<<<(unsigned)((len + BLML_PSD_NT - 1) / BLML_PSD_NT), BLML_PSD_NT, 0, st>>>(a, mode, len, out);
    HIPCHECK(hipGetLastError());
    return BLML_OK;
}

int blml_lbfgs_twoloop(int32_t m, const double *const *S, const double *const *Y, const double *rho, double gamma,
                       int64_t len, const double *g, double *q, void *stream) {
    if (m < 0 || m > 64 || len < 1 || !g || !q || (m > 0 && (!S || !Y || !rho)))
        return fail(BLML_EVALUE, "m: 0 .. 64 pairs, len, g, q");
    for (int i = 0; i < m; ++i)
        if (!S[i] || !Y[i]) return fail(BLML_EVALUE, "null pair");
    hipStream_t st = (hipStream_t)stream;
    const int nb = (int)std::min<long>((len + BLML_PSD_NT - 1) / BLML_PSD_NT, 256);
    const unsigned ge = (unsigned)((len + 255) / 256);
    StreamScratch sc(st);
    HIPCHECK(sc.alloc(sizeof(double) * ((size_t)nb + 2 * (size_t)m + 1)));
    double *part = (double *)sc.p, *al = part + nb, *cf = al + m;
    HIPCHECK(hipMemcpyAsync(q, g, sizeof(double) * len, hipMemcpyDeviceToDevice, st));
    for (int i = m - 1; i >= 0; --i) {              // a_i = rho_i s_i.q, q -= a_i y_i
        vec_dot_part_kernel<<<nb, BLML_PSD_NT, 0, st>>>(S[i], q, len, part);
        vec_dot_final_kernel<<<1, 64, 0, st>>>(part, nb, rho[i], 0, nullptr, al + i);
        vec_axpy_dev_kernel<<<ge, 256, 0, st>>>(len, al + i, -1.0, Y[i], q);
    }
    HIPCHECK(hipGetLastError());
    {
        VecArgs a;
        const double *X1[1] = {q};
        int rc = vec_args(1, X1, a);
        if (rc) return rc;
        a.c[0] = gamma;
        vec_lincomb_kernel<<<(unsigned)((len + BLML_PSD_NT - 1) / BLML_PSD_NT), BLML_PSD_NT, 0, st>>>(a, 0, len, q);
    }
    for (int i = 0; i < m; ++i) {                   // q += (a_i - rho_i y_i.q) s_i
        vec_dot_part_kernel<<<nb, BLML_PSD_NT, 0, st>>>(Y[i], q, len, part);
        vec_dot_final_kernel<<<1, 64, 0, st>>>(part, nb, rho[i], 1, al + i, cf + i);
        vec_axpy_dev_kernel<<<ge, 256, 0, st>>>(len, cf + i, 1.0, S[i], q);
    }
    HIPCHECK(hipGetLastError());
    HIPCHECK(sc.release());
    return BLML_OK;
}

// ------------------------------------------------------------------------------------------------ warm-start net
static int net_check(int N, int Cin, int Cout, int H, int W) {
    if (N < 1 || Cin < 1 || Cin > 32 || Cout < 1 || Cout > 32)
        return fail(BLML_EVALUE, "net: 1..32 input and output channels, N >= 1");
    if (W != BLML_NET_W || H < BLML_NET_RT || H % BLML_NET_RT)
        return fail(BLML_EVALUE, "net: images of width 64 and even height");
    return BLML_OK;
}
#define NET_GRID(n) (unsigned)(((n) + 255) / 256)

int blml_net_conv(const void *X, const void *Wk, const float *bias, float *Z, void *A, int32_t N, int32_t Cin,
                  int32_t Cout, int32_t H, int32_t W, int32_t gelu, void *stream) {
    int rc = net_check(N, Cin, Cout, H, W);
    if (rc) return rc;
    if (!X || !Wk || !Z || (gelu && !A)) return fail(BLML_EVALUE, "null argument");
    const int Kpad = (Cin * 9 + 7) / 8 * 8;
    const int lds = (((Cin * BLML_NET_PATCH + 1 + 7) / 8) * 8 + 32 * Kpad) * 2 + Kpad * 4;
    net_conv_kernel<<<dim3(H / BLML_NET_RT, N), BLML_NET_NT, lds, (hipStream_t)stream>>>(
        (const bf16_t *)X, (const bf16_t *)Wk, bias, Z, (bf16_t *)A, Cin, Cout, H, gelu);
    HIPCHECK(hipGetLastError());
    return BLML_OK;
}

int blml_net_conv_grad_w(const void *X, const void *dZ, float *dW, float *db, int32_t N, int32_t Cin, int32_t Cout,
                         int32_t H, int32_t W, void *stream) {
    int rc = net_check(N, Cin, Cout, H, W);
    if (rc) return rc;
    if (!X || !dZ || !dW || !db) return fail(BLML_EVALUE, "null argument");
    hipStream_t st = (hipStream_t)stream;
    const int K = Cin * 9, len = Cout * K;
    StreamScratch sc_w(st), sc_b(st);
    HIPCHECK(sc_w.alloc(sizeof(float) * (size_t)N * len));
    HIPCHECK(sc_b.alloc(sizeof(float) * (size_t)N * Cout));
    float *part = (float *)sc_w.p, *bpart = (float *)sc_b.p;
    const int lds = ((Cin * BLML_NET_PATCH + 7) / 8) * 8 * 2 + 32 * BLML_NET_RT * BLML_NET_W * 2;
    net_conv_dw_kernel<<<N, BLML_NET_NT, lds, st>>>((const bf16_t *)X, (const bf16_t *)dZ, part, Cin, Cout, H);
    HIPCHECK(hipGetLastError());
    net_sum_parts_kernel<<<NET_GRID(len), 256, 0, st>>>(part, N, len, dW);
    HIPCHECK(hipGetLastError());
    net_bias_grad_kernel<<<N * Cout, BLML_NET_NT, 0, st>>>((const bf16_t *)dZ, Cout, H * W, bpart);
    HIPCHECK(hipGetLastError());
    net_sum_parts_kernel<<<NET_GRID(Cout), 256, 0, st>>>(bpart, N, Cout, db);
    HIPCHECK(hipGetLastError());
    HIPCHECK(sc_w.release());
    HIPCHECK(sc_b.release());
    return BLML_OK;
}

int blml_net_flip(const void *Wk, void *Wf, int32_t Cin, int32_t Cout, void *stream) {
    if (!Wk || !Wf || Cin < 1 || Cout < 1) return fail(BLML_EVALUE, "bad argument");
    net_flip_kernel<<<NET_GRID(Cin * Cout * 9), 256, 0, (hipStream_t)stream>>>((const bf16_t *)Wk, (bf16_t *)Wf, Cin,
                                                                                Cout);
    HIPCHECK(hipGetLastError());
    return BLML_OK;
}

int blml_net_gelu_grad(const float *dA, const float *Z, void *dZ, int64_t n, void *stream) {
    if (!dA || !Z || !dZ || n < 0) return fail(BLML_EVALUE, "bad argument");
    if (n) net_gelu_bwd_kernel<<<NET_GRID(n), 256, 0, (hipStream_t)stream>>>(dA, Z, (bf16_t *)dZ, n);
    HIPCHECK(hipGetLastError());
    return BLML_OK;
}

int blml_net_mse(const float *Z, const float *Y, const float *M, void *dZ, float *loss_part, int64_t n, float inv_msum,
                 void *stream) {
    if (!Z || !Y || !M || !dZ || !loss_part || n < 1) return fail(BLML_EVALUE, "bad argument");
    net_mse_kernel<<<NET_GRID(n), BLML_NET_NT, 0, (hipStream_t)stream>>>(Z, Y, M, (bf16_t *)dZ, loss_part, n,
                                                                          inv_msum);
    HIPCHECK(hipGetLastError());
    return BLML_OK;
}

int blml_net_adam(float *p, const float *g, float *m, float *v, void *bf, int64_t n, float lr, float b1, float b2,
                  float eps, int32_t step, void *stream) {
    if (!p || !g || !m || !v || n < 1 || step < 1) return fail(BLML_EVALUE, "bad argument");
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    net_adam_kernel<<<NET_GRID(n), 256, 0, (hipStream_t)stream>>>(p, g, m, v, (bf16_t *)bf, n, lr, b1, b2, eps, c1, c2);
    HIPCHECK(hipGetLastError());
    return BLML_OK;
}

int blml_net_to_bf16(const float *x, void *y, int64_t n, void *stream) {
    if (!x || !y || n < 0) return fail(BLML_EVALUE, "bad argument");
    if (n) net_to_bf16_kernel<<<NET_GRID(n), 256, 0, (hipStream_t)stream>>>(x, (bf16_t *)y, n);
    HIPCHECK(hipGetLastError());
    return BLML_OK;
}

// ------------------------------------------------------------------------------------------------ start search
int blml_complex_starts(const blml_hio_params_t *prm, const uint8_t *band, const double *amp, const int64_t *occ,
                        const double *wocc, const double *phase, const double *kick, double *P, double *gocc,
                        double *score, void *stream) {
    if (!prm) return fail(BLML_EVALUE, "null parameters");
    const blml_hio_params_t q = *prm;
    if (q.d < 1 || q.d > BLML_MAXD) return fail(BLML_EVALUE, "d: 1 .. 5 axes");
    long ncell = 1;
    int Lmax = 2;
    for (int j = 0; j < q.d; ++j) {
        const int m = q.shape[j];
        if (m < 2 || m > 16 * BLML_NT || (m & (m - 1)))
            return fail(BLML_EVALUE, "grid sides: powers of two from 2 to 8192");
        if (q.box[j] < 1 || q.box[j] > m) return fail(BLML_EVALUE, "box: 1 .. the grid side per axis");
        ncell *= m;
        Lmax = std::max(Lmax, m);
    }
    if (ncell >= (1L << 31)) return fail(BLML_EVALUE, "grid: below 2^31 cells");
    if (q.n_rand < 0 || q.n_kick < 0 || q.n_rand + q.n_kick < 1) return fail(BLML_EVALUE, "runs: at least one");
    if (q.n_hio < 0 || q.n_er < 0) return fail(BLML_EVALUE, "steps must be >= 0");
    if (q.nocc < 1 || !(q.ab2 > 0)) return fail(BLML_EVALUE, "no counts on the box");
    if (!band || !amp || !occ || !wocc || !P || !gocc || !score || (q.n_rand && !phase) || (q.n_kick && !kick))
        return fail(BLML_EVALUE, "null argument");
    hipStream_t st = (hipStream_t)stream;
    int dev = 0, max_lds = 0, st_lds = 0, ncu = 0;
    HIPCHECK(hipGetDevice(&dev));
    HIPCHECK(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    HIPCHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    int rc = static_lds((const void *)hio_kernel, st_lds);
    if (rc) return rc;
    long lds = 0;      // the line passes' dynamic LDS, as run_pass lays it out
    for (int j = 0; j < q.d; ++j) {
        const PassShape sh = pass_shape(OP_FFT_FULL, q.shape[j], q.shape[j], 0, ncell / q.shape[j]);
        lds = std::max(lds, (long)(sh.roots + (long)sh.lines * sh.buf) * (long)sizeof(cpx<double>));
    }
    if (lds + st_lds > max_lds) return fail(BLML_EVALUE, "grid side too long for the line passes' LDS");
    const int runs = q.n_rand + q.n_kick;
    StreamScratch sc(st);
    HIPCHECK(sc.alloc(sizeof(cpx<double>) * ((size_t)runs * ncell + Lmax)));
    cpx<double> *X = (cpx<double> *)sc.p, *tw = X + (size_t)runs * ncell;
    twiddle_kernel<double><<<(Lmax + 255) / 256, 256, 0, st>>>(tw, Lmax);
    HIPCHECK(hipGetLastError());
    HioArgs a;
    a.d = q.d;
    for (int j = 0; j < BLML_MAXD; ++j) {
        a.shape[j] = j < q.d ? q.shape[j] : 1;
        a.box[j] = j < q.d ? q.box[j] : 1;
    }
    a.ncell = ncell;
    a.n_hio = q.n_hio;
    a.n_er = q.n_er;
    a.n_rand = q.n_rand;
    a.n_kick = q.n_kick;
    a.beta = q.beta;
    a.ab2 = q.ab2;
    a.nocc = q.nocc;
    a.band = band;
    a.amp = amp;
    a.occ = (const long *)occ;
    a.wocc = wocc;
    a.phase = phase;
    a.kick = kick;
    a.P = (cpx<double> *)P;
    a.X = X;
    a.gocc = (cpx<double> *)gocc;
    a.score = score;
    a.tw = tw;
    a.Lmax = Lmax;
    hio_kernel<<<std::min(runs, std::max(ncu, 1)), BLML_NT, (unsigned)lds, st>>>(a);
    HIPCHECK(hipGetLastError());
    return BLML_OK;
}

int blml_band_eval(int32_t d, const int32_t *mq, int64_t nb, int64_t nband, const int32_t *dz, const int32_t *kf,
                   const double *c, double *g, void *stream) {
    if (d < 1 || d > BLML_MAXD || !mq) return fail(BLML_EVALUE, "d: 1 .. 5 axes");
    if (nb < 0 || nband < 0) return fail(BLML_EVALUE, "counts must be >= 0");
    if (nb == 0) return BLML_OK;
    if (!dz || !g || (nband && (!kf || !c))) return fail(BLML_EVALUE, "null argument");
    BandEvalArgs a;
    a.d = d;
    for (int j = 0; j < BLML_MAXD; ++j) {
        a.mq[j] = j < d ? mq[j] : 1;
        if (j < d && mq[j] < 1) return fail(BLML_EVALUE, "mq: positive");
    }
    a.nb = nb;
    a.nband = nband;
    a.dz = dz;
    a.kf = kf;
    a.c = (const cpx<double> *)c;
    a.g = (cpx<double> *)g;
    band_eval_kernel<<<(unsigned)((nb + 255) / 256), 256, 0, (hipStream_t)stream>>>(a);
    HIPCHECK(hipGetLastError());
    return BLML_OK;
}

}  // extern "C"
