// CPU replay of a plan's fit path under AddressSanitizer (test infrastructure).
//
// Reads a plan serialized by tests/cpu_build/replay.py (the blml_fit_t / blml_factor_t rows, tables, parameters and
// the inputs of blml_fit), runs the engine's own host planning (layout() and, for a spread plan, sep_plan_host() of
// plan_layout.hpp) and then the device code of the engine's kernels on the CPU (hip/hip_runtime.h of this
// directory): the spectra (spec_factor), the whole fit (run_fit: start, coordinate descent or L-BFGS, outputs), the
// FP64 finishing pass of FP32 plans (run_finish64) and, on request, the certificate (run_cert).  A batched plan runs
// every fit as one workgroup of BLML_NT threads (fit_kernel's work for that fit); a spread plan runs its one fit on
// a cooperative grid of `ranks` workgroups (fit_grid_kernel's work, grid barriers included; the device launches as
// many as fit on it at once).  Every buffer the kernels receive is an allocation of its own and of the exact size
// blml_plan_create gives it, so are each workgroup's LDS objects and dynamic LDS, and the alignment gaps inside the
// fit workspaces are poisoned, so AddressSanitizer stops at the first access outside them.  Writes the results for
// the caller to compare with the oracle.
#define CPUHIP_DEFINE_SWITCH
#include "hip/hip_runtime.h"
#include "plan_layout.hpp"
#include "empirical.hpp"
#include "starts.hpp"

#include <stdio.h>
#include <stdlib.h>

#include <vector>

cpuhip_dim3 threadIdx, blockIdx, blockDim, gridDim;
cpuhip::Block *cpuhip::g_blk = nullptr;
long long cpuhip::g_clock = 0;

// the device's LDS per workgroup and the static LDS of the spread kernels (compiler resource report of the gfx950
// build: fit_grid_kernel 26896 bytes, matvec_grid_kernel and matvec_pipe_kernel at most 8192), which size the
// separable path's tiles
static constexpr int kMaxLds = 163840, kStaticFit = 26896, kStaticMv = 8192, kStaticPipe = 0;

namespace {

struct In {
    int32_t dtype = 0, nfits = 0, nfacs = 0, flags = 0, cert_iters = 0, ranks = 1;
    double cert_tol = 0;
    blml_params_t P{};
    std::vector<blml_fit_t> fits;
    std::vector<blml_factor_t> facs;
    std::vector<double> fdata, counts, draws, init, signs;
    std::vector<int32_t> freqs;
    blml_emp_params_t ep{};          // flags & 2: blml_empirical's fit in place of blml_fit's
    std::vector<double> aux;
};

template <typename V>
bool rd(FILE *f, V *p, size_t n) {
    return n == 0 || fread(p, sizeof(V), n, f) == n;
}

bool read_in(const char *path, In &in) {
    FILE *f = fopen(path, "rb");
    if (!f) return false;
    char magic[8];
    int64_t n_fdata, n_freqs, n_counts, n_draws, n_init, n_signs;
    bool ok = rd(f, magic, 8) && memcmp(magic, "BLMLRPL1", 8) == 0 && rd(f, &in.dtype, 1) && rd(f, &in.nfits, 1) &&
              rd(f, &in.nfacs, 1) && rd(f, &in.flags, 1) && rd(f, &in.cert_iters, 1) && rd(f, &in.ranks, 1) &&
              rd(f, &in.cert_tol, 1) &&
              rd(f, &n_fdata, 1) && rd(f, &n_freqs, 1) && rd(f, &n_counts, 1) && rd(f, &n_draws, 1) &&
              rd(f, &n_init, 1) && rd(f, &n_signs, 1) && rd(f, &in.P, 1);
    if (ok) {
        in.fits.resize(in.nfits);
        in.facs.resize(in.nfacs);
        in.fdata.resize(n_fdata);
        in.freqs.resize(n_freqs);
        in.counts.resize(n_counts);
        in.draws.resize(n_draws);
        in.init.resize(n_init);
        in.signs.resize(n_signs);
        ok = rd(f, in.fits.data(), in.fits.size()) && rd(f, in.facs.data(), in.facs.size()) &&
             rd(f, in.fdata.data(), in.fdata.size()) && rd(f, in.freqs.data(), in.freqs.size()) &&
             rd(f, in.counts.data(), in.counts.size()) && rd(f, in.draws.data(), in.draws.size()) &&
             rd(f, in.init.data(), in.init.size()) && rd(f, in.signs.data(), in.signs.size());
        if (ok && (in.flags & 2)) {
            in.aux.resize((size_t)in.nfits * BLML_EMP_AUX_LEN);
            ok = rd(f, &in.ep, 1) && rd(f, in.aux.data(), in.aux.size());
        }
    }
    fclose(f);
    return ok;
}

void poison(const void *p, size_t n) {
#ifdef CPUHIP_ASAN
    if (n) __asan_poison_memory_region(p, n);
#else
    (void)p;
    (void)n;
#endif
}
void unpoison(const void *p, size_t n) {
#ifdef CPUHIP_ASAN
    if (n) __asan_unpoison_memory_region(p, n);
#else
    (void)p;
    (void)n;
#endif
}

// a "device" allocation: exactly `bytes` (at least one element), its own heap block
template <typename V>
V *dev_alloc(size_t n, const V *src = nullptr) {
    const size_t bytes = sizeof(V) * (n ? n : 1);
    V *p = (V *)calloc(1, bytes);
    if (!p) {
        fprintf(stderr, "replay: out of memory (%zu bytes)\n", bytes);
        exit(3);
    }
    if (src && n) memcpy(p, src, sizeof(V) * n);
    return p;
}

// one launch of `blocks` workgroups of BLML_NT threads with `lds` bytes of dynamic LDS each (and the kernel's
// static LDS, `st` bytes, within the device's limit, as blml_plan_create checks)
long long launch(int lds, int st, int blocks, std::function<void()> body) {
    if (lds < 0 || lds + st > kMaxLds) {
        fprintf(stderr, "replay: a launch asks for %d + %d bytes of LDS (limit %d)\n", lds, st, kMaxLds);
        exit(4);
    }
    return cpuhip::run_grid(blocks, BLML_NT, (size_t)lds, std::move(body));
}

// the pipelined mat-vec replay's input: a fixed sequence in [-1, 1) (Weyl sequence of the golden ratio)
double pipe_value(long i) {
    const double x = (double)i * 0.6180339887498949;
    return 2.0 * (x - floor(x)) - 1.0;
}

template <typename T>
int replay(In &in, const char *out_path) {
    Plan pl;
    pl.dtype = in.dtype;
    pl.nfits = in.nfits;
    pl.nfacs = in.nfacs;
    pl.params = in.P;
    pl.hfits = in.fits;
    pl.hfacs = in.facs;
    pl.esize = sizeof(T);
    // the FFT passes' radix-stage order (the autotuner's knob; fft_line's sched)
    if (const char *sc = getenv("BLML_REPLAY_FFT_SCHED")) pl.fft_sched = atoi(sc);
    if (const char *tm = getenv("BLML_REPLAY_TOEP")) pl.toep_max = atoi(tm);
    std::vector<int> fac_m;
    std::vector<long> tmp_off;
    long tmp_total = 0;
    int64_t spec_total = 0;
    int rc = layout(pl, in.fdata.data(), (int64_t)in.fdata.size(), in.freqs.data(), (int64_t)in.freqs.size(), fac_m,
                    tmp_off, tmp_total, spec_total);
    if (rc) {
        fprintf(stderr, "replay: layout: %s\n", g_err.c_str());
        return 10 + rc;
    }
    printf("replay: %d fits, %d factors, workspace %ld bytes, LDS %d bytes (FP64 view %d), Lmax %d\n", pl.nfits,
           pl.nfacs, (long)pl.ws_elems, pl.lds_bytes, pl.lds64_bytes, pl.Lmax);
    fflush(stdout);
    // the plan's device buffers, as create_t allocates them
    pl.d_fits = dev_alloc<blml_fit_t>(pl.nfits, pl.hfits.data());
    pl.d_facs = dev_alloc<blml_factor_t>(pl.nfacs, pl.hfacs.data());
    pl.d_fdata = dev_alloc<double>(in.fdata.size(), in.fdata.data());
    pl.d_freqs = dev_alloc<int>(in.freqs.size(), in.freqs.data());
    pl.d_spec = dev_alloc<T>(spec_total);
    pl.d_tw = dev_alloc<cpx<T>>(pl.Lmax);
    const size_t wsb = (size_t)std::max<int64_t>(256, pl.ws_elems);
    pl.d_ws = aligned_alloc(256, (wsb + 255) / 256 * 256);
    memset(pl.d_ws, 0xFF, wsb);      // NaN doubles: a read of a value no pass wrote shows in the results
    int *fm = dev_alloc<int>(fac_m.size(), fac_m.data());
    long *toff = dev_alloc<long>(tmp_off.size(), tmp_off.data());
    for (int i = 0; i < pl.Lmax; ++i) {
        double sn, cs;
        sincospi(-2.0 * (double)i / (double)pl.Lmax, &sn, &cs);
        ((cpx<T> *)pl.d_tw)[i] = {(T)cs, (T)sn};
    }
    {
        cpx<T> *tmp = dev_alloc<cpx<T>>(tmp_total);
        const DevPlan<T> dp = dev_plan<T>(pl);
        for (int fi = 0; fi < pl.nfacs; ++fi)
            launch(pl.lds_bytes, 0, 1, [&] { spec_factor<T>(dp, fi, fm, tmp, toff); });
        free(tmp);
    }
    if (sizeof(T) == 4) {
        pl.d_spec64 = dev_alloc<double>(spec_total);
        pl.d_tw64 = dev_alloc<cpx<double>>(pl.Lmax);
        for (int i = 0; i < pl.Lmax; ++i) {
            double sn, cs;
            sincospi(-2.0 * (double)i / (double)pl.Lmax, &sn, &cs);
            ((cpx<double> *)pl.d_tw64)[i] = {cs, sn};
        }
        cpx<double> *tmp = dev_alloc<cpx<double>>(tmp_total);
        const DevPlan<double> dp = dev_plan64(pl);
        for (int fi = 0; fi < pl.nfacs; ++fi)
            launch(pl.lds64_bytes, 0, 1, [&] { spec_factor<double>(dp, fi, fm, tmp, toff); });
        free(tmp);
    }
    // a spread plan (blml_plan_create's rule): one fit over a cooperative grid, the separable path when it applies
    const long V0 = vmax_of(pl.hfits[0]);
    pl.spread = pl.nfits == 1 && (pl.params.spread > 0 || (pl.params.spread == 0 && V0 >= kSpreadMin));
    SepHost<T> sh;
    unsigned *bar = nullptr;
    double *part = nullptr;
    if (pl.spread) {
        pl.coop = std::max(1, in.ranks);
        bar = dev_alloc<unsigned>(4);
        part = dev_alloc<double>(2 * (size_t)pl.coop * BLML_NRED);
        pl.lds_mv = pl.lds_fit = pl.lds_bytes;
        if (sep_plan_host<T>(pl, in.fdata.data(), in.freqs.data(), kMaxLds, kStaticMv, kStaticFit, kStaticPipe, sh)) {
            pl.d_sep = dev_alloc<SepPlan>(2, sh.sp);
            pl.d_sep_tab = dev_alloc<T>(sh.tab.size(), sh.tab.data());
            pl.d_sep_off = dev_alloc<long>(sh.off.size(), sh.off.data());
        }
        printf("replay: spread over %d workgroups, %s, LDS of the fit launch %d bytes\n", pl.coop,
               pl.sep ? "separable tile passes" : "transform passes", pl.lds_fit);
        fflush(stdout);
    }
    // BLML_REPLAY_PIPE: matvec_pipe_kernel's work on an FP32 pipelined plan instead of the fit, on the input
    // E_i = pipe_value(i), once with the dense tables and once with the centrosymmetric ones; writes both U
    if (getenv("BLML_REPLAY_PIPE")) {
        if constexpr (sizeof(T) == 4) {
            if (!pl.spread || !pl.sep || !pl.pipe_mv || pl.sep_centro_off < 0) {
                fprintf(stderr, "replay: not an FP32 plan of the pipelined mat-vec with centrosymmetric tables\n");
                return 8;
            }
            const blml_fit_t &f = pl.hfits[0];
            const long n = (long)f.K * f.B * (f.complex_g ? 2 : 1);
            T *E = dev_alloc<T>(n);
            for (long i = 0; i < n; ++i) E[i] = (T)pipe_value(i);
            T *U[2] = {dev_alloc<T>(n), dev_alloc<T>(n)};
            const DevPlan<T> dp = dev_plan<T>(pl);
            for (int c = 0; c < 2; ++c) {
                memset(bar, 0, 4 * sizeof(unsigned));
                launch(pl.lds_pipe, kStaticPipe, pl.coop, [&] {
                    const Team tm{(int)blockIdx.x, (int)gridDim.x, bar, part, 0};
                    if (c) run_matvec_pipe<T, true>(dp, tm, E, U[1]);
                    else run_matvec_pipe<T, false>(dp, tm, E, U[0]);
                });
            }
            FILE *fo = fopen(out_path, "wb");
            if (!fo) return 5;
            fwrite(U[0], sizeof(T), (size_t)n, fo);
            fwrite(U[1], sizeof(T), (size_t)n, fo);
            fclose(fo);
            printf("replay: pipelined mat-vec of %ld values, dense and centrosymmetric\n", n);
            printf("replay: ok\n");
            return 0;
        } else {
            fprintf(stderr, "replay: the pipelined mat-vec replay takes FP32 plans\n");
            return 8;
        }
    }
    // BLML_REPLAY_MATVEC: matvec_kernel's work (each fit of a batched plan on one workgroup) on the input
    // E_i = pipe_value(i) over the plan's outputs; writes U
    if (getenv("BLML_REPLAY_MATVEC")) {
        if (pl.spread) {
            fprintf(stderr, "replay: the batched mat-vec replay takes batched plans\n");
            return 8;
        }
        const long n = pl.out_elems;
        T *E = dev_alloc<T>(n);
        for (long i = 0; i < n; ++i) E[i] = (T)pipe_value(i);
        T *U = dev_alloc<T>(n);
        const DevPlan<T> dp = dev_plan<T>(pl);
        for (int fi = 0; fi < pl.nfits; ++fi)
            launch(pl.lds_bytes, kStaticMv, 1, [&] { run_matvec<T>(dp, fi, block_team(), E, U, nullptr); });
        FILE *fo = fopen(out_path, "wb");
        if (!fo) return 5;
        fwrite(U, sizeof(T), (size_t)n, fo);
        fclose(fo);
        printf("replay: batched mat-vec of %ld values\n", n);
        printf("replay: ok\n");
        return 0;
    }
    // the alignment gaps of every workspace: [X Y | gap | pool | gap | (FP32: X64 Y64) | gap]
    unsigned char *ws = (unsigned char *)pl.d_ws;
    for (const blml_fit_t &f : pl.hfits) {
        const long x_end = 4L * f.scratch_len * (long)sizeof(T);
        const long pool = pool_bytes_off(f, (int)sizeof(T));
        const long pool_end = pool + 8L * n_slots(pl.params, f) * vmax_of(f);
        const long fin = fin_bytes_off(pl.params, f, (int)sizeof(T));
        const long fin_end = sizeof(T) == 4 ? fin + 32L * f.scratch_len : fin;
        poison(ws + f.ws_off + x_end, pool - x_end);
        poison(ws + f.ws_off + pool_end, fin - pool_end);
        poison(ws + f.ws_off + fin_end, f.ws_len - fin_end);
    }
    // the inputs and outputs of blml_fit (FitBatch.build's buffers)
    const int nc_hist = pl.params.hist_cap;
    double *counts = dev_alloc<double>(in.counts.size(), in.counts.data());
    double *draws = dev_alloc<double>(in.draws.size(), in.draws.data());
    double *init = in.init.empty() ? nullptr : dev_alloc<double>(in.init.size(), in.init.data());
    double *signs = in.signs.empty() ? nullptr : dev_alloc<double>(in.signs.size(), in.signs.data());
    T *e_out = dev_alloc<T>(pl.out_elems);
    T *se_out = dev_alloc<T>(pl.out_elems);
    double *dens = dev_alloc<double>(pl.dens_elems);
    blml_result_t *res = dev_alloc<blml_result_t>(pl.nfits);
    double *mh = dev_alloc<double>((size_t)pl.nfits * nc_hist);
    double *rh = dev_alloc<double>((size_t)pl.nfits * nc_hist);
    const DevPlan<T> dp = dev_plan<T>(pl);
    const bool emp = (in.flags & 2) != 0;
    double *aux = emp ? dev_alloc<double>(in.aux.size(), in.aux.data()) : nullptr;
    double *emp_out = emp ? dev_alloc<double>((size_t)pl.nfits * BLML_EMP_OUT_LEN) : nullptr;
    for (int fi = 0; emp && fi < pl.nfits; ++fi) {
        if (pl.spread) {
            fprintf(stderr, "replay: blml_empirical runs batched plans only\n");
            return 7;
        }
        launch(pl.lds_bytes, kStaticFit, 1, [&] {
            run_empirical<T>(dp, fi, block_team(), in.ep, counts, draws, init, aux, e_out, se_out, dens, res, mh,
                             emp_out, 1e8);
        });
        printf("replay: empirical fit %d: %d linear-fit iterations, %d alternations, loglik %.12f\n", fi,
               res[fi].n_sweeps, res[fi].start_sweeps, res[fi].loglik);
        fflush(stdout);
    }
    for (int fi = 0; !emp && fi < pl.nfits; ++fi) {
        if (pl.spread) memset(bar, 0, 4 * sizeof(unsigned));      // launch_coop clears the barrier words
        const long long nb = pl.spread
            ? launch(pl.lds_fit, kStaticFit, pl.coop, [&] {
                  run_fit<T>(dp, 0, Team{(int)blockIdx.x, (int)gridDim.x, bar, part, 0}, counts, draws, init, signs,
                             e_out, se_out, dens, res, mh, rh, 1e8);
              })
            : launch(pl.lds_bytes, kStaticFit, 1, [&] {
                  run_fit<T>(dp, fi, block_team(), counts, draws, init, signs, e_out, se_out, dens, res, mh, rh, 1e8);
              });
        if (pl.spread && res[fi].reserved) {
            fprintf(stderr, "replay: a grid barrier timed out\n");
            return 6;
        }
        printf("replay: fit %d: %d sweeps, stop %d, loglik %.12f, %lld barriers\n", fi, res[fi].n_sweeps,
               res[fi].stop_reason, res[fi].loglik, nb);
        fflush(stdout);
    }
    if (sizeof(T) == 4 && !emp) {
        const DevPlan<double> d64 = dev_plan64(pl);
        for (int fi = 0; fi < pl.nfits; ++fi) {
            if (pl.spread) memset(bar, 0, 4 * sizeof(unsigned));
            if (pl.spread)
                launch(pl.lds64_bytes, 9728, pl.coop, [&] {
                    run_finish64(d64, 0, Team{(int)blockIdx.x, (int)gridDim.x, bar, part, 0}, (float *)se_out, counts,
                                 dens, res);
                });
            else
                launch(pl.lds64_bytes, 9728, 1,
                       [&] { run_finish64(d64, fi, block_team(), (float *)se_out, counts, dens, res); });
        }
    }
    std::vector<double> lam(pl.nfits, 0.0);
    std::vector<int32_t> its(pl.nfits, 0);
    if (in.flags & 1) {
        double *dl = dev_alloc<double>(pl.nfits);
        int *di = dev_alloc<int>(pl.nfits);
        const DevPlan<double> dc = sizeof(T) == 4 ? dev_plan64(pl) : dev_plan<double>(pl);
        const int lds = sizeof(T) == 4 ? pl.lds64_bytes : pl.lds_bytes;
        for (int fi = 0; fi < pl.nfits; ++fi) {
            if (pl.spread) {
                memset(bar, 0, 4 * sizeof(unsigned));
                launch(lds, 9728, pl.coop, [&] {
                    run_cert<T>(dc, 0, Team{(int)blockIdx.x, (int)gridDim.x, bar, part, 0}, e_out, counts,
                                in.cert_iters, in.cert_tol, dl, di);
                });
            } else {
                launch(lds, 9728, 1, [&] {
                    run_cert<T>(dc, fi, block_team(), e_out, counts, in.cert_iters, in.cert_tol, dl, di);
                });
            }
        }
        for (int fi = 0; fi < pl.nfits; ++fi) {
            lam[fi] = dl[fi];
            its[fi] = di[fi];
        }
        free(dl);
        free(di);
    }
    FILE *fo = fopen(out_path, "wb");
    if (!fo) return 5;
    fwrite(res, sizeof(blml_result_t), pl.nfits, fo);
    fwrite(e_out, sizeof(T), (size_t)pl.out_elems, fo);
    fwrite(dens, sizeof(double), (size_t)pl.dens_elems, fo);
    fwrite(lam.data(), sizeof(double), lam.size(), fo);
    fwrite(its.data(), sizeof(int32_t), its.size(), fo);
    fwrite(mh, sizeof(double), (size_t)pl.nfits * nc_hist, fo);
    fwrite(rh, sizeof(double), (size_t)pl.nfits * nc_hist, fo);
    if (emp) fwrite(emp_out, sizeof(double), (size_t)pl.nfits * BLML_EMP_OUT_LEN, fo);
    fclose(fo);
    for (const blml_fit_t &f : pl.hfits) unpoison(ws + f.ws_off, f.ws_len);
    free(pl.d_ws);
    printf("replay: ok\n");
    return 0;
}

}  // namespace

// --hio <in> <out>: the start search's runs (hio_run, hio_kernel's work for each run) on one workgroup each, every
// array an allocation of its exact size.  Input: int32 d, shape[5], box[5], n_hio, n_er, n_rand, n_kick; double
// beta, ab2; int64 nocc; uint8 band (ncell); double amp (ncell); int64 occ (nocc); double wocc (nocc), phase
// (n_rand ncell), kick (n_kick ncell).  Output: P (runs ncell complex), gocc (runs nocc complex), score (runs).
int replay_hio(const char *in_path, const char *out_path) {
    FILE *f = fopen(in_path, "rb");
    if (!f) return 2;
    int32_t hdr[15];
    double bd[2];
    int64_t nocc = 0;
    bool ok = rd(f, hdr, 15) && rd(f, bd, 2) && rd(f, &nocc, 1);
    HioArgs a{};
    a.d = hdr[0];
    long M = 1;
    int Lmax = 2;
    for (int j = 0; j < BLML_MAXD; ++j) {
        a.shape[j] = j < a.d ? hdr[1 + j] : 1;
        a.box[j] = j < a.d ? hdr[6 + j] : 1;
        if (j < a.d) {
            M *= a.shape[j];
            Lmax = std::max(Lmax, a.shape[j]);
        }
    }
    a.n_hio = hdr[11];
    a.n_er = hdr[12];
    a.n_rand = hdr[13];
    a.n_kick = hdr[14];
    a.beta = bd[0];
    a.ab2 = bd[1];
    a.ncell = M;
    a.nocc = nocc;
    const int runs = a.n_rand + a.n_kick;
    unsigned char *band = dev_alloc<unsigned char>(M);
    double *amp = dev_alloc<double>(M);
    long *occ = dev_alloc<long>(nocc);
    double *wocc = dev_alloc<double>(nocc);
    double *phase = dev_alloc<double>((size_t)a.n_rand * M);
    double *kick = dev_alloc<double>((size_t)a.n_kick * M);
    ok = ok && rd(f, band, M) && rd(f, amp, M) && rd(f, occ, nocc) && rd(f, wocc, nocc) &&
         rd(f, phase, (size_t)a.n_rand * M) && rd(f, kick, (size_t)a.n_kick * M);
    fclose(f);
    if (!ok) return 2;
    cpx<double> *P = dev_alloc<cpx<double>>((size_t)runs * M);
    cpx<double> *X = dev_alloc<cpx<double>>((size_t)runs * M);
    cpx<double> *g = dev_alloc<cpx<double>>((size_t)runs * nocc);
    double *score = dev_alloc<double>(runs);
    cpx<double> *tw = dev_alloc<cpx<double>>(Lmax);
    for (int i = 0; i < Lmax; ++i) {
        double sn, cs;
        sincospi(-2.0 * (double)i / (double)Lmax, &sn, &cs);
        tw[i] = {cs, sn};
    }
    a.band = band;
    a.amp = amp;
    a.occ = occ;
    a.wocc = wocc;
    a.phase = phase;
    a.kick = kick;
    a.P = P;
    a.X = X;
    a.gocc = g;
    a.score = score;
    a.tw = tw;
    a.Lmax = Lmax;
    long lds = 0;
    for (int j = 0; j < a.d; ++j) {
        const PassShape sh = pass_shape(OP_FFT_FULL, a.shape[j], a.shape[j], 0, M / a.shape[j]);
        lds = std::max(lds, (long)(sh.roots + (long)sh.lines * sh.buf) * (long)sizeof(cpx<double>));
    }
    for (int r = 0; r < runs; ++r)
        launch((int)lds, 8192 + 64, 1, [&] {
            double *red = cpuhip::lds_obj<double>(6, BLML_NWAVE);
            hio_run(a, r, red);
        });
    FILE *fo = fopen(out_path, "wb");
    if (!fo) return 5;
    fwrite(P, sizeof(cpx<double>), (size_t)runs * M, fo);
    fwrite(g, sizeof(cpx<double>), (size_t)runs * nocc, fo);
    fwrite(score, sizeof(double), runs, fo);
    fclose(fo);
    printf("replay: %d start-search runs on a grid of %ld cells\n", runs, M);
    printf("replay: ok\n");
    return 0;
}

int main(int argc, char **argv) {
    if (argc == 4 && strcmp(argv[1], "--hio") == 0) return replay_hio(argv[2], argv[3]);
    if (argc != 3) {
        fprintf(stderr, "usage: replay <plan.bin> <out.bin>\n");
        return 2;
    }
    In in;
    if (!read_in(argv[1], in)) {
        fprintf(stderr, "replay: cannot read %s\n", argv[1]);
        return 2;
    }
    return in.dtype == BLML_F32 ? replay<float>(in, argv[2]) : replay<double>(in, argv[2]);
}
