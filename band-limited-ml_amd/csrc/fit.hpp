// The device side of one fit (shared by the engine's kernels and by the CPU build of the tests): the plan as the
// kernels see it, the workspace layout, the spectra of the factors, the mat-vec context, and the whole fit, the
// certificate and the FP64 finishing pass, each run by a team (one workgroup, or every workgroup of a cooperative
// launch).
#pragma once
#include <math.h>

#include "blml_hip.h"
#include "solver.hpp"

template <typename T>
struct DevPlan {
    const blml_fit_t *fits;
    const blml_factor_t *facs;
    const double *fdata;
    const int *freqs;
    T *spec;
    const cpx<T> *tw;
    int Lmax;
    unsigned char *ws;     // per-fit workspaces, byte offsets fit.ws_off
    int nfits;
    blml_params_t params;
    const SepPlan *sep;    // spread plans of separable modes: [0] for the mat-vec kernel, [1] for the fit kernel
    const T *sep_tab;
    const T *sep_tab_c;    // FP32 plans: the centrosymmetric split of every table (sep_centro_table), or null
    const long *sep_off;
    int res_off, res_cap;  // LDS-resident intermediates: byte offset in the dynamic LDS and its size (0: none)
    int fft_sched;         // the radix-stage order of the FFT passes (the plan's autotuned choice)
    int toep_max;          // padded line factors up to this length are summed directly (OP_TOEP; 0: none)
};

// A fit's workspace (bytes from ws_off): [X, Y: scratch_len complex values of the transform type]
// [pool: n_slots * Vmax FP64 state vectors][FP32 plans: X64, Y64, the finishing pass's FP64 scratch].
__host__ __device__ inline long align256(long v) { return (v + 255) / 256 * 256; }
__host__ __device__ inline int n_slots(const blml_params_t &P, const blml_fit_t &f) {
    const int cd = 13 + 2 * (P.anderson + 1);
    const int lb = 9 + 2 * (P.lbfgs_m + 1);
    const bool psd = f.complex_g || f.K > 1;
    return (P.solver == 1 && psd) ? (cd > lb ? cd : lb) : cd;
}
__host__ __device__ inline long pool_bytes_off(const blml_fit_t &f, int esize) { return align256(4L * f.scratch_len * esize); }
__host__ __device__ inline long vmax_of(const blml_fit_t &f) { return (long)f.K * f.B * (f.complex_g ? 2 : 1); }
// LDS bytes a fit's transform intermediates need to stay in LDS: X (scratch_len complex values), and Y as well when
// a mode has a multi-axis (GROUP) factor
__host__ __device__ inline long res_need(const blml_fit_t &f, const blml_factor_t *facs, int esize) {
    bool grp = false;
    for (int k = 0; k < f.K; ++k)
        for (int q = 0; q < f.fac_count[k]; ++q) grp = grp || facs[f.fac_begin[k] + q].kind != BLML_FAC_LINE;
    return 2L * esize * f.scratch_len * (grp ? 2 : 1);
}
__host__ __device__ inline long fin_bytes_off(const blml_params_t &P, const blml_fit_t &f, int esize) {
    return align256(pool_bytes_off(f, esize) + 8L * n_slots(P, f) * vmax_of(f));
}

// Spectrum of one factor.  LINE padded: Kf[q] = (t0 + 2 sum_{o=1}^{m-1} t_o cos(2 pi q o / L)) / L, the DFT of the
// even, zero-padded row, in FP64 with exact integer phase reduction.  Circular factors: the given values.
// GROUP padded: the forward FFT of the table over its power-of-two lattice (real, the table being even), times
// prod 1/L.
template <typename T>
__device__ void spec_factor(const DevPlan<T> &p, int fi, const int *fac_m, cpx<T> *tmp, const long *tmp_off) {
    const blml_factor_t fa = p.facs[fi];
    T *spec = p.spec + fa.spec_off;
    const double *data = p.fdata + fa.data_off;
    if (fa.kind == BLML_FAC_LINE && !fa.circ[0]) {
        const int m = fac_m[fi * BLML_MAXD], L = fa.tlen[0];
        for (int q = threadIdx.x; q < L; q += BLML_NT) {
            double acc = data[0];
            long ph = 0;
            for (int o = 1; o < m; ++o) {
                ph += q;
                if (ph >= L) ph -= L;
                acc += 2.0 * data[o] * cospi(2.0 * (double)ph / (double)L);
            }
            spec[q] = (T)(acc / (double)L);
        }
        return;
    }
    long total = 1;
    for (int a = 0; a < fa.naxes; ++a) total *= fa.tlen[a];
    bool allcirc = true;
    for (int a = 0; a < fa.naxes; ++a) allcirc = allcirc && fa.circ[a];
    if (allcirc) {
        for (long i = threadIdx.x; i < total; i += BLML_NT) spec[i] = (T)data[i];
        return;
    }
    // padded group: FFT of the table along each of its axes (the table's own lattice: ext = tlen)
    cpx<T> *X = tmp + tmp_off[fi], *Y = X + total;
    for (long i = threadIdx.x; i < total; i += BLML_NT) X[i] = {(T)data[i], (T)0};
    __syncthreads();
    int ext[BLML_MAXD];
    for (int a = 0; a < fa.naxes; ++a) ext[a] = fa.tlen[a];
    cpx<T> *cur = X, *oth = Y;
    for (int a = fa.naxes - 1; a >= 0; --a) {
        Pass<T> ps;
        ps.op = OP_FFT_FULL;
        ps.L = fa.tlen[a];
        ps.m = ps.L;
        ps.nF = 0;
        ps.F = nullptr;
        ps.spec = nullptr;
        ps.n_in = ps.n_out = ps.L;
        ps.pairs = false;
        ps.rank = 0;
        ps.nranks = 1;
        ps.sched = p.fft_sched;
        line_geom<T>(ext, fa.naxes, a, ps.outer, ps.inner);
        run_pass<T>(ps, cur, oth, p.tw, p.Lmax);
        cpx<T> *t = cur; cur = oth; oth = t;
    }
    double inv = 1.0;
    for (int a = 0; a < fa.naxes; ++a) inv /= (double)fa.tlen[a];
    for (long i = threadIdx.x; i < total; i += BLML_NT) spec[i] = (T)((double)cur[i].x * inv);
}

template <typename T>
__device__ void make_matvec_ctx(const DevPlan<T> &p, const blml_fit_t &f, MatvecCtx<T> &mv, const Team &tm,
                                const SepPlan *sep = nullptr) {
    mv.tm = tm;
    mv.sep = sep;
    mv.sep_last_sync = true;
    mv.sep_tab = p.sep_tab;
    mv.sep_off = p.sep_off;
    mv.d = f.d;
    for (int j = 0; j < BLML_MAXD; ++j) mv.shape[j] = j < f.d ? f.shape[j] : 1;
    mv.facs = p.facs;
    mv.spec = p.spec;
    mv.freqs = p.freqs;
    mv.fdata = p.fdata;
    mv.toep_max = p.toep_max;
    mv.tw = p.tw;
    mv.Lmax = p.Lmax;
    mv.X = (cpx<T> *)(p.ws + f.ws_off);
    mv.Y = mv.X + f.scratch_len;
    mv.fft_sched = p.fft_sched;
    // a fit run by one workgroup whose intermediates fit the launch's resident LDS keeps them there: every pass
    // then stages its lines LDS to LDS, and only E is read and U written in HBM
    if (p.res_cap > 0 && !sep && tm.nranks == 1 && res_need(f, p.facs, (int)sizeof(T)) <= p.res_cap) {
        mv.X = reinterpret_cast<cpx<T> *>(g_dyn_lds + p.res_off);
        mv.Y = mv.X + f.scratch_len;
    }
}

template <typename T>
__device__ void run_matvec(const DevPlan<T> &p, int fi, const Team &tm, const T *E, T *U, const SepPlan *sep) {
    const blml_fit_t f = p.fits[fi];
    MatvecCtx<T> mv;
    make_matvec_ctx<T>(p, f, mv, tm, sep);
    mv.sep_last_sync = f.K > 1;      // between modes only: the kernel's end orders the last one
    const int nc = f.complex_g ? 2 : 1;
    for (int k = 0; k < f.K; ++k) {
        const long off = f.out_off + (long)k * f.B * nc;
        matvec_mode<T, T, T>(mv, f.fac_begin[k], f.fac_count[k], E + off, U + off, f.B, nc);
    }
}

// One whole fit (start, solve, outputs), run by a team.
// The stand-alone mat-vec of a separable spread plan whose every pass is a pipelined whole-tile pass (the host
// checks it), one pass at a time: pass q of mode k over the tiles rank, rank + nranks, ...  Each pass stages every
// axis's table once (after the largest tile of any pass; the host sized the LDS), while its first tile loads.  FP32
// plans with the centrosymmetric tables run every axis as sep_axis_centro.
template <typename T, bool CENTRO>
__device__ void run_matvec_pipe_pass(const DevPlan<T> &p, int rank, int nranks, int k, int q, const T *E, T *U) {
    const blml_fit_t *fp = p.fits;                     // read in place: a local copy indexed by k would be scratch
    const int nc = fp->complex_g ? 2 : 1;
    const long B = fp->B;
    const SepPlan *sp = p.sep;
    const T *tabs = CENTRO ? p.sep_tab_c : p.sep_tab;
    T *Z = reinterpret_cast<T *>(p.ws + fp->ws_off);
    int shape[BLML_MAXD];
#pragma unroll
    for (int a = 0; a < BLML_MAXD; ++a) shape[a] = fp->shape[a];
    const long off = fp->out_off + (long)k * B * nc;
    const int fb = fp->fac_begin[k], fcnt = fp->fac_count[k];
    int fac_of[BLML_MAXD];                              // indexed by constants only
#pragma unroll
    for (int a = 0; a < BLML_MAXD; ++a) {
        int fa = -1;
        for (int j = 0; j < fcnt; ++j)
            if (p.facs[fb + j].axes[0] == a) fa = fb + j;
        fac_of[a] = fa;
    }
    const int np = sp->npass;
    int o = 0;
    for (int r = 0; r < np; ++r) {
        const int t = sp->pass[r].OB * sp->pass[r].R * sp->pass[r].w;
        o = max(o, sep_lds(t) + 1 + 64);
    }
    constexpr int BLK = SepMfma<T>::BLK;
    int aoff[BLML_MAXD];
#pragma unroll
    for (int a = 0; a < BLML_MAXD; ++a) {
        aoff[a] = -1;
        if (fac_of[a] >= 0 && shape[a] > 1) {
            aoff[a] = o;
            o += BLK * (BLK + 1);
        }
    }
    constexpr int NV = sizeof(T) == 4 ? 64 : 32;          // 128 KB tiles
    const SepPass ps = sp->pass[q];
    const T *src = q == 0 ? E + off : Z;
    T *dst = q == np - 1 ? U + off : Z;
    sep_pass_pipe<T, NV, CENTRO>(ps, shape, fac_of, aoff, src, dst, rank, nranks,
                                 SepStageAll<T>{fac_of, tabs, p.sep_off, aoff});
}

// The same mat-vec by one team: every pass of every mode in order, the team synchronised between them.
template <typename T, bool CENTRO>
__device__ void run_matvec_pipe(const DevPlan<T> &p, const Team &tm, const T *E, T *U) {
    const int K = p.fits->K, np = p.sep->npass;
    for (int k = 0; k < K; ++k)
        for (int q = 0; q < np; ++q) {
            run_matvec_pipe_pass<T, CENTRO>(p, tm.rank, tm.nranks, k, q, E, U);
            if (q < np - 1 || k < K - 1) tm.sync();
        }
}

template <typename T>
__device__ void run_fit(const DevPlan<T> &p, int fi, const Team &team, const double *counts, const double *draws,
                        const double *init, const double *signs, T *e_out, T *se_out, double *density,
                        blml_result_t *results, double *merit_hist, double *rho_hist, double tick_hz) {
    const blml_fit_t f = p.fits[fi];
    FitSolver<T> S;
    S.tm = team;
    make_matvec_ctx<T>(p, f, S.mv, team, team.nranks > 1 && p.sep ? p.sep + 1 : nullptr);
    S.f = &p.fits[fi];
    S.P = &p.params;
    S.B = f.B;
    S.n = f.n;
    for (int k = 0; k < BLML_MAXK; ++k) S.s[k] = f.s[k];
    S.equal_s = f.equal_s != 0;
    S.w = counts + f.counts_off;
    S.sg = f.has_signs ? signs + f.signs_off : nullptr;
    const int ncf = f.complex_g ? 2 : 1;
    S.Vmax = (long)f.K * f.B * ncf;
    S.pool = (double *)(p.ws + f.ws_off + pool_bytes_off(f, (int)sizeof(T)));
    S.inc = sizeof(T) == 4;
    S.gstep = 0.0;
    S.t_mv = S.t_dir = S.t_ls = 0;
    S.n_mv = 0;
    const long long tfit0 = wall_clock64();
    long long tstart = 0;
    S.merit_hist = merit_hist ? merit_hist + f.hist_off : nullptr;
    S.rho_hist = rho_hist ? rho_hist + f.hist_off : nullptr;
    S.hist_cap = p.params.hist_cap;
    S.nhist_m = S.nhist_r = 0;
    S.ring_n = 0;
    const bool psd_full = f.complex_g || f.K > 1;
    const double pn = p.params.phase_noise;
    const double *dr = draws + f.draws_off;
    int start_sweeps = 0;
    double *E = S.vec(0);
    if (f.init_K > 0) {
        // warm start (blml_binned.py:403-408): given modes, zero off the occupied bins, noise in the others
        const double *E0 = init + f.init_off;
        const int kk = f.init_K < f.K ? f.init_K : f.K;
        for (int b = (int)S.tm.first(); b < f.B; b += (int)S.tm.stride()) {
            const bool o = S.w[b] > 0;
            for (int k = 0; k < f.K; ++k) {
                const long i = ((long)k * f.B + b) * ncf;
                double re = 0.0, im = 0.0;
                if (k < kk && o) {
                    re = E0[((long)k * f.B + b) * 2];
                    im = f.complex_g ? E0[((long)k * f.B + b) * 2 + 1] : 0.0;
                }
                E[i] = re;
                if (ncf == 2) E[i + 1] = im;
            }
        }
        S.tm.sync();
        long dp = 0;
        for (int k = kk; k < f.K; ++k) {
            for (int b = (int)S.tm.first(); b < f.B; b += (int)S.tm.stride()) {
                const double a0 = ncf == 2 ? sqrt((double)E[2 * b] * E[2 * b] + (double)E[2 * b + 1] * E[2 * b + 1])
                                           : fabs((double)E[b]);
                const long i = ((long)k * f.B + b) * ncf;
                if (ncf == 2) {
                    E[i] = (pn * a0 * (dr[dp + b] / sqrt(2.0)));
                    E[i + 1] = (pn * a0 * (dr[dp + f.B + b] / sqrt(2.0)));
                } else {
                    E[i] = (pn * a0 * dr[dp + b]);
                }
            }
            dp += (long)f.B * ncf;
        }
        S.tm.sync();
    } else {
        // cold start: e0 = sqrt(c / s0) on occupied bins (times the signs); with init_K < 0 the caller's real
        // rank-1 vector stands in for e0 -- and, in a psd fit with init_K = -1, for the start solve below, which
        // with init_K = -2 starts from it instead (the same convex solve to the same tolerance, from nearer)
        const double *er0 = f.init_K < 0 ? init + f.init_off : nullptr;
        for (int b = (int)S.tm.first(); b < f.B; b += (int)S.tm.stride()) {
            const double sgb = S.sg ? S.sg[b] : 1.0;
            E[b] = S.w[b] > 0 ? (er0 ? er0[2 * b] : sqrt(f.n * S.w[b] / f.s[0]) * sgb) : 0.0;
        }
        S.tm.sync();
        if (psd_full) {
            // the real rank-1 solution under mode 0 (blml_binned.py:421-425)
            S.K = 1; S.nc = 1; S.psd = false; S.rec = false;
            if (!er0 || f.init_K == -2) {
                const FitOut so = S.solve_cd();
                start_sweeps = so.n_sweeps;
            }
            tstart = wall_clock64() - tfit0;
            // phases and the other modes (blml_binned.py:429-436), assembled in slot 2 then moved to slot 0
            double *E2 = S.vec(2);
            const double *er = S.vec(0);
            long dp = 0;
            for (int b = (int)S.tm.first(); b < f.B; b += (int)S.tm.stride()) {
                if (f.complex_g) {
                    double sn, cs;
                    sincos(pn * dr[b], &sn, &cs);
                    const double r = er[b];
                    E2[2 * b] = (r * cs);
                    E2[2 * b + 1] = (r * sn);
                } else {
                    E2[b] = er[b];
                }
            }
            if (f.complex_g) dp = f.B;
            S.tm.sync();
            for (int k = 1; k < f.K; ++k) {
                for (int b = (int)S.tm.first(); b < f.B; b += (int)S.tm.stride()) {
                    const double a0 = ncf == 2
                        ? sqrt((double)E2[2 * b] * E2[2 * b] + (double)E2[2 * b + 1] * E2[2 * b + 1])
                        : fabs((double)E2[b]);
                    const long i = ((long)k * f.B + b) * ncf;
                    if (ncf == 2) {
                        E2[i] = (pn * a0 * (dr[dp + b] / sqrt(2.0)));
                        E2[i + 1] = (pn * a0 * (dr[dp + f.B + b] / sqrt(2.0)));
                    } else {
                        E2[i] = (pn * a0 * dr[dp + b]);
                    }
                }
                dp += (long)f.B * ncf;
                S.tm.sync();
            }
            S.K = f.K; S.nc = ncf; S.psd = true;
            S.copy(S.vec(0), E2);
        }
    }
    S.K = f.K;
    S.nc = ncf;
    S.psd = psd_full;
    S.rec = true;
    S.nhist_m = S.nhist_r = 0;
    FitOut o;
    if (p.params.solver == 1 && psd_full)
        o = (S.inc && p.params.lbfgs_m <= BLML_CMAX) ? S.solve_lbfgs_compact() : S.solve_lbfgs();
    else
        o = S.solve_cd();
    // outputs (blml_binned.py:438-445)
    const double *Ef = S.vec(0), *Uf = S.vec(1);
    const long V = (long)f.K * f.B * ncf;
    for (long i = S.tm.first(); i < V; i += S.tm.stride()) {
        e_out[f.out_off + i] = (T)Ef[i];
        se_out[f.out_off + i] = (T)Uf[i];
    }
    double ll = 0.0;
    const double n2 = f.n * f.n;
    for (int b = (int)S.tm.first(); b < f.B; b += (int)S.tm.stride()) {
        const double dens = S.Fb(Uf, b) / n2;
        density[f.dens_off + b] = dens;
        if (S.w[b] > 0) ll += S.w[b] * log(dens);
    }
    ll = S.tm.sum1(ll, g_sl.red);
    const unsigned timed_out = S.tm.nranks > 1 ? __hip_atomic_load(&S.tm.bar[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0u;
    if (S.tm.lead()) {
        blml_result_t r;
        r.loglik = ll / f.n;
        r.residual = o.residual;
        r.n_sweeps = o.n_sweeps;
        r.stop_reason = o.stop;
        r.converged = o.converged;
        r.n_anderson = o.n_and;
        r.n_damped = o.n_damp;
        r.n_gradient = o.n_grad;
        r.start_sweeps = start_sweeps;
        r.n_hist = S.nhist_m;
        r.n_matvec = S.n_mv;
        r.reserved = (int)timed_out;   // a grid barrier timed out: the results are void
        r.seconds[0] = (double)(wall_clock64() - tfit0) / tick_hz;
        r.seconds[1] = (double)tstart / tick_hz;
        r.seconds[2] = (double)S.t_mv / tick_hz;
        r.seconds[3] = (double)S.t_dir / tick_hz;
        r.seconds[4] = (double)S.t_ls / tick_hz;
        results[fi] = r;
    }
    S.tm.sync();
}

// certificate(): lambda_max(D^1/2 S_0 D^1/2) / n, D_b = |e_b|^2 / w_b on occupied bins, by power iteration with FP64
// transforms and an FP64 Rayleigh quotient.  It stops when the eigen-residual of the unit iterate x and its quotient
// lam is small, ||A x - lam x|| <= tol |lam| (ARPACK's convergence test for a Ritz pair, the test eigsh applies at
// blml_binned.py:510-518; the quotient's error is then at most tol^2 lam^2 / gap), or after max_iter steps.  Te: the
// plan's coefficient type; an FP32 plan runs through its FP64 view, with the finishing pass's FP64 scratch (xoff
// bytes into the fit's workspace).  iters: the steps taken, negated when max_iter ended the iteration.
template <typename Te>
__device__ void run_cert(const DevPlan<double> &p, int fi, Team tm, const Te *e_hat, const double *counts,
                         int max_iter, double tol, double *lam_out, int *iters) {
    __shared__ double red[BLML_NWAVE * BLML_NRED];
    const blml_fit_t f = p.fits[fi];
    if (!f.shared) {
        if (tm.lead()) { lam_out[fi] = NAN; iters[fi] = 0; }
        return;
    }
    MatvecCtx<double> mv;
    make_matvec_ctx<double>(p, f, mv, tm);
    if (sizeof(Te) == 4) {
        mv.X = (cpx<double> *)(p.ws + f.ws_off + fin_bytes_off(p.params, f, 4));
        mv.Y = mv.X + f.scratch_len;
    }
    const int nc = f.complex_g ? 2 : 1;
    const int B = f.B;
    const long b0 = tm.first(), bs = tm.stride();
    double *dh = (double *)(p.ws + f.ws_off + pool_bytes_off(f, (int)sizeof(Te)));   // B doubles
    double *x = dh + B, *y = x + B;
    const double *w = counts + f.counts_off;
    const Te *E = e_hat + f.out_off;
    for (long b = b0; b < B; b += bs) {
        double D = 0.0;
        if (w[b] > 0) {
            for (int k = 0; k < f.K; ++k) {
                const long i = ((long)k * B + b) * nc;
                D += (double)E[i] * E[i] + (nc == 2 ? (double)E[i + 1] * E[i + 1] : 0.0);
            }
            D /= w[b];
        }
        dh[b] = sqrt(D);
    }
    tm.sync();
    double nrm = 0.0;
    for (long b = b0; b < B; b += bs) nrm += dh[b] * dh[b];
    nrm = sqrt(tm.sum1(nrm, red));
    if (!(nrm > 0)) {
        if (tm.lead()) { lam_out[fi] = 0.0; iters[fi] = 0; }
        return;
    }
    for (long b = b0; b < B; b += bs) x[b] = dh[b] / nrm;
    tm.sync();
    double lam = 0.0;
    int it = 0;
    bool conv = false;
    for (it = 1; it <= max_iter; ++it) {
        for (long b = b0; b < B; b += bs) y[b] = dh[b] * x[b];
        tm.sync();
        matvec_mode<double, double, double>(mv, f.fac_begin[0], f.fac_count[0], y, y, B, 1);
        double a[3] = {0.0, 0.0, 0.0};
        for (long b = b0; b < B; b += bs) {
            const double yb = dh[b] * y[b];
            a[0] += x[b] * yb;
            a[1] += yb * yb;
            a[2] += x[b] * x[b];
        }
        tm.sum<3>(a, red);
        lam = a[0] / a[2];
        const double ny = sqrt(a[1]);
        if (!(ny > 0)) break;
        // the residual of (x, lam), summed in the pass that forms the next iterate A x / ||A x|| (each thread reads
        // and writes its own bins only)
        double r2 = 0.0;
        for (long b = b0; b < B; b += bs) {
            const double yb = dh[b] * y[b];
            const double v = yb - lam * x[b];
            r2 += v * v;
            x[b] = yb / ny;
        }
        const double r = sqrt(tm.sum1(r2, red) / a[2]);
        if (r <= tol * fabs(lam)) { conv = true; break; }
    }
    if (tm.lead()) {
        lam_out[fi] = lam / f.n;
        iters[fi] = conv ? it : -(it > max_iter ? max_iter : it);
    }
}

// FP32 plans only: Se, density and loglik of every fit re-evaluated with an FP64 mat-vec of the FP32-converged
// coefficients (the iterations run in FP32; the outputs' likelihood sum is FP64 end to end).  p carries the FP64
// spectra and twiddles; its workspace is the FP32 plan's, reinterpreted (sized for it by layout()).
__device__ void run_finish64(const DevPlan<double> &p, int fi, Team tm, float *se_out, const double *counts,
                             double *density, blml_result_t *results) {
    __shared__ double red[BLML_NWAVE * BLML_NRED];
    const blml_fit_t f = p.fits[fi];
    MatvecCtx<double> mv;
    mv.tm = tm;
    mv.sep = nullptr;
    mv.sep_last_sync = true;
    mv.d = f.d;
    for (int j = 0; j < BLML_MAXD; ++j) mv.shape[j] = j < f.d ? f.shape[j] : 1;
    mv.facs = p.facs;
    mv.spec = p.spec;
    mv.freqs = p.freqs;
    mv.tw = p.tw;
    mv.Lmax = p.Lmax;
    mv.X = (cpx<double> *)(p.ws + f.ws_off + fin_bytes_off(p.params, f, 4));
    mv.Y = mv.X + f.scratch_len;
    mv.fft_sched = p.fft_sched;
    const int nc = f.complex_g ? 2 : 1;
    const long V = (long)f.K * f.B * nc;
    // the fit's FP64 state: E in pool slot 0; U64 overwrites slot 1
    const double *E64 = (const double *)(p.ws + f.ws_off + pool_bytes_off(f, 4));
    double *U64 = (double *)E64 + vmax_of(f);
    for (int k = 0; k < f.K; ++k) {
        const long off = (long)k * f.B * nc;
        matvec_mode<double, double, double>(mv, f.fac_begin[k], f.fac_count[k], E64 + off, U64 + off, f.B, nc);
    }
    for (long i = tm.first(); i < V; i += tm.stride()) se_out[f.out_off + i] = (float)U64[i];
    const double *w = counts + f.counts_off;
    const double n2 = f.n * f.n;
    double ll = 0.0;
    for (long b = tm.first(); b < f.B; b += tm.stride()) {
        double F = 0.0;
        for (int k = 0; k < f.K; ++k) {
            const long i = ((long)k * f.B + b) * nc;
            F += U64[i] * U64[i] + (nc == 2 ? U64[i + 1] * U64[i + 1] : 0.0);
        }
        const double dens = F / n2;
        density[f.dens_off + b] = dens;
        if (w[b] > 0) ll += w[b] * log(dens);
    }
    ll = tm.sum1(ll, red);
    if (tm.lead()) results[fi].loglik = ll / f.n;
}

