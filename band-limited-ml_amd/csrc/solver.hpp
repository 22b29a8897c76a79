// One binned fit, solved start to finish by one workgroup: the start (blml_binned.py:415-436), Jacobi coordinate
// descent with Anderson mixing, damping and gradient fallback (blml_binned.py:223-276, 357-380, 1463-1478), or
// L-BFGS with Armijo backtracking and the canonical form (blml_binned.py:278-355).
//
// State vectors (E, U, gradients, histories) are FP64; the transforms of the mat-vec run in T (FP32 or FP64) and
// every reduction accumulates in FP64.  Two mat-vec modes:
//   exact (FP64 plans): U_new = S E_new recomputed from scratch wherever the reference does, so the iteration is the
//         reference's step for step (up to FFT rounding);
//   incremental (FP32 plans): U_new = U + S (E_new - E), and across line-search trials U + t S q with one S q per
//         iteration.  The FP32 rounding then scales with the increment, not with U, so the likelihoods the solver
//         compares stay consistent to FP64 accuracy (a fresh FP32 S E carries an absolute error ~1e-7 |U|, which puts
//         noise of ~n 1e-7 into L = sum w log F - n log Q and defeats the Armijo and merit tests).
// The scalar control flow is executed identically by every thread: it depends only on reduction results that all
// threads hold bit-identically.
#pragma once
#include <limits.h>
#include <math.h>
#include "transform.hpp"

#define BLML_MAXHIST 64      // L-BFGS memory limit
#define BLML_MAXAND 5        // Anderson history limit
#define BLML_RING 512        // merit window ring (ll_window + 1 <= BLML_RING)
#define BLML_CMAX 24         // L-BFGS memory of the compact form (incremental mode); larger memories use two-loop
#define BLML_CCHUNK 8        // history pairs per sub-pass of the compact form's dot products
#define BLML_HARD_CAP 1000000L  // sweeps / iterations of one solve when max_sweeps is None (every launch ends)

struct SolverLds {
    double red[BLML_NWAVE * BLML_NRED];
    double ring[BLML_RING];
    double lb_r[BLML_MAXHIST + 1], lb_a[BLML_MAXHIST + 1], lb_sy[BLML_MAXHIST + 1], lb_yy[BLML_MAXHIST + 1];
    // compact L-BFGS: R = triu(S^T Y), Y^T Y (logical order, oldest first), S^T g, Y^T g, and the new dots
    double cR[BLML_CMAX * BLML_CMAX], cYY[BLML_CMAX * BLML_CMAX];
    double ca[BLML_CMAX + 1], cb[BLML_CMAX + 1], csg[BLML_CMAX + 1], cyg[BLML_CMAX + 1];
    double cp1[BLML_CMAX], cp2[BLML_CMAX];
    double cu[BLML_CMAX], crc[BLML_CMAX + 1], cyc[BLML_CMAX + 1];   // thread 0's work arrays for the small systems
};

// The solver's LDS (reduction scratch, merit ring, L-BFGS scalars): a file-scope __shared__ object, so every access
// compiles to ds_ instructions (see dyn_lds).
__shared__ SolverLds g_sl;

struct FitOut {
    int n_sweeps, stop, converged, n_and, n_damp, n_grad;
    double residual;
};

__device__ inline double cd_root(double b, double c, double s) {
    const double disc = sqrt(b * b + 4.0 * s * c);
    return b >= 0 ? 2.0 * c / fmax(b + disc, 1e-300) : (-b + disc) / (2.0 * s);
}

// The root lam > max s of lam * sum_k R2_k / (lam - s_k)^2 = c (blml_binned.py:1556-1573), 16 safeguarded
// Newton steps on phi^-1/2.
__device__ inline double secular_root(const double *R2, const double *s, int K, double c) {
    double smax = s[0], tot = 0.0;
    for (int k = 0; k < K; ++k) { smax = fmax(smax, s[k]); tot += R2[k]; }
    const double cpos = fmax(c, 1e-30);
    double hi = (tot + sqrt(tot * tot + 4 * cpos * smax * tot)) / (2 * cpos);
    double lo = 0.0, mu = hi;
    for (int it = 0; it < 16; ++it) {
        double A = 0.0, Bv = 0.0;
        for (int k = 0; k < K; ++k) {
            const double q = 1.0 / (mu - (s[k] - smax));
            const double W = R2[k] * (q * q);
            A += W;
            Bv += W * q;
        }
        const double lam = smax + mu, phi = lam * A;
        const bool big = phi > c;
        lo = big ? mu : lo;
        hi = big ? hi : mu;
        const double nt = mu + 2 * phi * (1 - sqrt(phi / cpos)) / (A - 2 * lam * Bv);
        mu = (nt >= lo && nt <= hi) ? nt : 0.5 * (lo + hi);
    }
    return smax + mu;
}

template <typename T>
struct FitSolver {
    Team tm;               // who runs the fit: one workgroup, or every workgroup of a cooperative launch
    MatvecCtx<T> mv;
    const blml_fit_t *f;
    const blml_params_t *P;
    int B, K, nc;          // active: the real start runs K = 1, nc = 1
    bool psd;
    bool inc;              // incremental mat-vecs (FP32 transforms)
    double n;
    double s[BLML_MAXK];
    bool equal_s;
    const double *w;
    const double *sg;      // signs or null
    double *pool;
    long Vmax;
    double gstep;          // the gradient fallback's step, kept between calls (0: unset)
    double *merit_hist, *rho_hist;
    int hist_cap, nhist_m, nhist_r;
    int ring_n;            // entries pushed to the merit ring
    bool rec;              // record merit / rho histories (the final solve only, as merit_history_ is)
    long long t_mv, t_dir, t_ls;   // wall-clock ticks in mat-vecs, L-BFGS direction, line search
    int n_mv;

    __device__ double *vec(int i) const { return pool + (long)i * Vmax; }
    // the compact L-BFGS form's history pairs, stored in FP32 in their FP64 slots (FP32 plans only): half the
    // bytes of the two passes over the history per iteration; s.y and y.y are formed in FP64 before the rounding
    __device__ float *hvec(int i) const { return reinterpret_cast<float *>(pool + (long)i * Vmax); }
    __device__ long V() const { return (long)K * B * nc; }

    // ------------------------------------------------------------------ mat-vecs
    __device__ void Sall(const double *E, double *U) {
        const long long t0 = wall_clock64();
        for (int k = 0; k < K; ++k) {
            const long o = (long)k * B * nc;
            matvec_mode<T, double, double>(mv, f->fac_begin[k], f->fac_count[k], E + o, U + o, B, nc);
        }
        t_mv += wall_clock64() - t0;
        n_mv += K;
    }
    // U_new = S E_new: recomputed (exact mode) or U + S (E_new - E) (incremental mode)
    __device__ void Snew(const double *En, double *Un, const double *E, const double *U) {
        if (!inc) { Sall(En, Un); return; }
        const long long t0 = wall_clock64();
        for (int k = 0; k < K; ++k) {
            const long o = (long)k * B * nc;
            matvec_mode<T, double, double>(mv, f->fac_begin[k], f->fac_count[k], En + o, Un + o, B, nc, E + o,
                                           U + o);
        }
        t_mv += wall_clock64() - t0;
        n_mv += K;
    }

    // ------------------------------------------------------------------ basic vector operations
    __device__ double dot(const double *a, const double *b) {
        double acc = 0.0;
        const long n_ = V();
        for (long i = tm.first(); i < n_; i += tm.stride()) acc += a[i] * b[i];
        return tm.sum1(acc, g_sl.red);
    }
    __device__ void copy(double *dst, const double *src) {
        const long n_ = V();
        for (long i = tm.first(); i < n_; i += tm.stride()) dst[i] = src[i];
        tm.sync();
    }
    __device__ void scale(double *a, double sc) {
        const long n_ = V();
        for (long i = tm.first(); i < n_; i += tm.stride()) a[i] = a[i] * sc;
        tm.sync();
    }
    // out = a + t * b
    __device__ void axpy(double *out, const double *a, double t, const double *b) {
        const long n_ = V();
        for (long i = tm.first(); i < n_; i += tm.stride()) out[i] = a[i] + t * b[i];
        tm.sync();
    }
    __device__ bool occ(int b) const { return w[b] > 0; }
    __device__ double cb(int b) const { return n * w[b]; }
    __device__ double absq(const double *X, int k, int b) const {
        const long i = ((long)k * B + b) * nc;
        return nc == 2 ? X[i] * X[i] + X[i + 1] * X[i + 1] : X[i] * X[i];
    }
    __device__ double Fb(const double *U, int b) const {
        double F = 0.0;
        for (int k = 0; k < K; ++k) F += absq(U, k, b);
        return F;
    }

    // ------------------------------------------------------------------ merit, update, rescale, residual
    // Qout (real rank one): sum e u of the pair, the same value dot(E, U) gives (same products, order and reduction)
    __device__ double merit(const double *E, const double *U, double *Qout = nullptr) {
        if (psd) {
            double a[2] = {0.0, 0.0};
            for (int b = (int)tm.first(); b < B; b += (int)tm.stride()) a[0] += cb(b) * log(fmax(Fb(U, b), 1e-30));
            const long n_ = V();
            for (long i = tm.first(); i < n_; i += tm.stride()) a[1] += E[i] * U[i];
            tm.sum<2>(a, g_sl.red);
            const double Q = a[1];
            if (!(Q > 0)) return -INFINITY;
            return a[0] / n - n * log(Q);
        }
        double a[3] = {0.0, 0.0, 0.0};
        for (int b = (int)tm.first(); b < B; b += (int)tm.stride()) {
            const double e = E[b], u = U[b];
            const double sgb = sg ? sg[b] : 1.0;
            a[0] += e * u;
            if (occ(b)) {
                a[1] += cb(b) * log(fabs(e));
                if (sgb * e <= 0) a[2] += 1.0;
            }
        }
        tm.sum<3>(a, g_sl.red);
        if (Qout) *Qout = a[0];
        if (a[2] > 0) return -INFINITY;
        return -(0.5 * a[0] - a[1]) / n;
    }

    // Real rank one (psd false: K = nc = 1, so V() = B): the merit terms of one bin, as merit() sums them.
    __device__ __forceinline__ void merit_terms(int b, double e, double u, double (&a)[3]) const {
        const double sgb = sg ? sg[b] : 1.0;
        a[0] += e * u;
        if (occ(b)) {
            a[1] += cb(b) * log(fabs(e));
            if (sgb * e <= 0) a[2] += 1.0;
        }
    }
    __device__ double merit_of(double (&a)[3]) {
        tm.sum<3>(a, g_sl.red);
        if (a[2] > 0) return -INFINITY;
        return -(0.5 * a[0] - a[1]) / n;
    }
    // Real rank one: EJ = E + t (EP - E), UJ = U + t (UP - U), and merit(EJ, UJ) summed in the same pass (the
    // values and the order of the sums are merit()'s, so the result is the same to the bit).
    __device__ double lerp_merit(double *EJ, double *UJ, const double *E, const double *U, const double *EP,
                                 const double *UP, double t, double *Qout) {
        double a[3] = {0.0, 0.0, 0.0};
        for (int b = (int)tm.first(); b < B; b += (int)tm.stride()) {
            const double e = E[b] + t * (EP[b] - E[b]);
            const double u = U[b] + t * (UP[b] - U[b]);
            EJ[b] = e;
            UJ[b] = u;
            merit_terms(b, e, u, a);
        }
        const double m = merit_of(a);
        *Qout = a[0];
        return m;
    }
    // Real rank one: rescale(E, U), then merit(E, U) and resid(E, U, 1) of the rescaled pair, both summed in the
    // pass that scales (one pass over the bins instead of three; values and sum orders are theirs).  Q: dot(E, U),
    // known from the merit evaluation of (E, U); Qout: the rescaled pair's.
    __device__ void rescale_merit_resid(double *E, double *U, double Q, double &mer, double &res, double &Qout) {
        const double sc = Q > 0 ? n / sqrt(Q) : 1.0;
        double a[3] = {0.0, 0.0, 0.0};
        double r = 0.0;
        for (int b = (int)tm.first(); b < B; b += (int)tm.stride()) {
            const double e = E[b] * sc, u = U[b] * sc;
            E[b] = e;
            U[b] = u;
            merit_terms(b, e, u, a);
            if (occ(b)) {
                const double nrm2 = e * e;
                const double v = u * (nrm2 / cb(b)) - e;
                r = fmax(r, sqrt(v * v) / sqrt(fmax(nrm2, 1e-300)));
            }
        }
        mer = merit_of(a);
        Qout = a[0];
        res = tm.max(r, g_sl.red);
    }

    __device__ void update(const double *E, const double *U, double *out) {
        if (!psd) {
            const double s0 = s[0];
            for (int b = (int)tm.first(); b < B; b += (int)tm.stride()) {
                const double sgb = sg ? sg[b] : 1.0;
                const double R = U[b] - s0 * E[b];
                const double t = occ(b) ? cd_root(sgb * R, cb(b), s0) : 0.0;
                out[b] = sgb * t;
            }
            tm.sync();
            return;
        }
        for (int b = (int)tm.first(); b < B; b += (int)tm.stride()) {
            double Rr[BLML_MAXK], Ri[BLML_MAXK], R2[BLML_MAXK];
            double a2 = 0.0;
            for (int k = 0; k < K; ++k) {
                const long i = ((long)k * B + b) * nc;
                Rr[k] = U[i] - s[k] * E[i];
                Ri[k] = nc == 2 ? U[i + 1] - s[k] * E[i + 1] : 0.0;
                R2[k] = Rr[k] * Rr[k] + Ri[k] * Ri[k];
                a2 += R2[k];
            }
            const double c = cb(b);
            const double a = sqrt(a2);
            if (equal_s) {
                const double s0 = s[0];
                const double t = 2 * c / (a + sqrt(a2 + 4 * s0 * c));
                const double g = t / (a > 0 ? a : 1.0);
                for (int k = 0; k < K; ++k) { Rr[k] *= g; Ri[k] *= g; }
                if (a == 0 && c > 0) { Rr[0] = sqrt(c / s0); Ri[0] = 0.0; }
            } else {
                const bool lone = (a2 == 0) && (c > 0);
                const double lam = secular_root(R2, s, K, c);
                for (int k = 0; k < K; ++k) {
                    if (c > 0 && !lone) {
                        const double dn = lam - s[k];
                        Rr[k] /= dn; Ri[k] /= dn;
                    } else {
                        Rr[k] = 0.0; Ri[k] = 0.0;
                    }
                }
                if (lone) { Rr[0] = sqrt(c / s[0]); Ri[0] = 0.0; }
            }
            for (int k = 0; k < K; ++k) {
                const long i = ((long)k * B + b) * nc;
                out[i] = Rr[k];
                if (nc == 2) out[i + 1] = Ri[k];
            }
        }
        tm.sync();
    }

    __device__ void rescale(double *E, double *U) {
        const double Q = dot(E, U);
        const double sc = Q > 0 ? n / sqrt(Q) : 1.0;
        const long n_ = V();
        for (long i = tm.first(); i < n_; i += tm.stride()) {
            E[i] *= sc;
            U[i] *= sc;
        }
        tm.sync();
    }

    // max over occupied bins of ||u_b (|e_b|^2 / c_b) - e_b|| / ||e_b|| of the rescaled pair (sc E, sc U)
    __device__ double resid(const double *E, const double *U, double sc) {
        double r = 0.0;
        for (int b = (int)tm.first(); b < B; b += (int)tm.stride()) {
            if (!occ(b)) continue;
            double nrm2 = 0.0;
            for (int k = 0; k < K; ++k) nrm2 += absq(E, k, b);
            nrm2 *= sc * sc;
            const double fac = nrm2 / cb(b);
            double acc = 0.0;
            for (int k = 0; k < K; ++k) {
                const long i = ((long)k * B + b) * nc;
                for (int q = 0; q < nc; ++q) {
                    const double v = sc * U[i + q] * fac - sc * E[i + q];
                    acc += v * v;
                }
            }
            r = fmax(r, sqrt(acc) / sqrt(fmax(nrm2, 1e-300)));
        }
        return tm.max(r, g_sl.red);
    }

    // L = sum w log F - n log Q (blml_binned.py:278-284); writes Vt = (w / F) U on occupied bins when Vt is given
    // (the gradient's mat-vec input) and Q to *Qout
    __device__ double lval(const double *E, const double *U, double *Vt, double *Qout) {
        double a[2] = {0.0, 0.0};
        for (int b = (int)tm.first(); b < B; b += (int)tm.stride()) {
            const bool o = occ(b);
            const double F = Fb(U, b);
            a[0] += w[b] * log(o ? F : 1.0);
            const double fac = o ? w[b] / F : 0.0;
            for (int k = 0; k < K; ++k) {
                const long i = ((long)k * B + b) * nc;
                for (int q = 0; q < nc; ++q) {
                    a[1] += E[i + q] * U[i + q];
                    if (Vt) Vt[i + q] = fac * U[i + q];
                }
            }
        }
        tm.sum<2>(a, g_sl.red);
        *Qout = a[1];
        if (!(a[1] > 0)) return -INFINITY;
        return a[0] - n * log(a[1]);
    }
    // A line-search trial in one pass over the bins: En = E + t q, Un = U + t Sq, then lval(En, Un, Vt, Qout) (the
    // same values in the same order: every value of bin b is this thread's own).  Complex fits take the NC = 2
    // body: a bin's (re, im) pair of every vector moves as one 16-byte access (every state slot of a complex fit
    // starts 16-byte aligned: its length K B 2 is even and the pool 256-byte aligned).
    template <int NC>
    __device__ __forceinline__ double lerp_lval_nc(double *En, double *Un, const double *E, const double *U,
                                                   const double *q, const double *Sq, double t, double *Vt,
                                                   double *Qout) {
        if (NC == 2) {
            En = (double *)__builtin_assume_aligned(En, 16);
            Un = (double *)__builtin_assume_aligned(Un, 16);
            Vt = (double *)__builtin_assume_aligned(Vt, 16);
            E = (const double *)__builtin_assume_aligned(E, 16);
            U = (const double *)__builtin_assume_aligned(U, 16);
            q = (const double *)__builtin_assume_aligned(q, 16);
            Sq = (const double *)__builtin_assume_aligned(Sq, 16);
        }
        double a[2] = {0.0, 0.0};
        for (int b = (int)tm.first(); b < B; b += (int)tm.stride()) {
            double F = 0.0;
            for (int k = 0; k < K; ++k) {
                const long i = ((long)k * B + b) * NC;
                double en[NC], un[NC];
#pragma unroll
                for (int c = 0; c < NC; ++c) {
                    en[c] = E[i + c] + t * q[i + c];
                    un[c] = U[i + c] + t * Sq[i + c];
                }
#pragma unroll
                for (int c = 0; c < NC; ++c) {
                    En[i + c] = en[c];
                    Un[i + c] = un[c];
                }
                F += NC == 2 ? un[0] * un[0] + un[1] * un[1] : un[0] * un[0];
            }
            const bool o = occ(b);
            a[0] += w[b] * log(o ? F : 1.0);
            const double fac = o ? w[b] / F : 0.0;
            for (int k = 0; k < K; ++k) {
                const long i = ((long)k * B + b) * NC;
                double vt[NC];
#pragma unroll
                for (int c = 0; c < NC; ++c) {
                    a[1] += En[i + c] * Un[i + c];
                    vt[c] = fac * Un[i + c];
                }
#pragma unroll
                for (int c = 0; c < NC; ++c) Vt[i + c] = vt[c];
            }
        }
        tm.sum<2>(a, g_sl.red);
        *Qout = a[1];
        if (!(a[1] > 0)) return -INFINITY;
        return a[0] - n * log(a[1]);
    }
    __device__ double lerp_lval(double *En, double *Un, const double *E, const double *U, const double *q,
                                const double *Sq, double t, double *Vt, double *Qout) {
        return nc == 2 ? lerp_lval_nc<2>(En, Un, E, U, q, Sq, t, Vt, Qout)
                       : lerp_lval_nc<1>(En, Un, E, U, q, Sq, t, Vt, Qout);
    }

    // g = 2 (S Vt - n U / Q) on occupied bins (blml_binned.py:285-288); SVt: scratch
    __device__ void lgrad_from(const double *U, double Q, const double *Vt, double *SVt, double *g) {
        Sall(Vt, SVt);
        const double nq = n / Q;
        for (int b = (int)tm.first(); b < B; b += (int)tm.stride()) {
            const bool o = occ(b);
            for (int k = 0; k < K; ++k) {
                const long i = ((long)k * B + b) * nc;
                for (int q = 0; q < nc; ++q) g[i + q] = o ? 2.0 * (SVt[i + q] - nq * U[i + q]) : 0.0;
            }
        }
        tm.sync();
    }

    // lgrad_from(U, Q, Vt, SVt, g), and resid(E, U, Q > 0 ? n / sqrt(Q) : 1) of the same point summed in its pass
    // over the bins (resid()'s values; a maximum does not depend on the order, so the residual is the same to the bit)
    // Complex fits take the NC = 2 body: a bin's (re, im) pairs move as 16-byte accesses (the slots of a complex
    // fit start 16-byte aligned), with the same values in the same order.
    template <int NC>
    __device__ __forceinline__ double lgrad_resid_nc(const double *E, const double *U, double Q, double *SVt,
                                                     double *g) {
        if (NC == 2) {
            E = (const double *)__builtin_assume_aligned(E, 16);
            U = (const double *)__builtin_assume_aligned(U, 16);
            SVt = (double *)__builtin_assume_aligned(SVt, 16);
            g = (double *)__builtin_assume_aligned(g, 16);
        }
        const double nq = n / Q;
        const double sc = Q > 0 ? n / sqrt(Q) : 1.0;
        double r = 0.0;
        for (int b = (int)tm.first(); b < B; b += (int)tm.stride()) {
            const bool o = occ(b);
            for (int k = 0; k < K; ++k) {
                const long i = ((long)k * B + b) * NC;
                double gv[NC];
#pragma unroll
                for (int q = 0; q < NC; ++q) gv[q] = o ? 2.0 * (SVt[i + q] - nq * U[i + q]) : 0.0;
#pragma unroll
                for (int q = 0; q < NC; ++q) g[i + q] = gv[q];
            }
            if (!o) continue;
            double nrm2 = 0.0;
            for (int k = 0; k < K; ++k) {
                const long i = ((long)k * B + b) * NC;
                nrm2 += NC == 2 ? E[i] * E[i] + E[i + 1] * E[i + 1] : E[i] * E[i];
            }
            nrm2 *= sc * sc;
            const double fac = nrm2 / cb(b);
            double acc = 0.0;
            for (int k = 0; k < K; ++k) {
                const long i = ((long)k * B + b) * NC;
#pragma unroll
                for (int q = 0; q < NC; ++q) {
                    const double v = sc * U[i + q] * fac - sc * E[i + q];
                    acc += v * v;
                }
            }
            r = fmax(r, sqrt(acc) / sqrt(fmax(nrm2, 1e-300)));
        }
        return tm.max(r, g_sl.red);
    }
    __device__ double lgrad_resid(const double *E, const double *U, double Q, const double *Vt, double *SVt,
                                  double *g) {
        Sall(Vt, SVt);
        return nc == 2 ? lgrad_resid_nc<2>(E, U, Q, SVt, g) : lgrad_resid_nc<1>(E, U, Q, SVt, g);
    }

    // ------------------------------------------------------------------ histories
    __device__ void push_merit(double v) {
        if (threadIdx.x == 0) g_sl.ring[ring_n % BLML_RING] = v;
        if (tm.lead() && rec && merit_hist && nhist_m < hist_cap) merit_hist[nhist_m] = v;
        ++ring_n;
        if (rec) ++nhist_m;
        __syncthreads();
    }
    __device__ double ring_back(int k) const { return g_sl.ring[(ring_n - 1 - k) % BLML_RING]; }
    __device__ void push_rho(double v) {
        if (tm.lead() && rec && rho_hist && nhist_r < hist_cap) rho_hist[nhist_r] = v;
        if (rec) ++nhist_r;
    }

    // ------------------------------------------------------------------ Anderson type II
    // slotX / slotF: the history slots in order (oldest first); returns false on a singular history.
    __device__ bool anderson(const int *slotX, const int *slotF, int cnt, const double *E, const double *f,
                             double *EA) {
        const int m = cnt - 1;
        double acc[20];
        for (int i = 0; i < 20; ++i) acc[i] = 0.0;
        const long n_ = V();
        for (long e = tm.first(); e < n_; e += tm.stride()) {
            double dF[BLML_MAXAND];
            for (int i = 0; i < m; ++i) dF[i] = vec(slotF[i + 1])[e] - vec(slotF[i])[e];
            const double fe = f[e];
            int p = 0;
            for (int i = 0; i < m; ++i)
                for (int j = i; j < m; ++j) acc[p++] += dF[i] * dF[j];
            for (int i = 0; i < m; ++i) acc[p++] += dF[i] * fe;
        }
        tm.sum<20>(acc, g_sl.red);
        double G[BLML_MAXAND][BLML_MAXAND + 1];
        int p = 0;
        double tr = 0.0;
        for (int i = 0; i < m; ++i)
            for (int j = i; j < m; ++j) { G[i][j] = acc[p]; G[j][i] = acc[p]; ++p; }
        for (int i = 0; i < m; ++i) { G[i][m] = acc[p++]; tr += G[i][i]; }
        for (int i = 0; i < m; ++i) G[i][i] += 1e-12 * tr;
        // Gaussian elimination with partial pivoting (every thread solves the same small system)
        for (int col = 0; col < m; ++col) {
            int piv = col;
            for (int r = col + 1; r < m; ++r)
                if (fabs(G[r][col]) > fabs(G[piv][col])) piv = r;
            if (G[piv][col] == 0.0) return false;
            if (piv != col)
                for (int q = 0; q <= m; ++q) { const double t = G[col][q]; G[col][q] = G[piv][q]; G[piv][q] = t; }
            for (int r = col + 1; r < m; ++r) {
                const double fct = G[r][col] / G[col][col];
                for (int q = col; q <= m; ++q) G[r][q] -= fct * G[col][q];
            }
        }
        double gam[BLML_MAXAND];
        for (int i = m - 1; i >= 0; --i) {
            double v = G[i][m];
            for (int q = i + 1; q < m; ++q) v -= G[i][q] * gam[q];
            gam[i] = v / G[i][i];
        }
        for (int i = 0; i < m; ++i)
            if (!isfinite(gam[i])) return false;
        for (long e = tm.first(); e < n_; e += tm.stride()) {
            double mix = 0.0;
            for (int i = 0; i < m; ++i) {
                const double dX = vec(slotX[i + 1])[e] - vec(slotX[i])[e];
                const double dF = vec(slotF[i + 1])[e] - vec(slotF[i])[e];
                mix += gam[i] * (dX + dF);
            }
            EA[e] = E[e] + f[e] - mix;
        }
        tm.sync();
        return true;
    }

    // ------------------------------------------------------------------ gradient fallback (psd)
    // From (E, U, cur): up to 5 Armijo-backtracked steepest-ascent steps on L; result in (Eo, Uo), merit returned.
    __device__ double grad_steps(const double *E0, const double *U0, double cur, double *Eo, double *Uo, double *Vb,
                                 double *Gb, double *SGb, double *En, double *Un) {
        copy(Eo, E0);
        copy(Uo, U0);
        for (int st = 0; st < 5; ++st) {
            double Q = 0.0;
            lval(Eo, Uo, Vb, &Q);
            Sall(Vb, Gb);
            const double nq = n / Q;
            for (int b = (int)tm.first(); b < B; b += (int)tm.stride()) {
                const bool o = occ(b);
                for (int k = 0; k < K; ++k) {
                    const long i = ((long)k * B + b) * nc;
                    for (int q = 0; q < nc; ++q) Gb[i + q] = o ? Gb[i + q] - nq * Uo[i + q] : 0.0;
                }
            }
            tm.sync();
            Sall(Gb, SGb);
            const double g2 = dot(Gb, Gb);
            if (!(g2 > 0)) break;
            double a = gstep;
            if (!(a != 0.0)) a = 1e-3 * sqrt(dot(Eo, Eo) / g2);
            bool ok = false;
            double mn = cur;
            for (int t = 0; t < 40; ++t) {
                axpy(En, Eo, a, Gb);
                axpy(Un, Uo, a, SGb);
                mn = merit(En, Un);
                if (mn > cur) { ok = true; break; }
                a *= 0.5;
            }
            if (!ok) break;
            copy(Eo, En);
            copy(Uo, Un);
            cur = mn;
            gstep = 2 * a;
        }
        return cur;
    }

    // ------------------------------------------------------------------ Jacobi CD + Anderson (blml_binned.py:223-276)
    // E in slot 0; the solution ends in slots 0 (E) and 1 (U).
    __device__ FitOut solve_cd() {
        // slots: 0 E, 1 U, 2 EJ, 3 UJ, 4 EA, 5 UA, 6 EP, 7 UP, 8 V, 9 G, 10 SG, 11 En, 12 Un, 13.. history
        int iE = 0, iU = 1, iEJ = 2, iUJ = 3, iEA = 4, iUA = 5;
        int iEP = 6, iUP = 7;
        const int mA = P->anderson;
        const int hcap = mA + 1;
        int hhead = 0, hcnt = 0;
        FitOut o{};
        Sall(vec(iE), vec(iU));
        rescale(vec(iE), vec(iU));
        double Qcur = 0.0;              // real rank one: dot(E, U) of the current pair, from its merit evaluation
        double cur = merit(vec(iE), vec(iU), &Qcur);
        long it = 0;
        double res = resid(vec(iE), vec(iU), 1.0);
        push_rho(res);
        const long cap = P->max_sweeps < 0 ? BLML_HARD_CAP : P->max_sweeps;
        double best_res = res, best_m = cur;
        long best_it = 0;
        ring_n = 0;
        push_merit(cur);
        int stop = BLML_STOP_TOL;
        const bool use_ll = !isnan(P->ll_tol);
        const int lw = P->ll_window;
        while (res >= P->tol && it < cap) {
            if (it - best_it >= P->stall) { stop = BLML_STOP_STALL; break; }
            if (use_ll && ring_n > lw && (ring_back(0) - ring_back(lw)) / n < P->ll_tol) {
                stop = BLML_STOP_LLTOL; break;
            }
            ++it;
            update(vec(iE), vec(iU), vec(iEJ));
            Snew(vec(iEJ), vec(iUJ), vec(iE), vec(iU));
            double QJ = 0.0;            // real rank one: dot(EJ, UJ)
            double mJ = merit(vec(iEJ), vec(iUJ), &QJ);
            bool plain = true;          // EJ, UJ still hold the Jacobi update of this sweep
            // a sweep that will damp (no ascent) with fewer than 3 pairs after this one tries no Anderson step, and the
            // damped step clears the history: its pair is not stored (the state after the sweep is the same)
            const bool skip_pair = hcnt + 1 < 3 && mJ < cur - 1e-12 * fabs(cur);
            if (mA && !skip_pair) {
                // f = EJ - E; hX.append(E); hF.append(f); keep the last mA + 1
                int slot;
                if (hcnt < hcap) {
                    slot = (hhead + hcnt) % hcap;
                    ++hcnt;
                } else {
                    slot = hhead;
                    hhead = (hhead + 1) % hcap;
                }
                double *xs = vec(13 + slot), *fs = vec(13 + hcap + slot);
                const double *Ec = vec(iE), *EJc = vec(iEJ);
                const long n_ = V();
                for (long e = tm.first(); e < n_; e += tm.stride()) {
                    xs[e] = Ec[e];
                    fs[e] = EJc[e] - Ec[e];
                }
                tm.sync();
                if (hcnt >= 3) {
                    int sx[BLML_MAXAND + 1], sf[BLML_MAXAND + 1];
                    for (int i = 0; i < hcnt; ++i) {
                        sx[i] = 13 + (hhead + i) % hcap;
                        sf[i] = 13 + hcap + (hhead + i) % hcap;
                    }
                    if (anderson(sx, sf, hcnt, vec(iE), fs, vec(iEA))) {
                        Snew(vec(iEA), vec(iUA), vec(iE), vec(iU));
                        double QA = 0.0;
                        const double mAc = merit(vec(iEA), vec(iUA), &QA);
                        if (mAc > mJ) {
                            int t = iEJ; iEJ = iEA; iEA = t;
                            t = iUJ; iUJ = iUA; iUA = t;
                            mJ = mAc;
                            QJ = QA;
                            plain = false;
                            ++o.n_and;
                        }
                    } else {
                        hcnt = 0; hhead = 0;
                    }
                }
            }
            if (mJ < cur - 1e-12 * fabs(cur)) {
                hcnt = 0; hhead = 0;
                ++o.n_damp;
                // the damped step runs along the Jacobi update (blml_binned.py:252-255 recomputes it; when EJ is
                // still that update its slots are taken over, the same values without the update and mat-vec)
                if (plain) {
                    int t = iEP; iEP = iEJ; iEJ = t;
                    t = iUP; iUP = iUJ; iUJ = t;
                } else {
                    update(vec(iE), vec(iU), vec(iEP));
                    Snew(vec(iEP), vec(iUP), vec(iE), vec(iU));
                }
                double a = 0.5;
                bool found = false;
                while (a > 1e-8) {
                    const long n_ = V();
                    double *EJ = vec(iEJ), *UJ = vec(iUJ);
                    const double *Ec = vec(iE), *Uc = vec(iU), *EPc = vec(iEP), *UPc = vec(iUP);
                    if (!psd) {
                        mJ = lerp_merit(EJ, UJ, Ec, Uc, EPc, UPc, a, &QJ);
                    } else {
                        for (long e = tm.first(); e < n_; e += tm.stride()) {
                            EJ[e] = Ec[e] + a * (EPc[e] - Ec[e]);
                            UJ[e] = Uc[e] + a * (UPc[e] - Uc[e]);
                        }
                        tm.sync();
                        mJ = merit(vec(iEJ), vec(iUJ));
                    }
                    if (mJ >= cur) { found = true; break; }
                    a *= 0.5;
                }
                if (!found) {
                    if (psd) {
                        mJ = grad_steps(vec(iE), vec(iU), cur, vec(iEJ), vec(iUJ), vec(8), vec(9), vec(10),
                                        vec(11), vec(12));
                        ++o.n_grad;
                    } else {
                        copy(vec(iEJ), vec(iE));
                        copy(vec(iUJ), vec(iU));
                        mJ = cur;
                        QJ = Qcur;
                    }
                }
            }
            if (!psd) {
                rescale_merit_resid(vec(iEJ), vec(iUJ), QJ, cur, res, Qcur);
                { int t = iE; iE = iEJ; iEJ = t; t = iU; iU = iUJ; iUJ = t; }
            } else {
                rescale(vec(iEJ), vec(iUJ));
                { int t = iE; iE = iEJ; iEJ = t; t = iU; iU = iUJ; iUJ = t; }
                cur = mJ;
                res = resid(vec(iE), vec(iU), 1.0);
            }
            push_merit(cur);
            push_rho(res);
            if (res < 0.9 * best_res || cur > best_m + 1e-13 * fabs(best_m)) {
                best_res = fmin(res, best_res);
                best_it = it;
                best_m = fmax(cur, best_m);
            }
        }
        if (res >= P->tol && stop == BLML_STOP_TOL) stop = BLML_STOP_MAXSWEEPS;
        if (iE != 0) { copy(vec(0), vec(iE)); copy(vec(1), vec(iU)); }
        o.n_sweeps = (int)it;
        o.stop = stop;
        o.residual = res;
        o.converged = res < P->tol;
        return o;
    }

    // ------------------------------------------------------------------ L-BFGS (blml_binned.py:290-355)
    // Each trial of the line search evaluates L only; the gradient is formed once, at the accepted point (the
    // reference forms it at every trial and uses only the accepted one: the same iterates).
    __device__ FitOut solve_lbfgs() {
        // slots: 0 E, 1 U, 2 g, 3 q, 4 En, 5 Un, 6 gn, 7 Vt, 8 SVt (also S q), 9.. S (m + 1), 10 + m.. Y (m + 1).
        // The pair memory is a ring of m + 1 physical slots holding at most m pairs, so a candidate pair never
        // overwrites a kept one before its curvature test.
        int iE = 0, iU = 1, ig = 2, iEn = 4, iUn = 5, ign = 6;
        const int iq = 3, iVt = 7, iSVt = 8;
        const int mh = P->lbfgs_m, ph = mh + 1;
        FitOut o{};
        Sall(vec(iE), vec(iU));
        rescale(vec(iE), vec(iU));
        double Q = 0.0;
        double L = lval(vec(iE), vec(iU), vec(iVt), &Q);
        lgrad_from(vec(iU), Q, vec(iVt), vec(iSVt), vec(ig));
        int head = 0, cnt = 0;
        long it = 0;
        double res = resid(vec(iE), vec(iU), Q > 0 ? n / sqrt(Q) : 1.0);
        push_rho(res);
        ring_n = 0;
        push_merit(L / n);
        const long cap = P->max_sweeps < 0 ? BLML_HARD_CAP : P->max_sweeps;
        int stop = BLML_STOP_TOL;
        bool reset = false;
        const double ftol = isnan(P->ll_tol) ? 1e-11 : P->ll_tol;
        const int fwin = 10;
        while (res >= P->tol && it < cap) {
            if (ring_n > fwin && ring_back(0) - ring_back(fwin) < ftol) { stop = BLML_STOP_LLTOL; break; }
            ++it;
            const long long td0 = wall_clock64();
            double *q = vec(iq);
            copy(q, vec(ig));
            for (int j = cnt - 1; j >= 0; --j) {
                const int sl_ = (head + j) % ph;
                const double a = g_sl.lb_r[sl_] * dot(vec(9 + sl_), q);
                if (threadIdx.x == 0) g_sl.lb_a[j] = a;
                axpy(q, q, -a, vec(9 + ph + sl_));
            }
            if (cnt > 0) {
                const int nw = (head + cnt - 1) % ph;
                scale(q, g_sl.lb_sy[nw] / g_sl.lb_yy[nw]);
            } else {
                const double ee = dot(vec(iE), vec(iE)), gg = dot(vec(ig), vec(ig));
                scale(q, 1e-3 * sqrt(ee / fmax(gg, 1e-300)));
            }
            for (int j = 0; j < cnt; ++j) {
                const int sl_ = (head + j) % ph;
                const double bb = g_sl.lb_r[sl_] * dot(vec(9 + ph + sl_), q);
                axpy(q, q, g_sl.lb_a[j] - bb, vec(9 + sl_));
            }
            double slope = dot(vec(ig), q);
            if (!(slope > 0)) {
                cnt = 0; head = 0;
                const double ee = dot(vec(iE), vec(iE)), gg = dot(vec(ig), vec(ig));
                const double sc = 1e-3 * sqrt(ee / fmax(gg, 1e-300));
                const long n_ = V();
                const double *g = vec(ig);
                for (long e = tm.first(); e < n_; e += tm.stride()) q[e] = g[e] * sc;
                tm.sync();
                slope = dot(vec(ig), q);
            }
            const long long td1 = wall_clock64();
            t_dir += td1 - td0;
            if (inc) Sall(q, vec(iSVt));             // S q: every trial's U is U + t S q
            double t = 1.0, Ln = -INFINITY, Qn = 0.0;
            bool ok = false;
            for (int tr = 0; tr < 40; ++tr) {
                const long n_ = V();
                double *En = vec(iEn), *Un = vec(iUn);
                const double *Ec = vec(iE), *Uc = vec(iU), *Sq = vec(iSVt);
                for (long e = tm.first(); e < n_; e += tm.stride()) {
                    En[e] = Ec[e] + t * q[e];
                    if (inc) Un[e] = Uc[e] + t * Sq[e];
                }
                tm.sync();
                if (!inc) Sall(vec(iEn), vec(iUn));
                Ln = lval(vec(iEn), vec(iUn), nullptr, &Qn);
                if (Ln >= L + 1e-4 * t * slope) { ok = true; break; }
                t *= 0.5;
            }
            t_ls += wall_clock64() - td1;
            if (!ok) {
                if (reset || cnt == 0) { stop = BLML_STOP_STALL; break; }
                cnt = 0; head = 0;
                reset = true;
                continue;
            }
            reset = false;
            lval(vec(iEn), vec(iUn), vec(iVt), &Qn);
            lgrad_from(vec(iUn), Qn, vec(iVt), vec(iSVt), vec(ign));
            // s = En - E, y = g - gn, into the next slot; curvature test
            const int slot = (head + cnt) % ph;
            {
                double *Sv = vec(9 + slot), *Yv = vec(9 + ph + slot);
                const double *En = vec(iEn), *Ec = vec(iE), *g = vec(ig), *gn = vec(ign);
                double a3[3] = {0.0, 0.0, 0.0};
                const long n_ = V();
                for (long e = tm.first(); e < n_; e += tm.stride()) {
                    const double sv = En[e] - Ec[e];
                    const double yv = g[e] - gn[e];
                    Sv[e] = sv;
                    Yv[e] = yv;
                    a3[0] += sv * yv;
                    a3[1] += sv * sv;
                    a3[2] += yv * yv;
                }
                tm.sum<3>(a3, g_sl.red);
                if (a3[0] > 1e-12 * sqrt(a3[1] * a3[2])) {
                    if (threadIdx.x == 0) {
                        g_sl.lb_r[slot] = 1.0 / a3[0];
                        g_sl.lb_sy[slot] = a3[0];
                        g_sl.lb_yy[slot] = a3[2];
                    }
                    if (cnt < mh) ++cnt;
                    else head = (head + 1) % ph;
                }
                __syncthreads();
            }
            { int tt = iE; iE = iEn; iEn = tt; tt = iU; iU = iUn; iUn = tt; tt = ig; ig = ign; ign = tt; }
            L = Ln;
            Q = Qn;
            res = resid(vec(iE), vec(iU), Q > 0 ? n / sqrt(Q) : 1.0);
            push_rho(res);
            push_merit(L / n);
        }
        if (res >= P->tol && stop == BLML_STOP_TOL) stop = BLML_STOP_MAXSWEEPS;
        // canonical form: e_k = (Q / n) (w / F) u_k on occupied bins, kept unless it loses likelihood
        {
            const double Qc = dot(vec(iE), vec(iU));
            double *Ecn = vec(iEn);
            const double *Uc = vec(iU);
            for (int b = (int)tm.first(); b < B; b += (int)tm.stride()) {
                const bool ob = occ(b);
                const double F = Fb(Uc, b);
                const double fac = ob ? (Qc / n) * w[b] / F : 0.0;
                for (int k = 0; k < K; ++k) {
                    const long i = ((long)k * B + b) * nc;
                    for (int qq = 0; qq < nc; ++qq) Ecn[i + qq] = fac * Uc[i + qq];
                }
            }
            tm.sync();
            Snew(vec(iEn), vec(iUn), vec(iE), vec(iU));
            double Qd = 0.0;
            const double Lc = lval(vec(iEn), vec(iUn), nullptr, &Qd);
            if (Lc >= L - 1e-9 * n) {
                int tt = iE; iE = iEn; iEn = tt; tt = iU; iU = iUn; iUn = tt;
            }
        }
        rescale(vec(iE), vec(iU));
        res = resid(vec(iE), vec(iU), 1.0);
        if (iE != 0) { copy(vec(0), vec(iE)); copy(vec(1), vec(iU)); }
        o.n_sweeps = (int)it;
        o.stop = stop;
        o.residual = res;
        o.converged = res < P->tol || stop == BLML_STOP_LLTOL;
        return o;
    }

    // ------------------------------------------------------------------ L-BFGS, compact form (incremental mode)
    // The same iteration as solve_lbfgs (the same BFGS matrix from the same accepted pairs, blml_binned.py:303-335),
    // with the direction H g from the compact representation H = gamma I + [S  gamma Y] M [S^T; gamma Y^T]
    // (Byrd, Nocedal & Schnabel 1994): with a = S^T g, b = Y^T g, R = triu(S^T Y), D = diag(s_i . y_i),
    //   u = R^-1 a,  p1 = R^-T ((D + gamma Y^T Y) u - gamma b),  p2 = -u,  H g = gamma g + S p1 + gamma Y p2.
    // One pass over the history per iteration produces S^T g_new and Y^T g_new, from which the new column of R
    // (s_i . y = s_i . g - s_i . g_new) and of Y^T Y follow; a second pass forms the direction.
    template <int W>
    __device__ __forceinline__ void dots_elems(const float *const *Sp, const float *const *Yp, int nc_,
                                               const double *gn, double *acc) {
        typedef typename f32xw<W>::type vw;
        const long n_ = V();
        const long nw = (n_ + W - 1) / W;
        for (long ew = tm.first(); ew < nw; ew += tm.stride()) {
            const long e = W * ew;
            if (e + W <= n_) {
                double gv[W];
#pragma unroll
                for (int u = 0; u < W; ++u) gv[u] = gn[e + u];
#pragma unroll
                for (int i = 0; i < BLML_CCHUNK; ++i) {
                    if (i < nc_) {
                        const vw sv = *reinterpret_cast<const vw *>(Sp[i] + e);
                        const vw yv = *reinterpret_cast<const vw *>(Yp[i] + e);
#pragma unroll
                        for (int u = 0; u < W; ++u) {
                            acc[i] += (double)sv.v[u] * gv[u];
                            acc[BLML_CCHUNK + i] += (double)yv.v[u] * gv[u];
                        }
                    }
                }
            } else {
                for (long ee = e; ee < n_; ++ee) {
                    const double ge = gn[ee];
#pragma unroll
                    for (int i = 0; i < BLML_CCHUNK; ++i) {
                        if (i < nc_) {
                            acc[i] += (double)Sp[i][ee] * ge;
                            acc[BLML_CCHUNK + i] += (double)Yp[i][ee] * ge;
                        }
                    }
                }
            }
        }
    }

    // this thread's share of the compact direction: q[e] = gam g[e] + sum_i (p1_i s_i[e] + p2_i y_i[e]) for W
    // consecutive elements per step (W-value loads of the FP32 history, its pairs in chunks of BLML_CCHUNK);
    // returns this thread's part of g . q
    template <int W>
    __device__ __forceinline__ double dir_elems(int head, int ph, int cnt, double gam, const double *g, double *q) {
        typedef typename f32xw<W>::type vw;
        const long n_ = V();
        const long stp = tm.stride();
        double acc = 0.0;
        for (long e0 = W * tm.first(); e0 < n_; e0 += W * stp) {
            const bool whole = e0 + W <= n_;
            double v[W], gv[W];
#pragma unroll
            for (int u = 0; u < W; ++u) {
                gv[u] = e0 + u < n_ ? g[e0 + u] : 0.0;
                v[u] = gam * gv[u];
            }
            for (int c0 = 0; c0 < cnt; c0 += BLML_CCHUNK) {
                const int nc_ = cnt - c0 < BLML_CCHUNK ? cnt - c0 : BLML_CCHUNK;
                const float *Sp[BLML_CCHUNK], *Yp[BLML_CCHUNK];
                double a1[BLML_CCHUNK], a2[BLML_CCHUNK];
#pragma unroll
                for (int i = 0; i < BLML_CCHUNK; ++i) {
                    int sl_ = head + c0 + (i < nc_ ? i : 0);
                    if (sl_ >= ph) sl_ -= ph;
                    Sp[i] = hvec(9 + sl_);
                    Yp[i] = hvec(9 + ph + sl_);
                    a1[i] = g_sl.cp1[c0 + (i < nc_ ? i : 0)];
                    a2[i] = g_sl.cp2[c0 + (i < nc_ ? i : 0)];
                }
                if (whole) {
#pragma unroll
                    for (int i = 0; i < BLML_CCHUNK; ++i)
                        if (i < nc_) {
                            const vw sv = *reinterpret_cast<const vw *>(Sp[i] + e0);
                            const vw yv = *reinterpret_cast<const vw *>(Yp[i] + e0);
#pragma unroll
                            for (int u = 0; u < W; ++u) v[u] += a1[i] * (double)sv.v[u] + a2[i] * (double)yv.v[u];
                        }
                } else {
#pragma unroll
                    for (int u = 0; u < W; ++u)
                        if (e0 + u < n_) {
#pragma unroll
                            for (int i = 0; i < BLML_CCHUNK; ++i)
                                if (i < nc_) v[u] += a1[i] * (double)Sp[i][e0 + u] + a2[i] * (double)Yp[i][e0 + u];
                        }
                }
            }
#pragma unroll
            for (int u = 0; u < W; ++u) {
                if (e0 + u < n_) {
                    q[e0 + u] = v[u];
                    acc += gv[u] * v[u];
                }
            }
        }
        return acc;
    }

    __device__ void dots_history(int head, int ph, int cnt, const double *gn) {
        // csg[i] = s_i . gn, cyg[i] = y_i . gn for the cnt pairs (logical order), in chunks of BLML_CCHUNK pairs
        const long n_ = V();
        for (int c0 = 0; c0 < cnt; c0 += BLML_CCHUNK) {
            double acc[2 * BLML_CCHUNK];
#pragma unroll
            for (int i = 0; i < 2 * BLML_CCHUNK; ++i) acc[i] = 0.0;
            const int nc_ = cnt - c0 < BLML_CCHUNK ? cnt - c0 : BLML_CCHUNK;
            const float *Sp[BLML_CCHUNK], *Yp[BLML_CCHUNK];
#pragma unroll
            for (int i = 0; i < BLML_CCHUNK; ++i) {
                int sl_ = head + c0 + (i < nc_ ? i : 0);     // head < ph and c0 + i < ph: one wrap at most
                if (sl_ >= ph) sl_ -= ph;
                Sp[i] = hvec(9 + sl_);
                Yp[i] = hvec(9 + ph + sl_);
            }
            // W consecutive elements per thread and step, each history vector read by W-value loads: 16-byte
            // loads when every history vector starts 16-byte aligned (an even slot length), else 8-byte ones
            if ((Vmax & 1) == 0) dots_elems<4>(Sp, Yp, nc_, gn, acc);
            else dots_elems<2>(Sp, Yp, nc_, gn, acc);
            tm.sum<2 * BLML_CCHUNK>(acc, g_sl.red);
            if (threadIdx.x == 0)
                for (int i = 0; i < nc_; ++i) { g_sl.csg[c0 + i] = acc[i]; g_sl.cyg[c0 + i] = acc[BLML_CCHUNK + i]; }
        }
        __syncthreads();
    }

    __device__ FitOut solve_lbfgs_compact() {
        int iE = 0, iU = 1, ig = 2, iEn = 4, iUn = 5, ign = 6;
        const int iq = 3, iVt = 7, iSVt = 8;
        const int mh = P->lbfgs_m, ph = mh + 1;
        FitOut o{};
        Sall(vec(iE), vec(iU));
        rescale(vec(iE), vec(iU));
        double Q = 0.0;
        double L = lval(vec(iE), vec(iU), vec(iVt), &Q);
        lgrad_from(vec(iU), Q, vec(iVt), vec(iSVt), vec(ig));
        double gg = dot(vec(ig), vec(ig)), ee = dot(vec(iE), vec(iE));
        int head = 0, cnt = 0;
        double gam = 1.0;
        long it = 0;
        double res = resid(vec(iE), vec(iU), Q > 0 ? n / sqrt(Q) : 1.0);
        push_rho(res);
        ring_n = 0;
        push_merit(L / n);
        const long cap = P->max_sweeps < 0 ? BLML_HARD_CAP : P->max_sweeps;
        int stop = BLML_STOP_TOL;
        bool reset = false;
        const double ftol = isnan(P->ll_tol) ? 1e-11 : P->ll_tol;
        const int fwin = 10;
        const long n_ = V();
        while (res >= P->tol && it < cap) {
            if (ring_n > fwin && ring_back(0) - ring_back(fwin) < ftol) { stop = BLML_STOP_LLTOL; break; }
            ++it;
            const long long td0 = wall_clock64();
            double *q = vec(iq);
            double slope = 0.0;
            if (cnt > 0) {
                // the small systems, solved by thread 0 with its work arrays in LDS (arrays indexed at run time
                // in private memory would live in scratch); p1 = w and p2 = -gamma u go to LDS for everyone
                if (threadIdx.x == 0) {
                    double *u = g_sl.cu, *w = g_sl.cp1;
                    for (int i = cnt - 1; i >= 0; --i) {            // R u = a (upper triangular)
                        double v = g_sl.ca[i];
                        for (int j = i + 1; j < cnt; ++j) v -= g_sl.cR[i * BLML_CMAX + j] * u[j];
                        u[i] = v / g_sl.cR[i * BLML_CMAX + i];
                    }
                    for (int i = 0; i < cnt; ++i) {                  // w = (D + gamma Y^T Y) u - gamma b
                        double v = g_sl.cR[i * BLML_CMAX + i] * u[i] - gam * g_sl.cb[i];
                        for (int j = 0; j < cnt; ++j) v += gam * g_sl.cYY[i * BLML_CMAX + j] * u[j];
                        w[i] = v;
                    }
                    for (int i = 0; i < cnt; ++i) {                  // R^T p1 = w (lower triangular)
                        double v = w[i];
                        for (int j = 0; j < i; ++j) v -= g_sl.cR[j * BLML_CMAX + i] * w[j];
                        w[i] = v / g_sl.cR[i * BLML_CMAX + i];
                    }
                    for (int i = 0; i < cnt; ++i) g_sl.cp2[i] = -gam * u[i];
                }
                __syncthreads();
                // q = gam g + sum_i (p1_i s_i + p2_i y_i), slope = g . q (the additions in pair order), with the
                // history read as in dots_history
                const double acc = (Vmax & 1) == 0 ? dir_elems<4>(head, ph, cnt, gam, vec(ig), q)
                                                   : dir_elems<2>(head, ph, cnt, gam, vec(ig), q);
                slope = tm.sum1(acc, g_sl.red);
            }
            if (!(slope > 0)) {
                cnt = 0; head = 0;
                const double sc = 1e-3 * sqrt(ee / fmax(gg, 1e-300));
                const double *g = vec(ig);
                double acc = 0.0;
                for (long e = tm.first(); e < n_; e += tm.stride()) {
                    q[e] = g[e] * sc;
                    acc += g[e] * q[e];
                }
                slope = tm.sum1(acc, g_sl.red);
            }
            const long long td1 = wall_clock64();
            t_dir += td1 - td0;
            Sall(q, vec(iSVt));                              // S q: every trial's U is U + t S q
            double t = 1.0, Ln = -INFINITY, Qn = 0.0;
            bool ok = false;
            for (int tr = 0; tr < 40; ++tr) {
                // the trial's Vt = (w / F) Un too: the accepted trial's is the gradient's mat-vec input
                Ln = lerp_lval(vec(iEn), vec(iUn), vec(iE), vec(iU), q, vec(iSVt), t, vec(iVt), &Qn);
                if (Ln >= L + 1e-4 * t * slope) { ok = true; break; }
                t *= 0.5;
            }
            t_ls += wall_clock64() - td1;
            if (!ok) {
                if (reset || cnt == 0) { stop = BLML_STOP_STALL; break; }
                cnt = 0; head = 0;
                reset = true;
                continue;
            }
            reset = false;
            const double res_n = lgrad_resid(vec(iEn), vec(iUn), Qn, vec(iVt), vec(iSVt), vec(ign));
            // the kept pairs' dots with the new gradient
            dots_history(head, ph, cnt, vec(ign));
            // the candidate pair s = En - E, y = g - gn and its scalars
            const int slot = (head + cnt) % ph;
            double a7[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
            {
                float *Sv = hvec(9 + slot), *Yv = hvec(9 + ph + slot);
                const double *En = vec(iEn), *Ec = vec(iE), *g = vec(ig), *gn = vec(ign);
                for (long e = tm.first(); e < n_; e += tm.stride()) {
                    const double sv = En[e] - Ec[e];
                    const double yv = g[e] - gn[e];
                    Sv[e] = (float)sv;
                    Yv[e] = (float)yv;
                    a7[0] += sv * yv;
                    a7[1] += sv * sv;
                    a7[2] += yv * yv;
                    a7[3] += sv * gn[e];
                    a7[4] += yv * gn[e];
                    a7[5] += gn[e] * gn[e];
                    a7[6] += En[e] * En[e];
                }
                tm.sum<7>(a7, g_sl.red);
            }
            // update the small matrices (thread 0) and a, b for the new gradient
            const bool accept = a7[0] > 1e-12 * sqrt(a7[1] * a7[2]);
            if (threadIdx.x == 0) {
                // R's new column and Y^T Y's new row / column for the kept pairs, from the previous a, b
                double *rc = g_sl.crc, *yc = g_sl.cyc;
                for (int i = 0; i < cnt; ++i) {
                    rc[i] = g_sl.ca[i] - g_sl.csg[i];
                    yc[i] = g_sl.cb[i] - g_sl.cyg[i];
                }
                for (int i = 0; i < cnt; ++i) { g_sl.ca[i] = g_sl.csg[i]; g_sl.cb[i] = g_sl.cyg[i]; }
                if (accept) {
                    int c = cnt;
                    if (cnt == mh) {                            // drop the oldest pair
                        for (int i = 1; i < cnt; ++i)
                            for (int j = 1; j < cnt; ++j) {
                                g_sl.cR[(i - 1) * BLML_CMAX + (j - 1)] = g_sl.cR[i * BLML_CMAX + j];
                                g_sl.cYY[(i - 1) * BLML_CMAX + (j - 1)] = g_sl.cYY[i * BLML_CMAX + j];
                            }
                        for (int i = 1; i < cnt; ++i) {
                            g_sl.ca[i - 1] = g_sl.ca[i]; g_sl.cb[i - 1] = g_sl.cb[i];
                            rc[i - 1] = rc[i]; yc[i - 1] = yc[i];
                        }
                        c = cnt - 1;
                    }
                    for (int i = 0; i < c; ++i) {
                        g_sl.cR[i * BLML_CMAX + c] = rc[i];
                        g_sl.cR[c * BLML_CMAX + i] = 0.0;
                        g_sl.cYY[i * BLML_CMAX + c] = yc[i];
                        g_sl.cYY[c * BLML_CMAX + i] = yc[i];
                    }
                    g_sl.cR[c * BLML_CMAX + c] = a7[0];
                    g_sl.cYY[c * BLML_CMAX + c] = a7[2];
                    g_sl.ca[c] = a7[3];
                    g_sl.cb[c] = a7[4];
                }
            }
            if (accept) {
                if (cnt < mh) ++cnt;
                else head = (head + 1) % ph;
                gam = a7[0] / a7[2];
            }
            __syncthreads();
            gg = a7[5];
            ee = a7[6];
            { int tt = iE; iE = iEn; iEn = tt; tt = iU; iU = iUn; iUn = tt; tt = ig; ig = ign; ign = tt; }
            L = Ln;
            Q = Qn;
            res = res_n;
            push_rho(res);
            push_merit(L / n);
        }
        if (res >= P->tol && stop == BLML_STOP_TOL) stop = BLML_STOP_MAXSWEEPS;
        {
            const double Qc = dot(vec(iE), vec(iU));
            double *Ecn = vec(iEn);
            const double *Uc = vec(iU);
            for (int b = (int)tm.first(); b < B; b += (int)tm.stride()) {
                const bool ob = occ(b);
                const double F = Fb(Uc, b);
                const double fac = ob ? (Qc / n) * w[b] / F : 0.0;
                for (int k = 0; k < K; ++k) {
                    const long i = ((long)k * B + b) * nc;
                    for (int qq = 0; qq < nc; ++qq) Ecn[i + qq] = fac * Uc[i + qq];
                }
            }
            tm.sync();
            Snew(vec(iEn), vec(iUn), vec(iE), vec(iU));
            double Qd = 0.0;
            const double Lc = lval(vec(iEn), vec(iUn), nullptr, &Qd);
            if (Lc >= L - 1e-9 * n) {
                int tt = iE; iE = iEn; iEn = tt; tt = iU; iU = iUn; iUn = tt;
            }
        }
        rescale(vec(iE), vec(iU));
        res = resid(vec(iE), vec(iU), 1.0);
        if (iE != 0) { copy(vec(0), vec(iE)); copy(vec(1), vec(iU)); }
        o.n_sweeps = (int)it;
        o.stop = stop;
        o.residual = res;
        o.converged = res < P->tol || stop == BLML_STOP_LLTOL;
        return o;
    }
};
