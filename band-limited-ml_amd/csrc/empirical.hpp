// BLMLEmpirical (blml_binned.py:1121-1291) on the device: the fit on the observed lattice's normalisation, one fit
// per team (one workgroup of the batched launch), on the same plan as the binned fit (its tables, spectra, line
// passes and FP64 state pool).
//
//   target      t = sqrt(h), r = sqrt(w) or sqrt(w + 3/8) rescaled to the same total, t = r / sqrt(n Delta)
//               (recomputed where it is read: the same operations give the same values every time);
//   _lin        the regularised fit of one mode to a real target: CG on (S + lam I) c = n t (ridge) or
//               (S^2 + lam I) c = n S t (l2), FISTA on min ||S a - t||^2 + lam ||a||_1 (l1, step 1 / (2 max|Kf|^2),
//               c = n a); complex targets as their real and imaginary parts (blml_binned.py:1168-1181, 1499-1526);
//   retrieve    the hidden-phase alternation (Gerchberg-Saxton with the regularised fit as the projection):
//               targets t u_b from the current directions, every mode's fit warm-started, Anderson type II on the
//               map (kept only when it lowers J_red), stop on |grad J_red| / its first value (l1: the relative
//               change of C) below outer_tol (blml_binned.py:1229-1289);
//   outputs     e = C sc, Se = U sc with sc = 1 / sqrt(Zw) ('window': sum_b g_b^2 Delta = 1) or 1 / sqrt(Zp)
//               ('plane': c'Sc / n^2 = 1), density, loglik (blml_binned.py:1280-1299).
// Vectors are FP64 (the pool); the mat-vecs run in the plan's transform type; every dot is an FP64 team sum.
#pragma once
#include <math.h>

#include "fit.hpp"

static_assert(BLML_EMP_AUX_LEN == BLML_MAXK + 1, "aux: max|Kf| of every mode, then Delta");

// Solves the m x m system in G[:, :m] with right-hand side G[:, m] by Gaussian elimination with partial pivoting
// (every thread solves the same small system); false on a zero pivot or a non-finite solution.
__device__ inline bool solve_small(double (&G)[BLML_MAXAND][BLML_MAXAND + 1], int m, double *x) {
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
    for (int i = m - 1; i >= 0; --i) {
        double v = G[i][m];
        for (int q = i + 1; q < m; ++q) v -= G[i][q] * x[q];
        x[i] = v / G[i][i];
    }
    for (int i = 0; i < m; ++i)
        if (!isfinite(x[i])) return false;
    return true;
}

template <typename T>
struct EmpSolver {
    Team tm;
    MatvecCtx<T> mv;
    const blml_fit_t *f;
    blml_emp_params_t ep;
    int B, K, nc;
    long V;
    double n, Delta, rs1, rs2;     // t_b = sqrt(w_b [+ 3/8]) * rs1 / rs2
    const double *w;
    const double *kfmax;
    double *pool;
    long Vmax;
    int it_total;
    double res_max;
    double *l2tmp;                 // a B-vector the l2 operator's intermediate S x goes through

    __device__ double *vec(int i) const { return pool + (long)i * Vmax; }
    __device__ double tb(long b) const {
        const double r = ep.anscombe ? sqrt(w[b] + 0.375) : sqrt(w[b]);
        return r * rs1 / rs2;
    }
    __device__ double sum1(double v) { return tm.sum1(v, g_sl.red); }
    template <int N>
    __device__ void sum(double (&v)[N]) { tm.sum<N>(v, g_sl.red); }

    // S_k on a real B-vector; the stacked Sall on a V-vector of the fit's value type
    __device__ void Sk(int k, const double *x, double *y) {
        matvec_mode<T, double, double>(mv, f->fac_begin[k], f->fac_count[k], x, y, B, 1);
    }
    __device__ void Sall(const double *E, double *U) {
        for (int k = 0; k < K; ++k) {
            const long o = (long)k * B * nc;
            matvec_mode<T, double, double>(mv, f->fac_begin[k], f->fac_count[k], E + o, U + o, B, nc);
        }
    }

    // ------------------------------------------------------------------ _lin: CG (blml_binned.py:1499-1510)
    // y = A x: S x + lam x (ridge), S (S x) + lam x (l2)
    __device__ void applyA(int k, const double *x, double *y, double *tmp) {
        if (ep.reg == 1) { Sk(k, x, tmp); Sk(k, tmp, y); }
        else Sk(k, x, y);
        for (long e = tm.first(); e < B; e += tm.stride()) y[e] = y[e] + ep.lam * x[e];
        tm.sync();
    }
    // x: the start on entry when has_x0 (else zero), the solution on exit; returns the iterations, res the final
    // relative residual
    __device__ int cg(int k, const double *b, double *x, bool has_x0, double *r, double *p, double *Ap, double *tmp,
                      double &res) {
        double a2[2] = {0.0, 0.0};
        if (has_x0) applyA(k, x, Ap, tmp);
        for (long e = tm.first(); e < B; e += tm.stride()) {
            if (!has_x0) x[e] = 0.0;
            const double re = has_x0 ? b[e] - Ap[e] : b[e];
            r[e] = re;
            p[e] = re;
            a2[0] += re * re;
            a2[1] += b[e] * b[e];
        }
        sum<2>(a2);
        double rs = a2[0];
        double bn = sqrt(a2[1]);
        if (!(bn != 0.0)) bn = 1.0;
        for (int it = 1; it <= ep.max_iter; ++it) {
            applyA(k, p, Ap, tmp);
            double pAp = 0.0;
            for (long e = tm.first(); e < B; e += tm.stride()) pAp += p[e] * Ap[e];
            pAp = sum1(pAp);
            const double a = rs / pAp;
            double rn = 0.0;
            for (long e = tm.first(); e < B; e += tm.stride()) {
                x[e] = x[e] + a * p[e];
                const double re = r[e] - a * Ap[e];
                r[e] = re;
                rn += re * re;
            }
            const double rs_new = sum1(rn);
            if (sqrt(rs_new) / bn < ep.tol) { res = sqrt(rs_new) / bn; return it; }
            const double beta = rs_new / rs;
            for (long e = tm.first(); e < B; e += tm.stride()) p[e] = r[e] + beta * p[e];
            tm.sync();
            rs = rs_new;
        }
        res = sqrt(rs) / bn;
        return ep.max_iter;
    }

    // ------------------------------------------------------------------ _lin: FISTA (blml_binned.py:1513-1526)
    // min ||S a - t||^2 + lam ||a||_1, step 1 / L; x: the start (c0 / n) when has_x0, the solution a on exit
    __device__ int fista(int k, const double *t, double L, double *x, bool has_x0, double *y, double *q, double *g,
                         double &res) {
        for (long e = tm.first(); e < B; e += tm.stride()) {
            if (!has_x0) x[e] = 0.0;
            y[e] = x[e];
        }
        tm.sync();
        double tk = 1.0, dx = 1.0, nx = 1.0;
        const double thr = ep.lam / L;
        for (int it = 1; it <= ep.max_iter; ++it) {
            Sk(k, y, q);
            for (long e = tm.first(); e < B; e += tm.stride()) q[e] = q[e] - t[e];
            tm.sync();
            Sk(k, q, g);
            const double tn = (1.0 + sqrt(1.0 + 4.0 * tk * tk)) / 2.0;
            const double mom = (tk - 1.0) / tn;
            double a2[2] = {0.0, 0.0};
            for (long e = tm.first(); e < B; e += tm.stride()) {
                const double z = y[e] - (2.0 * g[e]) / L;
                const double az = fabs(z) - thr;
                const double xn = az > 0.0 ? (z > 0.0 ? az : (z < 0.0 ? -az : 0.0)) : 0.0;
                const double xo = x[e];
                y[e] = xn + mom * (xn - xo);
                a2[0] += (xn - xo) * (xn - xo);
                a2[1] += xn * xn;
                x[e] = xn;
            }
            sum<2>(a2);
            dx = sqrt(a2[0]);
            nx = sqrt(a2[1]);
            if (!(nx != 0.0)) nx = 1.0;
            tk = tn;
            if (dx / nx < ep.tol) { res = dx / nx; return it; }
        }
        res = dx / nx;
        return ep.max_iter;
    }

    // ------------------------------------------------------------------ the alternation's pieces
    // T = t u (u = U / |U_b| per bin; mode 0's direction 1 where U_b = 0), blml_binned.py:1183-1189; complex values
    // are scaled by the reciprocal, as numpy divides a complex by a real
    __device__ void targets(const double *U, double *Tt) {
        for (long b = tm.first(); b < B; b += tm.stride()) {
            double F = 0.0;
            for (int k = 0; k < K; ++k) {
                const long i = ((long)k * B + b) * nc;
                F += nc == 2 ? hypot(U[i], U[i + 1]) * hypot(U[i], U[i + 1]) : U[i] * U[i];
            }
            const double a = sqrt(F), t = tb(b);
            const double inv = a > 0 ? 1.0 / a : 1.0;
            for (int k = 0; k < K; ++k) {
                const long i = ((long)k * B + b) * nc;
                double ur, ui = 0.0;
                if (nc == 2) { ur = U[i] * inv; ui = U[i + 1] * inv; }
                else ur = a > 0 ? U[i] / a : U[i];
                if (k == 0 && !(a > 0)) { ur = 1.0; ui = 0.0; }
                Tt[i] = t * ur;
                if (nc == 2) Tt[i + 1] = t * ui;
            }
        }
        tm.sync();
    }

    // J_red(C, U) = sum |U / n - T|^2 + lam pen(C), T = targets(U) formed per bin (blml_binned.py:1191-1216)
    __device__ double Jred(const double *C, const double *U) {
        double a2[2] = {0.0, 0.0};
        const double invn = 1.0 / n;
        for (long b = tm.first(); b < B; b += tm.stride()) {
            double F = 0.0;
            for (int k = 0; k < K; ++k) {
                const long i = ((long)k * B + b) * nc;
                F += nc == 2 ? hypot(U[i], U[i + 1]) * hypot(U[i], U[i + 1]) : U[i] * U[i];
            }
            const double a = sqrt(F), t = tb(b);
            const double inv = a > 0 ? 1.0 / a : 1.0;
            for (int k = 0; k < K; ++k) {
                const long i = ((long)k * B + b) * nc;
                double ur, ui = 0.0;
                if (nc == 2) { ur = U[i] * inv; ui = U[i + 1] * inv; }
                else ur = a > 0 ? U[i] / a : U[i];
                if (k == 0 && !(a > 0)) { ur = 1.0; ui = 0.0; }
                if (nc == 2) {
                    const double dr = U[i] * invn - t * ur, di = U[i + 1] * invn - t * ui;
                    const double hh = hypot(dr, di);
                    a2[0] += hh * hh;
                    if (ep.reg == 0) a2[1] += C[i] * U[i] + C[i + 1] * U[i + 1];
                    else if (ep.reg == 1) { const double hc = hypot(C[i], C[i + 1]); a2[1] += hc * hc; }
                    else a2[1] += fabs(C[i]) + fabs(C[i + 1]);
                } else {
                    const double dr = U[i] / n - t * ur;
                    a2[0] += dr * dr;
                    if (ep.reg == 0) a2[1] += C[i] * U[i];
                    else if (ep.reg == 1) a2[1] += C[i] * C[i];
                    else a2[1] += fabs(C[i]);
                }
            }
        }
        sum<2>(a2);
        const double n2 = n * n;
        return a2[0] + (ep.reg == 2 ? ep.lam * a2[1] / n : ep.lam * a2[1] / n2);
    }

    // |grad J_red| = |(2/n) Sall(((a/n - t) / As) U) + pen'| (blml_binned.py:1195-1204); G, SG: V-vector temps
    __device__ double grad_norm(const double *C, const double *U, double *G, double *SG) {
        double amax = 0.0;
        for (long b = tm.first(); b < B; b += tm.stride()) {
            double F = 0.0;
            for (int k = 0; k < K; ++k) {
                const long i = ((long)k * B + b) * nc;
                F += nc == 2 ? hypot(U[i], U[i + 1]) * hypot(U[i], U[i + 1]) : U[i] * U[i];
            }
            amax = fmax(amax, sqrt(F));
        }
        amax = tm.max(amax, g_sl.red);
        const double floor_ = 1e-30 * (amax != 0.0 ? amax : 1.0);
        for (long b = tm.first(); b < B; b += tm.stride()) {
            double F = 0.0;
            for (int k = 0; k < K; ++k) {
                const long i = ((long)k * B + b) * nc;
                F += nc == 2 ? hypot(U[i], U[i + 1]) * hypot(U[i], U[i + 1]) : U[i] * U[i];
            }
            const double a = sqrt(F);
            const double c = (a / n - tb(b)) / fmax(a, floor_);
            for (int k = 0; k < K; ++k) {
                const long i = ((long)k * B + b) * nc;
                G[i] = c * U[i];
                if (nc == 2) G[i + 1] = c * U[i + 1];
            }
        }
        tm.sync();
        Sall(G, SG);
        double acc = 0.0;
        const double n2 = n * n;
        for (long e = tm.first(); e < V; e += tm.stride()) {
            double g = (2.0 / n) * SG[e];
            if (ep.reg == 0) g = g + (2.0 * ep.lam * U[e]) / n2;
            else if (ep.reg == 1) g = g + (2.0 * ep.lam * C[e]) / n2;
            acc += g * g;
        }
        return sqrt(sum1(acc));
    }

    // _lin (blml_binned.py:1168-1181) of mode k to the real target tpart (B values), the start in x when has_x0
    // (raw coefficients), the result c into x.  lt: B-vector temporaries [0..2] CG / FISTA vectors, [3] unused
    // (x), [4] CG's right-hand side; l2tmp holds the l2 operator's intermediate.
    __device__ void lin_into(int k, const double *tpart, bool has_x0, double *x, double *const *lt) {
        double res = 0.0;
        int it;
        double *r = lt[0], *p = lt[1], *Ap = lt[2], *b = lt[4];
        if (ep.reg == 2) {
            const double L = 2.0 * kfmax[k] * kfmax[k];
            if (has_x0) {
                for (long e = tm.first(); e < B; e += tm.stride()) x[e] = x[e] / n;
                tm.sync();
            }
            it = fista(k, tpart, L, x, has_x0, r, p, Ap, res);
            for (long e = tm.first(); e < B; e += tm.stride()) x[e] = n * x[e];
            tm.sync();
        } else {
            if (ep.reg == 1) {
                Sk(k, tpart, Ap);                            // n S t (the l2 normal equations' right-hand side)
                for (long e = tm.first(); e < B; e += tm.stride()) b[e] = n * Ap[e];
            } else {
                for (long e = tm.first(); e < B; e += tm.stride()) b[e] = n * tpart[e];
            }
            tm.sync();
            it = cg(k, b, x, has_x0, r, p, Ap, l2tmp, res);
        }
        it_total += it;
        res_max = fmax(res_max, res);
    }

    // one alternation (blml_binned.py:1250-1255): the targets from U, every mode's fit (real and imaginary parts
    // separately, warm-started from C when given) into CP, then UP = Sall(CP).  Tt: a V-vector temporary; tmp: six
    // B-vector temporaries ([5] the target part, [3] the coefficients).
    __device__ void step(const double *C, const double *U, double *CP, double *UP, double *Tt, double *const *tmp) {
        targets(U, Tt);
        double *tpart = tmp[5], *x = tmp[3];
        double *lt[5] = {tmp[0], tmp[1], tmp[2], x, tmp[4]};
        for (int k = 0; k < K; ++k) {
            for (int part = 0; part < nc; ++part) {
                const long o = (long)k * B * nc + part;
                for (long b = tm.first(); b < B; b += tm.stride()) {
                    tpart[b] = Tt[o + b * nc];
                    if (C) x[b] = C[o + b * nc];
                }
                tm.sync();
                lin_into(k, tpart, C != nullptr, x, lt);
                for (long b = tm.first(); b < B; b += tm.stride()) CP[o + b * nc] = x[b];
                tm.sync();
            }
        }
        Sall(CP, UP);
    }

    // Anderson type II weights (blml_binned.py:1463-1478): gam = (G + 1e-12 tr(G) I)^-1 dF^T f with G = dF^T dF,
    // dF_i the differences of consecutive F history entries (slots sf, oldest first), f the newest F
    // entry; the dots are FP64 team sums (complex values as pairs of reals: the real part of conj(a) b).  False
    // when the system is singular or the weights are not finite (the reference's LinAlgError).
    __device__ bool anderson_gamma(const int *sf, int cnt, const double *fv, double *gam) {
        const int m = cnt - 1;
        double acc[20];
        for (int i = 0; i < 20; ++i) acc[i] = 0.0;
        for (long e = tm.first(); e < V; e += tm.stride()) {
            double dF[BLML_MAXAND];
            for (int i = 0; i < m; ++i) dF[i] = vec(sf[i + 1])[e] - vec(sf[i])[e];
            const double fe = fv[e];
            int q = 0;
            for (int i = 0; i < m; ++i)
                for (int j = i; j < m; ++j) acc[q++] += dF[i] * dF[j];
            for (int i = 0; i < m; ++i) acc[q++] += dF[i] * fe;
        }
        sum<20>(acc);
        double G[BLML_MAXAND][BLML_MAXAND + 1];
        int q = 0;
        double tr = 0.0;
        for (int i = 0; i < m; ++i)
            for (int j = i; j < m; ++j) { G[i][j] = acc[q]; G[j][i] = acc[q]; ++q; }
        for (int i = 0; i < m; ++i) { G[i][m] = acc[q++]; tr += G[i][i]; }
        for (int i = 0; i < m; ++i) G[i][i] += 1e-12 * tr;
        return solve_small(G, m, gam);
    }
};

// The whole BLMLEmpirical fit of plan fit fi by a team (one workgroup).
template <typename T>
__device__ void run_empirical(const DevPlan<T> &p, int fi, const Team &team, const blml_emp_params_t &ep,
                              const double *counts, const double *draws, const double *init, const double *aux,
                              T *e_out, T *se_out, double *density, blml_result_t *results, double *j_hist,
                              double *emp_out, double tick_hz) {
    const blml_fit_t f = p.fits[fi];
    const long long t0 = wall_clock64();
    EmpSolver<T> S;
    S.tm = team;
    make_matvec_ctx<T>(p, f, S.mv, team, nullptr);
    S.f = &p.fits[fi];
    S.ep = ep;
    S.B = f.B;
    S.K = f.K;
    S.nc = f.complex_g ? 2 : 1;
    S.V = (long)f.K * f.B * S.nc;
    S.Vmax = S.V;
    S.n = f.n;
    S.w = counts + f.counts_off;
    S.kfmax = aux + (long)fi * BLML_EMP_AUX_LEN;
    S.Delta = S.kfmax[BLML_MAXK];
    S.pool = (double *)(p.ws + f.ws_off + pool_bytes_off(f, (int)sizeof(T)));
    S.it_total = 0;
    S.res_max = 0.0;
    const int B = f.B, K = f.K, nc = S.nc;
    const long V = S.V;
    // r = sqrt(w [+ 3/8]) * sqrt(n / sum r^2), t = r / sqrt(n Delta) (blml_binned.py:1222-1224)
    {
        double r2 = 0.0;
        for (long b = S.tm.first(); b < B; b += S.tm.stride()) {
            const double r = ep.anscombe ? sqrt(S.w[b] + 0.375) : sqrt(S.w[b]);
            r2 += r * r;
        }
        r2 = S.sum1(r2);
        S.rs1 = sqrt(S.n / r2);
        S.rs2 = sqrt(S.n * S.Delta);
    }
    // slots: 0 C, 1 U, 2 CP, 3 UP, 4 CA, 5 UA, 6 T, 7..12 B-vector temps, 13.. Anderson's X and F histories
    int iC = 0, iU = 1, iCP = 2, iUP = 3, iCA = 4, iUA = 5;
    double *Tt = S.vec(6);
    double *tmp[6] = {S.vec(7), S.vec(8), S.vec(9), S.vec(10), S.vec(11), S.vec(12)};
    S.l2tmp = S.vec(iUA);            // free while the linear fits run (UA is written only after them)
    const int mA = ep.retrieve ? ep.anderson : 0;
    int n_outer = 0, n_and = 0, nJ = 0;
    double crit = 0.0;
    const int hcap = p.params.hist_cap;
    double *Jh = j_hist ? j_hist + f.hist_off : nullptr;
    auto push_J = [&](double J) {
        if (S.tm.lead() && Jh && nJ < hcap) Jh[nJ] = J;
        ++nJ;
    };
    // the target as a B-vector (lin's right-hand side source)
    double *tvec = tmp[5];
    if (!ep.retrieve) {
        // real rank 1: the +sqrt h fit (warm-started from init's raw coefficients when given)
        for (long b = S.tm.first(); b < B; b += S.tm.stride()) tvec[b] = S.tb(b);
        const bool has0 = f.init_K > 0;
        double *x = tmp[3];
        if (has0) {
            const double *E0 = init + f.init_off;
            for (long b = S.tm.first(); b < B; b += S.tm.stride()) x[b] = E0[2 * b];
        }
        S.tm.sync();
        double *lt[5] = {tmp[0], tmp[1], tmp[2], x, tmp[4]};
        S.lin_into(0, tvec, has0, x, lt);
        double *C = S.vec(iC), *U = S.vec(iU);
        for (long b = S.tm.first(); b < B; b += S.tm.stride()) C[b] = x[b];
        S.tm.sync();
        S.Sk(0, C, U);
    } else {
        double *C = S.vec(iC), *U = S.vec(iU);
        bool haveC = false;
        if (f.init_K > 0) {
            // a previous retrieve fit's raw coefficients (already at this n), every mode
            const double *E0 = init + f.init_off;
            for (long b = S.tm.first(); b < B; b += S.tm.stride())
                for (int k = 0; k < K; ++k) {
                    const long i = ((long)k * B + b) * nc;
                    C[i] = E0[((long)k * B + b) * 2];
                    if (nc == 2) C[i + 1] = E0[((long)k * B + b) * 2 + 1];
                }
            S.tm.sync();
            S.Sall(C, U);
            haveC = true;
        } else {
            // the rank-1 +sqrt h fit under mode 0, then phases and the other modes from the seeded draws
            // (blml_binned.py:1240-1249)
            for (long b = S.tm.first(); b < B; b += S.tm.stride()) tvec[b] = S.tb(b);
            S.tm.sync();
            double *x = tmp[3];
            double *lt[5] = {tmp[0], tmp[1], tmp[2], x, tmp[4]};
            const int it0 = S.it_total;
            const double r0 = S.res_max;
            S.lin_into(0, tvec, false, x, lt);
            S.it_total = it0;
            S.res_max = r0;
            double *u0 = tmp[0];
            S.Sk(0, x, u0);
            const double pn = p.params.phase_noise;
            const double *dr = draws + f.draws_off;
            long dp = 0;
            for (long b = S.tm.first(); b < B; b += S.tm.stride()) {
                if (nc == 2) {
                    double sn, cs;
                    sincos(pn * dr[b], &sn, &cs);
                    U[2 * b] = u0[b] * cs;
                    U[2 * b + 1] = u0[b] * sn;
                } else {
                    U[b] = u0[b];
                }
            }
            if (nc == 2) dp = B;
            for (int k = 1; k < K; ++k) {
                for (long b = S.tm.first(); b < B; b += S.tm.stride()) {
                    const long i = ((long)k * B + b) * nc;
                    const double a0 = fabs(u0[b]);
                    if (nc == 2) {
                        U[i] = pn * a0 * (dr[dp + b] / sqrt(2.0));
                        U[i + 1] = pn * a0 * (dr[dp + B + b] / sqrt(2.0));
                    } else {
                        U[i] = pn * a0 * dr[dp + b];
                    }
                }
                dp += (long)B * nc;
            }
            S.tm.sync();
        }
        // the first alternation, then the loop (blml_binned.py:1256-1289)
        S.step(haveC ? S.vec(iC) : nullptr, S.vec(iU), S.vec(iCP), S.vec(iUP), Tt, tmp);
        { int t_ = iC; iC = iCP; iCP = t_; t_ = iU; iU = iUP; iUP = t_; }
        double Jc = S.Jred(S.vec(iC), S.vec(iU));
        push_J(Jc);
        double g0 = S.grad_norm(S.vec(iC), S.vec(iU), S.vec(iCA), S.vec(iUA));
        if (!(g0 != 0.0)) g0 = 1.0;
        int cnt = 0;                       // history entries (X in slots hx0.., F in hf0..: a ring, oldest at ring_head)
        const int hx0 = 13, hf0 = 13 + (mA + 1);
        int ring_head = 0;
        for (int outer = 1; outer <= ep.max_outer; ++outer) {
            n_outer = outer;
            S.l2tmp = S.vec(iUA);
            S.step(S.vec(iC), S.vec(iU), S.vec(iCP), S.vec(iUP), Tt, tmp);
            double JP = S.Jred(S.vec(iCP), S.vec(iUP));
            if (mA > 0) {
                // f = CP - C and C join the histories (the last mA + 1 kept)
                int slot;
                if (cnt < mA + 1) { slot = (ring_head + cnt) % (mA + 1); ++cnt; }
                else { slot = ring_head; ring_head = (ring_head + 1) % (mA + 1); }
                double *hx = S.vec(hx0 + slot), *hf = S.vec(hf0 + slot);
                const double *Cc = S.vec(iC), *Cp = S.vec(iCP);
                for (long e = S.tm.first(); e < V; e += S.tm.stride()) {
                    hx[e] = Cc[e];
                    hf[e] = Cp[e] - Cc[e];
                }
                S.tm.sync();
                if (cnt >= 3) {
                    int sx[BLML_MAXAND + 1], sf[BLML_MAXAND + 1];
                    for (int i = 0; i < cnt; ++i) {
                        sx[i] = hx0 + (ring_head + i) % (mA + 1);
                        sf[i] = hf0 + (ring_head + i) % (mA + 1);
                    }
                    double gam[BLML_MAXAND];
                    if (!S.anderson_gamma(sf, cnt, hf, gam)) {
                        cnt = 0; ring_head = 0;
                    } else {
                        // CA = C + f - sum_i gam_i (dX_i + dF_i)
                        double *CA = S.vec(iCA);
                        const int m = cnt - 1;
                        for (long e = S.tm.first(); e < V; e += S.tm.stride()) {
                            double mix = 0.0;
                            for (int i = 0; i < m; ++i) {
                                const double dX = S.vec(sx[i + 1])[e] - S.vec(sx[i])[e];
                                const double dF = S.vec(sf[i + 1])[e] - S.vec(sf[i])[e];
                                mix += gam[i] * (dX + dF);
                            }
                            CA[e] = Cc[e] + hf[e] - mix;
                        }
                        S.tm.sync();
                        S.Sall(S.vec(iCA), S.vec(iUA));
                        const double JA = S.Jred(S.vec(iCA), S.vec(iUA));
                        if (JA < JP) {
                            int t_ = iCP; iCP = iCA; iCA = t_; t_ = iUP; iUP = iUA; iUA = t_;
                            JP = JA;
                            ++n_and;
                        }
                    }
                }
            }
            double a2[2] = {0.0, 0.0};
            {
                const double *Cc = S.vec(iC), *Cp = S.vec(iCP);
                for (long e = S.tm.first(); e < V; e += S.tm.stride()) {
                    const double d = Cp[e] - Cc[e];
                    a2[0] += d * d;
                    a2[1] += Cp[e] * Cp[e];
                }
            }
            S.template sum<2>(a2);
            double nC = sqrt(a2[1]);
            if (!(nC != 0.0)) nC = 1.0;
            const double dC = sqrt(a2[0]) / nC;
            { int t_ = iC; iC = iCP; iCP = t_; t_ = iU; iU = iUP; iUP = t_; }
            Jc = JP;
            push_J(Jc);
            crit = ep.reg == 2 ? dC : S.grad_norm(S.vec(iC), S.vec(iU), S.vec(iCA), S.vec(iUA)) / g0;
            if (crit < ep.outer_tol) break;
        }
    }
    // outputs (blml_binned.py:1280-1299)
    const double *C = S.vec(iC), *U = S.vec(iU);
    double z2[2] = {0.0, 0.0};
    for (long b = S.tm.first(); b < B; b += S.tm.stride()) {
        for (int k = 0; k < K; ++k) {
            const long i = ((long)k * B + b) * nc;
            if (nc == 2) {
                const double hh = hypot(U[i], U[i + 1]);
                z2[0] += hh * hh;
                z2[1] += C[i] * U[i] + C[i + 1] * U[i + 1];
            } else {
                z2[0] += U[i] * U[i];
                z2[1] += C[i] * U[i];
            }
        }
    }
    S.template sum<2>(z2);
    const double n2 = S.n * S.n;
    const double Zw = z2[0] / n2 * S.Delta, Zp = z2[1] / n2;
    const bool ok = Zw > 0 && Zp > 0;
    const double sc = ok ? 1.0 / sqrt(ep.normalise ? Zp : Zw) : 0.0;
    double ll = 0.0;
    for (long b = S.tm.first(); b < B; b += S.tm.stride()) {
        double F = 0.0;
        for (int k = 0; k < K; ++k) {
            const long i = ((long)k * B + b) * nc;
            const double ur = U[i] * sc, ui = nc == 2 ? U[i + 1] * sc : 0.0;
            e_out[f.out_off + i] = (T)(C[i] * sc);
            se_out[f.out_off + i] = (T)ur;
            if (nc == 2) {
                e_out[f.out_off + i + 1] = (T)(C[i + 1] * sc);
                se_out[f.out_off + i + 1] = (T)ui;
                const double hh = hypot(ur, ui);
                F += hh * hh;
            } else {
                F += ur * ur;
            }
        }
        const double dens = F / n2;
        density[f.dens_off + b] = dens;
        if (S.w[b] > 0) ll += S.w[b] * log(fmax(dens, 1e-300));
    }
    ll = S.sum1(ll);
    if (S.tm.lead()) {
        blml_result_t r;
        r.loglik = ll / S.n;
        r.residual = S.res_max;
        r.n_sweeps = S.it_total;
        r.stop_reason = 0;
        r.converged = S.res_max < ep.tol;
        r.n_anderson = n_and;
        r.n_damped = 0;
        r.n_gradient = 0;
        r.start_sweeps = n_outer;
        r.n_hist = nJ;
        r.n_matvec = 0;
        r.reserved = 0;
        r.seconds[0] = (double)(wall_clock64() - t0) / tick_hz;
        for (int q = 1; q < 5; ++q) r.seconds[q] = 0.0;
        results[fi] = r;
        double *eo = emp_out + (long)fi * BLML_EMP_OUT_LEN;
        eo[0] = Zw;
        eo[1] = Zp;
        eo[2] = crit;
        eo[3] = sc;
    }
    S.tm.sync();
}
