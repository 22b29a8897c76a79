/* C-ABI of the MI355X engine for the binned BLML fit (libblml_hip.so).
 *
 * The reference has no native interface for this path: BLMLBinned reaches its device through a Python seam,
 * `_backend(engine) -> (xp, fft)` (blml_binned.py:40-46), and calls cupyx rfftn/irfftn (blml_binned.py:156-165),
 * CuPy elementwise/reduction ops (blml_binned.py:171-220, 278-288) and cupyx eigsh (blml_binned.py:510-518).
 * Each entry point below replaces one group of those calls; the Python host mirror (blml_amd.binned.BLMLBinned)
 * binds them with ctypes.  All array arguments are device pointers unless marked (host); no torch types appear.
 *
 * Status codes map to the reference's exception types at the Python layer:
 *   BLML_OK 0, BLML_EVALUE 1 -> ValueError, BLML_ERUNTIME 2 -> RuntimeError, BLML_EMEMORY 3 -> MemoryError.
 * blml_last_error() returns the message of the last failing call on the calling thread.
 */
#ifndef BLML_HIP_H
#define BLML_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BLML_MAXD 5   /* lattice dimensions */
#define BLML_MAXK 8   /* modes per fit */

enum { BLML_OK = 0, BLML_EVALUE = 1, BLML_ERUNTIME = 2, BLML_EMEMORY = 3 };
enum { BLML_F32 = 0, BLML_F64 = 1 };
enum { BLML_FAC_LINE = 0, BLML_FAC_GROUP = 1 };
enum { BLML_STOP_TOL = 0, BLML_STOP_STALL = 1, BLML_STOP_LLTOL = 2, BLML_STOP_MAXSWEEPS = 3 };

/* One factor (kernel unit) of a mode's product kernel, as an operator on the lattice.
 * LINE: one axis; padded (circ=0): linear convolution with the even row data[0..m-1] = K(off*h), off = 0..m-1,
 *       done as a zero-padded power-of-two FFT of length tlen >= 2m-1 (any such length gives the same operator
 *       as the reference's 5-smooth pad, blml_binned.py:63-90);
 *       circular (circ=1): the periodic kernel's spectrum restricted to its band: data[0..tlen-1] are the
 *       spectrum values (already divided by the period) at the frequencies freqs[freq_off[0] + q].
 * GROUP: several axes (a multi-column unit: ball, pball).  All circular: data holds the dense spectrum over the
 *       product of the per-axis frequency lists (row-major in increasing axis order).  All padded: data holds the
 *       kernel table over the power-of-two lattice prod tlen (wrapped offsets, zero outside |off| <= m-1); the
 *       library transforms it to the spectrum at plan creation.                                                */
typedef struct {
    int32_t kind;
    int32_t naxes;
    int32_t axes[BLML_MAXD];
    int32_t circ[BLML_MAXD];
    int32_t tlen[BLML_MAXD];
    int64_t freq_off[BLML_MAXD];
    int64_t data_off;
    int64_t spec_off;          /* set by the library */
} blml_factor_t;

/* One fit of a batch (the arguments of BLMLBinned.fit, blml_binned.py:383). */
typedef struct {
    int32_t d, K, complex_g, shared;
    int32_t equal_s;                 /* np.allclose(s_, s_[0]) (blml_binned.py:197) */
    int32_t B;
    int32_t shape[BLML_MAXD];
    int32_t fac_begin[BLML_MAXK];
    int32_t fac_count[BLML_MAXK];
    int32_t init_K;                  /* 0: cold start (blml_binned.py:415-427); >0: warm start with that many modes
                                        (:403-408); -1: the cold start with init's real rank-1 solution (layout
                                        (1, B, 2)) in place of its start solve, the draws as the cold start's;
                                        -2: the same, init being the start solve's initial point (the solve runs) */
    int32_t has_signs;
    double n;                        /* total count */
    double s[BLML_MAXK];             /* K_k(0) */
    int64_t counts_off;              /* element offsets into the batch arrays */
    int64_t draws_off;
    int64_t init_off;
    int64_t signs_off;
    int64_t out_off;                 /* e_hat / Se: K*B*(1+complex_g) elements */
    int64_t dens_off;                /* density: B elements */
    int64_t hist_off;                /* merit / rho history: hist_cap elements */
    int64_t ws_off;                  /* bytes into the plan workspace (set by the library) */
    int64_t ws_len;                  /* bytes (set by the library) */
    int64_t scratch_len;             /* complex elements of each transform scratch buffer (set by the library) */
} blml_fit_t;

/* Solver knobs (BLMLBinned.__init__, blml_binned.py:107-116); one set per batch. */
typedef struct {
    double tol;
    int64_t max_sweeps;              /* < 0: None */
    int32_t anderson;                /* history, <= 5 */
    int32_t stall;
    double phase_noise;
    double ll_tol;                   /* NaN: None */
    int32_t ll_window;
    int32_t solver;                  /* 0 'cd', 1 'lbfgs' */
    int32_t lbfgs_m;                 /* <= 64 */
    int32_t hist_cap;                /* merit / rho history entries kept per fit */
    int32_t spread;                  /* a plan of one fit: 0 spreads it over every CU when it has >= 2^18 values,
                                        1 always, -1 never (one workgroup per fit) */
} blml_params_t;

/* Per-fit outputs (the scalar attributes of a fitted BLMLBinned, blml_binned.py:273-275, 351-354, 438-445). */
typedef struct {
    double loglik;
    double residual;
    int32_t n_sweeps;
    int32_t stop_reason;
    int32_t converged;
    int32_t n_anderson;
    int32_t n_damped;
    int32_t n_gradient;
    int32_t start_sweeps;
    int32_t n_hist;
    int32_t n_matvec;                /* single-mode Gram mat-vecs the fit performed */
    int32_t reserved;
    double seconds[5];               /* device time: whole fit, start solve, mat-vecs, L-BFGS direction, line search */
} blml_result_t;

/* Plan: uploads the fit and factor tables (host arrays), computes every spectrum on the device, allocates the
 * workspace.  Replaces _setup/_build_kernels/_ModeKernel (blml_binned.py:49-154).  dtype: BLML_F32 (FP32
 * transforms, FP64 reductions) or BLML_F64 (the plain FP64 method).                                          */
int blml_plan_create(const blml_fit_t *fits, int32_t nfits, const blml_factor_t *facs, int32_t nfacs,
                     const double *fdata, int64_t n_fdata, const int32_t *freqs, int64_t n_freqs,
                     int32_t dtype, const blml_params_t *params, void **plan);
int blml_plan_destroy(void *plan);

/* _Sall over every fit of the plan: U = [S_k E_k]_k (blml_binned.py:156-168).  E and U hold every fit's
 * K*B*(1+complex_g) values of the plan dtype at the fit's out_off.                                           */
int blml_matvec(void *plan, const void *E, void *U, void *stream);

/* The whole fit of every fit of the plan in one persistent launch: start (real CD under mode 0 and the phase
 * draws, or the warm start), CD + Anderson or L-BFGS, canonical form, outputs (blml_binned.py:223-445).
 * counts, draws, init, signs: double arrays indexed by the fits' offsets (init / signs may be NULL when no fit
 * uses them).  e_out, se_out: plan dtype.  results: device array of nfits blml_result_t.  merit_hist, rho_hist:
 * device doubles, hist_cap per fit (may be NULL).                                                             */
int blml_fit(void *plan, const double *counts, const double *draws, const double *init, const double *signs,
             void *e_out, void *se_out, double *density, blml_result_t *results, double *merit_hist,
             double *rho_hist, void *stream);

/* The plan's autotuner (no counterpart in the reference, whose transforms are library FFTs): times blml_matvec
 * `reps` times under each candidate configuration of the plan -- the radix-stage order of its FFT passes and, for a
 * batched plan, whether a fit's transform intermediates stay in LDS -- and keeps the fastest for every later launch
 * (mat-vec, fit, certificate).  times: BLML_TUNE_MAX doubles, the mean seconds per mat-vec of candidate code
 * (sched * 2 + (resident ? 0 : 1)), NaN for a candidate the plan does not have; choice: the code kept.  Allocates
 * two mat-vec buffers for the duration of the call.                                                           */
#define BLML_TUNE_MAX 4
int blml_plan_autotune(void *plan, int32_t reps, double *times, int32_t *choice, void *stream);

/* certificate(): lambda_max(D^1/2 S_0 D^1/2) / n per fit (blml_binned.py:501-518, which calls eigsh(k=1, tol=tol)) by
 * power iteration on the device, FP64 transforms (an FP32 plan uses its FP64 spectra) and an FP64 Rayleigh
 * quotient; a fit stops when its eigen-residual ||A x - lam x|| <= tol * lam for the unit iterate x (ARPACK's test
 * for a Ritz pair) or after max_iter steps.  e_hat: plan dtype (as written by blml_fit).  lam: nfits doubles (NaN
 * for a fit without a shared kernel).  iters: nfits int32, the steps used, negative when max_iter ended them.   */
int blml_certificate(void *plan, const void *e_hat, const double *counts, int32_t max_iter, double tol,
                     double *lam, int32_t *iters, void *stream);

/* BLMLEmpirical (blml_binned.py:1121-1291): the fit on the observed lattice's normalisation, sum_b g_b^2 Delta = 1,
 * by regularised linear fits of sqrt(h) (CG for reg ridge / l2, FISTA for l1) and, for complex or rank-K fits, the
 * hidden-phase alternation with Anderson mixing.  Knobs of BLMLEmpirical.__init__ (blml_binned.py:1145-1153). */
typedef struct {
    int32_t reg;                     /* 0 'ridge', 1 'l2', 2 'l1' */
    int32_t normalise;               /* 0 'window', 1 'plane' */
    int32_t anscombe;                /* sqrt(w + 3/8) targets */
    int32_t retrieve;                /* 1: the alternation (required for complex_g or K > 1) */
    double lam;
    double tol;                      /* CG / FISTA relative tolerance */
    double outer_tol;
    int32_t max_iter;                /* per linear fit */
    int32_t max_outer;
    int32_t anderson;                /* alternation history (<= the plan's anderson, which sizes its state pool) */
    int32_t reserved;
} blml_emp_params_t;

#define BLML_EMP_AUX_LEN 9   /* aux doubles per fit: max|Kf| of modes 0..BLML_MAXK-1, then Delta = prod(h) */
#define BLML_EMP_OUT_LEN 4   /* emp_out doubles per fit: window norm Zw, plane norm Zp, the outer criterion, scale */

/* The BLMLEmpirical fit of every fit of a batched plan (blml_plan_create with spread = -1 and solver 'cd'), one
 * persistent launch: replaces _fit_counts (blml_binned.py:1218-1291) with its _lin / _cg / _fista / _anderson_mix
 * (:1168-1216, 1463-1526).  counts / draws as for blml_fit (draws: the start's phases and modes, in the order
 * blml_binned.py:1244-1249 takes them); init (may be NULL): for a fit with init_K > 0 its start coefficients as
 * (K, B, 2) doubles -- the raw c of a previous fit (retrieve: every mode, already scaled to this n; else mode 0).
 * aux: BLML_EMP_AUX_LEN doubles per fit.  e_out = C sc, se_out = S C sc (plan dtype), density, results (loglik,
 * residual = the largest linear-fit residual, n_sweeps = CG / FISTA iterations, converged, n_anderson,
 * start_sweeps = alternations, n_hist = J entries); j_hist: hist_cap doubles per fit (J_history_, may be NULL);
 * emp_out: BLML_EMP_OUT_LEN doubles per fit.  A fit whose Zw or Zp is not positive gets scale 0.               */
int blml_empirical(void *plan, const blml_emp_params_t *ep, const double *counts, const double *draws,
                   const double *init, const double *aux, void *e_out, void *se_out, double *density,
                   blml_result_t *results, double *j_hist, double *emp_out, void *stream);

/* ---------------------------------------------------------------------------------------------------------------
 * Pair accumulation over lattice bins (the reference's blml_local_gpu.rows / .cols, blml_local_gpu.py:42-105,
 * 320-361, and their CPU twins blml_local.py:203-282): the kernel value of a pair of bins is the product of its
 * columns' table entries -- an offset row (kind 0: T[base + za - zb + J]), a circular row (1: the min circular
 * distance), a rank table (3: T[base + za * mk + zb], ranks as coordinates), a radial group's row on the r grid
 * (2: linear interpolation at sqrt(sum off^2) * 4) or a periodic ball's torus table (4: the row-major circular
 * offset) -- the layout of blml_kernels.ProductKernel.tables (blml_kernels.py:133-200).  Coordinates are int32
 * (ranks for rank columns); tables REAL = float (f32 = 1) or double.  Sums are FP64. */
#define BLML_PAIR_MAXD 8      /* columns */
#define BLML_PAIR_MAXG 8      /* radial / torus groups */

typedef struct {
    int32_t d, G, f32, reserved;
    int32_t kind[BLML_PAIR_MAXD];
    int32_t grp[BLML_PAIR_MAXD];     /* group of a radial / torus column, else -1 */
    int64_t base[BLML_PAIR_MAXD];
    int64_t J[BLML_PAIR_MAXD];       /* offset rows: half width */
    int64_t mk[BLML_PAIR_MAXD];      /* rank tables: values per side */
    int64_t nbins[BLML_PAIR_MAXD];   /* circular rows and torus columns: bins per period */
    int64_t nT;                      /* values of T */
    int64_t LR;                      /* values per row of RT */
    const void *T;                   /* device: the columns' rows, flat */
    const void *RT;                  /* device: G rows of LR values (radial rows / torus tables) */
} blml_pairtab_t;

/* s_r = sum_j S(ZA_r, ZB_j) w_j (pairs with w_j == 0 skipped), FP64, in a fixed order (deterministic).  ZA (R x d),
 * ZB (M x d) int32, w (M) double, s (R) double: device.  Coordinates must lie inside the tables' spans (checked on
 * the host side of the binding; an index outside its table is clamped on the device and reported). */
int blml_pair_rows(const blml_pairtab_t *tab, const int32_t *ZA, const int32_t *ZB, const double *w, int64_t R,
                   int64_t M, double *s, void *stream);

/* blml_pair_rows without any device allocation.  No counterpart in the reference: it is for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  The same kernels
 * and the same results as blml_pair_rows, with the FP64 partial sums in a fixed array of the library's own on each
 * GPU (BLML_PAIR_NOALLOC_MAX doubles, 16 MiB, held until the process ends): a call whose partials need more -- more
 * than 2^21 rows, or fewer rows whose slices need more -- fails with BLML_EVALUE.  Every call on one GPU shares that
 * array, so calls that overlap on one GPU (other streams, other host threads) race: serialize them. */
#define BLML_PAIR_NOALLOC_MAX (1L << 21)
int blml_pair_rows_noalloc(const blml_pairtab_t *tab, const int32_t *ZA, const int32_t *ZB, const double *w,
                           int64_t R, int64_t M, double *s, void *stream);

/* Derivative rows (the reference's rows(..., grad=True): rows_kernel's dval and its prefix / suffix products,
 * blml_local_gpu.py:21-40, 94-103; blml_kernels.py:165-182 builds the rows).  Entry q is a derivative of unit
 * wunit[q]'s factor: for a column (wgrp[q] < 0) the row at dT[dbase[q]:] of column wcol[q], read at
 * dbase[q] + za * mk + zb for a rank column, else at dbase[q] + za - zb + J (raw offsets, circular columns too); for
 * a group (wgrp[q] = g) the row dR[q] (LR values, as tab->RT): a torus table read as is at the group's row-major
 * circular offset, or a row on the r grid interpolated as the value rows and, when wcol[q] >= 0 (a coordinate
 * shift), times the offset za - zb of column wcol[q].  Units: the columns' factors (dim_unit; -1 for a group's
 * column) and the groups' (grp_unit), U in all.  dT, dR: device, REAL as tab->T. */
#define BLML_PAIR_MAXW 16     /* derivative entries per launch */

typedef struct {
    int32_t nW, U, reserved[2];
    int32_t wunit[BLML_PAIR_MAXW];
    int32_t wcol[BLML_PAIR_MAXW];
    int32_t wgrp[BLML_PAIR_MAXW];
    int32_t dim_unit[BLML_PAIR_MAXD];
    int32_t grp_unit[BLML_PAIR_MAXG];
    int64_t dbase[BLML_PAIR_MAXW];
    int64_t ndT;                     /* values of dT */
    const void *dT;
    const void *dR;                  /* nW rows of tab->LR values (rows of column entries unused) */
} blml_pairgrad_t;

/* blml_pair_rows with derivatives: s as blml_pair_rows, ds[r * nW + q] = sum_j dS_q(ZA_r, ZB_j) w_j (FP64), dS_q
 * the product of entry q's derivative value and the other units' factors (blml_local_gpu.py:94-103). */
int blml_pair_rows_grad(const blml_pairtab_t *tab, const blml_pairgrad_t *grad, const int32_t *ZA, const int32_t *ZB,
                        const double *w, int64_t R, int64_t M, double *s, double *ds, void *stream);

/* Several densities in one pass (the reference's rows_multi(ZA, ZB, Wt, tab, umask, grad),
 * blml_local_gpu.py:107-180, 432-439; the fused joint and marginals): Wt (M x Mm, row-major, device) the densities'
 * weights, umask (Mm x grad->U int32, HOST) 1 where a unit's factor enters density m (the others count as 1).
 * s (R x Mm) = sum_j S_m(ZA_r, ZB_j) Wt[j, m]; with derivative entries (grad->nW > 0, Mm * nW <= BLML_PAIR_MAXW)
 * ds (R x Mm x nW) as blml_pair_rows_grad per density, 0 where the density's mask leaves the entry's unit out.
 * grad->nW = 0: no derivatives (ds unused; grad still gives the units). */
#define BLML_PAIR_MAXM 4      /* densities per fused pass */
int blml_pair_rows_multi(const blml_pairtab_t *tab, const blml_pairgrad_t *grad, const int32_t *ZA, const int32_t *ZB,
                         const double *Wt, const int32_t *umask, int32_t Mm, int64_t R, int64_t M, double *s, double *ds,
                         void *stream);

/* The explicit block out[r * M + j] = S(ZA_r, ZB_j) (REAL); with depmask != 0 also out_m = the product over the
 * columns not in depmask (the reference's joint + marginal pass). */
int blml_pair_cols(const blml_pairtab_t *tab, const int32_t *ZA, const int32_t *ZB, int64_t R, int64_t M,
                   int64_t depmask, void *out, void *out_m, void *stream);

/* The reductions of BLMLBinned.param_grad and log_density (blml_binned.py:543-563, 685-693) over one fit's K modes:
 * E, U = S E (K * B * nc values of the plan dtype, mode-major), derivative mat-vecs dU (nE * B * nc values, entry j a
 * derivative of mode emode[j], HOST array) and counts m (B doubles).  out (device, 4 + 2 nE doubles) = [Z = Re sum
 * conj(E) U, M = sum m, sum wn F, sum_{m > 0} m log max(F / Z + eps, 1e-300), then per entry sum wn Re(conj(U_k) dU_j)
 * and Re sum conj(E_k) dU_j], F = sum_k |U_k|^2, wn = m / max(F + eps Z, 1e-300) on m > 0; FP64, fixed order. */
#define BLML_GRAD_MAXE 16
int blml_grad_dots(int32_t dtype, int32_t K, int32_t nc, int64_t B, const void *E, const void *U, const void *dU,
                   int32_t nE, const int32_t *emode, const double *m, double eps, double *out, void *stream);

/* ---------------------------------------------------------------------------------------------------------------
 * The continuous kernel at arbitrary points: BLMLBinned.sqrt_density / predict / predict_grid (blml_binned.py:452-499,
 * which evaluate the units' closed forms at x - z on the host) as s_q = sum_j K(x_q - z_j) w_j.  A unit: kind 0 sinc
 * fc sinc(fc t) (blml_trivial.py:368-370), 1 slits fc sinc(fc t) - fi sinc(fi t) (:442-444), 2 dirichlet
 * scale (sin(pi M t/P) - sin(pi Mi t/P)) / sin(pi t/P) on the folded distance (:509-515), 3 ball
 * prod fc (1/2)^(D/2) J_{D/2}(pi s) / s^(D/2), s = ||fc t||, D = ncol <= 5 (:135-143, 588-591), 4 periodic ball
 * scale sum_k cos(2 pi sum_i k_i t_i / P_i) over its nharm harmonics (:656-665). */
#define BLML_POINT_MAXU 8     /* units per kernel */
typedef struct {
    int32_t kind;
    int32_t ncol;
    int32_t col[BLML_MAXD];
    int32_t M, Mi;           /* dirichlet: 2N + 1 and the removed inner 2Ni + 1 (0: none) */
    int32_t nharm;           /* periodic ball: harmonics at harm[hoff], ncol int32 each */
    int32_t reserved;
    int64_t hoff;
    double fc[BLML_MAXD];    /* per column of the unit (one-column kinds: fc[0]) */
    double fi;               /* slits: (1 - fill) fc */
    double period[BLML_MAXD];
    double scale;            /* dirichlet 1/P, periodic ball 1/prod P */
} blml_pointunit_t;

/* S (Nq x nw) = sum_j K(X_q - Z_j) W_j for the product of the U units (host array), X (Nq x d), Z (M x d) and W
 * (M x nw, nw = 1 real or 2 complex) doubles and harm int32 (device); FP64 sums in a fixed order. */
int blml_point_rows(const blml_pointunit_t *units, int32_t U, int32_t d, const int32_t *harm, const double *X,
                    const double *Z, const double *W, int32_t nw, int64_t Nq, int64_t M, double *S, void *stream);

/* ---------------------------------------------------------------------------------------------------------------
 * The start search of a complex counts fit (blml_trivial.py:1605-1676 complex_starts, which _BLMLQuickCUDA.fit
 * calls at blml_gpu.py:470-486): hybrid input-output and error-reduction runs on the FFT grid of the bin box.  The
 * host builds the grid, the band, the draws and the kicks as the reference does; every run is one workgroup. */
typedef struct {
    int32_t d;
    int32_t shape[BLML_MAXD];   /* grid sides: powers of two, 2 .. 8192 */
    int32_t box[BLML_MAXD];     /* the data box: cells [0, box_j) of each axis */
    int32_t n_hio, n_er;        /* steps of a random-phase run (HIO, then ER); a symmetry-breaking run: n_er ER */
    int32_t n_rand, n_kick;     /* runs: n_rand random-phase runs first, then n_kick symmetry-breaking runs */
    double beta;                /* the HIO feedback */
    double ab2;                 /* sum of the counts on the box */
    int64_t nocc;               /* occupied cells */
} blml_hio_params_t;

/* Runs every start (all device pointers, cells row-major with the last axis fastest): band (ncell) uint8, amp
 * (ncell) sqrt(counts) on the box, occ (nocc) int64 cells with counts wocc (nocc), phase (n_rand x ncell) uniform
 * draws, kick (n_kick x ncell) the phase kicks; out: P (runs x ncell complex, interleaved doubles) the final
 * band-limited spectra (numpy's fftn convention), gocc (runs x nocc complex) the fields ifftn(P) at the occupied
 * cells, score (runs) sum_occ w log(|g|^2 / sum_box |g|^2). */
int blml_complex_starts(const blml_hio_params_t *prm, const uint8_t *band, const double *amp, const int64_t *occ,
                        const double *wocc, const double *phase, const double *kick, double *P, double *gocc,
                        double *score, void *stream);

/* The exact band-limited field of a coarse-grid start at the original bins (blml_trivial.py:1660-1664):
 * g_b = sum_n c_n exp(2 pi i sum_j dz_bj kf_nj / mq_j) with dz (nb x d) int32 bin offsets from the box corner, kf
 * (nband x d) int32 signed frequency indices, c (nband complex) and g (nb complex) device arrays; mq (d) host
 * int32, the grid side times the coarsening of each axis. */
int blml_band_eval(int32_t d, const int32_t *mq, int64_t nb, int64_t nband, const int32_t *dz, const int32_t *kf,
                   const double *c, double *g, void *stream);

/* ---------------------------------------------------------------------------------------------------------------
 * The warm-start network (no counterpart in the reference: the learned start of the fit, entering through
 * BLMLBinned.fit(init=...), blml_binned.py:403-408).  NCHW images of width 64, even height; bf16 = 16-bit brain
 * floats (uint16 storage); FP32 accumulation.  All pointers are device pointers. */

/* Z = conv3x3(X, W) + b (same padding), W [Cout][Cin][3][3] bf16, X bf16, Z FP32; gelu != 0: A = bf16(GELU(Z)). */
int blml_net_conv(const void *X, const void *Wk, const float *bias, float *Z, void *A, int32_t N, int32_t Cin,
                  int32_t Cout, int32_t H, int32_t W, int32_t gelu, void *stream);
/* dW [Cout][Cin][3][3] = sum over images and pixels of dZ (x) X's 3x3 windows; db [Cout] = sum of dZ (FP32). */
int blml_net_conv_grad_w(const void *X, const void *dZ, float *dW, float *db, int32_t N, int32_t Cin, int32_t Cout,
                         int32_t H, int32_t W, void *stream);
/* Wf [Cin][Cout][3][3] = W [Cout][Cin] flipped in space: the data gradient is blml_net_conv(dZ, Wf). */
int blml_net_flip(const void *Wk, void *Wf, int32_t Cin, int32_t Cout, void *stream);
/* dZ = bf16(dA * GELU'(Z)) (n values). */
int blml_net_gelu_grad(const float *dA, const float *Z, void *dZ, int64_t n, void *stream);
/* masked squared loss: dZ = bf16(2 M (Z - Y) inv_msum), loss_part[b] = partial sums of M (Z - Y)^2 inv_msum
 * (ceil(n / 256) floats). */
int blml_net_mse(const float *Z, const float *Y, const float *M, void *dZ, float *loss_part, int64_t n, float inv_msum,
                 void *stream);
/* Adam step `step` (1-based) on FP32 weights p with gradient g and moments m, v; bf (may be NULL) gets bf16(p). */
int blml_net_adam(float *p, const float *g, float *m, float *v, void *bf, int64_t n, float lr, float b1, float b2,
                  float eps, int32_t step, void *stream);
int blml_net_to_bf16(const float *x, void *y, int64_t n, void *stream);

/* ---- BLMLQuick on counts on a fixed lattice, real rank 1 (blml_gpu.py:249-429; blml_trivial.py:3516-3598) ----
 * Bins Z (B x tab->d int32, ranked as for blml_pair_rows), counts w (B), n = sum w, sbb = S(z, z), optional signs
 * sg (B, +-1; NULL: none); e, Se (B, FP64, device) are updated in place. */
/* One sweep of block-Jacobi CD (blml_gpu.py:340-373): blocks of Bs bins in the order perm (B int32, the sweep's
 * permutation), each moved to its one-coordinate minimisers; safeguard != 0: the block-level step-size rule. */
int blml_quick_sweep(const blml_pairtab_t *tab, const int32_t *Z, int64_t B, const double *w, const double *sg,
                     const int32_t *perm, int64_t Bs, double n, double sbb, int32_t safeguard, double *e, double *Se,
                     void *stream);
/* Se = S e from scratch (exact_Se, blml_gpu.py:249-258): FP64 sums. */
int blml_quick_exact(const blml_pairtab_t *tab, const int32_t *Z, int64_t B, const double *e, const double *sg,
                     double *Se, void *stream);
/* out (4 doubles, device) = (e . Se, sum w log max(e, 1e-300), max residual, count of non-finite e)
 * (objective and residual, blml_gpu.py:262-276). */
int blml_quick_stats(const double *e, const double *Se, const double *w, int64_t B, double n, double *out,
                     void *stream);

/* ---- BLMLQuick's complex / rank-K counts fits (blml_gpu.py:565-744 _psd_solve; blml_trivial.py:2590-2651) ----
 * E, SE: (B, K, nc) FP64 on the device (nc = 2 interleaved complex, 1 real; K <= BLML_QUICK_MAXK), the layout of
 * the reference's (B, K) arrays.  tabs: ntab pair tables, 1 (one kernel shared by every mode) or K (per-mode
 * kernels, mode k's S_k in tabs[k]); Z: ntab host-held pointers to each table's ranked bins (B x d int32, device).
 * w: counts (B), n = sum w; s: S_k(z, z) per mode (K host doubles). */
#define BLML_QUICK_MAXK 8
/* One sweep of _psd_solve (blml_gpu.py:650-679): blocks of Bs bins in the order perm (B int32), each row moved to
 * its own stationary point (closed form for one table, the secular root for per-mode tables), SE += S[:, I] d;
 * safeguard != 0: the step the likelihood allows (the first of 2^-k, k < 31, not lowering L). */
int blml_quick_psd_sweep(const blml_pairtab_t *tabs, int32_t ntab, const int32_t *const *Z, int64_t B, int32_t K,
                         int32_t nc, const double *w, const int32_t *perm, int64_t Bs, double n, const double *s,
                         int32_t safeguard, double *E, double *SE, void *stream);
/* SE = S E from scratch (exact_Se, blml_gpu.py:598-603): FP64 sums. */
int blml_quick_psd_exact(const blml_pairtab_t *tabs, int32_t ntab, const int32_t *const *Z, int64_t B, int32_t K,
                         int32_t nc, const double *E, double *SE, void *stream);
/* out (4 doubles, device) = (Re<E, SE>, sum w log max(F, 1e-300), max residual, non-finite count of E) at
 * (scale E, scale SE) (loglik and residual, blml_gpu.py:606-620). */
int blml_quick_psd_stats(const double *E, const double *SE, const double *w, int64_t B, int32_t K, int32_t nc,
                         double n, double scale, double *out, void *stream);
/* G = 2 (S ((w / max(F, 1e-300)) SE) - (n / Q) SE), Q = Re<E, SE> (lbfgs_psd's grad, blml_trivial.py:2606-2608). */
int blml_quick_psd_grad(const blml_pairtab_t *tabs, int32_t ntab, const int32_t *const *Z, int64_t B, int32_t K,
                        int32_t nc, const double *w, double n, const double *E, const double *SE, double *G,
                        void *stream);
/* out = (coef w / max(F, 1e-300)) SE (the canonical form e = (Q / n) (w / F) S e, blml_trivial.py:2645-2647). */
int blml_quick_psd_wscale(const double *SE, const double *w, int64_t B, int32_t K, int32_t nc, double coef,
                          double *out, void *stream);

/* The elementwise terms of implicit_grad's Hessian-vector product and mixed derivative (blml_binned.py:583-628):
 * FP64 (K, B, nc) vectors, mode-major like blml_fit's outputs; w the fit's counts (occupied: w > 0).
 *   op 0  out = occ ? (w / Fs) dU - (w dF / Fs^2) U : 0, F = sum_k |U_k|^2, Fs = occ ? F : 1, dF = 2 Re sum conj(U) dU
 *         (_hvp's inner term)
 *   op 1  out = dU + (occ ? n w / e^2 : 0) m, e = X (K = 1: the real rank-1 Hessian, _hvp, m the direction D)
 *   op 2  out = -wd U_j + [j == k] ws dU, out2 = ws U_k with ws = occ ? w / Fs : 0, wd = occ ? w dF_k / Fs^2 : 0,
 *         dF_k = 2 Re conj(U_k) dU, dU mode k's B nc values (_mixed's sources)
 *   op 3  out = (m / max(F + eps Q, 1e-300) / M) U (_dP_de's sources; m: B counts)
 *   op 4  out = occ ? X : 0 (_free's mask) */
int blml_ift_elem(int32_t op, int32_t K, int32_t nc, int64_t B, int32_t k, double n, double eps, double Q, double M,
                  const double *U, const double *dU, const double *X, const double *w, const double *m, double *out,
                  double *out2, void *stream);

/* FP64 vector work of the Anderson mixing and L-BFGS above (the reference's CuPy array expressions,
 * blml_gpu.py:705-718, blml_trivial.py:2617-2641).  X: nv <= BLML_VEC_MAXV host-held device pointers of len values.
 * blml_vec_gram: out (ny x ny, device) = the Gram matrix of Y_i = sum_j A[i * nv + j] X_j (A host, ny <= 8).
 * blml_vec_lincomb: mode 0 out = sum_j c_j X_j; mode 1 out = X_{nv-1} - sum_{j<nv-1} c_j (X_{j+1} - X_j)
 * (Anderson's combination); out may be one of the X.
 * blml_lbfgs_twoloop: q = H g by the two-loop recursion over m pairs (S, Y, rho = 1 / s.y, oldest first, host
 * rho), the middle scaling gamma; the dots stay on the device. */
#define BLML_VEC_MAXV 8
int blml_vec_gram(int32_t nv, const double *const *X, int32_t ny, const double *A, int64_t len, double *out,
                  void *stream);
int blml_vec_lincomb(int32_t nv, const double *const *X, const double *c, int32_t mode, int64_t len, double *out,
                     void *stream);
int blml_lbfgs_twoloop(int32_t m, const double *const *S, const double *const *Y, const double *rho, double gamma,
                       int64_t len, const double *g, double *q, void *stream);

/* Which kernel blml_matvec launches for the plan (reporting; no counterpart in the reference): 0 matvec_kernel (one
 * workgroup per fit), 1 matvec_grid_kernel (one fit over the device), 2 matvec_pipe_kernel (one separable fit whose
 * every tile pass is pipelined; it moves 16-byte vectors, so a call whose E or U slices are not 16-byte aligned
 * runs matvec_grid_kernel instead, with the same results). */
int blml_plan_matvec_kernel(void *plan, int32_t *which);

/* Sizes the plan chose (for the caller's buffers and for reporting).  grid: workgroups of the batched launch, or
 * minus the workgroup count when the plan's one fit is spread over the device. */
int blml_plan_info(void *plan, int64_t *out_elems, int64_t *ws_bytes, int32_t *lds_bytes, int32_t *grid);

const char *blml_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* BLML_HIP_H */
