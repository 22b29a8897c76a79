// The host side of a plan's layout (shared by the engine and by the CPU build of the tests): the host record of a
// plan, the error channel, and layout(), which validates the tables and places every fit's workspace.
#pragma once
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "blml_hip.h"
#include "fit.hpp"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

struct Plan {
    int dtype;
    int nfits, nfacs;
    blml_params_t params;
    std::vector<blml_fit_t> hfits;
    std::vector<blml_factor_t> hfacs;
    blml_fit_t *d_fits = nullptr;
    blml_factor_t *d_facs = nullptr;
    double *d_fdata = nullptr;
    int *d_freqs = nullptr;
    void *d_spec = nullptr;
    void *d_tw = nullptr;
    void *d_spec64 = nullptr;   // FP32 plans: FP64 spectra and twiddles for the finishing pass
    void *d_tw64 = nullptr;
    int lds64_bytes = 0;
    double tick_hz = 1e8;      // wall_clock64() rate
    void *d_ws = nullptr;
    int *d_order = nullptr;
    int *d_queue = nullptr;
    int Lmax = 2;
    int lds_bytes = 0;
    int grid = 1;
    bool spread = false;       // one fit over every CU: cooperative launches of coop workgroups
    int coop = 0;
    unsigned *d_bar = nullptr;  // grid barrier words
    double *d_part = nullptr;   // team partial sums (2 * coop * BLML_NRED)
    bool sep = false;           // spread plan of separable modes: tile passes with dense line operators
    SepPlan *d_sep = nullptr;   // [0] mat-vec kernel, [1] fit kernel
    bool pipe_mv = false;       // the stand-alone mat-vec takes matvec_pipe_kernel (every pass pipelined)
    void *d_sep_tab = nullptr;
    long sep_centro_off = -1;   // FP32 pipelined plans: where sep_centro_table's tables start in d_sep_tab (-1: none)
    long *d_sep_off = nullptr;
    int lds_mv = 0, lds_fit = 0;   // dynamic LDS of the spread mat-vec and fit launches
    int lds_pipe = 0;              // dynamic LDS of matvec_pipe_kernel (largest tile, then every axis's table)
    int sep_npass = 0;             // passes of the stand-alone mat-vec's separable plan (one launch each)
    int res_off = 0;               // LDS-resident intermediates of the batched kernels: byte offset (after the lines)
    int res_mv = 0, res_fit = 0;   // ... and the resident bytes matvec_kernel / fit_kernel launch with (0: none)
    int res_need_mv = 0, res_need_fit = 0;   // the resident bytes the plan's fits allow (the autotuner's options)
    int fft_sched = 0;             // the radix-stage order of the FFT passes (fft_line's sched; autotuned)
    int toep_max = 0;              // FP32 plans: padded line factors up to this length summed directly (OP_TOEP)
    int64_t out_elems = 0, dens_elems = 0;
    int64_t ws_elems = 0;
    size_t esize = 4;
};

int ceil_pow2(int v) {
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

template <typename T>
DevPlan<T> dev_plan(const Plan &pl) {
    DevPlan<T> p;
    p.fits = pl.d_fits;
    p.facs = pl.d_facs;
    p.fdata = pl.d_fdata;
    p.freqs = pl.d_freqs;
    p.spec = (T *)pl.d_spec;
    p.tw = (const cpx<T> *)pl.d_tw;
    p.Lmax = pl.Lmax;
    p.ws = (unsigned char *)pl.d_ws;
    p.nfits = pl.nfits;
    p.params = pl.params;
    p.sep = pl.sep ? pl.d_sep : nullptr;
    p.sep_tab = (const T *)pl.d_sep_tab;
    p.sep_tab_c = pl.sep_centro_off >= 0 ? (const T *)pl.d_sep_tab + pl.sep_centro_off : nullptr;
    p.sep_off = pl.d_sep_off;
    p.res_off = p.res_cap = 0;
    p.fft_sched = pl.fft_sched;
    p.toep_max = sizeof(T) == 4 ? pl.toep_max : 0;
    return p;
}

// the FP64 view of an FP32 plan used by the finishing pass
DevPlan<double> dev_plan64(const Plan &pl) {
    DevPlan<double> p;
    p.fits = pl.d_fits;
    p.facs = pl.d_facs;
    p.fdata = pl.d_fdata;
    p.freqs = pl.d_freqs;
    p.spec = (double *)pl.d_spec64;
    p.tw = (const cpx<double> *)pl.d_tw64;
    p.Lmax = pl.Lmax;
    p.ws = (unsigned char *)pl.d_ws;
    p.nfits = pl.nfits;
    p.params = pl.params;
    p.sep = nullptr;
    p.sep_tab = nullptr;
    p.sep_tab_c = nullptr;
    p.sep_off = nullptr;
    p.res_off = p.res_cap = 0;
    p.fft_sched = pl.fft_sched;
    p.toep_max = 0;
    return p;
}

// validates the tables and fills in spec_off, ws_off, ws_len, scratch_len; returns the LDS each launch needs
int layout(Plan &pl, const double *fdata, int64_t n_fdata, const int32_t *freqs, int64_t n_freqs,
           std::vector<int> &fac_m, std::vector<long> &tmp_off, long &tmp_total, int64_t &spec_total) {
    const blml_params_t &P = pl.params;
    if (!(P.anderson >= 0 && P.anderson <= BLML_MAXAND))
        return fail(BLML_EVALUE, "anderson: 0.." + std::to_string(BLML_MAXAND) + " on this engine");
    if (!(P.lbfgs_m >= 1 && P.lbfgs_m <= BLML_MAXHIST))
        return fail(BLML_EVALUE, "lbfgs_m: 1.." + std::to_string(BLML_MAXHIST) + " on this engine");
    if (!(P.ll_window >= 1 && P.ll_window < BLML_RING))
        return fail(BLML_EVALUE, "ll_window: 1.." + std::to_string(BLML_RING - 1) + " on this engine");
    if (P.hist_cap < 0) return fail(BLML_EVALUE, "hist_cap must be >= 0");
    fac_m.assign((size_t)pl.nfacs * BLML_MAXD, 0);
    std::vector<int> fac_owner(pl.nfacs, -1);
    int lds_elems = 0;
    int Lmax = 2;
    spec_total = 0;
    // factors: sizes and spectra offsets (owner fit gives the axis lengths)
    for (int fi = 0; fi < pl.nfits; ++fi) {
        const blml_fit_t &f = pl.hfits[fi];
        if (f.d < 1 || f.d > BLML_MAXD) return fail(BLML_EVALUE, "fit " + std::to_string(fi) + ": 1 <= d <= 5");
        if (f.K < 1 || f.K > BLML_MAXK) return fail(BLML_EVALUE, "fit " + std::to_string(fi) + ": 1 <= K <= 8");
        if (f.init_K < -2) return fail(BLML_EVALUE, "fit " + std::to_string(fi) + ": init_K >= -2");
        long B = 1;
        for (int j = 0; j < f.d; ++j) {
            if (f.shape[j] < 1) return fail(BLML_EVALUE, "fit " + std::to_string(fi) + ": empty lattice axis");
            B *= f.shape[j];
        }
        if (B != f.B || B > (1L << 30)) return fail(BLML_EVALUE, "fit " + std::to_string(fi) + ": B mismatch");
        for (int k = 0; k < f.K; ++k) {
            if (f.fac_begin[k] < 0 || f.fac_count[k] < 0 || f.fac_begin[k] + f.fac_count[k] > pl.nfacs)
                return fail(BLML_EVALUE, "fit " + std::to_string(fi) + ": factor range");
            for (int q = 0; q < f.fac_count[k]; ++q) {
                const int id = f.fac_begin[k] + q;
                if (fac_owner[id] >= 0 && fac_owner[id] != fi)
                    return fail(BLML_EVALUE, "factor shared between fits");
                fac_owner[id] = fi;
            }
        }
    }
    tmp_off.assign(pl.nfacs, 0);
    tmp_total = 0;
    for (int id = 0; id < pl.nfacs; ++id) {
        blml_factor_t &fa = pl.hfacs[id];
        const int fi = fac_owner[id];
        if (fi < 0) return fail(BLML_EVALUE, "factor " + std::to_string(id) + " belongs to no fit");
        const blml_fit_t &f = pl.hfits[fi];
        if (fa.naxes < 1 || fa.naxes > f.d) return fail(BLML_EVALUE, "factor axes");
        if (fa.kind == BLML_FAC_LINE && fa.naxes != 1) return fail(BLML_EVALUE, "LINE factor with several axes");
        long total = 1;
        int ncirc = 0;
        for (int a = 0; a < fa.naxes; ++a) {
            const int ax = fa.axes[a];
            if (ax < 0 || ax >= f.d || (a > 0 && ax <= fa.axes[a - 1])) return fail(BLML_EVALUE, "factor axes");
            const int m = f.shape[ax];
            fac_m[(size_t)id * BLML_MAXD + a] = m;
            const int tl = fa.tlen[a];
            if (fa.circ[a]) {
                ++ncirc;
                if (tl < 1 || tl > m) return fail(BLML_EVALUE, "circular factor: 1 <= frequencies <= m");
                if (fa.freq_off[a] < 0 || fa.freq_off[a] + tl > n_freqs) return fail(BLML_EVALUE, "freq range");
                for (int q = 0; q < tl; ++q) {
                    const int v = freqs[fa.freq_off[a] + q];
                    if (v < 0 || v >= m) return fail(BLML_EVALUE, "frequency outside [0, m)");
                }
            } else {
                if (tl < 2 * m - 1 || tl != ceil_pow2(tl) || tl < 2)
                    return fail(BLML_EVALUE, "padded factor: power-of-two length >= 2m - 1");
                if (tl > 8192)
                    return fail(BLML_EVALUE, "padded lattice axis longer than 4096 bins: not supported by this engine");
                Lmax = std::max(Lmax, tl);
            }
            total *= tl;
        }
        if (ncirc != 0 && ncirc != fa.naxes)
            return fail(BLML_EVALUE, "a multi-column unit with both circular and padded axes is not supported");
        const bool line_pad = fa.kind == BLML_FAC_LINE && !fa.circ[0];
        const long need = line_pad ? fac_m[(size_t)id * BLML_MAXD] : total;
        if (fa.data_off < 0 || fa.data_off + need > n_fdata) return fail(BLML_EVALUE, "factor data range");
        fa.spec_off = spec_total;
        spec_total += total;
        if (fa.kind == BLML_FAC_GROUP && ncirc == 0) {
            tmp_off[id] = tmp_total;
            tmp_total += 2 * total;
            // the spectrum's forward FFTs over the table's own lattice (spec_kernel)
            for (int a = 0; a < fa.naxes; ++a)
                lds_elems = std::max(lds_elems, pass_lds_elems(OP_FFT_FULL, fa.tlen[a], fa.tlen[a], 0,
                                                               total / fa.tlen[a]));
        }
    }
    // fits: scratch (largest intermediate of any factor chain) and workspace
    int64_t ws = 0;
    for (int fi = 0; fi < pl.nfits; ++fi) {
        blml_fit_t &f = pl.hfits[fi];
        long scratch = f.B;
        // replay every pass of the fit's mat-vecs: scratch = the largest intermediate, LDS = the largest round
        for (int k = 0; k < f.K; ++k) {
            for (int q = 0; q < f.fac_count[k]; ++q) {
                const blml_factor_t &fa = pl.hfacs[f.fac_begin[k] + q];
                if (fa.kind == BLML_FAC_LINE) {
                    const int m = f.shape[fa.axes[0]];
                    const long nl = f.B / m;
                    lds_elems = std::max(lds_elems, fa.circ[0] ? pass_lds_elems(OP_CONV_CIRC, m, 0, fa.tlen[0], nl)
                                                               : pass_lds_elems(OP_CONV_PAD, m, fa.tlen[0], 0, nl));
                    continue;
                }
                long ext[BLML_MAXD];
                for (int j = 0; j < f.d; ++j) ext[j] = f.shape[j];
                auto lines_of = [&](int axis) {
                    long t = 1;
                    for (int j = 0; j < f.d; ++j)
                        if (j != axis) t *= ext[j];
                    return t;
                };
                for (int a = fa.naxes - 1; a >= 0; --a) {
                    const int ax = fa.axes[a], m = f.shape[ax];
                    lds_elems = std::max(lds_elems, fa.circ[a]
                        ? pass_lds_elems(OP_DFT_FWD, m, 0, fa.tlen[a], lines_of(ax))
                        : pass_lds_elems(OP_FFT_FWD, m, fa.tlen[a], 0, lines_of(ax)));
                    ext[ax] = fa.tlen[a];
                    long t = 1;
                    for (int j = 0; j < f.d; ++j) t *= ext[j];
                    scratch = std::max(scratch, t);
                }
                for (int a = 0; a < fa.naxes; ++a) {
                    const int ax = fa.axes[a], m = f.shape[ax];
                    lds_elems = std::max(lds_elems, fa.circ[a]
                        ? pass_lds_elems(OP_DFT_INV, m, 0, fa.tlen[a], lines_of(ax))
                        : pass_lds_elems(OP_FFT_INV, m, fa.tlen[a], 0, lines_of(ax)));
                    ext[ax] = m;
                }
            }
        }
        if (scratch >= (1L << 31) || (long)f.B * 2 >= (1L << 31))
            return fail(BLML_EVALUE, "a fit's transform intermediate exceeds 2^31 values");
        f.scratch_len = scratch;
        const long Vmax = vmax_of(f);
        // bytes: transform scratch, FP64 state pool (also the certificate's 3 B doubles: n_slots >= 3), and for
        // FP32 plans the finishing pass's FP64 scratch
        long len = fin_bytes_off(P, f, (int)pl.esize);
        if (pl.esize == 4) len += 32L * scratch;
        len = align256(len);
        f.ws_off = ws;
        f.ws_len = len;
        ws += len;
        pl.out_elems = std::max<int64_t>(pl.out_elems, f.out_off + Vmax);
        pl.dens_elems = std::max<int64_t>(pl.dens_elems, f.dens_off + f.B);
    }
    pl.ws_elems = ws;
    pl.Lmax = Lmax;
    lds_elems = std::max(lds_elems, 16);
    pl.lds_bytes = lds_elems * (int)(2 * pl.esize);
    pl.lds64_bytes = lds_elems * 16;
    return BLML_OK;
}

constexpr long kSpreadMin = 1L << 18;

// The dense symmetric operator of one LINE factor on m points (FP64), zero-padded to the MFMA block multiple
// P x P of separable.hpp.
void sep_tables(const blml_factor_t &fa, int m, int blk, const double *fdata, const int32_t *freqs,
                std::vector<double> &out) {
    std::vector<double> M((size_t)m * m);
    if (!fa.circ[0]) {
        const double *row = fdata + fa.data_off;
        for (int i = 0; i < m; ++i)
            for (int j = 0; j < m; ++j) M[(size_t)i * m + j] = row[i > j ? i - j : j - i];
    } else {
        std::vector<double> c(m, 0.0);
        const double *sp = fdata + fa.data_off;
        const int32_t *F = freqs + fa.freq_off[0];
        for (int k = 0; k <= m / 2; ++k) {
            double acc = 0.0;
            for (int q = 0; q < fa.tlen[0]; ++q) {
                const long ph = ((long)F[q] * k) % m;
                acc += sp[q] * std::cos(2.0 * M_PI * (double)ph / (double)m);
            }
            c[k] = acc;
            c[(m - k) % m] = acc;
        }
        for (int i = 0; i < m; ++i)
            for (int j = 0; j < m; ++j) M[(size_t)i * m + j] = c[((i - j) % m + m) % m];
    }
    const int P = sep_table_dim(m, blk);
    for (int i = 0; i < P; ++i)
        for (int j = 0; j < P; ++j) out.push_back(i < m && j < m ? M[(size_t)i * m + j] : 0.0);
}

// The centrosymmetric split of one factor's 32 x 32 table M (sep_axis_centro): rows i < 16 hold
// [C_i0 .. C_i15, D_i0 .. D_i15], C_ij = (M_ij + M_{i,31-j}) / 2 and D_ij = (M_ij - M_{i,31-j}) / 2 from the FP64
// operator, the other rows zero.  false (nothing appended) unless the table is 32 x 32 and M[31-i][31-j] == M[i][j]
// to the bit.
bool sep_centro_table(const double *M, int P, std::vector<double> &out) {
    if (P != 32) return false;
    for (int i = 0; i < 32; ++i)
        for (int j = 0; j < 32; ++j)
            if (M[(size_t)(31 - i) * 32 + (31 - j)] != M[(size_t)i * 32 + j]) return false;
    for (int i = 0; i < 32; ++i)
        for (int j = 0; j < 32; ++j) {
            double v = 0.0;
            if (i < 16) {
                const int jj = j & 15;
                const double a = M[(size_t)i * 32 + jj], b = M[(size_t)i * 32 + 31 - jj];
                v = 0.5 * (j < 16 ? a + b : a - b);
            }
            out.push_back(v);
        }
    return true;
}

// One tile pass over axes [lo, hi) in `avail` LDS values, which hold the tile (with its bank padding and dump
// slots) and the staged tables of the pass's axes (stage[a] values each); rows of at least 16 values, or the whole
// inner extent.  false if the group does not fit.
bool sep_make_pass(const int *m, const int *stage, int d, int nc, int lo, int hi, long avail, SepPass &ps) {
    long st_sum = 0;
    for (int a = lo; a < hi; ++a) st_sum += stage[a];
    const long tile_max = (avail - st_sum - 66) * 32 / 33;
    if (tile_max < 64) return false;
    long rs = nc;
    for (int j = hi; j < d; ++j) rs *= m[j];
    long R = 1;
    for (int j = lo; j < hi; ++j) R *= m[j];
    const long wmin = hi == d ? rs : std::min<long>(rs, 16);
    if (R * wmin > tile_max) return false;
    long w;
    if (hi == d || rs * R <= tile_max) {
        w = rs;
    } else {
        w = 1;
        while (w * 2 <= rs && R * w * 2 <= tile_max) w *= 2;
    }
    long nout = 1;
    for (int j = 0; j < lo; ++j) nout *= m[j];
    const long OB = std::min(std::max<long>(1, tile_max / (R * w)), nout);
    ps = SepPass{};
    ps.lo = lo;
    ps.hi = hi;
    ps.OB = (int)OB;
    ps.R = (int)R;
    ps.w = (int)w;
    ps.rs = rs;
    ps.nout = nout;
    ps.nchunks = (rs + w - 1) / w;
    ps.ntiles = ((nout + OB - 1) / OB) * ps.nchunks;
    return true;
}

// Cuts the lattice into groups of consecutive axes, one tile pass each (innermost group first), with as few
// passes as the tile size allows (every cut of d <= 5 axes is tried); false if some axis alone does not fit.
bool sep_plan_passes(const int *m, const int *stage, int d, int nc, long avail, SepPlan &sp) {
    int best = -1, best_np = BLML_SEP_MAXPASS + 1;
    long best_tile = 0;
    for (int cuts = 0; cuts < (1 << (d - 1)); ++cuts) {     // bit j set: a cut between axes j and j + 1
        int np = 0, hi = d;
        bool ok = true;
        long mint = -1;
        SepPass ps;
        for (int lo = d - 1; lo >= 0 && ok; --lo) {
            if (lo == 0 || (cuts >> (lo - 1) & 1)) {
                ok = sep_make_pass(m, stage, d, nc, lo, hi, avail, ps);
                const long t = (long)ps.OB * ps.R * ps.w;
                mint = mint < 0 ? t : std::min(mint, t);
                ++np;
                hi = lo;
            }
        }
        if (ok && (np < best_np || (np == best_np && mint > best_tile))) {
            best = cuts;
            best_np = np;
            best_tile = mint;
        }
    }
    if (best < 0) return false;
    sp = SepPlan{};
    long big = 0;
    int hi = d;
    for (int lo = d - 1; lo >= 0; --lo) {
        if (lo == 0 || (best >> (lo - 1) & 1)) {
            SepPass &ps = sp.pass[sp.npass++];
            sep_make_pass(m, stage, d, nc, lo, hi, avail, ps);
            long st_sum = 0;
            for (int a = lo; a < hi; ++a) st_sum += stage[a];
            const long t = (long)ps.OB * ps.R * ps.w;
            big = std::max(big, t + t / 32 + 1 + 64 + st_sum);
            hi = lo;
        }
    }
    sp.tile = (int)big;     // LDS values of the largest pass: tile, padding, dump slots, staged tables
    return true;
}

// Whether every pass of a mat-vec plan takes the pipelined whole-tile path (the device's sep_pipe_ok, asked of every
// axis of the lattice): then the stand-alone mat-vec runs matvec_pipe_kernel, which compiles in that path alone.
template <typename T>
bool sep_pipe_host(const SepPlan &sp, const int *shape) {
    constexpr int BLK = SepMfma<T>::BLK;
    const long NV = sizeof(T) == 4 ? 64 : 32;
    for (int q = 0; q < sp.npass; ++q) {
        const SepPass &ps = sp.pass[q];
        const long tsize = (long)ps.OB * ps.R * ps.w;
        if (tsize != NV * BLML_NT || ps.nout % ps.OB || ps.rs % ps.w) return false;
        if (!(ps.w == ps.rs || (ps.w <= BLML_NT && BLML_NT % ps.w == 0 && ps.w % (16 / (long)sizeof(T)) == 0)))
            return false;
        if (tsize % ((long)BLML_SEP_FULL_NH * BLK * BLK)) return false;
        long st = ps.w;
        for (int a = ps.hi - 1; a >= ps.lo; --a) {
            if (shape[a] > 1 && !(shape[a] == BLK && (st == 1 || st % 32 == 0))) return false;
            st *= shape[a];
        }
    }
    return true;
}

// The separable path of a spread plan, planned on the host: every factor's dense line operator (zero-padded to the
// MFMA block multiple, in T), the pass plans of the mat-vec [0] and fit [1] kernels sized for the LDS their static
// arrays (st_mv, st_fit) leave free, and whether the stand-alone mat-vec takes matvec_pipe_kernel (static LDS
// st_pipe).  Sets pl.sep, lds_mv, lds_fit, lds_pipe and pipe_mv; false (pl unchanged) when the plan's modes are not
// all LINE factors on axes of at most SepMfma<T>::MMAX points or some axis does not fit a tile.
template <typename T>
struct SepHost {
    std::vector<T> tab;
    std::vector<long> off;
    SepPlan sp[2];
};

template <typename T>
bool sep_plan_host(Plan &pl, const double *fdata, const int32_t *freqs, int max_lds, int st_mv, int st_fit,
                   int st_pipe, SepHost<T> &out) {
    const blml_fit_t &f = pl.hfits[0];
    for (const blml_factor_t &fa : pl.hfacs)
        if (fa.kind != BLML_FAC_LINE || f.shape[fa.axes[0]] > SepMfma<T>::MMAX) return false;
    std::vector<double> tab;
    std::vector<long> off(std::max(1, pl.nfacs), 0);
    int stage[BLML_MAXD] = {0, 0, 0, 0, 0};
    for (int a = 0; a < f.d; ++a) stage[a] = sep_stage_len<T>(f.shape[a]);
    for (int i = 0; i < pl.nfacs; ++i) {
        off[i] = (long)tab.size();
        sep_tables(pl.hfacs[i], f.shape[pl.hfacs[i].axes[0]], SepMfma<T>::BLK, fdata, freqs, tab);
    }
    const int nc = f.complex_g ? 2 : 1;
    const size_t es = sizeof(T);
    const int stat[2] = {st_mv, st_fit};
    for (int k = 0; k < 2; ++k) {
        const long avail = ((long)max_lds - stat[k] - 64) / (long)es - 2;
        if (!sep_plan_passes(f.shape, stage, f.d, nc, avail, out.sp[k])) return false;
        out.sp[k].pq = 0;
    }
    out.tab.assign(tab.begin(), tab.end());
    out.off = off;
    const SepPlan *sp = out.sp;
    pl.lds_mv = std::max(pl.lds_bytes, (int)((sp[0].tile + 2) * es));
    pl.lds_fit = std::max(pl.lds_bytes, (int)((sp[1].tile + 2) * es));
    pl.sep = true;
    long pipe_vals = 0, all_tabs = 0;
    for (int q = 0; q < sp[0].npass; ++q) {
        const long t = (long)sp[0].pass[q].OB * sp[0].pass[q].R * sp[0].pass[q].w;
        pipe_vals = std::max(pipe_vals, t + t / 32 + 1 + 64);
    }
    for (int a = 0; a < f.d; ++a)
        if (f.shape[a] > 1) all_tabs += (long)SepMfma<T>::BLK * (SepMfma<T>::BLK + 1);
    const long lds_pipe = (pipe_vals + all_tabs + 2) * (long)es;
    pl.lds_pipe = (int)std::min<long>(lds_pipe, 1L << 30);
    pl.pipe_mv = st_pipe <= st_mv && lds_pipe + st_pipe <= max_lds && sep_pipe_host<T>(sp[0], f.shape);
    pl.sep_npass = sp[0].npass;
    // FP32 pipelined plans: the centrosymmetric tables after the dense ones, at the same offsets (every axis the
    // pipelined passes act on has 32 points; tables of other factors are never staged there and stay zero)
    pl.sep_centro_off = -1;
    if (sizeof(T) == 4 && pl.pipe_mv) {
        std::vector<double> ct;
        bool ok = true;
        for (int i = 0; i < pl.nfacs && ok; ++i) {
            const int m = f.shape[pl.hfacs[i].axes[0]];
            const int P = sep_table_dim(m, SepMfma<T>::BLK);
            if (m == SepMfma<T>::BLK)
                ok = sep_centro_table(tab.data() + off[i], P, ct);
            else
                ct.insert(ct.end(), (size_t)P * P, 0.0);
        }
        if (ok) {
            pl.sep_centro_off = (long)out.tab.size();
            out.tab.insert(out.tab.end(), ct.begin(), ct.end());
        }
    }
    return true;
}

}  // namespace
