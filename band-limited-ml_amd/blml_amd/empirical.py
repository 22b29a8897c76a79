"""BLMLEmpirical on the MI355X engine (blml_binned.py:1121-1391): the band-limited fit of the square-root histogram
under the observed lattice's normalisation, sum_b g(z_b)^2 Delta = 1, regularised (reg 'ridge' / 'l2' by CG,
'l1' by FISTA), with the hidden-phase alternation for complex or rank-K fits (retrieve).

    m = BLMLEmpirical(reg='ridge', lam=1e-2).fit(counts, modes)
    m.density_, m.e_hat_, m.Se_, m.loglik_, m.c_raw_, m.J_history_, m.score(held_out)
    best = BLMLEmpirical.select(train, val, modes, lams=[1e-1, 1e-2, 1e-3])

The whole fit -- targets, every CG / FISTA iteration, the alternation and its Anderson mixing, the normalisation --
runs in one launch of the engine's blml_empirical on the same plan as BLMLBinned's fit (its tables, spectra and
mat-vecs).  The host validates, builds the tables, the start's draws (numpy default_rng(seed), blml_binned.py:1244-
1249) and the per-mode FISTA bound max|Kf| on the reference's padded lattice."""
import numpy as np

from . import _lib
from .binned import BLMLBinned, FitBatch, _FitSpec
from .units import mode_kf_max

__all__ = ['BLMLEmpirical', 'empirical_many']

_REG = {'ridge': 0, 'l2': 1, 'l1': 2}


class BLMLEmpirical(BLMLBinned):
    """blml_binned.BLMLEmpirical's constructor, fit signature, attributes and errors (blml_binned.py:1145-1391).
    engine: 'hip' (alias 'gpu'); dtype np.float64 (FP64 transforms) or np.float32 (FP32 transforms, FP64 vectors
    and sums)."""

    def __init__(self, reg='ridge', lam=1e-2, anscombe=False, normalise='window', tol=1e-8, max_iter=20000,
                 retrieve=None, outer_tol=1e-6, max_outer=500, anderson=5, phase_noise=0.1, seed=0,
                 engine='hip', dtype=np.float64, workers=-1, verbose=False):
        super().__init__(tol=tol, engine=engine, dtype=dtype, workers=workers, verbose=verbose,
                         phase_noise=phase_noise, seed=seed, anderson=anderson)
        if reg not in ('ridge', 'l2', 'l1') or normalise not in ('window', 'plane'):
            raise ValueError("reg: 'ridge' | 'l2' | 'l1'; normalise: 'window' | 'plane'")
        if int(anderson or 0) > 5:
            raise ValueError("anderson: at most 5 history entries on this engine")
        self.reg, self.lam, self.anscombe, self.normalise, self.max_iter = reg, float(lam), anscombe, normalise, max_iter
        self.retrieve, self.outer_tol, self.max_outer = retrieve, outer_tol, max_outer

    def fit(self, counts, modes, h=1.0, origin=0.0, init=None, complex_g=False):
        """blml_binned.py:1155-1166, 1218-1291."""
        fs, params, ep, aux = self._prepare(counts, modes, h, origin, init, complex_g)
        batch = FitBatch([fs], params, self.dtype, hist_cap=int(self.max_outer) + 2, spread=-1).build()
        batch.run_empirical(ep, aux[None])
        return self._adopt_empirical(fs, batch, batch.results()[0], batch.empirical_outputs()[0], aux, 0)

    def _prepare(self, counts, modes, h, origin, init, complex_g):
        """Host work of a fit: validation, the fit spec (tables, the start's draws), the warm start, the engine's
        knobs (blml_emp_params_t) and aux (max|Kf| per mode, Delta)."""
        fs = _FitSpec(counts, modes, h, origin, complex_g, None, None, self.seed)
        rt = self.retrieve if self.retrieve is not None else (fs.K > 1 or fs.complex_g)
        if not rt and (fs.K > 1 or fs.complex_g):
            raise ValueError("BLMLEmpirical: rank K > 1 or complex_g need retrieve=True (the hidden phase)")
        self.retrieve_ = bool(rt)
        self.complex_g_ = fs.complex_g
        n = fs.n
        # the start: a previous fit's raw coefficients (blml_binned.py:1226-1228, 1235-1237), else the cold start
        # (the rank-1 fit, then the seeded phases and modes: the draws _FitSpec made in that order)
        fs.init, fs.init_K = None, 0
        c_prev = getattr(init, 'c_raw_', None) if init is not None else None
        if c_prev is not None:
            c_prev = np.asarray(c_prev)
            if not self.retrieve_ and c_prev.shape == fs.shape:
                e = np.zeros((1, fs.B, 2))
                e[0, :, 0] = np.real(c_prev).ravel()
                fs.init, fs.init_K = e, 1
            elif self.retrieve_ and getattr(init, 'retrieve_', False) and c_prev.shape == (fs.K,) + fs.shape:
                C = c_prev.astype(complex if fs.complex_g else float) * (n / init.n_raw_)
                e = np.zeros((fs.K, fs.B, 2))
                e[:, :, 0] = np.real(C).reshape(fs.K, -1)
                e[:, :, 1] = np.imag(C).reshape(fs.K, -1) if fs.complex_g else 0.0
                fs.init, fs.init_K = e, fs.K
        aux = np.zeros(_lib.EMP_AUX_LEN)
        for k, sp in enumerate(fs.specs):
            aux[k] = mode_kf_max(sp, fs.shape, fs.h, fs.pad)
        aux[_lib.MAXK] = float(np.prod(fs.h))
        mA = int(self.anderson or 0) if self.retrieve_ else 0
        params = dict(tol=self.tol, max_sweeps=None, anderson=max(mA, int(self.anderson or 0)), stall=1000,
                      phase_noise=self.phase_noise, ll_tol=None, ll_window=100, solver='cd', lbfgs_m=20)
        ep = _lib.EmpParams()
        ep.reg, ep.normalise, ep.anscombe, ep.retrieve = _REG[self.reg], int(self.normalise == 'plane'), \
            int(bool(self.anscombe)), int(self.retrieve_)
        ep.lam, ep.tol, ep.outer_tol = float(self.lam), float(self.tol), float(self.outer_tol)
        ep.max_iter, ep.max_outer, ep.anderson = int(self.max_iter), int(self.max_outer), mA
        return fs, params, ep, aux

    def _adopt_empirical(self, fs, batch, r, emp_out, aux, index):
        """The fitted attributes from the engine's outputs (blml_binned.py:1277-1299); index: the fit in batch."""
        n = fs.n
        Zw, Zp, crit, sc = (float(v) for v in emp_out)
        self._adopt(fs, batch, r, index)
        self.Delta_ = float(aux[_lib.MAXK])
        self.window_norm_, self.plane_norm_ = Zw, Zp
        if not (Zw > 0 and Zp > 0):
            raise ValueError(f"BLMLEmpirical: the fit is zero (reg={self.reg!r}, lam={self.lam:g} removes every "
                             "coefficient); lower lam")
        self.psd_ = False
        self.n_raw_ = n
        self.n_iter_, self.residual_, self.converged_ = r['n_sweeps'], r['residual'], r['residual'] < self.tol
        self.sweeps_ = self.n_iter_
        self.n_outer_ = r['start_sweeps'] if self.retrieve_ else 0
        self.J_history_ = list(r['merit_history']) if self.retrieve_ else []
        if self.retrieve_:
            self.n_anderson_, self.outer_crit_ = r['n_anderson'], crit
        E = np.asarray(self.e_hat_)
        self.c_raw_ = (E / sc) if self.retrieve_ else (E[0] / sc)
        self.sparsity_ = float(np.mean(np.abs(self.c_raw_) > 0)) if self.reg == 'l1' else 1.0
        self.n_samples_ = n
        return self

    def predict_grid(self, r):
        """f on the lattice refined r times per dim (blml_binned.py:472-499): the reference transforms every bin's
        coefficient -- the empirical fit has them on empty bins too -- so the sums here run over every bin."""
        r = int(r)
        if r < 1:
            raise ValueError("predict_grid: r >= 1")
        fine = tuple((m - 1) * r + 1 for m in self.shape_)
        g = np.meshgrid(*[self.origin_[j] + self.h_[j] * np.arange(s) / r for j, s in enumerate(fine)], indexing='ij')
        X = np.stack([x.ravel() for x in g], 1)
        G = self._g_at(X, np.ones(int(np.prod(self.shape_)), bool))
        f = np.abs(G) ** 2 if G.ndim == 1 else (np.abs(G) ** 2).sum(1)
        return f.reshape(fine)

    def score(self, counts):
        """Mean log-probability per count of held-out counts on the same lattice, with f normalised over the box
        (blml_binned.py:1375-1380)."""
        p = np.asarray(self.density_, np.float64)
        p = p / p.sum()
        w = np.asarray(counts, float)
        m = w > 0
        return float((w[m] * np.log(np.maximum(p[m], 1e-300))).sum() / w.sum())

    @classmethod
    def select(cls, train, val, modes, lams, fit_kw=None, **kw):
        """Fit for each lam (warm-started along the grid, largest first) and keep the best held-out score on val
        (blml_binned.py:1382-1391)."""
        best, prev, rows = None, None, []
        for lam in sorted(lams, reverse=True):
            m = cls(lam=lam, **kw).fit(train, modes, init=prev, **(fit_kw or {}))
            s = m.score(val)
            rows.append((lam, s))
            prev = m
            if best is None or s > best[0]:
                best = (s, m)
        best[1].path_ = sorted(rows)
        return best[1]


def empirical_many(jobs, dtype=np.float64, **kw):
    """Many BLMLEmpirical fits (frames, windows, a grid of lattices) in one launch of blml_empirical per kind of fit:
    jobs = [(counts, modes, fit keywords of BLMLEmpirical.fit)], kw the estimator's knobs (BLMLEmpirical.__init__).
    Linear fits and hidden-phase (retrieve) fits run as two launches when both occur.  Returns one fitted
    BLMLEmpirical per job, in order; each equals the same fit made alone."""
    prepared = []
    for counts, modes, fk in jobs:
        fk = dict(fk)
        e = BLMLEmpirical(dtype=dtype, **kw)
        fs, params, ep, aux = e._prepare(counts, modes, fk.pop('h', 1.0), fk.pop('origin', 0.0), fk.pop('init', None),
                                         fk.pop('complex_g', False))
        if fk:
            raise TypeError(f"unknown fit keywords {sorted(fk)}")
        prepared.append((e, fs, params, ep, aux))
    out = [None] * len(prepared)
    for rt in (False, True):
        idx = [i for i, p in enumerate(prepared) if p[0].retrieve_ == rt]
        if not idx:
            continue
        e0, _, params, ep, _ = prepared[idx[0]]
        fss = [prepared[i][1] for i in idx]
        aux = np.stack([prepared[i][4] for i in idx])
        batch = FitBatch(fss, params, dtype, hist_cap=int(e0.max_outer) + 2, spread=-1).build()
        batch.run_empirical(ep, aux)
        res, emp = batch.results(), batch.empirical_outputs()
        for j, i in enumerate(idx):
            e, fs = prepared[i][0], prepared[i][1]
            out[i] = e._adopt_empirical(fs, batch, res[j], emp[j], prepared[i][4], j)
    return out
