"""BLMLQuick(engine='hip') on counts recorded on a fixed lattice: the counts path of the reference's GPU estimator
(blml_gpu.py:61-520 and 565-744, _BLMLQuickCUDA.fit with h in the modes; CPU twin blml_trivial.py:3516-3598).

    q = BLMLQuick(tol=1e-4).fit(Z, make_modes(...) with h per column, sample_weight=counts)
    q.e_hat_, q.Se_, q.n_sweeps_, q.converged_, q.residual_, q.rho_history_, q.obj_history_

Real rank 1: with e_b = w_b c_b the BLML stationarity condition is (S e)_b e_b = n w_b over the occupied bins
(bins with w = 0 are dropped); g(x) = (1/n) sum_b e_b K(x, z_b), and e'Se = n^2 at the solution.  The solver is
the reference's block-Jacobi coordinate descent (blml_gpu.py:320-441): each sweep visits the bins in the order of
default_rng(seed).permutation(B), in blocks of min(block_size, ceil(B / 8)); plain sweeps until one raises the
objective, which is then undone and redone with the block-level step-size safeguard for good; Se is recomputed
exactly (FP64 sums) when the residual first drops under tol; a stall guard (< 1% over 50 sweeps) and max_sweeps
(then e is scaled to e'Se = n^2) end it otherwise.

Complex / rank K (complex_g, or several modes; blml_gpu.py:461-486, 565-744): the real fit above under mode 0's
kernel, then E (B, K) from its solution -- phases exp(i phase_noise N(0, 1)) on column 0, the other columns
phase_noise |e| CN(0, 1), drawn from the same generator -- and block-Jacobi CD of L = sum w log F - n log Q with
each row at its own stationary point (closed form for one shared kernel, the secular root for per-mode kernels),
Anderson mixing over a fixed visiting order and the likelihood safeguard; or solver='lbfgs' (lbfgs_psd,
blml_trivial.py:2590-2651).  Every vector of the solver lives on the device; the host keeps the control flow and
the Anderson step's small least-squares system, as the reference does.

Each block's S[:, I] d is one pair-rows pass of libblml_hip.so (the kernel tables of the reference's
ProductKernel.tables layout in LDS, FP64 sums, no B x |I| block in memory).  e and Se are FP64 on the device;
dtype picks the tables' precision (float32: the reference's f32 policy, products in FP32).

Conditional fits (columns marked dep, blml_gpu.py:67-72, 105-107, 899-953): the joint fit on every column and the
marginal fit of the other columns, predict_marginal and predict_ratio = joint / marginal.  Real rank 1: the
reference fits the marginal on the joint's bins with the kernel of the units off the dep side; bins that differ
only in dep columns then share one marginal kernel row, so at its optimum a bin's coefficient is its merged bin's
split in proportion to the counts (e_b = e_m w_b / W_m).  This engine fits the marginal on the merged counts and
returns that split as e_marg_ (the same fixed point, with fewer bins).  Complex / rank K: two separate fits, as the
reference does.

init= warm-starts a real rank-1 fit from a fit on the same lattice (the previous e at shared bins, each new bin at
its 1-D minimiser) or on another lattice (re-binned: e_b = w_b / g_prev(z_b) where g_prev is not ~0,
blml_gpu.py:213-219); a conditional fit warm-starts its joint and, from a conditional fit over the same dep
columns, its marginal.  sqrt_density / predict evaluate the exact continuous kernel at any points
(blml_point_rows).  A complex fit with n_starts = (n_hio, n_sb) other than (0, 0) and no periodic column searches
its starts first (complex_starts: hybrid input-output runs on the FFT grid of the bin box, every run one workgroup
of blml_complex_starts, blml_trivial.py:1605-1676) and keeps the fit of the largest likelihood (starts_).  Not
built on this engine (they raise NotImplementedError): warm starts of complex / rank-K fits (the reference's branch
for them, blml_gpu.py:196-219, builds FP64 arrays from the complex (B, K) rows of e_hat_ and cannot run either).

Points (modes without h, blml_gpu.py:128-176): binned on the fit's own lattice on the host (bin_points: target
widths bin_scale n^-1/4 / fc, a folded period tiled by whole bins, the lattice kept for a refit whose targets moved
less than (1 + rebin_tol)^(1/4), the signs resolved per bin by majority), then fitted as counts on it."""
import ctypes
import os
import warnings

import numpy as np

from . import _lib, local
from .modes import marginal_modes, parse_modes, strip_dep

__all__ = ['BLMLQuick']


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("BLMLQuick(engine='hip') needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch


def modes_lattice(modes, d):
    """The lattice steps (d,) the modes declare (key 'h' on every column), or None (blml_trivial.py:1361-1374)."""
    hs = []
    for k, m in enumerate(modes):
        per = {}
        for key, c in m.items():
            for j in (key if isinstance(key, tuple) else (key,)):
                per[j] = c
        hk = [per[j].get('h') if j in per else None for j in range(d)]
        if any(v is None for v in hk) and any(v is not None for v in hk):
            raise ValueError(f"mode {k}: h (a fixed lattice) must be given for every column or none")
        hs.append(None if hk[0] is None else np.array(hk, float))
    if any((a is None) != (hs[0] is None) or (a is not None and not np.allclose(a, hs[0])) for a in hs):
        raise ValueError("h (the lattice) must be the same in every mode")
    if hs[0] is not None and not np.all(hs[0] > 0):
        raise ValueError("h must be positive")
    return hs[0]


def strip_lattice(modes):
    return [{key: {a: v for a, v in c.items() if a != 'h'} for key, c in m.items()} for m in modes]


def with_lattice(modes, h):
    """The modes on the fixed lattice h (one step per column; a multi-column unit takes one step for all of its
    columns, blml_trivial.py:1377-1389)."""
    out = []
    for m in modes:
        mm = {}
        for key, c in m.items():
            cols = key if isinstance(key, tuple) else (key,)
            hv = np.broadcast_to(np.asarray(h, float), (max(max(cols) + 1, np.size(h)),))
            if isinstance(key, tuple) and len({float(hv[j]) for j in cols}) > 1:
                raise ValueError(f"columns {key}: one unit, one lattice step")
            mm[key] = dict(c, h=float(hv[cols[0]]))
        out.append(mm)
    return out


def _kind_name(kind):
    return str(kind).partition(':')[0]


def lattice_fold(specs):
    """Per column, whether the shared lattice folds it onto its period (blml_trivial.py:1326-1358): a periodic column
    (dirichlet, or a periodic ball's) whose every mode is periodic there with one period and one kind; any other
    column is binned raw."""
    d = len(specs[0]['kinds'])
    out = []
    for j in range(d):
        per = [_kind_name(sp['kinds'][j]) in ('dirichlet', 'pball') for sp in specs]
        if not any(per):
            out.append(True)
            continue
        out.append(all(per) and len({float(sp['periods'][j]) for sp in specs}) == 1
                   and len({_kind_name(sp['kinds'][j]) for sp in specs}) == 1)
    return out


def lattice_steps(kinds, periods, fold, h_target):
    """The lattice step of each column for the target widths (ProductKernel.steps, blml_kernels.py:55-66): a
    folded periodic column (dirichlet, or a periodic ball's column) tiles its period P with P / max(8,
    ceil(P / h - 1e-9)); every other column takes the target itself."""
    h_t = np.asarray(h_target, float)
    if not np.all(h_t > 0):
        raise ValueError("bin width h must be positive")
    h = np.empty(len(kinds))
    for k, kind in enumerate(kinds):
        if _kind_name(kind) in ('dirichlet', 'pball') and fold[k]:
            P = float(periods[k])
            h[k] = P / max(8, int(np.ceil(P / h_t[k] - 1e-9)))
        else:
            h[k] = h_t[k]
    return h


def quantize(X, kinds, periods, fold, h):
    """Every point snapped to its bin centre (ProductKernel.quantize, blml_kernels.py:74-82): round(x / h) h, and on
    a folded periodic column mod(round(mod(x, P) / h) h, P); -0.0 becomes +0.0 (bins are unique by their bytes)."""
    X = np.atleast_2d(np.asarray(X, float))
    Xq = np.empty_like(X)
    for k, kind in enumerate(kinds):
        hk = float(h[k])
        if _kind_name(kind) in ('dirichlet', 'pball') and fold[k]:
            P = float(periods[k])
            Xq[:, k] = np.mod(np.round(np.mod(X[:, k], P) / hk) * hk, P)
        else:
            Xq[:, k] = np.round(X[:, k] / hk) * hk
    return Xq + 0.0


def resolve_signs(signs, Zc, inverse=None, n_samples=None):
    """+-1 per bin centre, or None (blml_trivial.py:1137-1164): a callable of the centres, one value per centre, or,
    with inverse (sample -> centre), one value per sample reduced to the majority sign of each centre."""
    if signs is None:
        return None
    B = len(Zc)
    if callable(signs):
        v = np.asarray(signs(Zc), float).reshape(-1)
        if len(v) != B:
            raise ValueError("signs(Z) must return one value per centre")
    else:
        v = np.asarray(signs, float).reshape(-1)
        if inverse is not None and n_samples is not None and len(v) == n_samples and len(v) != B:
            sv = np.where(v < 0, -1.0, 1.0)
            tot = np.bincount(inverse, weights=sv, minlength=B)
            cnt = np.bincount(inverse, minlength=B)
            mixed = int((np.abs(tot) < cnt).sum())
            if mixed:
                warnings.warn(f"signs disagree within {mixed} of {B} bins; using the majority sign")
            v = tot
        elif len(v) != B:
            raise ValueError(f"signs has length {len(v)}; expected {B} (centres)"
                             + (f" or {n_samples} (samples)" if n_samples is not None else ""))
    return np.where(v < 0, -1.0, 1.0)


def bin_points(X, specs, sample_weight=None, signs=None, init=None, rebin_tol=1.0, bin_scale=0.1):
    """Points to counts on the fit's lattice (blml_gpu.py:128-176): target widths bin_scale n^-1/4 / fc (the finest
    fc over the modes per column; a radial unit's columns share their finest width, blml_trivial.py:1333-1341), the
    lattice kept from init while every target stays within (1 + rebin_tol)^(1/4) of init's (blml_trivial.py:1403-1410),
    the columns' steps and folds, the points snapped to bin centres and merged (unique centres by their bytes, counts
    or summed weights), the signs resolved per centre.  Returns dict(Z, w, h, signs, h_target, n_lattice, rebinned,
    fc_lattice)."""
    from .modes import units_of
    X = np.atleast_2d(np.asarray(X, np.float64))
    n_rows, d = X.shape
    sw = None if sample_weight is None else np.asarray(sample_weight, np.float64)
    if sw is not None and (sw.shape != (n_rows,) or not np.all(sw > 0)):
        raise ValueError("sample_weight: (n,) positive weights")
    n = float(n_rows) if sw is None else float(sw.sum())
    shared = all(_same_spec(q, specs[0]) for q in specs)
    fc_lat = np.asarray(specs[0]['fc'], float) if shared else np.max([np.asarray(q['fc'], float) for q in specs], 0)
    h_t = bin_scale * n ** -0.25 / fc_lat
    for q in specs:
        for name, dims in units_of(q['kinds']):
            if len(dims) > 1:
                h_t[list(dims)] = h_t[list(dims)].min()
    h0 = getattr(init, 'h_target_', None) if init is not None else None
    keep = (h0 is not None and np.shape(h0) == np.shape(h_t)
            and bool(np.all(np.abs(np.log(h_t / np.asarray(h0))) <= 0.25 * np.log1p(rebin_tol) + 1e-12)))
    rebin = init is not None and not keep
    h_target, n_lattice, h_use = h_t, n, h_t
    if init is not None and not rebin:
        h_target, n_lattice, h_use = np.asarray(init.h_target_), init.n_lattice_, np.asarray(init._h_vec, float)
    kinds, periods = specs[0]['kinds'], specs[0]['periods']
    fold = lattice_fold(specs)
    h = lattice_steps(kinds, periods, fold, h_use)
    Xq = np.ascontiguousarray(quantize(X, kinds, periods, fold, h))
    vdt = np.dtype((np.void, Xq.dtype.itemsize * d))
    _, idx, inv, cnt = np.unique(Xq.view(vdt), return_index=True, return_inverse=True, return_counts=True)
    inv = np.ravel(inv)
    Z = Xq[idx]
    w = cnt.astype(float) if sw is None else np.bincount(inv, weights=sw, minlength=len(idx))
    sg = resolve_signs(signs, Z, inverse=inv, n_samples=n_rows)
    return dict(Z=Z, w=w, h=h, signs=sg, h_target=np.asarray(h_target, float), n_lattice=n_lattice,
                rebinned=rebin, fc_lattice=fc_lat)


class BLMLQuick:
    """The reference's BLMLQuick(engine='gpu') constructor (blml_gpu.py:41-58) and fit signature (:61)."""

    def __init__(self, tol=1e-4, block_size=4096, seed=0, verbose=False, dtype=None, bin_scale=0.1,
                 max_sweeps=None, solver='cd', lbfgs_m=20, engine='hip'):
        if solver not in ('cd', 'lbfgs'):
            raise ValueError("solver: 'cd' (block-Jacobi CD with Anderson mixing) or 'lbfgs'")
        if engine not in ('hip', 'gpu'):
            raise ValueError("engine: 'hip' (alias 'gpu'): this estimator runs on the MI355X engine only")
        self.solver, self.lbfgs_m = solver, int(lbfgs_m)
        self.tol, self.max_sweeps, self.block_size = tol, max_sweeps, block_size
        self.seed, self.verbose, self.bin_scale, self.engine = seed, verbose, bin_scale, engine
        self.dtype = np.dtype(dtype or os.environ.get('BLML_DTYPE') or 'float32')
        # the complex / rank-K fit's knobs the reference reads with getattr defaults (blml_gpu.py:576, 656, 472)
        self.phase_noise, self.anderson, self.n_starts = 0.1, 5, (3, 2)

    # ---------------------------------------------------------------------------------------------------- fit
    def fit(self, X, modes, init=None, rebin_tol=1.0, signs=None, complex_g=False, sample_weight=None):
        Z = np.atleast_2d(np.asarray(X, dtype=np.float64))
        d = Z.shape[1]
        h = modes_lattice(modes, d)
        if h is None:
            # points: binned on the fit's lattice first (blml_gpu.py:128-176), then the counts fit on it
            specs, _ = parse_modes(modes, d, allow_dep=True)
            bp = bin_points(Z, specs, sample_weight, signs, init, rebin_tol, self.bin_scale)
            self._fit_counts(bp['Z'], bp['w'], bp['h'], modes, init, bp['signs'], complex_g)
            self.h_target_, self.n_lattice_, self.fc_lattice_ = bp['h_target'], bp['n_lattice'], bp['fc_lattice']
            self.rebinned_ = bp['rebinned']
            return self
        w = np.ones(len(Z)) if sample_weight is None else np.asarray(sample_weight, float)
        if w.shape != (len(Z),) or not np.all(w >= 0):
            raise ValueError("sample_weight: (n,) non-negative counts")
        return self._fit_counts(Z, w, h, strip_lattice(modes), init, signs, complex_g)

    def _fit_counts(self, Z, w, h, modes, init, signs, complex_g):
        """The fit of counts w at bin centres Z on the lattice of steps h (d,); modes without h."""
        d = Z.shape[1]
        h = np.asarray(h, float)
        specs, dep = parse_modes(modes, d, allow_dep=True)
        if dep is not None:
            return self._fit_conditional(Z, w, h, modes, specs, dep, d, init, signs, complex_g)
        psd = bool(complex_g) or len(specs) > 1
        if psd:
            if signs is not None:
                raise ValueError("complex_g / rank > 1 / modes estimate the phases; do not pass signs as well")
            if init is not None:
                raise NotImplementedError("BLMLQuick(engine='hip'): warm starts of complex / rank-K fits are not "
                                          "built")
            if len(specs) > _lib.MAXK:
                raise ValueError(f"at most {_lib.MAXK} modes on this engine")
        keep = w > 0
        sg = None
        if signs is not None:
            sg = np.sign(np.asarray(signs(Z) if callable(signs) else signs, float)).reshape(-1)
            if sg.shape != (len(Z),):
                raise ValueError("signs: one value per bin (or a callable of the bin centres)")
            sg = sg[keep]
            if np.any(sg == 0):
                raise ValueError("signs: nonzero on every occupied bin")
        Z, w = Z[keep], w[keep]
        B = len(Z)
        if B == 0:
            raise ValueError("sample_weight: no occupied bin")
        n = float(w.sum())
        Zi = np.round(Z / h).astype(np.int64)
        if np.abs(Zi * h - Z).max() > 1e-6 * h.max():
            raise ValueError("bin centers must lie on the lattice h * integers")
        e0, new = self._warm(init, h, Zi, Z, w, sg)
        sp = specs[0]
        tab = local.lattice_tables(sp['kinds'], sp['fc'], sp['periods'], h, Zi)
        f32 = self.dtype == np.float32
        rng = np.random.default_rng(self.seed)
        out = self._solve(Zi, w, n, sg, tab, f32, e0, new, rng)
        self.mode_specs_, self.aa_accepted_, self.starts_ = None, None, None
        K = len(specs)
        if psd:
            shared = all(_same_spec(q, specs[0]) for q in specs)
            starts = None
            nh, ns = self.n_starts
            if complex_g and nh + ns > 0:
                # the start search on mode 0's band (blml_gpu.py:470-486)
                if any(str(k).partition(':')[0] in ('dirichlet', 'pball') for k in specs[0]['kinds']):
                    warnings.warn("complex_g: start search not implemented for periodic dims; using one start")
                else:
                    starts = complex_starts(Zi, w, h, specs[0]['fc'], specs[0]['kinds'], nh, ns, seed=self.seed)
                    if starts is None:
                        warnings.warn("complex_g: FFT grid of the bin box too large for the start search; using "
                                      "one start")
            out = self._psd(Zi, w, n, specs if not shared else specs[:1], h, f32, out[0], K, bool(complex_g), rng,
                            starts)
            self.mode_specs_ = None if shared else [dict(q) for q in specs]
        e, Se, n_sweeps, converged, max_rho, rho_hist, obj_hist, safeguard = out
        if sg is not None:
            e = e * sg
        self.Se_ = None if Se is None else Se * (sg if sg is not None else 1.0)
        self.signs_ = sg
        self.e_hat_, self.Z_, self.w_ = e, Z, w
        self.n_samples_, self.n_bins_, self.dim_ = n, B, d
        self.n_sweeps_, self.converged_, self.residual_ = n_sweeps, converged, max_rho
        self.safeguard_used_ = safeguard
        self.rho_history_, self.obj_history_ = rho_hist, obj_hist
        self.complex_g_, self.rank_, self.dep_ = bool(complex_g), K, None
        self._joint = self._marg = self.e_marg_ = None
        self.kinds_, self.periods_, self._fc_arr = sp['kinds'], sp['periods'], sp['fc']
        self._specs = specs
        self._h_vec, self._Zi, self._tab = h, Zi, tab
        # the lattice a later refit of points keeps or leaves (blml_gpu.py:158-168): a fixed lattice is its own target
        same = init is not None and getattr(init, 'h_target_', None) is not None and not self.rebinned_
        self.h_target_ = np.asarray(init.h_target_, float) if same else np.asarray(h, float).copy()
        self.n_lattice_ = init.n_lattice_ if same else n
        self.fc_lattice_ = np.asarray(sp['fc'], float) if shared_spec(specs) else np.max(
            [np.asarray(q['fc'], float) for q in specs], 0)
        return self

    def _fit_conditional(self, Z, w, h, modes, specs, dep, d, init, signs, complex_g):
        """The joint fit on every column and the marginal fit of the others (blml_gpu.py:105-107, 197-217, 899-921):
        the marginal's counts are the joint's summed over the dep columns; its modes the joint's restricted to the
        other columns.  Sets the joint's attributes on self, dep_, _joint, _marg and, for real rank 1, e_marg_ on
        the joint's bins (the merged bin's coefficient split by the counts) with marg_converged_, marg_sweeps_ and
        marg_residual_."""
        from .modes import units_of
        if len(dep) == d:
            raise ValueError(f"dep {dep} must name some, but not all, of the {d} dims")
        for _, dims in units_of(specs[0]['kinds']):
            if 0 < sum(j in dep for j in dims) < len(dims):
                raise ValueError(f"a radial group over columns {dims} straddles dep and conditioning columns")
        if signs is not None:
            raise ValueError("signs with dep=... (the marginal's g has its own signs)")
        rest = [j for j in range(d) if j not in dep]
        kw = dict(tol=self.tol, block_size=self.block_size, seed=self.seed, verbose=self.verbose, dtype=self.dtype,
                  bin_scale=self.bin_scale, max_sweeps=self.max_sweeps, solver=self.solver, lbfgs_m=self.lbfgs_m)

        def make():
            q = BLMLQuick(**kw)
            q.phase_noise, q.anderson, q.n_starts = self.phase_noise, self.anderson, self.n_starts
            return q

        ij = init if getattr(init, '_joint', None) is None else init._joint
        im = init._marg if init is not None and getattr(init, 'dep_', None) == list(dep) else None
        joint = make()._fit_counts(Z, w, h, strip_dep(modes), ij, None, complex_g)
        Zm, inv = np.unique(np.ascontiguousarray(Z[:, rest]), axis=0, return_inverse=True)
        inv = np.ravel(inv)
        wm = np.bincount(inv, weights=w, minlength=len(Zm))
        marg = make()._fit_counts(Zm, wm, h[rest], marginal_modes(modes, rest), im, None, complex_g)
        self.__dict__.update(joint.__dict__)
        self._joint, self._marg, self.dep_ = joint, marg, list(dep)
        if np.ndim(joint.e_hat_) == 1 and not np.iscomplexobj(joint.e_hat_):
            keep = w > 0                                 # the joint's bins, in its order
            mrow = inv[keep]
            occ = np.flatnonzero(wm > 0)                 # the marginal's bins, in its order
            pos = np.full(len(Zm), -1)
            pos[occ] = np.arange(len(occ))
            em = np.asarray(marg.e_hat_, np.float64)[pos[mrow]]
            self.e_marg_ = em * w[keep] / wm[mrow]
            self.marg_converged_, self.marg_sweeps_ = marg.converged_, marg.n_sweeps_
            self.marg_residual_ = marg.residual_
        return self

    def predict_marginal(self, Xrest):
        """The marginal density of the non-dep columns at points given in those columns (blml_gpu.py:923-932)."""
        if getattr(self, 'dep_', None) is None:
            raise RuntimeError("fit with dep=... to get the marginal")
        return self._marg.predict(Xrest)

    def predict_ratio(self, Xq):
        """The conditional density of the dep columns given the others, joint / marginal (blml_gpu.py:939-943)."""
        if getattr(self, 'dep_', None) is None:
            raise RuntimeError("fit with dep=... to get the marginal")
        Xq = np.atleast_2d(np.asarray(Xq, float))
        rest = [k for k in range(self.dim_) if k not in self.dep_]
        return self._joint.predict(Xq) / np.maximum(self.predict_marginal(Xq[:, rest]), 1e-300)

    def _warm(self, init, h, Zi, Z, w, sg):
        """init= for a real rank-1 fit.  On the same lattice (blml_gpu.py:145-153, 200-208): the previous fit's e at
        the bins both share (by lattice coordinates; |e| with signs), 0 at new bins (mask `new`, started by _solve).
        On another lattice, re-binned (blml_gpu.py:213-219): e_b = w_b / g_prev(z_b) (s g_prev with signs) from the
        previous fit's exact kernel at the new bins, the bins where g_prev ~ 0 (below 1e-3 of its largest size)
        started by _solve."""
        self.rebinned_, self.n_new_bins_ = False, 0
        if init is None:
            return None, None
        hp = getattr(init, '_h_vec', None)
        if hp is None or getattr(init, 'e_hat_', None) is None:
            raise ValueError("init: a fitted BLMLQuick")
        if np.ndim(init.e_hat_) != 1 or np.iscomplexobj(init.e_hat_):
            raise NotImplementedError("BLMLQuick(engine='hip'): complex / rank-K warm starts are not built")
        if np.shape(hp) != np.shape(h) or not np.allclose(hp, h):
            g = np.asarray(init.sqrt_density(Z), float)
            if sg is not None:
                g = g * sg
            new = ~(g > 1e-3 * max(float(np.abs(g).max()), 1e-300))
            self.rebinned_, self.n_new_bins_ = True, int(new.sum())
            return np.where(new, 0.0, w / np.where(new, 1.0, g)), new
        key = lambda Z: [z.tobytes() for z in np.ascontiguousarray(Z, np.int64)]
        prev = dict(zip(key(init._Zi), np.asarray(init.e_hat_, float)))
        e0 = np.array([prev.get(k, np.nan) for k in key(Zi)])
        new = np.isnan(e0)
        if sg is not None:
            e0 = np.abs(e0)
        self.n_new_bins_ = int(new.sum())
        return np.where(new, 0.0, e0), new

    def _solve(self, Zi, w, n, sg, tab, f32, e0=None, new=None, rng=None):
        torch = _torch()
        lib = _lib.load()
        B = len(Zi)
        t, dZ, _, keep_tab = local._prepare(Zi, Zi, tab, f32, False)      # checks every pair's table span
        st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        tp = ctypes.byref(t)
        P = lambda x: None if x is None else ctypes.c_void_p(x.data_ptr())
        # S(z, z): the tables' own value at a zero offset
        one = torch.ones(1, dtype=torch.float64, device='cuda')
        s1 = torch.zeros(1, dtype=torch.float64, device='cuda')
        _lib.check(lib.blml_pair_rows(tp, P(dZ), P(dZ), P(one), 1, 1, P(s1), st))
        sbb = float(s1.item())
        if not sbb > 0:
            raise RuntimeError(f"S(z, z) = {sbb}: the kernel's value at zero must be positive")
        dw = torch.from_numpy(np.ascontiguousarray(w, np.float64)).cuda()
        dsg = None if sg is None else torch.from_numpy(np.ascontiguousarray(sg, np.float64)).cuda()
        e = (torch.ones(B, dtype=torch.float64, device='cuda') if e0 is None else
             torch.from_numpy(np.ascontiguousarray(e0, np.float64)).cuda())
        Se = torch.empty_like(e)
        stats = torch.zeros(4, dtype=torch.float64, device='cuda')
        rng = np.random.default_rng(self.seed) if rng is None else rng
        Bs = max(1, min(int(self.block_size), -(-B // 8)))

        def exact():
            _lib.check(lib.blml_quick_exact(tp, P(dZ), B, P(e), P(dsg), P(Se), st))

        def measure():
            _lib.check(lib.blml_quick_stats(P(e), P(Se), P(dw), B, n, P(stats), st))
            a, lw, rho, bad = stats.cpu().numpy()
            return (0.5 * a - n * lw) / n, float(rho), bad == 0

        def sweep(safe):
            perm = torch.from_numpy(rng.permutation(B).astype(np.int32)).cuda()
            _lib.check(lib.blml_quick_sweep(tp, P(dZ), B, P(dw), P(dsg), P(perm), Bs, n, sbb, int(safe), P(e),
                                            P(Se), st))

        exact()
        if new is not None and new.any():
            # new bins of a warm start: the exact 1-D minimiser of the objective in e_b with the others fixed,
            # root of S_bb e^2 + b e - n w_b with b = (S e)_b (blml_gpu.py:304-320)
            J = torch.from_numpy(np.flatnonzero(new)).cuda()
            b, cc = Se[J], n * dw[J]
            disc = torch.sqrt(b * b + 4.0 * sbb * cc)
            e[J] = torch.where(b >= 0, 2.0 * cc / (b + disc), (-b + disc) / (2.0 * sbb))
            exact()
        obj_prev = measure()[0]
        safeguard, done, conv = False, False, False
        n_sweeps, max_rho = 0, np.inf
        rho_hist, obj_hist = [], []
        cap = self.max_sweeps
        while not done:
            if cap is not None and n_sweeps >= cap:
                break
            n_sweeps += 1
            save = (e.clone(), Se.clone())
            sweep(safeguard)
            obj_now, rho, finite = measure()
            if not safeguard and obj_now > obj_prev + 1e-10 * abs(obj_prev):
                e.copy_(save[0])
                Se.copy_(save[1])
                safeguard = True
                sweep(True)
                obj_now, rho, finite = measure()
            obj_prev = obj_now
            if not finite:
                max_rho = np.inf
                rho_hist.append(np.inf)
                break
            if rho < self.tol:
                exact()
                obj_ex, rho, _ = measure()
                obj_now = obj_ex if rho < self.tol else obj_now
            max_rho = rho
            rho_hist.append(rho)
            obj_hist.append(obj_now)
            if self.verbose:
                print(f"  sweep {n_sweeps}: max|rho|={rho:.2e} Phi/n={obj_hist[-1]:.6f}", flush=True)
            if rho < self.tol:
                conv = done = True
            elif len(rho_hist) > 50 and rho_hist[-1] > 0.99 * min(rho_hist[:-50]):
                done = True
        exact_se = conv
        if cap is not None and not conv and np.isfinite(max_rho):
            exact()                                 # capped: scaled to e'Se = n^2 (blml_gpu.py:432-439)
            _lib.check(lib.blml_quick_stats(P(e), P(Se), P(dw), B, n, P(stats), st))
            q = float(stats[0].item())
            sc = n / np.sqrt(max(q, 1e-300))
            e.mul_(sc)
            Se.mul_(sc)
            max_rho = measure()[1]
            conv = max_rho < self.tol
            exact_se = True
        del keep_tab
        return (e.cpu().numpy(), Se.cpu().numpy() if exact_se else None, n_sweeps, conv, max_rho, rho_hist,
                obj_hist, safeguard)

    # ------------------------------------------------------------------------------------- complex / rank K
    def _psd(self, Zi, w, n, tspecs, h, f32, e_real, K, complex_g, rng, starts=None):
        """The rank-K / complex fit from the real solution e_real (blml_gpu.py:565-744): tspecs holds one spec
        (a kernel shared by every mode) or one per mode.  Returns _solve's tuple with E (B, K) (B, for K = 1).
        starts: [(family, score, e0)] of the start search; each replaces the real start (blml_gpu.py:476-484), the
        fits run in turn on the same generator, and the one of the largest likelihood is returned (self.starts_
        records each)."""
        torch = _torch()
        lib = _lib.load()
        B = len(Zi)
        nc = 2 if complex_g else 1
        V = B * K * nc
        st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        P = lambda x: ctypes.c_void_p(x.data_ptr())
        tabs, Zs, keep = [], [], []
        for q in tspecs:
            tq = local.lattice_tables(q['kinds'], q['fc'], q['periods'], h, Zi)
            t, dZ, _, kt = local._prepare(Zi, Zi, tq, f32, False)
            tabs.append(t)
            Zs.append(dZ)
            keep.append(kt)
        ntab = len(tabs)
        tarr = (local.PairTab * ntab)(*tabs)
        zarr = (ctypes.c_void_p * ntab)(*[z.data_ptr() for z in Zs])
        sbb = []
        one = torch.ones(1, dtype=torch.float64, device='cuda')
        s1 = torch.zeros(1, dtype=torch.float64, device='cuda')
        for t, dZ in zip(tabs, Zs):
            _lib.check(lib.blml_pair_rows(ctypes.byref(t), P(dZ), P(dZ), P(one), 1, 1, P(s1), st))
            sbb.append(float(s1.item()))
        if not all(v > 0 for v in sbb):
            raise RuntimeError(f"S(z, z) = {sbb}: the kernel's value at zero must be positive")
        sbbs = np.array(sbb * (K if ntab == 1 else 1), np.float64)
        sarr = (ctypes.c_double * K)(*sbbs)
        dw = torch.from_numpy(np.ascontiguousarray(w, np.float64)).cuda()
        stats = torch.zeros(4, dtype=torch.float64, device='cuda')
        Bs = max(1, min(int(self.block_size), -(-B // 8)))

        def ptrs(xs):
            return (ctypes.c_void_p * len(xs))(*[x.data_ptr() for x in xs])

        def exact(X, Y):
            _lib.check(lib.blml_quick_psd_exact(tarr, ntab, zarr, B, K, nc, P(X), P(Y), st))

        def measure(X, Y, scale=1.0):
            _lib.check(lib.blml_quick_psd_stats(P(X), P(Y), P(dw), B, K, nc, n, scale, P(stats), st))
            tr, wl, rho, bad = stats.cpu().numpy()
            return float(tr), (float(wl) - n * np.log(max(float(tr), 1e-300))) / n, float(rho), bad == 0

        def comb(xs, c, out, mode=0):
            _lib.check(lib.blml_vec_lincomb(len(xs), ptrs(xs), (ctypes.c_double * len(c))(*c), mode, V, P(out), st))
            return out

        def gram(xs, A):
            A = np.ascontiguousarray(A, np.float64)
            out = torch.empty(A.shape[0] ** 2, dtype=torch.float64, device='cuda')
            _lib.check(lib.blml_vec_gram(len(xs), ptrs(xs), A.shape[0], A.ctypes.data_as(ctypes.c_void_p), V,
                                         P(out), st))
            return out.cpu().numpy().reshape(A.shape[0], A.shape[0])

        def solve(ws):
            e0 = _psd_start(np.asarray(e_real, np.float64), K, complex_g, rng, float(self.phase_noise), ws)
            host = np.stack([e0.real, e0.imag], -1) if complex_g else e0[..., None]
            E = torch.from_numpy(np.ascontiguousarray(host, np.float64)).cuda().reshape(-1)
            SE = torch.empty_like(E)
            exact(E, SE)
            tr = measure(E, SE)[0]
            sc = n / np.sqrt(max(tr, 1e-300))
            comb([E], [sc], E)
            comb([SE], [sc], SE)
            if self.solver == 'lbfgs':
                return self._lbfgs_psd(E, SE, n, dw, tarr, ntab, zarr, B, K, nc, st, exact, measure, comb, gram)
            return self._psd_cd(E, SE, n, dw, tarr, ntab, zarr, B, K, nc, Bs, sarr, st, rng, exact, measure, comb,
                                gram)

        if starts:
            best = None
            self.starts_ = []
            for family, score, e0 in starts:
                out = solve(np.asarray(e0)[:, None])
                oh = out[6]
                ll = -n * oh[-1] if oh and np.isfinite(oh[-1]) else -np.inf
                self.starts_.append(dict(family=family, start_score=float(score), loglik=float(ll),
                                         converged=bool(out[3]), n_sweeps=int(out[2])))
                if best is None or ll > best[0]:
                    best = (ll, out)
            out = best[1]
        else:
            out = solve(None)
        Ef, SEf = out[0], out[1]
        del keep
        to_np = lambda X: _to_host(X, B, K, nc)
        e = to_np(Ef)
        Se = None if SEf is None else to_np(SEf)
        if K == 1:
            e = e[:, 0]
            Se = None if Se is None else Se[:, 0]
        return (e, Se) + tuple(out[2:])

    def _psd_cd(self, E, SE, n, dw, tarr, ntab, zarr, B, K, nc, Bs, sarr, st, rng, exact, measure, comb, gram):
        """Block-Jacobi CD with Anderson mixing and the likelihood safeguard (blml_gpu.py:650-744)."""
        lib = _lib.load()
        torch = _torch()
        P = lambda x: ctypes.c_void_p(x.data_ptr())
        aa_m = int(self.anderson)
        perm_fixed = rng.permutation(B) if aa_m else None

        def sweep(safe):
            perm = perm_fixed if perm_fixed is not None else rng.permutation(B)
            dp = torch.from_numpy(perm.astype(np.int32)).cuda()
            _lib.check(lib.blml_quick_psd_sweep(tarr, ntab, zarr, B, K, nc, P(dw), P(dp), Bs, n, sarr, int(safe),
                                                P(E), P(SE), st))

        hG, hSG, hR = [], [], []
        self.aa_accepted_ = 0
        safeguard, n_sweeps, converged = False, 0, False
        rho_hist, obj_hist = [], []
        L_prev = measure(E, SE)[1]
        cap, max_rho = self.max_sweeps, measure(E, SE)[2]
        SE_end = None
        while True:
            if cap is not None and n_sweeps >= cap:
                break
            n_sweeps += 1
            E_save, SE_save = E.clone(), SE.clone()
            sweep(safeguard)
            L_now = measure(E, SE)[1]
            if not safeguard and L_now < L_prev - 1e-10 * abs(L_prev):
                E.copy_(E_save)
                SE.copy_(SE_save)
                safeguard = True
                sweep(True)
                L_now = measure(E, SE)[1]
            if aa_m:
                hG.append(E.clone())
                hSG.append(SE.clone())
                hR.append(comb([E, E_save], [1.0, -1.0], torch.empty_like(E)))
                if len(hG) > aa_m + 1:
                    hG.pop(0)
                    hSG.pop(0)
                    hR.pop(0)
                if len(hG) >= 2:
                    L = len(hR)
                    A = np.zeros((L, L))
                    for j in range(L - 1):
                        A[j, j + 1], A[j, j] = 1.0, -1.0
                    A[L - 1, L - 1] = 1.0
                    Gm = gram(hR, A)
                    G, b = Gm[:L - 1, :L - 1], Gm[:L - 1, L - 1]
                    gam = np.linalg.lstsq(G + 1e-12 * np.trace(G) * np.eye(len(G)), b, rcond=None)[0]
                    E_aa = comb(hG, list(gam) + [0.0], torch.empty_like(E), mode=1)
                    SE_aa = comb(hSG, list(gam) + [0.0], torch.empty_like(SE), mode=1)
                    _, L_aa, _, fin = measure(E_aa, SE_aa)
                    L_aa = L_aa if fin else -np.inf
                    if L_aa >= L_now:
                        E, SE, L_now = E_aa, SE_aa, L_aa
                        self.aa_accepted_ += 1
                    else:
                        hG, hSG, hR = hG[-1:], hSG[-1:], hR[-1:]
            L_prev = L_now
            _, _, rho, fin = measure(E, SE)
            if not fin:
                max_rho = np.inf
                rho_hist.append(max_rho)
                break
            max_rho = rho
            if max_rho < self.tol:
                exact(E, SE)
                max_rho = measure(E, SE)[2]
            rho_hist.append(max_rho)
            obj_hist.append(-L_now)
            if self.verbose and n_sweeps % 50 == 0:
                print(f"  psd sweep {n_sweeps}: max|rho|={max_rho:.2e} -L/n={-L_now:.6f}", flush=True)
            if max_rho < self.tol:
                converged = True
                break
            if (len(rho_hist) > 500 and rho_hist[-1] > 0.99 * min(rho_hist[:-500])
                    and obj_hist[-501] - obj_hist[-1] < 1e-9 * abs(obj_hist[-1])):
                break
        if not converged and cap is not None and measure(E, SE)[3]:     # capped: e'Se = n^2
            exact(E, SE)
            tr = measure(E, SE)[0]
            sc = float(n / np.sqrt(max(tr, 1e-300)))
            comb([E], [sc], E)
            comb([SE], [sc], SE)
            _, L_end, max_rho, _ = measure(E, SE)
            converged = max_rho < self.tol
            obj_hist.append(-L_end)
            rho_hist.append(max_rho)
            SE_end = SE
        elif converged:
            SE_end = SE
        return E, SE_end, n_sweeps, converged, max_rho, rho_hist, obj_hist, safeguard

    def _lbfgs_psd(self, e, Se, n, dw, tarr, ntab, zarr, B, K, nc, st, exact, measure, comb, gram):
        """solver='lbfgs' (lbfgs_psd, blml_trivial.py:2590-2651): the two-loop recursion and every dot on the
        device, the line search's trial points through exact S e passes."""
        lib = _lib.load()
        torch = _torch()
        P = lambda x: ctypes.c_void_p(x.data_ptr())
        V = B * K * nc
        I2 = np.eye(2)

        def grad(x, Sx):
            G = torch.empty_like(x)
            _lib.check(lib.blml_quick_psd_grad(tarr, ntab, zarr, B, K, nc, P(dw), n, P(x), P(Sx), P(G), st))
            return G

        def rho_of(x, Sx):
            tr = measure(x, Sx)[0]
            return measure(x, Sx, n / np.sqrt(max(tr, 1e-300)))[2]

        L = measure(e, Se)[1] * n
        g = grad(e, Se)
        S_, Y_, R_, SY = [], [], [], []
        it = 0
        cap = np.inf if self.max_sweeps is None else self.max_sweeps
        max_rho = rho_of(e, Se)
        rho_hist, obj_hist = [max_rho], [-L / n]
        reset = llstop = False
        mem = int(self.lbfgs_m)
        q = torch.empty_like(e)
        while max_rho >= self.tol and it < cap:
            if len(obj_hist) > 10 and obj_hist[-11] - obj_hist[-1] < 1e-11:
                llstop = True
                break
            it += 1
            if S_:
                gamma = SY[-1][0] / SY[-1][1]
                _lib.check(lib.blml_lbfgs_twoloop(len(S_), (ctypes.c_void_p * len(S_))(*[x.data_ptr() for x in S_]),
                                                  (ctypes.c_void_p * len(Y_))(*[x.data_ptr() for x in Y_]),
                                                  (ctypes.c_double * len(R_))(*R_), gamma, V, P(g), P(q), st))
            else:
                eg = gram([e, g], I2)
                comb([g], [1e-3 * np.sqrt(eg[0, 0] / max(eg[1, 1], 1e-300))], q)
            slope = float(gram([g, q], I2)[0, 1])
            if not slope > 0:
                S_, Y_, R_, SY = [], [], [], []
                eg = gram([e, g], I2)
                comb([g], [1e-3 * np.sqrt(eg[0, 0] / max(eg[1, 1], 1e-300))], q)
                slope = float(gram([g, q], I2)[0, 1])
            t = 1.0
            ok = False
            en, Sen = torch.empty_like(e), torch.empty_like(e)
            for _ in range(40):
                comb([e, q], [1.0, t], en)
                exact(en, Sen)
                Ln = measure(en, Sen)[1] * n
                if Ln >= L + 1e-4 * t * slope:
                    ok = True
                    break
                t *= 0.5
            if not ok:
                if reset or not S_:
                    break
                S_, Y_, R_, SY = [], [], [], []
                reset = True
                continue
            reset = False
            gn = grad(en, Sen)
            s_ = comb([en, e], [1.0, -1.0], torch.empty_like(e))
            y_ = comb([g, gn], [1.0, -1.0], torch.empty_like(e))
            Gm = gram([s_, y_], I2)
            ss, sy, yy = Gm[0, 0], Gm[0, 1], Gm[1, 1]
            if sy > 1e-12 * np.sqrt(ss * yy):
                S_.append(s_)
                Y_.append(y_)
                R_.append(1.0 / sy)
                SY.append((sy, yy))
                if len(S_) > mem:
                    S_.pop(0)
                    Y_.pop(0)
                    R_.pop(0)
                    SY.pop(0)
            e, Se, L, g = en, Sen, Ln, gn
            max_rho = rho_of(e, Se)
            rho_hist.append(max_rho)
            obj_hist.append(-L / n)
            if self.verbose and it % 100 == 0:
                print(f"  lbfgs {it}: max|rho|={max_rho:.2e} -L/n={-L / n:.9f}", flush=True)
        Q = measure(e, Se)[0]
        ec, Sec = torch.empty_like(e), torch.empty_like(e)
        _lib.check(lib.blml_quick_psd_wscale(P(Se), P(dw), B, K, nc, Q / n, P(ec), st))
        exact(ec, Sec)
        if measure(ec, Sec)[1] * n >= L - 1e-9 * n:
            e, Se = ec, Sec
        tr = measure(e, Se)[0]
        sc = n / np.sqrt(max(tr, 1e-300))
        comb([e], [sc], e)
        comb([Se], [sc], Se)
        _, L_end, max_rho, _ = measure(e, Se)
        converged = max_rho < self.tol or llstop
        rho_hist.append(max_rho)
        obj_hist.append(-L_end)
        return e, Se, it, converged, max_rho, rho_hist, obj_hist, False

    # ---------------------------------------------------------------------------------------------- queries
    def sqrt_density(self, Z):
        """g(x) = (1/n) sum_b e_b K(x, z_b) at points Z (any points: the exact continuous kernel from the units'
        closed forms, blml_point_rows, as the reference's sqrt_density, blml_gpu.py:955-989): (Nq,) for rank 1,
        else (Nq, K); complex for complex fits."""
        from .points import point_rows, point_units
        Zq = np.atleast_2d(np.asarray(Z, float))
        if Zq.shape[1] != self.dim_:
            raise ValueError(f"Z: points with {self.dim_} coordinates")
        E = np.asarray(self.e_hat_)
        E = E[:, None] if E.ndim == 1 else E
        K = E.shape[1]
        cplx = np.iscomplexobj(E)
        out = np.zeros((len(Zq), K), dtype=complex if cplx else float)
        for k in range(K):
            sp = self._specs[k] if self.mode_specs_ is not None else self._specs[0]
            units, harm = point_units(sp)
            W = np.stack([E[:, k].real, E[:, k].imag], 1) if cplx else np.asarray(E[:, k], float)
            S = point_rows(units, harm, Zq, self.Z_, W)
            out[:, k] = (S[:, 0] + 1j * S[:, 1] if cplx else S[:, 0]) / self.n_samples_
        return out[:, 0] if K == 1 else out

    def predict(self, Z):
        """The density f(x) = sum_k |g_k(x)|^2 at points Z (blml_gpu.py:892-897)."""
        g = self.sqrt_density(Z)
        return (np.abs(g) ** 2).sum(1) if g.ndim == 2 else np.abs(g) ** 2


def _radial_dims(kinds):
    groups = {}
    for j, k in enumerate(kinds):
        name, _, tag = str(k).partition(':')
        if name in ('ball', 'pball'):
            groups.setdefault(tag, []).append(j)
    return list(groups.values())


def complex_starts(Zi, w, h, fc, kinds, n_hio=3, n_sb=2, hio_pool=30, sb_delta=1.0, n_hio_iter=400, n_er_iter=100,
                   beta=0.9, seed=0, max_cells=1 << 22):
    """Starts of a complex counts fit (blml_trivial.py:1605-1676), every run on the device (blml_complex_starts):
    [(family, score, e0)] -- the n_hio best of hio_pool random-phase runs (n_hio_iter HIO + n_er_iter ER steps) and
    the n_sb best symmetry-breaking runs (the band-limited sqrt(counts) kicked by exp(+-i sb_delta p) along each
    even phase direction, n_er_iter ER steps), by the Poisson score of |g|^2 normalised on the box, each as
    e0 = w / conj(g) at the bins; None when the search grid has more than max_cells cells or a side above 8192.
    The host lays out the grid, the band, the draws (default_rng(seed).random per random run) and the kicks as the
    reference does; the runs' arithmetic is FP64 on the device."""
    torch = _torch()
    lib = _lib.load()
    Zi0 = np.asarray(Zi, np.int64)
    w0 = np.asarray(w, np.float64)
    h0 = np.asarray(h, float)
    fc = np.asarray(fc, float)
    d = Zi0.shape[1]
    q = np.maximum(1, np.floor(0.5 / (fc * h0))).astype(np.int64)
    Zc = np.floor_divide(Zi0 - Zi0.min(0), q)
    if np.all(q == 1):
        Zg, wg = Zc, w0
    else:
        Zg, inv = np.unique(Zc, axis=0, return_inverse=True)
        wg = np.bincount(np.ravel(inv), weights=w0)
    hc = h0 * q
    lo = Zg.min(0)
    L = Zg.max(0) - lo + 1
    shape = tuple(1 << int(np.ceil(np.log2(2 * int(l)))) for l in L)
    M = int(np.prod(shape, dtype=np.int64))
    if M > max_cells or max(shape) > 8192:
        return None
    # the band: frequencies over fc / 2, |u_j| <= 1 per axis, ||u_G|| <= 1 per radial group
    U = np.stack(np.meshgrid(*[np.fft.fftfreq(m, hj) / (0.5 * f) for m, hj, f in zip(shape, hc, fc)],
                             indexing='ij'), -1)
    band = np.ones(shape, bool)
    rad = set()
    for dims in _radial_dims(kinds):
        band &= (U[..., dims] ** 2).sum(-1) <= 1
        rad |= set(dims)
    for j in range(d):
        if j not in rad:
            band &= np.abs(U[..., j]) <= 1
    kicks = []
    if n_sb > 0:
        r2 = (U ** 2).sum(-1)
        dirs = [2 * r2 - 1] + [U[..., 0] ** 2 - U[..., j] ** 2 for j in range(1, d)]
        dirs += [2 * U[..., i] * U[..., j] for i in range(d) for j in range(i + 1, d)] + [6 * r2 ** 2 - 6 * r2 + 1]
        kicks = [sg * sb_delta * p for p in dirs for sg in (1, -1)]
    rng = np.random.default_rng(seed)
    n_rand = hio_pool if n_hio > 0 else 0
    phase = np.stack([rng.random(shape).ravel() for _ in range(n_rand)]) if n_rand else np.zeros((0, M))
    runs = n_rand + len(kicks)
    if runs == 0:
        return []
    amp = np.zeros(shape)
    idx = tuple((Zg - lo).T)
    amp[idx] = np.sqrt(wg)
    occ = np.ravel_multi_index(idx, shape).astype(np.int64)
    dev = lambda a, dt=torch.float64: torch.from_numpy(np.ascontiguousarray(a)).to('cuda', dt)
    d_band = dev(band.ravel().astype(np.uint8), torch.uint8)
    d_amp, d_occ, d_w = dev(amp.ravel()), dev(occ, torch.int64), dev(wg)
    d_phase = dev(phase.reshape(-1)) if n_rand else None
    d_kick = dev(np.stack([k.ravel() for k in kicks]).reshape(-1)) if kicks else None
    dP = torch.empty(runs * M * 2, dtype=torch.float64, device='cuda')
    dg = torch.empty(runs * len(occ) * 2, dtype=torch.float64, device='cuda')
    dscore = torch.empty(runs, dtype=torch.float64, device='cuda')
    prm = _lib.HioParams()
    prm.d = d
    for j in range(d):
        prm.shape[j], prm.box[j] = shape[j], int(L[j])
    prm.n_hio, prm.n_er, prm.n_rand, prm.n_kick = int(n_hio_iter), int(n_er_iter), n_rand, len(kicks)
    prm.beta, prm.ab2, prm.nocc = float(beta), float(wg.sum()), len(occ)
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    V = lambda x: None if x is None else ctypes.c_void_p(x.data_ptr())
    _lib.check(lib.blml_complex_starts(ctypes.byref(prm), V(d_band), V(d_amp), V(d_occ), V(d_w), V(d_phase),
                                       V(d_kick), V(dP), V(dg), V(dscore), st))
    score = dscore.cpu().numpy()
    fam = ['hio'] * n_rand + ['symbreak'] * len(kicks)
    out = []
    for family, k in (('hio', n_hio), ('symbreak', n_sb)):
        cand = [r for r in range(runs) if fam[r] == family]
        for r in sorted(cand, key=lambda r: -score[r])[:k]:
            if np.all(q == 1):
                gv = dg[r * len(occ) * 2:(r + 1) * len(occ) * 2].cpu().numpy()
                g = gv[0::2] + 1j * gv[1::2]
            else:
                nz = np.flatnonzero(band.ravel())
                Pr = dP[r * M * 2:(r + 1) * M * 2].cpu().numpy()
                c = (Pr[0::2] + 1j * Pr[1::2])[nz] / M
                kf = np.stack(np.unravel_index(nz, shape), 1)
                kf = np.where(kf < np.array(shape) // 2, kf, kf - np.array(shape)).astype(np.int32)
                dz = (Zi0 - Zi0.min(0)).astype(np.int32)
                mq = (np.array(shape) * q).astype(np.int32)
                dc = dev(np.stack([c.real, c.imag], 1).ravel())
                ddz, dkf = dev(dz, torch.int32), dev(kf, torch.int32)
                dgb = torch.empty(len(Zi0) * 2, dtype=torch.float64, device='cuda')
                _lib.check(lib.blml_band_eval(d, mq.ctypes.data_as(ctypes.c_void_p), len(Zi0), len(nz), V(ddz),
                                              V(dkf), V(dc), V(dgb), st))
                gv = dgb.cpu().numpy()
                g = gv[0::2] + 1j * gv[1::2]
            out.append((family, float(score[r]), w0 / np.conj(np.where(np.abs(g) == 0, 1e-12, g))))
    return out


def shared_spec(specs):
    return all(_same_spec(q, specs[0]) for q in specs)


def _same_spec(a, b):
    """One kernel: the same kinds, periods and fc in every column (the reference's _shared_kernel)."""
    return (list(a['kinds']) == list(b['kinds']) and list(a['periods']) == list(b['periods'])
            and np.allclose(np.asarray(a['fc'], float), np.asarray(b['fc'], float)))


def _psd_start(e_real, K, complex_g, rng, pn, ws=None):
    """E0 (B, K) of the rank-K / complex fit (blml_gpu.py:576-590): column 0 the real solution, with phases
    exp(i pn N(0, 1)) when complex; the other columns pn |e| times CN(0, 1) (real: N(0, 1)), drawn in that order.
    ws (B, K'): a start that replaces the first K' columns, the rest drawn as above around |ws[:, 0]|."""
    B = len(e_real)
    e0 = np.zeros((B, K), dtype=complex if complex_g else np.float64)
    if ws is not None:
        ws = np.asarray(ws, dtype=complex if complex_g else np.float64)
        k0 = min(K, ws.shape[1])
        e0[:, :k0] = ws[:, :k0]
        er = np.abs(ws[:, 0])
    else:
        k0, er = 1, e_real
        e0[:, 0] = e_real * (np.exp(1j * pn * rng.standard_normal(B)) if complex_g else 1.0)
    for k in range(k0, K):
        noise = ((rng.standard_normal(B) + 1j * rng.standard_normal(B)) / np.sqrt(2) if complex_g
                 else rng.standard_normal(B))
        e0[:, k] = pn * er * noise
    return e0


def _to_host(X, B, K, nc):
    a = X.reshape(B, K, nc).cpu().numpy()
    return a[..., 0] + 1j * a[..., 1] if nc == 2 else a[..., 0].copy()
