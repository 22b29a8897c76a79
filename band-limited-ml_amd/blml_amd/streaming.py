"""Counts arriving in batches on one fixed lattice: the reference's BLMLBinnedStreaming and its kernel-parameter
learner (blml_binned.py:648-1118, _KernelLearner), on the MI355X engine.

    s = BLMLBinnedStreaming(decay=1.0, batch_sweeps=50, fc_update='adam', lr=0.05, param_every=1)
    for counts in batches:
        s.partial_fit(counts, modes)          # modes on the first call; fc NaN: selected at the burn-in
    s.fit(batches, modes, epochs=12, tol=1e-3, tol_check='exact')
    s.preq_, s.history_, s.converged_, s.epochs_, s.residual_

Every fit is the engine's (blml_fit): the burn-in's fit, each batch's warm refit (decay times the held counts plus
the batch, started from the current coefficients, to tol or batch_sweeps), the leave-batch-out refits of revisited
batches and the hill-climb's candidate refits.  The prequential score (log_density), its gradient (param_grad, or
implicit_grad with the refit response) and the burn-in's fc search (the batched FcSearch) are the engine's too.  The
host keeps the learner's bookkeeping: the held counts, Adam's moments, the hill-climb's paired gains and the epoch
order -- scalars and one count array per batch, as in the reference.

Kernel parameters: 'scale' ones (sinc, slits, ball fc) by Adam on log fc; 'integer' ones (dirichlet, pball fc:
whole harmonics), and every one with fc_update='hillclimb', by one step once its pooled paired gain on the batches
passes z_se standard errors.  A parameter move keeps the coefficients and recomputes S e under the new kernel
(rescaled to e'Se = n^2); the next batches' sweeps catch up."""
import contextlib

import numpy as np

from . import _lib
from .binned import BLMLBinned, FitBatch, _FitSpec, _torch
from .modes import parse_modes, units_of
from .select import FcSearch, val_score

__all__ = ['BLMLBinnedStreaming', 'adam_step']

_SCALE_KINDS = ('sinc', 'slits', 'ball')          # fc a 'scale' parameter; dirichlet / pball: 'integer'


def adam_step(value, grad, state, lr, betas):
    """One Adam ascent step on log(value) from grad = d / dlog value (blml_maps.adam_step, ptype 'scale',
    dlog=True); state = [m, v, t], updated in place."""
    b1, b2 = betas
    state[2] += 1
    t = state[2]
    state[0] = b1 * state[0] + (1 - b1) * grad
    state[1] = b2 * state[1] + (1 - b2) * grad * grad
    step = lr * (state[0] / (1 - b1 ** t)) / (np.sqrt(state[1] / (1 - b2 ** t)) + 1e-12)
    return float(value * np.exp(step))


def _fc_value(c):
    v = c.get('fc')
    return v.get('value') if isinstance(v, dict) else v


def _fc_trainable(c):
    v = c.get('fc')
    t = v.get('trainable') if isinstance(v, dict) else None
    return True if t is None else bool(t)          # every kind's fc is trainable by default


def _percol(m):
    out = {}
    for key, c in m.items():
        for j in (key if isinstance(key, tuple) else (key,)):
            out[j] = c
    return out


def modes_fc(modes, d):
    """(K, d) cut-offs of a modes list (NaN where fc is left out: a selector template)."""
    return np.array([[np.nan if _fc_value(pc[j]) is None else float(_fc_value(pc[j])) for j in range(d)]
                     for pc in map(_percol, modes)], dtype=float)


def modes_trainable(modes, d):
    return np.array([[_fc_trainable(pc[j]) for j in range(d)] for pc in map(_percol, modes)], dtype=bool)


def _modes_with_fc(modes, F):
    """modes with mode k's column-group fc set to F[k, first column] (templates filled, trainability kept;
    blml_binned.py:999-1013)."""
    out = []
    for k, m in enumerate(modes):
        mm = {}
        for key, c in m.items():
            c = dict(c)
            cols = key if isinstance(key, tuple) else (key,)
            val = float(F[k][cols[0]])
            c['fc'] = dict(c['fc'], value=val) if isinstance(c.get('fc'), dict) else val
            mm[key] = c
        out.append(mm)
    return out


def _plain_modes(modes):
    """modes with every fc a bare value (the fit's parser takes values; trainability is kept by the learner)."""
    out = []
    for m in modes:
        mm = {}
        for key, c in m.items():
            c = dict(c)
            if isinstance(c.get('fc'), dict):
                c['fc'] = c['fc'].get('value')
            mm[key] = c
        out.append(mm)
    return out


def _batch_order(nb, groups, rng):
    """Batch indices in random group order (no groups: a random permutation), whole groups together
    (blml_binned.py:1016-1024)."""
    if groups is None:
        return list(rng.permutation(nb))
    u, inv = np.unique(np.asarray(groups), return_inverse=True)
    members = [[] for _ in u]
    for i, k in enumerate(np.ravel(inv)):
        members[k].append(i)
    return [i for k in rng.permutation(len(u)) for i in members[k]]


class _Init:
    """A warm start for BLMLBinned.fit: coefficients e_hat_ (K, *shape) on the lattice shape_."""

    def __init__(self, e, shape):
        self.e_hat_, self.shape_ = e, tuple(shape)


def _secular(R2, s, c, iters=16):
    """Per bin the root lam > max s of lam sum_k R2_k / (lam - s_k)^2 = c: Newton on phi^-1/2 - c^-1/2 inside the
    bisection bracket (blml_binned.py:1556-1573).  R2 (K, B), s (K, 1), c (B,)."""
    smax = float(np.max(s))
    tot = R2.sum(0)
    d = s - smax
    cpos = np.maximum(c, 1e-30)
    hi = (tot + np.sqrt(tot * tot + 4 * cpos * smax * tot)) / (2 * cpos)
    lo = np.zeros_like(hi)
    mu = hi
    for _ in range(iters):
        with np.errstate(divide='ignore', over='ignore', invalid='ignore'):
            q = 1.0 / (mu[None] - d)
            W = R2 * (q * q)
            A = W.sum(0)
            Bq = (W * q).sum(0)
            lam = smax + mu
            phi = lam * A
            big = phi > c
            lo = np.where(big, mu, lo)
            hi = np.where(big, hi, mu)
            nt = mu + 2 * phi * (1 - np.sqrt(phi / cpos)) / (A - 2 * lam * Bq)
        mu = np.where((nt >= lo) & (nt <= hi), nt, 0.5 * (lo + hi))
    return smax + mu


def _row_update(E, U, s, c, occ, psd):
    """Every bin's coordinate update from U = S E (the fit's Jacobi update, blml_binned.py:184-206; no signs):
    E, U (K, B), s (K,), c = n w (B,)."""
    R = U - s[:, None] * E
    if not psd:
        b = R[0]
        disc = np.sqrt(b * b + 4.0 * float(s[0]) * c)
        t = np.where(b >= 0, 2.0 * c / np.maximum(b + disc, 1e-300), (-b + disc) / (2.0 * float(s[0])))
        return np.where(occ, t, 0.0)[None]
    a2 = (np.abs(R) ** 2).sum(0)
    a = np.sqrt(a2)
    if np.allclose(s, s[0]):
        s0 = float(s[0])
        t = 2 * c / (a + np.sqrt(a2 + 4 * s0 * c))
        out = R * (t / np.where(a > 0, a, 1.0))[None]
        lone = (a == 0) & (c > 0)
        out[0] = np.where(lone, np.sqrt(c / s0), out[0])
        return out
    lone = (a2 == 0) & (c > 0)
    lam = _secular(np.abs(R) ** 2, s[:, None], c)
    out = np.where((c > 0)[None] & ~lone[None], R / (lam[None] - s[:, None]), 0)
    out[0] = np.where(lone, np.sqrt(c / float(s[0])), out[0])
    return out


class BLMLBinnedStreaming(BLMLBinned):
    """BLMLBinnedStreaming (blml_binned.py:1027-1118): partial_fit(counts) holds counts until burn_in, then adds each
    batch (the held counts times decay, plus the batch) and refits warm-started from the current coefficients, to
    tol or for batch_sweeps sweeps (to tol during the burn-in); fit(batches, modes, epochs, ...) revisits the batches
    over epochs.  The learner's knobs are the reference's (see its _KernelLearner docstring, blml_binned.py:648-683);
    the other keywords are BLMLBinned's."""

    def __init__(self, decay=1.0, batch_sweeps=None, burn_in=0, fc_update='adam', lr=0.1, betas=(0.9, 0.999),
                 step=1.05, z_se=1.0, param_every=3, grad_eps=1e-3, implicit_grad=True, search='shared', bcd_frac=1.0,
                 freeze_frac=0.0, trace_trainable=False, loo_sweeps=None, **kw):
        if 'implicit' in kw:
            implicit_grad = kw.pop('implicit')
        super().__init__(**kw)
        self._base_kw = dict(kw)
        self.decay, self.batch_sweeps, self.loo_sweeps = decay, batch_sweeps, loo_sweeps
        if fc_update not in (None, 'adam', 'hillclimb'):
            raise ValueError("fc_update: 'adam', 'hillclimb' or None")
        if search not in ('shared', 'refine'):
            raise ValueError("search: 'shared' or 'refine'")
        if not (0.0 < bcd_frac <= 1.0 and 0.0 <= freeze_frac < 1.0):
            raise ValueError("bcd_frac in (0, 1], freeze_frac in [0, 1)")
        self.burn_in, self.fc_update, self.lr, self.betas, self.step, self.z_se = burn_in, fc_update, lr, betas, step, z_se
        self.param_every, self.grad_eps, self._implicit = max(1, int(param_every)), grad_eps, bool(implicit_grad)
        self.search, self.bcd_frac, self.freeze_frac = search, float(bcd_frac), float(freeze_frac)
        self.trace_trainable = trace_trainable
        self.preq_, self.history_ = [], []
        self._n_learn, self._pool, self._hc, self._adam, self._epoch_g = 0, [], {}, {}, []
        self._held, self._held_groups, self.n_batches_ = None, [], 0
        self.converged_ = self.epochs_ = self.residual_ = None

    @property
    def implicit(self):
        return self._implicit

    # ------------------------------------------------------------------------------------------- the fits
    def _modes_now(self):
        return _modes_with_fc(self._modes0, self._F)

    def _warm_fit(self, counts, E, sweeps):
        """The engine's fit of counts from the coefficients E (K, *shape) under the current kernels, sweeps at most
        (None: to tol) -- the reference's _solve(E, sweeps) then _finish."""
        cap = self.max_sweeps
        self.max_sweeps = sweeps
        try:
            _, h, origin, cg = self._modes_args
            BLMLBinned.fit(self, counts, self._modes_now(), h, origin, cg, init=_Init(E, self.shape_))
        finally:
            self.max_sweeps = cap
        return self

    def _set_state(self, E, U):
        """Adopts coefficients E and U = S E (K, *shape) under the current kernels without a solve: the device
        buffers of a fresh one-fit plan hold them, and the outputs follow (_finish, blml_binned.py:438-445)."""
        torch = _torch()
        _, h, origin, cg = self._modes_args
        fs = _FitSpec(self.w_.astype(np.float64), self._modes_now(), h, origin, cg, None, None, self.seed)
        b = FitBatch([fs], self._params(), self.dtype, hist_cap=0).build()
        nc = fs.nc
        flat = lambda X: (np.stack([X.real, X.imag], -1) if nc == 2 else np.real(X)[..., None]).reshape(-1)
        tdt = b.d_e.dtype
        b.d_e.copy_(torch.from_numpy(np.ascontiguousarray(flat(E.reshape(fs.K, -1)))).to('cuda', tdt))
        b.d_se.copy_(torch.from_numpy(np.ascontiguousarray(flat(U.reshape(fs.K, -1)))).to('cuda', tdt))
        self._batch, self._index = b, 0
        self._gplans = self._iplan = None
        self.specs_, self.modes_, self.shared_, self.s_ = fs.specs, fs.modes, fs.shared, fs.s
        self.e_hat_, self.Se_ = E, U
        F = (np.abs(U) ** 2).sum(0)
        self.density_ = F / self.n_ ** 2
        occ = self.occ_
        self.loglik_ = float((self.w_[occ] * np.log(np.maximum(self.density_[occ], 1e-300))).sum() / self.n_)
        return self

    def _matvec_now(self, E):
        """S E (K, *shape) under the current kernels (a one-fit plan's blml_matvec)."""
        torch = _torch()
        _, h, origin, cg = self._modes_args
        fs = _FitSpec(self.w_.astype(np.float64), self._modes_now(), h, origin, cg, None, None, self.seed)
        b = FitBatch([fs], self._params(), self.dtype, hist_cap=0).build()
        nc = fs.nc
        X = E.reshape(fs.K, -1)
        flat = (np.stack([X.real, X.imag], -1) if nc == 2 else np.real(X)[..., None]).reshape(-1)
        U = b.matvec(torch.from_numpy(np.ascontiguousarray(flat)).to('cuda', b.d_e.dtype))
        u = U.to(torch.float64).cpu().numpy().reshape(fs.K, fs.B, nc)
        b.close()
        out = u[..., 0] + 1j * u[..., 1] if nc == 2 else u[..., 0]
        return out.reshape(E.shape), fs.s

    def _refresh(self):
        """After a parameter move: S e under the new kernels, rescaled to e'Se = n^2 (blml_binned.py:1052-1054)."""
        E = np.asarray(self.e_hat_)
        U, _ = self._matvec_now(E)
        Q = float(np.real(np.sum(np.conj(E) * U)))
        sc = float(self.n_ / np.sqrt(Q)) if Q > 0 else 1.0
        return self._set_state(E * sc, U * sc)

    def _residual(self):
        """max over occupied bins of |U |e|^2 / c - e| / |e| at the current state (blml_binned.py:216-220)."""
        E, U = np.asarray(self.e_hat_), np.asarray(self.Se_)
        c = self.n_ * self.w_
        occ = self.occ_
        nrm2 = (np.abs(E) ** 2).sum(0)
        r = np.sqrt((np.abs(U * (nrm2 / np.where(occ, c, 1.0))[None] - E) ** 2).sum(0)) / np.sqrt(np.maximum(nrm2, 1e-300))
        return float(np.max(np.where(occ, r, 0.0)))

    def _hill_sweeps(self):
        return self.loo_sweeps if self.loo_sweeps is not None else self.max_sweeps

    def _base(self):
        return BLMLBinned(**self._base_kw)

    # ---------------------------------------------------------------------------------------------- learner
    def _registry(self):
        reg = []
        if self.fc_update is None:
            return reg
        for k, sp in enumerate(self.specs_):
            for u, (kind, dims) in enumerate(units_of(sp['kinds'])):
                if not bool(self._T[k][dims[0]]):
                    continue
                name = str(kind).partition(':')[0]
                ptype = 'scale' if name in _SCALE_KINDS else 'integer'
                rule = 'adam' if ptype == 'scale' and self.fc_update == 'adam' else 'hill'
                reg.append(dict(mode=k, unit=u, dims=tuple(dims), ptype=ptype, rule=rule))
        return reg

    def _fc_now(self):
        return [np.array(r, float) for r in self._F]

    def _nyq(self, e):
        """fc <= 0.5 / h: the lattice's Nyquist limit (blml_binned.py:761-764)."""
        return 0.5 / float(self.h_[e['dims'][0]])

    def _in_range(self, e, value):
        return 0.0 < value <= self._nyq(e) + 1e-12

    def _set_fc(self, k, dims, value):
        F = self._F.copy()
        F[k, list(dims)] = value
        self._F = F

    def _z0_eps(self, m):
        U = np.asarray(self.Se_)
        F = (np.abs(U) ** 2).sum(0)
        occ = m > 0
        Z0 = float(np.real(np.sum(np.conj(np.asarray(self.e_hat_)) * U)))
        eps = self.grad_eps * float(np.median(F[occ])) / Z0 if bool(np.any(occ)) else 0.0
        return F, Z0, eps, occ

    def _learn(self, counts):
        """The batch's gradient (pooled) and hill-climb gains, before it is added (blml_binned.py:727-759)."""
        m = np.asarray(counts, float)
        F, Z0, eps, occ = self._z0_eps(m)
        reg = self._registry()
        if any(e['rule'] == 'adam' for e in reg):
            G = self.implicit_grad(m, eps) if self._implicit else self.param_grad(m, eps)
            self._pool.append((float(m.sum()), [G[e['mode']][e['unit']] for e in reg if e['rule'] == 'adam']))
            self._epoch_g.append(self._pool[-1])
        hill = [e for e in reg if e['rule'] == 'hill']
        if hill:
            base = np.log(F / Z0 + eps)
            for e in hill:
                k, dims = e['mode'], list(e['dims'])
                v = float(self._F[k][dims[0]])
                for sgn in (+1, -1):
                    if e['ptype'] == 'integer':      # one harmonic: N = floor(fc P / 2) +- 1, its interval's middle
                        P = float(self.specs_[k]['periods'][dims[0]])
                        N = int(np.floor(v * P / 2.0)) + sgn
                        if N < 0:
                            continue
                        cand = (N + 0.5) * 2.0 / P
                    else:
                        cand = v * self.step ** sgn
                    if not self._in_range(e, cand):
                        continue
                    Fc, Zc = self._candidate_fit(e, cand)
                    dl = (np.log(Fc / Zc + eps) - base)[occ]
                    mw = m[occ]
                    acc = self._hc.setdefault((k, e['unit'], sgn), dict(s=0.0, ss=0.0, W=0.0, cand=cand))
                    acc['s'] += float((mw * dl).sum())
                    acc['ss'] += float((mw * dl * dl).sum())
                    acc['W'] += float(mw.sum())
                    acc['cand'] = cand

    def _keep_pools(self, saved):
        pools = (self._pool, self._hc, self._adam, self._epoch_g)
        self.__dict__.clear()
        self.__dict__.update(saved)
        self._pool, self._hc, self._adam, self._epoch_g = pools

    def _candidate_fit(self, e, cand):
        """(F, Z) of a warm refit of the held counts with e's parameter at cand (blml_binned.py:769-782)."""
        saved = dict(self.__dict__)
        try:
            self._set_fc(e['mode'], e['dims'], cand)
            self._warm_fit(self.w_, np.asarray(self.e_hat_), self._hill_sweeps())
            U = np.asarray(self.Se_)
            F = (np.abs(U) ** 2).sum(0)
            Z = float(np.real(np.sum(np.conj(np.asarray(self.e_hat_)) * U)))
            return F, Z
        finally:
            self._keep_pools(saved)

    def _param_step(self):
        """The parameter step: [(mode, unit, change)] (blml_binned.py:788-816)."""
        reg = self._registry()
        moved = []
        adam = [e for e in reg if e['rule'] == 'adam']
        if adam and self._pool:
            Wt = sum(p[0] for p in self._pool)
            G = sum(p[0] * np.asarray(p[1]) for p in self._pool) / Wt
            if not np.all(np.isfinite(G)):
                raise FloatingPointError(f"streaming: non-finite gradient {G}")
            for q, e in enumerate(adam):
                k, dims = e['mode'], e['dims']
                old = float(self._F[k][dims[0]])
                st = self._adam.setdefault((k, e['unit']), [0.0, 0.0, 0])
                new = min(adam_step(old, float(G[q]), st, self.lr, self.betas), self._nyq(e))
                self._set_fc(k, dims, new)
                moved.append((k, e['unit'], round(float(np.log(new / old)), 5)))
        self._pool = []
        for e in (e for e in reg if e['rule'] == 'hill'):
            for sgn in (+1, -1):
                acc = self._hc.get((e['mode'], e['unit'], sgn))
                if not acc or acc['W'] <= 1:
                    continue
                mu = acc['s'] / acc['W']
                se = np.sqrt(max(acc['ss'] / acc['W'] - mu * mu, 0.0) / (acc['W'] - 1))
                if mu > 0 and mu > self.z_se * se:
                    self._set_fc(e['mode'], e['dims'], acc['cand'])
                    moved.append((e['mode'], e['unit'], sgn))
                    self._hc = {kk: v for kk, v in self._hc.items() if kk[:2] != (e['mode'], e['unit'])}
                    break
        if moved:
            self._refresh()
        return moved

    def _param_stationary(self):
        """(stationary, max |mean| / se): the epoch's Adam gradients within z_se standard errors of 0 (or pushing
        into the Nyquist cap), no pending hill-climb move (blml_binned.py:818-841)."""
        reg = self._registry()
        adam = [e for e in reg if e['rule'] == 'adam']
        zmax = 0.0
        if adam and self._epoch_g:
            W = np.array([p[0] for p in self._epoch_g])
            G = np.array([np.ravel(np.asarray(p[1], float)) for p in self._epoch_g])
            if len(W) < 2:
                return False, np.inf
            mu = (W[:, None] * G).sum(0) / W.sum()
            var = (W[:, None] * (G - mu) ** 2).sum(0) / W.sum()
            se = np.sqrt(var / (len(W) - 1))
            ok = True
            for q, e in enumerate(adam):
                at_cap = float(self._F[e['mode']][e['dims'][0]]) >= self._nyq(e) - 1e-12
                z = abs(mu[q]) / se[q] if se[q] > 0 else (0.0 if mu[q] == 0 else np.inf)
                if not (z <= self.z_se or (at_cap and mu[q] > 0)):
                    ok = False
                    zmax = max(zmax, z)
        else:
            ok = not adam
        for acc in self._hc.values():
            if acc['W'] > 1:
                mu = acc['s'] / acc['W']
                se = np.sqrt(max(acc['ss'] / acc['W'] - mu * mu, 0.0) / (acc['W'] - 1))
                if mu > 0 and mu > self.z_se * se:
                    ok = False
        return ok, zmax

    def _trainable_now(self):
        return {k: {'fc': [float(v) for v, t in zip(self._F[k], self._T[k]) if t]} for k in range(len(self._F))}

    def _stream_step(self, counts, absorb, revisit=False, learn=True):
        """One batch: scored (leave-batch-out on a revisit) and learned from, absorbed, then a parameter step every
        param_every learning batches (blml_binned.py:846-866)."""
        burned = self.n_ >= self.burn_in
        do_learn = learn and burned and self.fc_update is not None
        with self._left_out(counts if revisit else None):
            preq = self.log_density(counts)
            if do_learn:
                self._learn(counts)
        if absorb is not None:
            absorb(burned)
        moved = []
        if do_learn:
            self._n_learn += 1
            if self._n_learn % self.param_every == 0:
                moved = self._param_step()
        self.preq_.append(preq)
        h = dict(n=self.n_, preq=preq, fc=[f.tolist() for f in self._fc_now()], moved=moved, revisit=bool(revisit))
        if self.trace_trainable:
            h['trainable'] = self._trainable_now()
        self.history_.append(h)

    def _left_out(self, counts):
        """Context: the fit with counts' share of the held counts removed, restored after (blml_binned.py:869-885)."""
        @contextlib.contextmanager
        def ctx():
            if counts is None:
                yield
                return
            saved = dict(self.__dict__)
            try:
                self._loo_enter(np.asarray(counts, float))
                yield
            finally:
                self._keep_pools(saved)
        return ctx()

    def _loo_enter(self, m):
        """Each bin's coefficient scaled by its remaining share (w - m) / w, then a warm refit on the held counts
        less m to tol (loo_sweeps: a cap; blml_binned.py:1065-1078)."""
        w = np.maximum(self.w_ - m, 0.0)
        if not float(w.sum()) > 0:
            return
        f = np.where(self.w_ > 0, w / np.where(self.w_ > 0, self.w_, 1.0), 0.0)
        self._warm_fit(w, np.asarray(self.e_hat_) * f[None], self._hill_sweeps())

    # ---------------------------------------------------------------------------------------------- burn-in
    def _hold(self, counts, modes, h, origin, complex_g, group):
        if self._held is None:
            if modes is None:
                raise ValueError("partial_fit: modes on the first call")
            self._modes_args = (modes, h, origin, complex_g)
            self._held = counts.copy()
            self._held_list = [counts]
        else:
            if counts.shape != self._held.shape:
                raise ValueError("partial_fit: the batch must be on the lattice of the first one")
            self._held = self._held + counts
            self._held_list.append(counts)
        self._held_groups.append(group)
        return float(self._held.sum()) >= self.burn_in

    def _select(self):
        """The fc left out of the modes (NaN) selected on the held counts: an 80/20 binomial split (or 20% of whole
        groups with split_groups) and the selector's search, every round of candidates one batched launch
        (blml_binned.py:899-936).  Returns the filled F (K, d)."""
        modes, h, origin, cg = self._modes_args
        d = self._held.ndim
        F = modes_fc(modes, d)
        free = ~np.isfinite(F)
        if not free.any():
            return F
        rng = np.random.default_rng(self.seed)
        groups = list(self._held_groups)
        if all(g is not None for g in groups):
            u = np.unique(groups)
            V = rng.choice(u, max(1, int(round(0.2 * len(u)))), replace=False)
            tr = sum((b for b, g in zip(self._held_list, groups) if g not in V), np.zeros(self._held.shape))
            va = sum((b for b, g in zip(self._held_list, groups) if g in V), np.zeros(self._held.shape))
        else:
            if not np.allclose(self._held, np.round(self._held)):
                raise ValueError("burn-in selection: integer counts for the binomial split (or pass split_groups)")
            tr = rng.binomial(np.round(self._held).astype(np.int64), 0.8).astype(float)
            va = self._held - tr
        specs, _ = parse_modes(_plain_modes(_modes_with_fc(modes, np.where(free, 0.1, F))), d)
        idx = np.argwhere(free)
        kinds = [specs[k]['kinds'][j].split(':')[0] for k, j in idx]
        hh = np.broadcast_to(np.asarray(h, float), (d,))
        periods = [None if specs[k]['periods'][j] is None else specs[k]['periods'][j] / hh[j] for k, j in idx]
        nyq = 0.5 / hh[idx[:, 1]]
        kw = dict(self._base_kw)
        kw.pop('engine', None)
        kw.pop('workers', None)
        kw.pop('verbose', None)
        from .binned import fit_many

        def score_many(vs):
            out = [-np.inf] * len(vs)
            ok = [i for i, v in enumerate(vs) if not np.any(np.asarray(v) > nyq + 1e-12)]
            jobs = []
            for i in ok:
                Fi = F.copy()
                Fi[free] = np.asarray(vs[i], float)
                jobs.append((tr, _plain_modes(_modes_with_fc(modes, Fi)), dict(h=h, origin=origin, complex_g=cg)))
            for i, m in zip(ok, fit_many(jobs, **kw) if jobs else []):
                out[i] = val_score(m.density_, va) if np.isfinite(m.loglik_) else -np.inf
            return out
        fc0 = 2.0 / (max(self._held.shape) * float(np.max(hh)))
        s = FcSearch(len(idx), kinds, periods, fc0=fc0, patience=2, max_sweeps=0 if self.search == 'shared' else 6,
                     factors=(np.sqrt(2.0), 2.0 ** 0.25, 2.0 ** 0.125))
        v, _ = s.run(score_many)
        F = F.copy()
        F[free] = np.asarray(v, float)
        return F

    # ---------------------------------------------------------------------------------------------- batches
    def partial_fit(self, counts, modes=None, h=1.0, origin=0.0, complex_g=False, split_groups=None, _learn=True):
        counts = np.asarray(counts, float)
        if self.n_batches_ == 0:
            if not self._hold(counts, modes, h, origin, complex_g, split_groups):
                return self
            d = self._held.ndim
            self._modes0 = _plain_modes(self._modes_args[0])
            self._T = modes_trainable(self._modes_args[0], d)
            self._F = self._select()
            _, h, origin, complex_g = self._modes_args
            cap = self.max_sweeps
            self.max_sweeps = self.batch_sweeps if self.batch_sweeps is not None and not self.burn_in else cap
            try:
                BLMLBinned.fit(self, self._held, self._modes_now(), h, origin, complex_g)
            finally:
                self.max_sweeps = cap
            self.n_batches_ = 1
            self._held_list = None
            self.history_.append(dict(n=self.n_, preq=np.nan, fc=[f.tolist() for f in self._fc_now()], moved=[],
                                      burn_in=True))
            self.preq_.append(np.nan)
            return self
        if counts.shape != tuple(self.shape_):
            raise ValueError("partial_fit: the batch must be on the lattice of the first one")

        def absorb(burned):
            w = self.decay * self.w_ + counts
            n_old, n_new = self.n_, float(w.sum())
            occ = w > 0
            E = np.asarray(self.e_hat_) * (n_new / n_old)
            E = np.where(occ[None], E, 0)
            new = occ & (np.abs(E).sum(0) == 0)               # newly occupied bins: their own root from S e
            if bool(np.any(new)):
                U, s = self._matvec_now(E)
                upd = _row_update(E.reshape(E.shape[0], -1), U.reshape(E.shape[0], -1), np.asarray(s, float),
                                  (n_new * w).ravel(), occ.ravel(), bool(self.psd_)).reshape(E.shape)
                E = np.where(new[None], upd, E)
            sweeps = self.batch_sweeps if (self.batch_sweeps is not None and burned) else self.max_sweeps
            self._warm_fit(w, E, sweeps)

        self._stream_step(counts, absorb, learn=_learn)
        self.n_batches_ += 1
        return self

    def _again(self, counts, learn=True):
        """A held batch revisited: scored leave-batch-out, its gradient pooled, the coefficients swept (not added;
        blml_binned.py:1113-1118)."""
        def absorb(burned):
            sweeps = self.batch_sweeps if (self.batch_sweeps is not None and burned) else self.max_sweeps
            self._warm_fit(self.w_, np.asarray(self.e_hat_), sweeps)
        self._stream_step(counts, absorb, revisit=True, learn=learn)

    def fit(self, batches, modes=None, epochs=2, h=1.0, origin=0.0, complex_g=False, split_groups=None, on_epoch=None,
            tol=None, tol_check='running'):
        """batches: count arrays on one lattice.  Epoch 1 adds them in random group order; later epochs revisit them
        (scored leave-batch-out, swept, the parameters moving); tol stops after the first epoch whose fit meets the
        solver's residual test with the parameters stationary (blml_binned.py:939-992)."""
        batches = [np.asarray(b, float) for b in batches]
        nb = len(batches)
        grp = None if split_groups is None else list(np.asarray(split_groups).tolist())
        if grp is not None and len(grp) != nb:
            raise ValueError("split_groups: one label per batch")
        if tol is not None and not tol > 0:
            raise ValueError("tol: None (fixed epochs) or > 0")
        if tol is not None and (self.bcd_frac < 1.0 or self.freeze_frac > 0.0):
            raise ValueError("tol needs bcd_frac = 1 and freeze_frac = 0")
        if tol_check not in ('running', 'exact'):
            raise ValueError("tol_check: 'running' or 'exact'")
        split = self.bcd_frac < 1.0
        if split:
            if not all(np.allclose(b, np.round(b)) for b in batches):
                raise ValueError("bcd_frac < 1: integer counts (the learn share is a binomial thinning)")
            r0 = np.random.default_rng([self.seed, 11])
            L = [r0.binomial(np.round(b).astype(np.int64), 1.0 - self.bcd_frac).astype(float) for b in batches]
        held, joined = [False] * nb, [False] * nb
        T_learn = int(round((1.0 - self.freeze_frac) * epochs * nb))
        t = 0
        self.converged_, self.epochs_, self.residual_ = None, 0, None
        first = dict(modes=modes, h=h, origin=origin, complex_g=complex_g)
        for ep in range(epochs):
            run_max = 0.0
            self._epoch_g = []
            for i in _batch_order(nb, grp, np.random.default_rng(self.seed + ep)):
                frozen = t >= T_learn
                t += 1
                both = not split and not frozen
                Fi = batches[i] if (not split or frozen) else batches[i] - L[i]
                if not held[i]:
                    self.partial_fit(Fi, split_groups=None if grp is None else grp[i], _learn=both and not frozen,
                                     **first)
                    held[i] = True
                    first = {}
                elif self.n_batches_ > 0:
                    run_max = max(run_max, self._residual())
                    self._again(Fi, learn=both)
                if split and not frozen and self.n_batches_ > 0:
                    self._stream_step(L[i], None, learn=True)
                if split and frozen and not joined[i] and self.n_batches_ > 0:
                    self.partial_fit(L[i], _learn=False)
                    joined[i] = True
            if self.n_batches_ == 0:
                raise ValueError(f"fit: burn_in = {self.burn_in} exceeds the counts given "
                                 f"({sum(float(b.sum()) for b in batches):g})")
            self.epochs_ = ep + 1
            if tol is not None and (tol_check == 'exact' or ep > 0):
                self.residual_ = self._residual() if tol_check == 'exact' else run_max
                still, z = self._param_stationary() if self.fc_update is not None else (True, 0.0)
                self.converged_ = bool(self.residual_ < tol and still)
                self.history_.append(dict(epoch=ep + 1, residual=self.residual_, params_stationary=still, grad_z=z))
            if on_epoch is not None:
                on_epoch(ep, self)
            if tol is not None and self.converged_:
                break
        return self
