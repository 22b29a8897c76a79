"""CPU FP64 oracle for the binned BLML fit -- TEST INFRASTRUCTURE ONLY.

This module restates, in numpy + scipy (float64 throughout), the algorithm of the reference estimator
`BLMLBinned` with engine='cpu' (blml_binned.py:40-518, 1463-1573) and the kernel math it reads
(blml_trivial.py:79-122, 135-216, 330-716, 949-972, 1247-1323, 1467-1505).  It is the checker of the HIP
engine: only tests, `__graft_entry__.smoke()` and bench.py's cpu_baseline leg import it, and never as the thing
measured or shipped.  Nothing in the product imports this file.

How each part is pinned (tests/test_oracle_*.py):
  kernels       closed forms: pball and periodic two-slit bands = explicit harmonic sums, pball on one column =
                dirichlet, fill=1 = plain dirichlet, ball in 1-D = sinc, slits = difference of two sincs
                (reference checks: results/pball_parity.py:28-39, results/pslits_parity.py:24-35)
  _S            brute force: the dense Gram sum over the lattice (results/pball_parity.py:66-76 style)
  real rank 1   a dense sequential coordinate-descent solve to 1e-10 (results/binned_parity.py:78-86)
  certificate   dense eigvalsh of D^1/2 S D^1/2 / n (results/binned_checks.py:42-47)
  _secular      the secular equation residual; equal-s closed form
  complex / K>1 golden vectors written by tests/golden/make_golden.py from the reference itself
  RNG draws     numpy default_rng(seed) in the reference's draw order (blml_binned.py:429-436): parity unpinned
                by any reference test, reproduced by calling numpy in the same order.
"""
import math
import numpy as np
import scipy.fft as sfft

__all__ = ['BLMLBinnedOracle', 'make_modes', 'parse_modes', 'kernel_eval', 'kernel_units', 'pball_harmonics',
           'dense_gram', 'cd_root', 'secular', 'next_fast_len5']


# ----------------------------------------------------------------------------------------------- kernel math
def _dirichlet_N(fc, P):
    """Harmonics kept on a P-periodic column (blml_trivial.py:79-81)."""
    return max(0, int(np.floor(fc * P / 2.0 + 1e-9)))


def _dirichlet_inner_M(fc, P, fill):
    """2 Ni + 1 harmonics removed by a periodic two-slit band (blml_trivial.py:107-115)."""
    if fill is None or fill >= 1.0:
        return 0
    Ni = int(np.ceil((1.0 - fill) * fc * P / 2.0 - 1e-9)) - 1
    return max(2 * Ni + 1, 0)


def pball_harmonics(fc, periods):
    """Integer vectors k with sum (k_i / P_i)^2 <= (fc / 2)^2 and the scale 1 / prod P (blml_trivial.py:632-639)."""
    P = np.atleast_1d(np.asarray(periods, float)); R = float(fc) / 2.0
    N = [max(0, int(np.floor(R * p + 1e-9))) for p in P]
    grids = np.meshgrid(*[np.arange(-n, n + 1) for n in N], indexing='ij')
    Kv = np.stack([g.ravel() for g in grids], 1)
    keep = ((Kv / P) ** 2).sum(1) <= R * R * (1 + 1e-9) + 1e-12
    return Kv[keep], 1.0 / float(np.prod(P))


def _ball_profile(s, D):
    """FT of the ball |nu| <= 1/2 in D dims (blml_trivial.py:135-143)."""
    from scipy.special import jv, gamma
    s = np.asarray(s, float)
    out = np.full(s.shape, np.pi ** (D / 2) * 0.5 ** D / gamma(D / 2 + 1))
    nz = s > 0
    out[nz] = 0.5 ** (D / 2) * jv(D / 2, np.pi * s[nz]) / s[nz] ** (D / 2)
    return out


def _split(kind):
    name, _, tag = str(kind).partition(':')
    return name, tag


def kernel_eval(kind, fc_cols, period_cols, diffs):
    """K of one kernel unit at coordinate differences (one array per column of the unit).
    sinc: blml_trivial.py:368-370; slits: :442-444; dirichlet(+fill): :509-515; ball: :588-591; pball: :656-665."""
    name, tag = _split(kind)
    if name == 'sinc':
        f = float(fc_cols[0]); return f * np.sinc(f * diffs[0])
    if name == 'slits':
        f = float(fc_cols[0]); fi = (1.0 - float(tag)) * f
        return f * np.sinc(f * diffs[0]) - fi * np.sinc(fi * diffs[0])
    if name == 'dirichlet':
        P, f = float(period_cols[0]), float(fc_cols[0]); fill = float(tag) if tag else None
        t = np.mod(np.abs(diffs[0]), P); t = np.minimum(t, P - t)
        M = 2 * _dirichlet_N(f, P) + 1; Mi = _dirichlet_inner_M(f, P, fill)
        with np.errstate(divide='ignore', invalid='ignore'):
            v = (np.sin(np.pi * M * t / P) - np.sin(np.pi * Mi * t / P)) / np.sin(np.pi * t / P)
        return (1.0 / P) * np.where(t < 1e-10, float(M - Mi), v)
    if name == 'ball':
        fc = np.asarray(fc_cols, float)
        s = np.sqrt(sum((f * dk) ** 2 for f, dk in zip(fc, diffs)))
        return float(np.prod(fc)) * _ball_profile(s, len(fc))
    if name == 'pball':
        P = np.asarray(period_cols, float)
        Kv, sc = pball_harmonics(float(fc_cols[0]), P)
        acc = 0.0 * sum(diffs)
        for k in Kv:
            ph = 0.0
            for i in range(len(P)):
                if k[i]:
                    ph = ph + (2 * np.pi * k[i] / P[i]) * diffs[i]
            acc = acc + np.cos(ph)
        return sc * acc
    raise ValueError(f"unknown kernel kind {kind!r}")


def _is_multi(kind):
    return _split(kind)[0] in ('ball', 'pball')


def kernel_units(kinds, d):
    """[(kind, dims)]: one unit per one-column kind, then one per multi-column group in order of first appearance
    (blml_trivial.py:252-263, 330-335)."""
    units = [(kinds[j], [j]) for j in range(d) if not _is_multi(kinds[j])]
    groups = {}
    for j, kind in enumerate(kinds):
        if _is_multi(kind):
            name, tag = _split(kind)
            groups.setdefault(tag, (name, []))[1].append(j)
    return units + [(name, dims) for name, dims in groups.values()]


def unit_params(kind, spec, dims):
    """(fc of the unit's columns, periods of the unit's columns) (blml_trivial.py:965-972)."""
    name = _split(kind)[0]
    fc = np.asarray(spec['fc'], float)[list(dims)]
    if name == 'pball':
        per = [spec['periods'][j] for j in dims]
    elif name == 'dirichlet':
        per = [spec['periods'][dims[0]]]
    else:
        per = [None]
    return fc, per


def spec_eval(spec, d, diffs):
    """The product kernel of one mode at per-dim coordinate differences."""
    K = 1.0
    for kind, dims in kernel_units(spec['kinds'], d):
        fc, per = unit_params(kind, spec, dims)
        K = K * kernel_eval(kind, fc, per, [diffs[j] for j in dims])
    return K


# ----------------------------------------------------------------------------------------------- model spec
_KINDS = ('sinc', 'dirichlet', 'ball', 'slits', 'pball')


def parse_modes(modes, d):
    """The subset of blml_trivial.parse_modes (:1247-1323) the binned fit reads: per mode fc (d,), kind tokens (d,)
    and periods (d,).  Multi-column units are tuple keys; their kind token is '<name>:<c0>-<c1>-...'."""
    if isinstance(modes, dict) or not isinstance(modes, (list, tuple)) or not modes:
        raise TypeError("modes: a non-empty list of mode dicts")
    specs = []
    for k, m in enumerate(modes):
        cols = {}
        for key, c in m.items():
            kind = c.get('kind', 'sinc')
            if kind not in _KINDS:
                raise ValueError(f"mode {k}: kind {kind!r}")
            if isinstance(key, tuple):
                if not _is_multi(kind):
                    raise ValueError(f"columns {key}: a tuple key needs a multi-column kind")
                per = c.get('period')
                for i, j in enumerate(key):
                    cj = dict(c, _unit=key)
                    if isinstance(per, (list, tuple, np.ndarray)):
                        cj['period'] = float(per[i])
                    cols[j] = cj
            else:
                if _is_multi(kind):
                    raise ValueError(f"column {key}: kind {kind!r} spans columns")
                cols[key] = c
        if set(cols) != set(range(d)):
            raise ValueError(f"mode {k}: columns {sorted(cols)}, expected 0..{d - 1}")
        fc, kinds, periods = [], [], []
        for j in range(d):
            c = cols[j]; kind = c.get('kind', 'sinc')
            if c.get('fc') is None or not float(c['fc']) > 0:
                raise ValueError(f"mode {k} column {j}: fc must be positive")
            fc.append(float(c['fc']))
            if kind in ('dirichlet', 'pball') and c.get('period') is None:
                raise ValueError(f"mode {k} column {j}: period missing")
            periods.append(None if kind not in ('dirichlet', 'pball') else float(c['period']))
            if kind == 'slits' and c.get('fill') is None:
                raise ValueError(f"mode {k} column {j}: slits fill missing")
            if c.get('fill') is not None and not 0 < float(c['fill']) <= 1:
                raise ValueError(f"mode {k} column {j}: fill must be in (0, 1]")
            if '_unit' in c:
                kinds.append(f"{kind}:{'-'.join(map(str, c['_unit']))}")
            elif kind in ('slits', 'dirichlet') and c.get('fill') is not None:
                kinds.append(f"{kind}:{float(c['fill'])!r}")
            else:
                kinds.append(kind)
        specs.append(dict(fc=np.array(fc), kinds=kinds, periods=periods))
    return specs


def make_modes(fc, kinds=None, periods=None, rank=1, d=None):
    """blml_trivial.make_modes (:1467-1505) without dep / NaN templates."""
    F = np.asarray(fc, float)
    if d is None:
        d = len(kinds) if kinds is not None else F.shape[-1]
    F = F if F.ndim == 2 else np.broadcast_to(F, (d,))[None]
    F = np.broadcast_to(F, (int(rank) if len(F) == 1 else len(F), d))
    kinds = list(kinds) if kinds is not None else ['sinc'] * d
    groups = {}
    for j, kd in enumerate(kinds):
        if _is_multi(kd):
            groups.setdefault(_split(kd)[1], []).append(j)
    units = [tuple(v) for v in groups.values()]

    def col(k, j):
        kind, _, tag = str(kinds[j]).partition(':')
        c = dict(kind=kind, fc=float(F[k, j]))
        if kind in ('slits', 'dirichlet') and tag:
            c['fill'] = float(tag)
        if kind == 'dirichlet':
            c['period'] = periods[j]
        if kind == 'pball':
            key = next(u for u in units if j in u)
            c['period'] = [periods[i] for i in key]
        return c

    out = []
    for k in range(len(F)):
        m = {}
        for j in range(d):
            key = next((u for u in units if j in u), j)
            if key == j or j == key[0]:
                m[key] = col(k, j)
        out.append(m)
    return out


def next_fast_len5(n):
    """Smallest 5-smooth integer >= n (scipy.fft.next_fast_len(n, real=True), blml_binned.py:134, 1533-1546)."""
    return int(sfft.next_fast_len(int(n), real=True))


def dense_gram(spec, shape, h, idx_a=None, idx_b=None):
    """S_ab = K(z_a - z_b) over the lattice bins (flat C order), by direct evaluation of the kernel."""
    d = len(shape); h = np.broadcast_to(np.asarray(h, float), (d,))
    Z = np.stack([g.ravel() for g in np.meshgrid(*[np.arange(m) * h[j] for j, m in enumerate(shape)],
                                                  indexing='ij')], 1)
    ZA = Z if idx_a is None else Z[idx_a]; ZB = Z if idx_b is None else Z[idx_b]
    diffs = [ZA[:, j, None] - ZB[None, :, j] for j in range(d)]
    return np.broadcast_to(spec_eval(spec, d, diffs), (len(ZA), len(ZB))).astype(float)


# ----------------------------------------------------------------------------------------------- scalar updates
def cd_root(b, c, s):
    """Positive root of s t^2 + b t - c = 0, cancellation free (blml_trivial.py:84-89)."""
    disc = np.sqrt(b * b + 4.0 * s * c)
    with np.errstate(divide='ignore', invalid='ignore'):
        return np.where(b >= 0, 2.0 * c / np.maximum(b + disc, 1e-300), (-b + disc) / (2.0 * s))


def secular(R2, s, c, iters=16):
    """Root lam > max s of lam sum_k R2_k / (lam - s_k)^2 = c per bin: safeguarded Newton on phi^-1/2
    (blml_binned.py:1556-1573).  R2 (K, ...), s (K, 1, ...), c (...)."""
    smax = float(np.max(s)); tot = R2.sum(0); d = s - smax
    cpos = np.maximum(c, 1e-30)
    hi = (tot + np.sqrt(tot * tot + 4 * cpos * smax * tot)) / (2 * cpos)
    lo = np.zeros_like(hi); mu = hi
    for _ in range(iters):
        with np.errstate(divide='ignore', over='ignore', invalid='ignore'):
            q = 1.0 / (mu[None] - d); W = R2 * (q * q); A = W.sum(0); B = (W * q).sum(0)
            lam = smax + mu; phi = lam * A
            big = phi > c
            lo = np.where(big, mu, lo); hi = np.where(big, hi, mu)
            nt = mu + 2 * phi * (1 - np.sqrt(phi / cpos)) / (A - 2 * lam * B)
        mu = np.where((nt >= lo) & (nt <= hi), nt, 0.5 * (lo + hi))
    return smax + mu


def _rdot(a, b):
    if np.iscomplexobj(a) or np.iscomplexobj(b):
        return float(np.real(np.sum(np.conj(a) * b, dtype=np.complex128)))
    return float(np.sum(a * b, dtype=np.float64))


def _anderson(hX, hF, f):
    """Anderson type II combination (dX + dF) gamma from FP64 normal equations (blml_binned.py:1463-1478)."""
    dX = [hX[i + 1] - hX[i] for i in range(len(hX) - 1)]
    dF = [hF[i + 1] - hF[i] for i in range(len(hF) - 1)]
    m = len(dF); G = np.empty((m, m)); rhs = np.array([_rdot(a, f) for a in dF])
    for i in range(m):
        for j in range(i, m):
            G[i, j] = G[j, i] = _rdot(dF[i], dF[j])
    g = np.linalg.solve(G + 1e-12 * np.trace(G) * np.eye(m), rhs)
    if not np.all(np.isfinite(g)):
        raise np.linalg.LinAlgError
    out = float(g[0]) * (dX[0] + dF[0])
    for i in range(1, m):
        out = out + float(g[i]) * (dX[i] + dF[i])
    return out


# ----------------------------------------------------------------------------------------------- the estimator
class _Kern:
    """One distinct mode kernel on the padded lattice: table T, s = T[0], Kf = rfftn(T) (blml_binned.py:49-90)."""

    def __init__(self, spec, shape, h, pad):
        d = len(shape); diffs, masks = [], []
        for j, (m, L) in enumerate(zip(shape, pad)):
            idx = np.arange(L); off = np.where(idx < (L + 1) // 2, idx, idx - L).astype(float)
            mk = np.ones(L) if L == m else (np.abs(off) <= m - 1).astype(float)
            sh = [1] * d; sh[j] = L
            diffs.append((off * h[j]).reshape(sh)); masks.append(mk.reshape(sh))
        T = spec_eval(spec, d, diffs)
        for mk in masks:
            T = T * mk
        T = np.broadcast_to(T, tuple(pad)).astype(np.float64)
        self.spec, self.diffs, self.masks = spec, diffs, masks
        self.T = T
        self.s = float(T[(0,) * d])
        self.Kf = sfft.rfftn(T, s=tuple(pad), workers=-1)


class BLMLBinnedOracle:
    """FP64 restatement of blml_binned.BLMLBinned (engine='cpu').  Same constructor knobs, fit signature and
    attributes (blml_binned.py:107-116, 383-445)."""

    def __init__(self, tol=1e-5, max_sweeps=None, anderson=5, phase_noise=0.1, seed=0, stall=1000, ll_tol=None,
                 ll_window=100, solver='cd', lbfgs_m=20, workers=-1):
        if solver not in ('cd', 'lbfgs'):
            raise ValueError("solver: 'cd' or 'lbfgs'")
        self.tol, self.max_sweeps, self.anderson = tol, max_sweeps, anderson
        self.phase_noise, self.seed, self.stall = phase_noise, seed, stall
        self.ll_tol, self.ll_window, self.solver, self.lbfgs_m = ll_tol, ll_window, solver, int(lbfgs_m)
        self.workers = workers

    # setup (blml_binned.py:119-154)
    def _setup(self, counts, modes, h, origin):
        counts = np.asarray(counts)
        if counts.ndim < 1 or np.any(counts < 0) or not np.all(np.isfinite(counts)):
            raise ValueError("counts: a non-negative finite array on the lattice")
        d = counts.ndim; shape = counts.shape
        specs = parse_modes(modes, d)
        self.h_ = np.broadcast_to(np.asarray(h, float), (d,)).copy()
        self.origin_ = np.broadcast_to(np.asarray(origin, float), (d,)).copy()
        pad = []
        for j in range(d):
            circ = all(_split(sp['kinds'][j])[0] in ('dirichlet', 'pball') and sp['periods'][j] is not None and
                       abs(shape[j] * self.h_[j] - sp['periods'][j]) <= 1e-9 * sp['periods'][j] for sp in specs)
            pad.append(shape[j] if circ else next_fast_len5(2 * shape[j] - 1))
        self.pad_, self.shape_, self.d_, self.specs_ = tuple(pad), shape, d, specs
        key = lambda sp: (tuple(sp['fc']), tuple(sp['kinds']), tuple(sp['periods']))
        cache = {}; self.kern_ = []
        for sp in specs:
            k = key(sp)
            if k not in cache:
                cache[k] = _Kern(sp, shape, self.h_, self.pad_)
            self.kern_.append(cache[k])
        self.shared_ = len(cache) == 1
        self.s_ = np.array([kk.s for kk in self.kern_])
        self.K_ = len(specs)

    # mat-vec (blml_binned.py:156-168)
    def _S(self, k, e):
        kk = self.kern_[k]; sl = tuple(slice(0, m) for m in self.shape_)
        conv = lambda x: sfft.irfftn(sfft.rfftn(x, s=self.pad_, workers=self.workers) * kk.Kf, s=self.pad_,
                                     workers=self.workers)[sl]
        if np.iscomplexobj(e):
            return conv(e.real) + 1j * conv(e.imag)
        return conv(e)

    def _Sall(self, E):
        return np.stack([self._S(k, E[k]) for k in range(self.K_)])

    # merit, update, rescale, residual (blml_binned.py:171-220)
    def _merit(self, E, U):
        c, occ, n = self.c_, self.occ_, self.n_
        if self.psd_:
            F = (np.abs(U) ** 2).sum(0); Q = _rdot(E, U)
            if not Q > 0:
                return -np.inf
            return float((c * np.log(np.maximum(F, 1e-30))).sum(dtype=np.float64)) / n - n * np.log(Q)
        e, u = E[0], U[0]
        if np.any(occ & (self.sg_ * e <= 0)):
            return -np.inf
        return -(0.5 * float(np.sum(e * u, dtype=np.float64))
                 - float((c * np.log(np.where(occ, np.abs(e), 1.0))).sum(dtype=np.float64))) / n

    def _update(self, E, U):
        c = self.c_
        s = self.s_.reshape((-1,) + (1,) * self.d_)
        R = U - s * E
        if not self.psd_:
            sg = self.sg_
            t = np.where(self.occ_, cd_root(sg * R[0], c, float(self.s_[0])), 0.0)
            return (sg * t)[None].astype(E.dtype)
        a2 = (np.abs(R) ** 2).sum(0); a = np.sqrt(a2)
        if np.allclose(self.s_, self.s_[0]):
            s0 = float(self.s_[0]); t = 2 * c / (a + np.sqrt(a2 + 4 * s0 * c))
            out = R * (t / np.where(a > 0, a, 1.0))[None]
            lone = (a == 0) & (c > 0)
            if np.any(lone):
                out[0] = np.where(lone, np.sqrt(c / s0), out[0])
            return out.astype(E.dtype)
        lone = (a2 == 0) & (c > 0)
        lam = secular(np.abs(R) ** 2, s, c)
        with np.errstate(divide='ignore', invalid='ignore'):
            out = np.where((c > 0)[None] & ~lone[None], R / (lam[None] - s), 0)
        if np.any(lone):
            out[0] = np.where(lone, np.sqrt(c / float(self.s_[0])), out[0])
        return out.astype(E.dtype)

    def _rescale(self, E, U):
        Q = _rdot(E, U); sc = float(self.n_ / np.sqrt(Q)) if Q > 0 else 1.0
        return E * sc, U * sc

    def _resid(self, E, U):
        c, occ = self.c_, self.occ_
        nrm2 = (np.abs(E) ** 2).sum(0)
        r = np.sqrt((np.abs(U * (nrm2 / np.where(occ, c, 1.0))[None] - E) ** 2).sum(0)) / \
            np.sqrt(np.maximum(nrm2, 1e-300))
        return float(np.max(np.where(occ, r, 0.0)))

    # Jacobi CD + Anderson + damping + gradient fallback (blml_binned.py:223-276)
    def _solve(self, E, max_sweeps):
        if self.solver == 'lbfgs' and self.psd_:
            return self._solve_lbfgs(E, max_sweeps)
        U = self._Sall(E); E, U = self._rescale(E, U)
        cur = self._merit(E, U); hX, hF = [], []; mA = self.anderson
        n_and = n_damp = n_grad = 0; it = 0; res = self._resid(E, U); hist = [res]
        cap = np.inf if max_sweeps is None else max_sweeps
        best_res, best_it, best_m = res, 0, cur
        mhist = [cur]; self.stop_reason_ = 'tol'
        while res >= self.tol and it < cap:
            if it - best_it >= self.stall:
                self.stop_reason_ = 'stall'; break
            if self.ll_tol is not None and len(mhist) > self.ll_window and \
                    (mhist[-1] - mhist[-1 - self.ll_window]) / self.n_ < self.ll_tol:
                self.stop_reason_ = 'll_tol'; break
            it += 1
            EJ = self._update(E, U); UJ = self._Sall(EJ); mJ = self._merit(EJ, UJ)
            if mA:
                f = (EJ - E).ravel()
                hX.append(E.ravel()); hF.append(f); hX, hF = hX[-(mA + 1):], hF[-(mA + 1):]
                if len(hX) >= 3:
                    try:
                        EA = E + f.reshape(E.shape) - _anderson(hX, hF, f).reshape(E.shape)
                        UA = self._Sall(EA); mAc = self._merit(EA, UA)
                        if mAc > mJ:
                            EJ, UJ, mJ = EA, UA, mAc; n_and += 1
                    except Exception:
                        hX, hF = [], []
            if mJ < cur - 1e-12 * abs(cur):
                hX, hF = [], []; n_damp += 1
                EP = self._update(E, U); UP = self._Sall(EP); a = 0.5
                while a > 1e-8:
                    EJ, UJ = E + a * (EP - E), U + a * (UP - U); mJ = self._merit(EJ, UJ)
                    if mJ >= cur:
                        break
                    a *= 0.5
                else:
                    EJ, UJ, mJ = E, U, cur
                    if self.psd_:
                        EJ, UJ, mJ = self._grad_steps(E, U, cur)
                        n_grad += 1
            E, U = self._rescale(EJ, UJ); cur = mJ if self.psd_ else self._merit(E, U); mhist.append(cur)
            res = self._resid(E, U); hist.append(res)
            if res < 0.9 * best_res or cur > best_m + 1e-13 * abs(best_m):
                best_res, best_it, best_m = min(res, best_res), it, max(cur, best_m)
        if res >= self.tol and self.stop_reason_ == 'tol':
            self.stop_reason_ = 'max_sweeps'
        self.n_sweeps_, self.residual_, self.converged_ = it, res, res < self.tol
        self.merit_history_, self.rho_history_ = mhist, hist
        self.n_anderson_, self.n_damped_, self.n_gradient_ = n_and, n_damp, n_grad
        return E, U

    def _lgrad(self, E, U):
        """L and its gradient 2 (S_k (w u_k / F) - n u_k / Q) on occupied bins (blml_binned.py:278-288)."""
        F = (np.abs(U) ** 2).sum(0); Q = _rdot(E, U)
        if not Q > 0:
            return -np.inf, None
        L = float((self.w_ * np.log(np.where(self.occ_, F, 1.0))).sum(dtype=np.float64)) - self.n_ * np.log(Q)
        V = np.where(self.occ_, self.w_ / np.where(self.occ_, F, 1.0), 0.0)[None] * U
        G = self._Sall(V) - (self.n_ / Q) * U
        return L, 2 * np.where(self.occ_[None], G, 0)

    def _solve_lbfgs(self, E, max_iter):
        """L-BFGS ascent of L (blml_binned.py:290-355)."""
        n = self.n_; m_ = self.lbfgs_m
        U = self._Sall(E); E, U = self._rescale(E, U)
        L, g = self._lgrad(E, U)
        S_, Y_, R_ = [], [], []; it = 0; res = self._resid(*self._rescale(E, U)); hist = [res]; mhist = [L / n]
        cap = np.inf if max_iter is None else max_iter; self.stop_reason_ = 'tol'; reset = False
        ftol, fwin = (self.ll_tol if self.ll_tol is not None else 1e-11), 10
        while res >= self.tol and it < cap:
            if len(mhist) > fwin and mhist[-1] - mhist[-1 - fwin] < ftol:
                self.stop_reason_ = 'll_tol'; break
            it += 1
            q = g.copy(); al = []
            for s_, y_, r_ in zip(reversed(S_), reversed(Y_), reversed(R_)):
                a = r_ * _rdot(s_, q); al.append(a); q = q - a * y_
            if S_:
                q = q * (_rdot(S_[-1], Y_[-1]) / _rdot(Y_[-1], Y_[-1]))
            else:
                q = q * (1e-3 * np.sqrt(_rdot(E, E) / max(_rdot(g, g), 1e-300)))
            for (s_, y_, r_), a in zip(zip(S_, Y_, R_), reversed(al)):
                b = r_ * _rdot(y_, q); q = q + (a - b) * s_
            slope = _rdot(g, q)
            if not slope > 0:
                S_, Y_, R_ = [], [], []
                q = g * (1e-3 * np.sqrt(_rdot(E, E) / max(_rdot(g, g), 1e-300))); slope = _rdot(g, q)
            t = 1.0; ok = False
            for _ in range(40):
                En = E + t * q; Un = self._Sall(En); Ln, gn = self._lgrad(En, Un)
                if Ln >= L + 1e-4 * t * slope:
                    ok = True; break
                t *= 0.5
            if not ok:
                if reset or not S_:
                    self.stop_reason_ = 'stall'; break
                S_, Y_, R_ = [], [], []; reset = True; continue
            reset = False
            s_, y_ = En - E, g - gn
            sy = _rdot(s_, y_)
            if sy > 1e-12 * np.sqrt(_rdot(s_, s_) * _rdot(y_, y_)):
                S_.append(s_); Y_.append(y_); R_.append(1.0 / sy)
                if len(S_) > m_:
                    S_.pop(0); Y_.pop(0); R_.pop(0)
            E, U, L, g = En, Un, Ln, gn
            res = self._resid(*self._rescale(E, U)); hist.append(res); mhist.append(L / n)
        if res >= self.tol and self.stop_reason_ == 'tol':
            self.stop_reason_ = 'max_sweeps'
        F = (np.abs(U) ** 2).sum(0); Q = _rdot(E, U)
        Ec = np.where(self.occ_, (Q / n) * self.w_ / np.where(self.occ_, F, 1.0), 0.0)[None] * U
        Uc = self._Sall(Ec); Lc, _ = self._lgrad(Ec, Uc)
        if Lc >= L - 1e-9 * n:
            E, U = Ec, Uc
        E, U = self._rescale(E, U); res = self._resid(E, U)
        self.n_sweeps_, self.residual_ = it, res
        self.converged_ = res < self.tol or self.stop_reason_ == 'll_tol'
        self.merit_history_, self.rho_history_ = mhist, hist
        self.n_anderson_ = self.n_damped_ = self.n_gradient_ = 0
        return E, U

    def _grad_steps(self, E, U, cur, steps=5):
        """Steepest ascent fallback with Armijo backtracking (blml_binned.py:357-380)."""
        for _ in range(steps):
            F = (np.abs(U) ** 2).sum(0); Q = _rdot(E, U)
            V = np.where(self.occ_, self.w_ / np.where(self.occ_, F, 1.0), 0.0)[None] * U
            G = self._Sall(V) - (self.n_ / Q) * U
            G = np.where(self.occ_[None], G, 0)
            SG = self._Sall(G); g2 = _rdot(G, G)
            if not g2 > 0:
                break
            a = getattr(self, '_gstep', None) or 1e-3 * np.sqrt(float(np.sum(np.abs(E) ** 2, dtype=np.float64)) / g2)
            ok = False
            for _ in range(40):
                En, Un = E + a * G, U + a * SG; mn = self._merit(En, Un)
                if mn > cur:
                    ok = True; break
                a *= 0.5
            if not ok:
                break
            E, U, cur = En, Un, mn; self._gstep = 2 * a
        return E, U, cur

    # fit (blml_binned.py:383-445)
    def fit(self, counts, modes, h=1.0, origin=0.0, complex_g=False, signs=None, init=None):
        self._setup(counts, modes, h, origin)
        w = np.asarray(counts, float)
        self.w_ = w; self.n_ = float(w.sum(dtype=np.float64))
        if not self.n_ > 0:
            raise ValueError("counts: no observations")
        self.c_ = self.n_ * w; self.occ_ = w > 0
        self.complex_g_ = bool(complex_g); self.psd_ = bool(complex_g) or self.K_ > 1
        if signs is not None and self.psd_:
            raise ValueError("complex_g / rank > 1 estimate the phases; do not pass signs as well")
        sg = np.ones(self.shape_) if signs is None else np.sign(np.asarray(signs, float))
        if np.any(sg[np.asarray(counts) > 0] == 0):
            raise ValueError("signs: nonzero on every occupied bin")
        self.sg_ = sg.astype(float)
        dt = np.complex128 if self.complex_g_ else np.float64
        if init is not None and getattr(init, 'e_hat_', None) is not None and tuple(init.shape_) == self.shape_:
            E0 = np.asarray(init.e_hat_)
            E = np.zeros((self.K_,) + self.shape_, dtype=dt)
            kk = min(self.K_, E0.shape[0]); E[:kk] = E0[:kk] if self.complex_g_ else np.real(E0[:kk])
            E[:, ~self.occ_] = 0
            E = self._fill_modes(E, kk, phases=False)
        else:
            E = self._start(dt)
        E, U = self._solve(E, self.max_sweeps)
        self._finish(E, U)
        return self

    def _start(self, dt):
        c, s0 = self.c_, float(self.s_[0])
        e0 = np.where(self.occ_, np.sqrt(c / s0), 0.0) * self.sg_
        if not self.psd_:
            return e0[None]
        save = (self.psd_, self.K_, self.kern_, self.s_)
        self.psd_, self.K_, self.kern_, self.s_ = False, 1, self.kern_[:1], self.s_[:1]
        er, _ = self._solve(e0[None], self.max_sweeps)
        self.start_sweeps_ = self.n_sweeps_
        self.psd_, self.K_, self.kern_, self.s_ = save
        E = np.zeros((self.K_,) + self.shape_, dtype=dt); E[0] = er[0]
        return self._fill_modes(E, 1)

    def _fill_modes(self, E, kk, phases=True):
        rng = np.random.default_rng(self.seed); pn = self.phase_noise; sh = self.shape_
        if self.complex_g_ and phases:
            E[0] = E[0] * np.exp(1j * pn * rng.standard_normal(sh))
        for k in range(kk, self.K_):
            z = rng.standard_normal(sh)
            if self.complex_g_:
                z = z + 1j * rng.standard_normal(sh)
            E[k] = pn * np.abs(E[0]) * (z / (np.sqrt(2) if self.complex_g_ else 1))
        return E

    def _finish(self, E, U):
        self.e_hat_, self.Se_ = E, U
        F = (np.abs(U) ** 2).sum(0)
        self.density_ = F / self.n_ ** 2
        occ = self.occ_
        self.loglik_ = float((self.w_[occ] * np.log(self.density_[occ])).sum(dtype=np.float64)) / self.n_

    def centres(self):
        g = np.meshgrid(*[self.origin_[j] + self.h_[j] * np.arange(m) for j, m in enumerate(self.shape_)],
                        indexing='ij')
        return np.stack([x.ravel() for x in g], 1)

    def _dKf_fd(self, k, u, rel=1e-5):
        """The spectrum of dK/dlog fc of unit u of mode k (the unit's columns scaled together), the table by a central
        difference in log fc (error ~ rel^2); None where fc moves in whole harmonics (dirichlet, pball)."""
        sp = self.specs_[k]
        kind, dims = kernel_units(sp['kinds'], self.d_)[u]
        if _split(kind)[0] in ('dirichlet', 'pball'):
            return None
        kk = self.kern_[k]
        T = []
        for sgn in (1.0, -1.0):
            fc = np.array(sp['fc'], float)
            fc[dims] = fc[dims] * np.exp(sgn * rel)
            T2 = spec_eval(dict(sp, fc=list(fc)), self.d_, kk.diffs)
            for mk in kk.masks:
                T2 = T2 * mk
            T.append(np.broadcast_to(T2, self.pad_).astype(np.float64))
        return sfft.rfftn((T[0] - T[1]) / (2 * rel), s=self.pad_, workers=self.workers)

    def _apply(self, Kf, e):
        sl = tuple(slice(0, m) for m in self.shape_)
        conv = lambda x: sfft.irfftn(sfft.rfftn(x, s=self.pad_, workers=self.workers) * Kf, s=self.pad_,
                                     workers=self.workers)[sl]
        return conv(e.real) + 1j * conv(e.imag) if np.iscomplexobj(e) else conv(e)

    def param_grad(self, counts, eps=0.0):
        """d/dlog fc_u of sum m log(F / Z + eps) / sum m at the current coefficients (blml_binned.py:543-563), with the
        derivative spectra from central differences of the tables (_dKf_fd)."""
        E, U = np.asarray(self.e_hat_), np.asarray(self.Se_)
        m = np.asarray(counts, float); M = float(m.sum())
        F = (np.abs(U) ** 2).sum(0); Z = float(np.real(np.sum(np.conj(E) * U)))
        wn = np.where(m > 0, m / np.maximum(F + eps * Z, 1e-300), 0.0)
        fz = float(np.sum(wn * F)) / M
        out = []
        for k, sp in enumerate(self.specs_):
            gk = []
            for u in range(len(kernel_units(sp['kinds'], self.d_))):
                dKf = self._dKf_fd(k, u)
                if dKf is None:
                    gk.append(np.nan)
                    continue
                dU = self._apply(dKf, E[k])
                dZ = float(np.real(np.sum(np.conj(E[k]) * dU)))
                gk.append(2 * float(np.sum(wn * np.real(np.conj(U[k]) * dU))) / M - fz * dZ / Z)
            out.append(np.array(gk))
        return out

    def implicit_grad(self, counts, eps=0.0, tol=1e-6, maxit=300):
        """param_grad with the coefficients' refit response (blml_binned.py:566-628): G - <v, d grad / dlog fc>,
        H v = dP/de by CG (_icg, blml_binned.py:1485-1496) over the occupied bins."""
        G = self.param_grad(counts, eps)
        E, U = np.asarray(self.e_hat_), np.asarray(self.Se_)
        occ = np.asarray(self.occ_)
        w, n = np.asarray(self.w_, float), float(self.n_)
        F = (np.abs(U) ** 2).sum(0)
        Q = float(np.real(np.sum(np.conj(E) * U)))
        Fs = np.where(occ, F, 1.0)
        m = np.asarray(counts, float); M = float(m.sum())
        rdot = lambda a, b: float(np.real(np.sum(np.conj(a) * b)))
        mask = occ[None].astype(float)
        wn = m / np.maximum(F + eps * Q, 1e-300)
        fz = float(np.sum(wn * F)) / M
        b = (2 * self._Sall((wn / M)[None] * U) - fz * 2 * U / Q) * mask

        def hvp(D):
            dU = self._Sall(D)
            if not self.psd_:
                e = E[0]; c = n * w
                return dU + np.where(occ, c / np.where(occ, e * e, 1.0), 0.0)[None] * D
            dF = 2 * np.real(np.conj(U) * dU).sum(0); dQ = 2 * rdot(U, D)
            inner = (w / Fs)[None] * dU - (w * dF / Fs ** 2)[None] * U
            return -2 * self._Sall(np.where(occ[None], inner, 0)) + 2 * n * dU / Q - 2 * n * U * dQ / Q ** 2

        x = np.zeros_like(b); r = b.copy(); p = r.copy(); rs = rdot(r, r); bn = np.sqrt(rs) or 1.0
        self.ift_iters_, self.ift_res_ = maxit, np.sqrt(rs) / bn
        for it in range(1, maxit + 1):
            Ap = hvp(p) * mask; pAp = rdot(p, Ap)
            if not pAp > 0:
                self.ift_iters_, self.ift_res_ = it, np.sqrt(rs) / bn
                break
            a = rs / pAp; x = x + a * p; r = r - a * Ap; rs_new = rdot(r, r)
            if np.sqrt(rs_new) / bn < tol:
                self.ift_iters_, self.ift_res_ = it, np.sqrt(rs_new) / bn
                break
            p = r + (rs_new / rs) * p; rs = rs_new
            self.ift_iters_, self.ift_res_ = it, np.sqrt(rs) / bn
        v = x
        ws = np.where(occ, w / Fs, 0.0)
        for k, sp in enumerate(self.specs_):
            for u in range(len(kernel_units(sp['kinds'], self.d_))):
                dKf = self._dKf_fd(k, u)
                if dKf is None:
                    continue
                dUk = self._apply(dKf, E[k])
                if not self.psd_:
                    mixed = dUk[None]
                else:
                    dF = 2 * np.real(np.conj(U[k]) * dUk); dQ = rdot(E[k], dUk)
                    wd = np.where(occ, w * dF / Fs ** 2, 0.0)
                    mixed = []
                    for j in range(self.K_):
                        t = -wd * U[j]
                        if j == k:
                            t = t + ws * dUk
                        g = -2 * self._S(j, t) - 2 * n * U[j] * dQ / Q ** 2
                        if j == k:
                            g = g - 2 * self._apply(dKf, ws * U[k]) + 2 * n * dUk / Q
                        mixed.append(g)
                    mixed = np.stack(mixed)
                G[k][u] -= rdot(v, mixed * mask)
        return G

    def log_density(self, counts, eps=0.0):
        """Mean over counts m of log(F / Z + eps) (blml_binned.py:685-693)."""
        U = np.asarray(self.Se_); F = (np.abs(U) ** 2).sum(0)
        Z = float(np.real(np.sum(np.conj(np.asarray(self.e_hat_)) * U)))
        m = np.asarray(counts, float); occ = m > 0
        return float((m[occ] * np.log(np.maximum(F[occ] / Z + eps, 1e-300))).sum()) / float(m.sum())

    def sqrt_density(self, X, chunk=4096):
        """g_k(x) = (1/n) sum_b e_bk K_k(x - z_b) at points X (Nq, d), direct sums over the occupied bins
        (blml_binned.py:452-466): (Nq,) for rank 1, else (Nq, K)."""
        X = np.atleast_2d(np.asarray(X, float)); Zb = self.centres()
        E = np.asarray(self.e_hat_).reshape(self.K_, -1); occ = np.asarray(self.occ_).ravel()
        Zb, E = Zb[occ], E[:, occ]
        out = np.zeros((len(X), self.K_), dtype=complex if self.complex_g_ else float)
        for i in range(0, len(X), chunk):
            diffs = X[i:i + chunk, None, :] - Zb[None, :, :]
            for k, sp in enumerate(self.specs_):
                Kv = spec_eval(sp, self.d_, [diffs[..., j] for j in range(self.d_)])
                out[i:i + chunk, k] = Kv @ E[k] / self.n_
        return out[:, 0] if self.K_ == 1 else out

    def predict(self, X):
        G = self.sqrt_density(X)
        return np.abs(G) ** 2 if G.ndim == 1 else (np.abs(G) ** 2).sum(1)

    def predict_grid(self, r):
        """f on the lattice refined r times per dim, one FFT pair per sub-lattice shift and mode: the mode's table at
        the offsets shifted by h sh / r (blml_binned.py:468-499)."""
        import itertools
        d = self.d_; sl = tuple(slice(0, m) for m in self.shape_)
        fine = tuple((m - 1) * r + 1 for m in self.shape_)
        out = np.zeros(fine)
        for sh in itertools.product(range(r), repeat=d):
            dens = 0.0
            for k, kk in enumerate(self.kern_):
                diffs = [kk.diffs[j] + self.h_[j] * sh[j] / r for j in range(d)]
                T = spec_eval(kk.spec, d, diffs)
                for mk in kk.masks:
                    T = T * mk
                Kf = sfft.rfftn(np.broadcast_to(T, self.pad_).astype(np.float64), s=self.pad_, workers=self.workers)
                e = np.asarray(self.e_hat_[k])
                parts = [e.real, e.imag] if np.iscomplexobj(e) else [e]
                g = [sfft.irfftn(sfft.rfftn(p, s=self.pad_, workers=self.workers) * Kf, s=self.pad_,
                                 workers=self.workers)[sl] for p in parts]
                gk = g[0] + 1j * g[1] if len(g) == 2 else g[0]
                dens = dens + np.abs(gk / self.n_) ** 2
            idx = tuple(slice(s_, None, r) for s_ in sh)
            tgt = out[idx]
            out[idx] = dens[tuple(slice(0, n_) for n_ in tgt.shape)]
        return out

    def certificate(self, k=1, tol=1e-6):
        """lambda_max(D^1/2 S D^1/2) / n by ARPACK on FP64 FFT mat-vecs (blml_binned.py:501-518)."""
        if not self.shared_:
            raise ValueError("certificate: shared kernel only (per-mode kernels: no convex lift)")
        from scipy.sparse.linalg import LinearOperator, eigsh
        D = np.where(self.occ_, (np.abs(self.e_hat_) ** 2).sum(0) / np.where(self.occ_, self.w_, 1.0), 0.0)
        dh = np.sqrt(D); B = int(np.prod(self.shape_))
        op = lambda x: (dh * self._S(0, dh * x.reshape(self.shape_))).ravel()
        lam = eigsh(LinearOperator((B, B), matvec=op, dtype=np.float64), k=k, which='LA', tol=tol,
                    return_eigenvectors=False)
        return float(np.max(lam)) / self.n_
