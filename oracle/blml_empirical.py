"""CPU FP64 oracle for BLMLEmpirical -- TEST INFRASTRUCTURE ONLY.

A numpy restatement of the reference estimator `BLMLEmpirical` (blml_binned.py:1121-1391: the fit on the observed
lattice's normalisation, the regularised linear fits _lin by CG / FISTA, the hidden-phase alternation with
Anderson mixing, the outputs, score and select) and its helpers `_cg` / `_fista` (:1499-1526), on the mat-vec of
oracle/blml_oracle.py.  Only tests import it; nothing in the product does.

How each part is pinned (tests/test_oracle_empirical.py):
  _lin ridge / l2   the dense linear system (S + lam I) c = n t, (S^2 + lam I) c = n S t solved by numpy
  _lin l1           the optimality (KKT) conditions of min ||S a - t||^2 + lam ||a||_1 on the dense Gram
  FISTA step 1/L    L = 2 max|Kf|^2 >= the gradient's Lipschitz constant 2 lambda_max(S)^2 (dense eigvalsh)
  retrieve          J_history never rises (each half-step minimises J over its own variables); the end point is a
                    stationary point of J_red (|grad J_red| / its first value < outer_tol); golden vectors of the
                    reference itself (tests/golden/make_golden_empirical.py) for the trajectory-dependent outputs
  normalisation     sum_b g_b^2 Delta = 1 (window) and c'Sc / n^2 = 1 (plane) of the outputs, directly
"""
import numpy as np

from .blml_oracle import BLMLBinnedOracle, _anderson, _rdot

__all__ = ['BLMLEmpiricalOracle', 'cg', 'fista', 'kf_max']


def cg(A, b, x0, tol, maxit):
    """Conjugate gradients for SPD A, relative residual tol (blml_binned.py:1499-1510)."""
    x = np.zeros_like(b) if x0 is None else x0.copy()
    r = b - A(x); p = r.copy(); rs = float(np.sum(r * r, dtype=np.float64))
    bn = np.sqrt(float(np.sum(b * b, dtype=np.float64))) or 1.0
    for it in range(1, maxit + 1):
        Ap = A(p); a = rs / float(np.sum(p * Ap, dtype=np.float64))
        x = x + a * p; r = r - a * Ap
        rs_new = float(np.sum(r * r, dtype=np.float64))
        if np.sqrt(rs_new) / bn < tol:
            return x, it, np.sqrt(rs_new) / bn
        p = r + (rs_new / rs) * p; rs = rs_new
    return x, maxit, np.sqrt(rs) / bn


def fista(S, t, lam, L, x0, tol, maxit):
    """min ||S a - t||^2 + lam ||a||_1 by FISTA, step 1/L; stops on the relative change of a
    (blml_binned.py:1513-1526)."""
    x = np.zeros_like(t) if x0 is None else x0.copy(); y = x.copy(); tk = 1.0
    dx = nx = 1.0
    for it in range(1, maxit + 1):
        g = 2 * S(S(y) - t)
        z = y - g / L; xn = np.sign(z) * np.maximum(np.abs(z) - lam / L, 0)
        tn = float(1 + np.sqrt(1 + 4 * tk * tk)) / 2; y = xn + ((tk - 1) / tn) * (xn - x)
        dx = float(np.sqrt(np.sum((xn - x) ** 2, dtype=np.float64)))
        nx = float(np.sqrt(np.sum(xn * xn, dtype=np.float64))) or 1.0
        x, tk = xn, tn
        if dx / nx < tol:
            return x, it, dx / nx
    return x, maxit, dx / nx


def kf_max(kern):
    """max |Kf| of one mode kernel (the reference's FISTA bound, blml_binned.py:1179)."""
    return float(np.max(np.abs(kern.Kf)))


class BLMLEmpiricalOracle(BLMLBinnedOracle):
    """FP64 restatement of blml_binned.BLMLEmpirical (engine='cpu'): same constructor knobs, fit signature and
    attributes (blml_binned.py:1145-1311)."""

    def __init__(self, reg='ridge', lam=1e-2, anscombe=False, normalise='window', tol=1e-8, max_iter=20000,
                 retrieve=None, outer_tol=1e-6, max_outer=500, anderson=5, phase_noise=0.1, seed=0, workers=-1):
        super().__init__(tol=tol, anderson=anderson, phase_noise=phase_noise, seed=seed, workers=workers)
        if reg not in ('ridge', 'l2', 'l1') or normalise not in ('window', 'plane'):
            raise ValueError("reg: 'ridge' | 'l2' | 'l1'; normalise: 'window' | 'plane'")
        self.reg, self.lam, self.anscombe, self.normalise, self.max_iter = reg, float(lam), anscombe, normalise, max_iter
        self.retrieve, self.outer_tol, self.max_outer = retrieve, outer_tol, max_outer

    def fit(self, counts, modes, h=1.0, origin=0.0, init=None, complex_g=False):
        self._setup(counts, modes, h, origin)
        w = np.asarray(counts, float)
        if not w.sum() > 0:
            raise ValueError("counts: no observations")
        self.complex_g_ = bool(complex_g)
        rt = self.retrieve if self.retrieve is not None else (self.K_ > 1 or self.complex_g_)
        if not rt and (self.K_ > 1 or self.complex_g_):
            raise ValueError("BLMLEmpirical: rank K > 1 or complex_g need retrieve=True (the hidden phase)")
        self.retrieve_ = bool(rt)
        return self._fit_counts(w, init)

    # the regularised linear fit of one mode (blml_binned.py:1168-1181)
    def _lin(self, k, t, c0):
        if np.iscomplexobj(t):
            r = self._lin(k, t.real, None if c0 is None else c0.real)
            i = self._lin(k, t.imag, None if c0 is None else c0.imag)
            return r[0] + 1j * i[0], r[1] + i[1], max(r[2], i[2])
        n = self.n_; S = lambda x: self._S(k, x)
        if self.reg == 'ridge':
            return cg(lambda x: S(x) + self.lam * x, n * t, c0, self.tol, self.max_iter)
        if self.reg == 'l2':
            return cg(lambda x: S(S(x)) + self.lam * x, n * S(t), c0, self.tol, self.max_iter)
        Lmax = kf_max(self.kern_[k])
        a, it, res = fista(S, t, self.lam, 2 * Lmax ** 2, None if c0 is None else c0 / n, self.tol, self.max_iter)
        return n * a, it, res

    def _targets(self, U, t):
        a = np.sqrt((np.abs(U) ** 2).sum(0))
        Uu = U / np.where(a > 0, a, 1.0)[None]
        Uu[0] = np.where(a > 0, Uu[0], 1.0)
        return t[None] * Uu

    def _Jred(self, C, U, t):
        return self._objective(C, U, self._targets(U, t))

    def _grad_norm(self, C, U, t):
        n = self.n_
        a = np.sqrt((np.abs(U) ** 2).sum(0)); As = np.maximum(a, 1e-30 * (float(np.max(a)) or 1.0))
        g = (2 / n) * self._Sall(((a / n - t) / As)[None] * U)
        if self.reg == 'ridge':
            g = g + 2 * self.lam * U / n ** 2
        elif self.reg == 'l2':
            g = g + 2 * self.lam * C / n ** 2
        return float(np.sqrt(np.sum(np.abs(g) ** 2)))

    def _objective(self, C, U, T):
        n = self.n_
        J = float(np.sum(np.abs(U / n - T) ** 2, dtype=np.float64))
        if self.reg == 'ridge':
            J += self.lam * _rdot(C, U) / n ** 2
        elif self.reg == 'l2':
            J += self.lam * float(np.sum(np.abs(C) ** 2, dtype=np.float64)) / n ** 2
        else:
            ab = (np.abs(C.real) + np.abs(C.imag)) if np.iscomplexobj(C) else np.abs(C)
            J += self.lam * float(np.sum(ab, dtype=np.float64)) / n
        return J

    def _fit_counts(self, w, init=None):
        """blml_binned.py:1218-1291."""
        n = float(w.sum())
        self.Delta_ = float(np.prod(self.h_))
        r = np.sqrt(w + 0.375) if self.anscombe else np.sqrt(w)
        r = r * np.sqrt(n / (r ** 2).sum())
        t = r / np.sqrt(n * self.Delta_)
        self.w_ = w; self.n_ = n; self.occ_ = w > 0
        self.psd_ = False
        crit = 0.0
        if not self.retrieve_:
            c0 = None if init is None or getattr(init, 'c_raw_', None) is None or init.c_raw_.shape != self.shape_ \
                else np.asarray(init.c_raw_, float)
            c, it, res = self._lin(0, t, c0)
            C, U = c[None], self._S(0, c)[None]; self.c_raw_ = c
            self.n_outer_, self.J_history_ = 0, []
        else:
            dt = np.complex128 if self.complex_g_ else np.float64
            if init is not None and getattr(init, 'retrieve_', False) and \
                    init.c_raw_.shape == (self.K_,) + self.shape_:
                C = np.asarray(init.c_raw_).astype(dt) * (n / init.n_raw_)
                U = self._Sall(C)
            else:
                c, _, _ = self._lin(0, t, None); u0 = self._S(0, c)
                rng = np.random.default_rng(self.seed); pn = self.phase_noise; sh = self.shape_
                U = np.zeros((self.K_,) + sh, dtype=dt); U[0] = u0
                if self.complex_g_:
                    U[0] = u0 * np.exp(1j * pn * rng.standard_normal(sh))
                for k in range(1, self.K_):
                    z = rng.standard_normal(sh) + (1j * rng.standard_normal(sh) if self.complex_g_ else 0)
                    U[k] = pn * np.abs(u0) * (z / (np.sqrt(2) if self.complex_g_ else 1))
                C = None
            state = dict(it=0, res=0.0)

            def step(C, U):
                T = self._targets(U, t); Cn = []
                for k in range(self.K_):
                    ck, i_, r_ = self._lin(k, T[k], None if C is None else C[k])
                    Cn.append(ck); state['it'] += i_; state['res'] = max(state['res'], r_)
                Cn = np.stack(Cn)
                return Cn, self._Sall(Cn)

            C, U = step(C, U); Jc = self._Jred(C, U, t); Jh = [Jc]
            g0 = self._grad_norm(C, U, t) or 1.0
            hX, hF = [], []; mA = self.anderson; n_and = 0
            outer = 0
            for outer in range(1, self.max_outer + 1):
                CP, UP = step(C, U); JP = self._Jred(CP, UP, t)
                if mA:
                    f = (CP - C).ravel(); hX.append(C.ravel()); hF.append(f)
                    hX, hF = hX[-(mA + 1):], hF[-(mA + 1):]
                    if len(hX) >= 3:
                        try:
                            CA = (C.ravel() + f - _anderson(hX, hF, f)).reshape(C.shape); UA = self._Sall(CA)
                            JA = self._Jred(CA, UA, t)
                            if JA < JP:
                                CP, UP, JP = CA, UA, JA; n_and += 1
                        except np.linalg.LinAlgError:
                            hX, hF = [], []
                dC = float(np.sqrt(np.sum(np.abs(CP - C) ** 2))) / (float(np.sqrt(np.sum(np.abs(CP) ** 2))) or 1.0)
                C, U, Jc = CP, UP, JP; Jh.append(Jc)
                crit = dC if self.reg == 'l1' else self._grad_norm(C, U, t) / g0
                if crit < self.outer_tol:
                    break
            self.n_anderson_ = n_and
            self.c_raw_ = C; self.n_outer_, self.J_history_ = outer, Jh
            it, res = state['it'], state['res']
        self.outer_crit_ = crit
        self.n_raw_ = n
        self.n_iter_, self.residual_, self.converged_ = it, res, res < self.tol
        F = (np.abs(U) ** 2).sum(0)
        Zw = float(np.sum(F, dtype=np.float64)) / n ** 2 * self.Delta_
        Zp = _rdot(C, U) / n ** 2
        self.window_norm_, self.plane_norm_ = Zw, Zp
        if not (Zw > 0 and Zp > 0):
            raise ValueError(f"BLMLEmpirical: the fit is zero (reg={self.reg!r}, lam={self.lam:g} removes every "
                             "coefficient); lower lam")
        sc = float(1 / np.sqrt(Zw if self.normalise == 'window' else Zp))
        self._set_state(C * sc, U * sc)
        self.sparsity_ = float(np.mean(np.abs(C) > 0)) if self.reg == 'l1' else 1.0
        self.n_samples_ = n; self.sweeps_ = it
        return self

    def _set_state(self, E, U):
        n = self.n_
        self.e_hat_, self.Se_ = E, U
        self.density_ = (np.abs(U) ** 2).sum(0) / n ** 2
        occ = self.occ_
        self.loglik_ = float((self.w_[occ] * np.log(np.maximum(self.density_[occ], 1e-300))).sum(
            dtype=np.float64)) / n

    def score(self, counts):
        """blml_binned.py:1375-1380."""
        p = np.asarray(self.density_, np.float64); p = p / p.sum()
        w = np.asarray(counts, float); m = w > 0
        return float((w[m] * np.log(np.maximum(p[m], 1e-300))).sum() / w.sum())
