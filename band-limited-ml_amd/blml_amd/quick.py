"""BLMLQuick(engine='hip') on counts recorded on a fixed lattice: the counts path of the reference's GPU estimator
(blml_gpu.py:61-520, _BLMLQuickCUDA.fit with h in the modes; CPU twin blml_trivial.py:3516-3598), real rank 1.

    q = BLMLQuick(tol=1e-4).fit(Z, make_modes(...) with h per column, sample_weight=counts)
    q.e_hat_, q.Se_, q.n_sweeps_, q.converged_, q.residual_, q.rho_history_, q.obj_history_

With e_b = w_b c_b the BLML stationarity condition is (S e)_b e_b = n w_b over the occupied bins (bins with w = 0
are dropped); g(x) = (1/n) sum_b e_b K(x, z_b), and e'Se = n^2 at the solution.  The solver is the reference's
block-Jacobi coordinate descent (blml_gpu.py:320-441): each sweep visits the bins in the order of
default_rng(seed).permutation(B), in blocks of min(block_size, ceil(B / 8)); plain sweeps until one raises the
objective, which is then undone and redone with the block-level step-size safeguard for good; Se is recomputed
exactly (FP64 sums) when the residual first drops under tol; a stall guard (< 1% over 50 sweeps) and max_sweeps
(then e is scaled to e'Se = n^2) end it otherwise.

Each block's S[:, I] dlt is one pair-rows pass of libblml_hip.so (blml_quick_sweep): the kernel tables of the
reference's ProductKernel.tables layout in LDS, FP64 sums, and no B x |I| block in memory.  e and Se are FP64 on
the device; dtype picks the tables' precision (float32: the reference's f32 policy, products in FP32).

init= warm-starts from a fit on the same lattice (the previous e at shared bins, each new bin at its 1-D
minimiser).  Not built on this engine (they raise NotImplementedError): complex_g / rank K > 1 (the reference's
_psd_solve), conditional dep columns, init from a fit on another lattice (re-binning), and points not already
binned (modes without h)."""
import ctypes
import os

import numpy as np

from . import _lib, local
from .modes import parse_modes

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

    # ---------------------------------------------------------------------------------------------------- fit
    def fit(self, X, modes, init=None, rebin_tol=1.0, signs=None, complex_g=False, sample_weight=None):
        Z = np.atleast_2d(np.asarray(X, dtype=np.float64))
        d = Z.shape[1]
        h = modes_lattice(modes, d)
        if h is None:
            raise NotImplementedError("BLMLQuick(engine='hip'): points are fitted on counts on a fixed lattice "
                                      "only (give h in the modes and the counts as sample_weight)")
        w = np.ones(len(Z)) if sample_weight is None else np.asarray(sample_weight, float)
        if w.shape != (len(Z),) or not np.all(w >= 0):
            raise ValueError("sample_weight: (n,) non-negative counts")
        specs, dep = parse_modes(strip_lattice(modes), d)
        if dep is not None:
            raise NotImplementedError("BLMLQuick(engine='hip'): conditional (dep) fits are not built")
        psd = bool(complex_g) or len(specs) > 1
        if psd:
            if signs is not None:
                raise ValueError("complex_g / rank > 1 / modes estimate the phases; do not pass signs as well")
            raise NotImplementedError("BLMLQuick(engine='hip'): complex / rank-K counts fits are not built")
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
        e0, new = self._warm(init, h, Zi, sg)
        sp = specs[0]
        tab = local.lattice_tables(sp['kinds'], sp['fc'], sp['periods'], h, Zi)
        f32 = self.dtype == np.float32
        out = self._solve(Zi, w, n, sg, tab, f32, e0, new)
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
        self.complex_g_, self.rank_, self.dep_ = False, 1, None
        self.kinds_, self.periods_, self._fc_arr = sp['kinds'], sp['periods'], sp['fc']
        self._h_vec, self._Zi, self._tab = h, Zi, tab
        return self

    def _warm(self, init, h, Zi, sg):
        """init= on the same lattice (blml_gpu.py:145-153, 200-208): the previous fit's e at the bins both share
        (by lattice coordinates; |e| with signs), 0 at new bins (mask `new`, started by _solve).  A fit on another
        lattice would be re-binned through init.sqrt_density (blml_gpu.py:213-219): not built here."""
        self.rebinned_, self.n_new_bins_ = False, 0
        if init is None:
            return None, None
        hp = getattr(init, '_h_vec', None)
        if hp is None or getattr(init, 'e_hat_', None) is None:
            raise ValueError("init: a fitted BLMLQuick")
        if np.shape(hp) != np.shape(h) or not np.allclose(hp, h):
            raise NotImplementedError("BLMLQuick(engine='hip'): warm starts from a fit on another lattice "
                                      "(re-binning through init.sqrt_density) are not built")
        if np.ndim(init.e_hat_) != 1 or np.iscomplexobj(init.e_hat_):
            raise NotImplementedError("BLMLQuick(engine='hip'): complex / rank-K warm starts are not built")
        key = lambda Z: [z.tobytes() for z in np.ascontiguousarray(Z, np.int64)]
        prev = dict(zip(key(init._Zi), np.asarray(init.e_hat_, float)))
        e0 = np.array([prev.get(k, np.nan) for k in key(Zi)])
        new = np.isnan(e0)
        if sg is not None:
            e0 = np.abs(e0)
        self.n_new_bins_ = int(new.sum())
        return np.where(new, 0.0, e0), new

    def _solve(self, Zi, w, n, sg, tab, f32, e0=None, new=None):
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
        rng = np.random.default_rng(self.seed)
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

    # ---------------------------------------------------------------------------------------------- queries
    def sqrt_density(self, Z):
        """g at lattice points Z (on the fit's lattice): (1/n) sum_b e_b K(z, z_b), by one pair-rows pass."""
        Zq = np.atleast_2d(np.asarray(Z, float))
        Zqi = np.round(Zq / self._h_vec).astype(np.int64)
        if np.abs(Zqi * self._h_vec - Zq).max() > 1e-6 * self._h_vec.max():
            raise NotImplementedError("sqrt_density: points off the fit's lattice are not built on this engine")
        tab = local.lattice_tables(self.kinds_, self._fc_arr, self.periods_, self._h_vec, np.vstack([self._Zi, Zqi]))
        s, _ = local.rows(Zqi, self._Zi, self.e_hat_, tab, f32=self.dtype == np.float32)
        return s / self.n_samples_

    def predict(self, Z):
        """The density at lattice points Z: g(z)^2."""
        return self.sqrt_density(Z) ** 2
