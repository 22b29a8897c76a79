"""BLMLBinned on the MI355X engine: the reference estimator's constructor, fit signature, attributes and errors
(blml_binned.py:104-518), with every fit computed by libblml_hip.so (one persistent launch per batch of fits).

    m = BLMLBinned(solver='lbfgs', dtype=np.float32).fit(counts, modes, complex_g=True)
    m.density_, m.Se_, m.e_hat_, m.n_, m.loglik_, m.certificate()
    fits = fit_many([(counts, modes, dict(complex_g=True)), ...], solver='lbfgs')      # one launch for all

Host work is validation, the model spec, the kernel tables (units.py) and the start's random draws (numpy
default_rng(seed) in the reference's order, blml_binned.py:429-436); the solve runs on the device.  Device memory
is allocated through torch (plumbing only); the engine receives raw pointers."""
import ctypes
import math

import numpy as np

from . import _lib
from .modes import parse_modes
from .units import circular_dims, mode_factor_dlogs, mode_factors, next_fast_len5

__all__ = ['BLMLBinned', 'fit_many', 'FitBatch']


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("BLMLBinned(engine='hip') needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch


class _FitSpec:
    """One validated fit: inputs and its factor tables (blml_binned.py:119-154, 383-413)."""

    def __init__(self, counts, modes, h, origin, complex_g, signs, init, seed):
        if hasattr(counts, 'detach'):
            counts = counts.detach().cpu().numpy()
        counts = np.asarray(counts.get() if hasattr(counts, 'get') else counts)
        if counts.ndim < 1 or np.any(counts < 0) or not np.all(np.isfinite(counts)):
            raise ValueError("counts: a non-negative finite array on the lattice, one axis per dimension")
        d = counts.ndim
        if d > _lib.MAXD:
            raise ValueError(f"counts: at most {_lib.MAXD} dimensions on this engine")
        shape = counts.shape
        specs, _ = parse_modes(modes, d)
        if len(specs) > _lib.MAXK:
            raise ValueError(f"at most {_lib.MAXK} modes on this engine")
        self.h = np.broadcast_to(np.asarray(h, float), (d,)).copy()
        self.origin = np.broadcast_to(np.asarray(origin, float), (d,)).copy()
        circ = circular_dims(specs, shape, self.h)
        self.pad = tuple(m if c else next_fast_len5(2 * m - 1) for m, c in zip(shape, circ))
        self.counts, self.shape, self.d, self.specs, self.modes = counts.astype(np.float64), shape, d, specs, modes
        self.K = len(specs)
        self.B = int(np.prod(shape))
        self.n = float(self.counts.sum(dtype=np.float64))
        if not self.n > 0:
            raise ValueError("counts: no observations")
        self.complex_g = bool(complex_g)
        self.psd = self.complex_g or self.K > 1
        if signs is not None and self.psd:
            raise ValueError("complex_g / rank > 1 estimate the phases; do not pass signs as well")
        self.signs = None
        if signs is not None:
            if callable(signs):
                sg = np.sign(signs(self.centres())).reshape(shape)
            else:
                sg = np.sign(np.asarray(signs, float))
            if np.any(sg[self.counts > 0] == 0):
                raise ValueError("signs: nonzero on every occupied bin")
            self.signs = np.ascontiguousarray(sg, dtype=np.float64)
        # one factor chain per distinct kernel (modes with the same kernel share it, blml_binned.py:142-154)
        key = lambda sp: (tuple(sp['fc']), tuple(sp['kinds']), tuple(sp['periods']))
        chains, self.mode_chain, s = {}, [], []
        for sp in specs:
            k = key(sp)
            if k not in chains:
                chains[k] = (len(chains),) + mode_factors(sp, shape, self.h, circ)
            self.mode_chain.append(chains[k][0])
            s.append(chains[k][1])
        self.chains = [v[2] for v in sorted(chains.values(), key=lambda v: v[0])]
        self.s = np.array(s)
        self.shared = len(chains) == 1
        self.equal_s = bool(np.allclose(self.s, self.s[0]))
        self.warm = False
        self.set_start(init, seed)

    def set_start_solution(self, er, solve=False):
        """The cold start (its draws unchanged) with er, on the lattice, in place of the real rank-1 start solve
        under mode 0 (blml_binned.py:421-425) -- or of e0 when the fit is not psd -- read by the engine as
        init_K = -1; with solve=True er is the initial point of that start solve instead, which still runs to its
        tolerance (init_K = -2: the convex solve's optimum, reached from nearer)."""
        e = np.zeros((1, self.B, 2))
        e[0, :, 0] = np.asarray(er, float).ravel()
        self.init = e
        self.init_K = -2 if solve else -1

    def set_start(self, init, seed):
        """The start: init's e_hat_ (blml_binned.py:403-408) when its shape_ matches, else the cold start; and the
        start's draws in the reference's order (blml_binned.py:415-436)."""
        shape = self.shape
        self.init = None
        kk = 0
        if init is not None and getattr(init, 'e_hat_', None) is not None and tuple(init.shape_) == tuple(shape):
            E0 = np.asarray(init.e_hat_)
            kk = min(self.K, E0.shape[0])
            e = np.zeros((kk, self.B, 2))
            e[:, :, 0] = np.real(E0[:kk]).reshape(kk, -1)
            e[:, :, 1] = np.imag(E0[:kk]).reshape(kk, -1) if np.iscomplexobj(E0) else 0.0
            self.init = e
        self.init_K = kk
        rng = np.random.default_rng(seed)
        draws = []
        if self.init is None:
            if self.psd:
                if self.complex_g:
                    draws.append(rng.standard_normal(shape).ravel())
                for _ in range(1, self.K):
                    draws.append(rng.standard_normal(shape).ravel())
                    if self.complex_g:
                        draws.append(rng.standard_normal(shape).ravel())
        else:
            for _ in range(kk, self.K):
                draws.append(rng.standard_normal(shape).ravel())
                if self.complex_g:
                    draws.append(rng.standard_normal(shape).ravel())
        self.draws = np.concatenate(draws) if draws else np.zeros(1)

    def centres(self):
        g = np.meshgrid(*[self.origin[j] + self.h[j] * np.arange(m) for j, m in enumerate(self.shape)],
                        indexing='ij')
        return np.stack([x.ravel() for x in g], 1)

    @property
    def nc(self):
        return 2 if self.complex_g else 1


class FitBatch:
    """A plan for a batch of fits and its device buffers.  build() uploads everything; run() is one persistent
    launch (asynchronous on the current stream); results() copies the outputs back."""

    def __init__(self, fitspecs, params, dtype, hist_cap=4096, spread=0):
        """spread (a batch of one fit): 0 spreads the fit over every CU when it has >= 2^18 values, 1 always,
        -1 never."""
        self.fits = list(fitspecs)
        self.spread = int(spread)
        self.params = params
        self.dtype = np.dtype(dtype)
        if self.dtype not in (np.float32, np.float64):
            raise ValueError("dtype: np.float64 (the plain FP64 method) or np.float32 (FP32 transforms, FP64 "
                             "reductions)")
        self.code = _lib.F32 if self.dtype == np.float32 else _lib.F64
        self.hist_cap = int(hist_cap)
        self.plan = None
        self._build_tables()

    def _build_tables(self):
        nf = len(self.fits)
        Fits = (_lib.Fit * nf)()
        fac_list, fdata, freqs = [], [], []
        nd = nfq = 0
        off = dict(counts=0, draws=0, init=0, signs=0, out=0, dens=0, hist=0)
        for i, fs in enumerate(self.fits):
            F = Fits[i]
            F.d, F.K, F.complex_g, F.shared, F.equal_s, F.B = fs.d, fs.K, int(fs.complex_g), int(fs.shared), \
                int(fs.equal_s), fs.B
            for j in range(_lib.MAXD):
                F.shape[j] = fs.shape[j] if j < fs.d else 1
            chain_begin = []
            for chain in fs.chains:
                chain_begin.append(len(fac_list))
                for fa in chain:
                    C = _lib.Factor()
                    C.kind = _lib.FAC_LINE if fa['kind'] == 'line' else _lib.FAC_GROUP
                    C.naxes = len(fa['axes'])
                    for a in range(C.naxes):
                        C.axes[a], C.circ[a], C.tlen[a] = fa['axes'][a], fa['circ'][a], fa['tlen'][a]
                        if fa['freqs'][a] is not None:
                            C.freq_off[a] = nfq
                            freqs.append(np.asarray(fa['freqs'][a], np.int32))
                            nfq += len(fa['freqs'][a])
                    C.data_off = nd
                    fdata.append(np.asarray(fa['data'], np.float64))
                    nd += len(fa['data'])
                    fac_list.append(C)
            for k in range(fs.K):
                ch = fs.mode_chain[k]
                F.fac_begin[k] = chain_begin[ch]
                F.fac_count[k] = len(fs.chains[ch])
                F.s[k] = fs.s[k]
            F.init_K = fs.init_K
            F.has_signs = int(fs.signs is not None)
            F.n = fs.n
            F.counts_off = off['counts']; off['counts'] += fs.B
            F.draws_off = off['draws']; off['draws'] += len(fs.draws)
            F.init_off = off['init']; off['init'] += 0 if fs.init is None else fs.init.size
            F.signs_off = off['signs']; off['signs'] += 0 if fs.signs is None else fs.B
            F.out_off = off['out']; off['out'] += fs.K * fs.B * fs.nc
            F.dens_off = off['dens']; off['dens'] += fs.B
            F.hist_off = off['hist']; off['hist'] += self.hist_cap
        self._Fits = Fits
        Facs = (_lib.Factor * max(1, len(fac_list)))(*fac_list)
        self._Facs, self._nfacs = Facs, len(fac_list)
        self._fdata = np.ascontiguousarray(np.concatenate(fdata) if fdata else np.zeros(1))
        self._freqs = np.ascontiguousarray(np.concatenate(freqs) if freqs else np.zeros(1, np.int32))
        self.sizes = off
        self.counts = np.concatenate([fs.counts.ravel() for fs in self.fits])
        self.draws = np.concatenate([fs.draws for fs in self.fits])
        inits = [fs.init.ravel() for fs in self.fits if fs.init is not None]
        self.inits = np.concatenate(inits) if inits else None
        signs = [fs.signs.ravel() for fs in self.fits if fs.signs is not None]
        self.signs = np.concatenate(signs) if signs else None

    def params_struct(self):
        """The batch's solver knobs as the C-ABI's blml_params_t."""
        P = self.params
        prm = _lib.Params()
        prm.tol = float(P['tol'])
        prm.max_sweeps = -1 if P['max_sweeps'] is None else int(P['max_sweeps'])
        prm.anderson = int(P['anderson'] or 0)
        prm.stall = int(P['stall'])
        prm.phase_noise = float(P['phase_noise'])
        prm.ll_tol = float('nan') if P['ll_tol'] is None else float(P['ll_tol'])
        prm.ll_window = int(P['ll_window'])
        prm.solver = 1 if P['solver'] == 'lbfgs' else 0
        prm.lbfgs_m = int(P['lbfgs_m'])
        prm.hist_cap = self.hist_cap
        prm.spread = self.spread
        return prm

    def build(self):
        """Creates the plan (spectra on the device, workspace) and uploads the inputs."""
        torch = _torch()
        lib = _lib.load()
        prm = self._prm = self.params_struct()
        plan = ctypes.c_void_p()
        _lib.check(lib.blml_plan_create(self._Fits, len(self.fits), self._Facs, self._nfacs,
                                        self._fdata.ctypes.data, len(self._fdata), self._freqs.ctypes.data,
                                        len(self._freqs), self.code, ctypes.byref(prm), ctypes.byref(plan)))
        self.plan = plan
        dev = 'cuda'
        tdt = torch.float32 if self.code == _lib.F32 else torch.float64
        f64 = torch.float64
        self.d_counts = torch.from_numpy(self.counts).to(dev)
        self.d_draws = torch.from_numpy(self.draws).to(dev)
        self.d_init = None if self.inits is None else torch.from_numpy(self.inits).to(dev)
        self.d_signs = None if self.signs is None else torch.from_numpy(self.signs).to(dev)
        nout = max(1, self.sizes['out'])
        self.d_e = torch.zeros(nout, dtype=tdt, device=dev)
        self.d_se = torch.zeros(nout, dtype=tdt, device=dev)
        self.d_dens = torch.zeros(max(1, self.sizes['dens']), dtype=f64, device=dev)
        self.d_res = torch.zeros(len(self.fits) * _lib.RESULT_DTYPE.itemsize, dtype=torch.uint8, device=dev)
        self.d_mh = torch.zeros(max(1, self.sizes['hist']), dtype=f64, device=dev)
        self.d_rh = torch.zeros(max(1, self.sizes['hist']), dtype=f64, device=dev)
        return self

    def autotune(self, reps=3):
        """The plan's autotuner (blml_plan_autotune): times the mat-vec under each candidate configuration (FFT
        radix-stage order; LDS-resident intermediates for batched plans) and keeps the fastest.  Sets self.tuned =
        dict(choice=(sched, resident), seconds={candidate: mean mat-vec seconds})."""
        torch = _torch()
        times = np.full(_lib.TUNE_MAX, np.nan)
        choice = ctypes.c_int32()
        _lib.check(_lib.load().blml_plan_autotune(self.plan, int(reps), times.ctypes.data, ctypes.byref(choice),
                                                  self.stream()))
        torch.cuda.synchronize()
        code = choice.value
        self.tuned = dict(choice=dict(fft_sched=code // 2, lds_resident=code % 2 == 0),
                          seconds={f'sched{c // 2}_{"resident" if c % 2 == 0 else "global"}': float(t)
                                   for c, t in enumerate(times) if np.isfinite(t)})
        return self

    def refresh_init(self):
        """Copies the specs' current inits to the device (the same layout as at build(): every spec keeps its
        init_K, so only the values change; a warm-start refresh between launches)."""
        inits = [fs.init.ravel() for fs in self.fits if fs.init is not None]
        new = np.concatenate(inits) if inits else None
        if (new is None) != (self.d_init is None) or (new is not None and new.size != self.d_init.numel()):
            raise ValueError("refresh_init: the inits' layout changed since build()")
        if new is not None:
            self.d_init.copy_(_torch().from_numpy(new))
        return self

    def info(self):
        lib = _lib.load()
        a, b = ctypes.c_int64(), ctypes.c_int64()
        c, d = ctypes.c_int32(), ctypes.c_int32()
        _lib.check(lib.blml_plan_info(self.plan, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c), ctypes.byref(d)))
        k = ctypes.c_int32()
        _lib.check(lib.blml_plan_matvec_kernel(self.plan, ctypes.byref(k)))
        return dict(out_elems=a.value, ws_bytes=b.value, lds_bytes=c.value, grid=d.value,
                    matvec_kernel=('matvec_kernel', 'matvec_grid_kernel', 'matvec_pipe_kernel')[k.value])

    def stream(self):
        return ctypes.c_void_p(_torch().cuda.current_stream().cuda_stream)

    def run(self):
        """One launch that fits every fit of the batch (asynchronous)."""
        lib = _lib.load()
        ptr = lambda t: None if t is None else ctypes.c_void_p(t.data_ptr())
        _lib.check(lib.blml_fit(self.plan, ptr(self.d_counts), ptr(self.d_draws), ptr(self.d_init),
                                ptr(self.d_signs), ptr(self.d_e), ptr(self.d_se), ptr(self.d_dens), ptr(self.d_res),
                                ptr(self.d_mh), ptr(self.d_rh), self.stream()))
        return self

    def run_empirical(self, ep, aux):
        """One launch of BLMLEmpirical's fit (blml_empirical) over every fit of a batched plan: ep the knobs
        (_lib.EmpParams), aux the per-fit max|Kf| of each mode and Delta ((nfits, EMP_AUX_LEN) doubles)."""
        torch = _torch()
        lib = _lib.load()
        ptr = lambda t: None if t is None else ctypes.c_void_p(t.data_ptr())
        self.d_aux = torch.from_numpy(np.ascontiguousarray(aux, np.float64).ravel()).to('cuda')
        self.d_emp = torch.zeros(len(self.fits) * _lib.EMP_OUT_LEN, dtype=torch.float64, device='cuda')
        _lib.check(lib.blml_empirical(self.plan, ctypes.byref(ep), ptr(self.d_counts), ptr(self.d_draws),
                                      ptr(self.d_init), ptr(self.d_aux), ptr(self.d_e), ptr(self.d_se),
                                      ptr(self.d_dens), ptr(self.d_res), ptr(self.d_mh), ptr(self.d_emp),
                                      self.stream()))
        return self

    def empirical_outputs(self):
        """(nfits, EMP_OUT_LEN) host doubles of the last run_empirical: Zw, Zp, the outer criterion, scale."""
        return self.d_emp.cpu().numpy().reshape(len(self.fits), _lib.EMP_OUT_LEN)

    def matvec(self, E):
        """U = _Sall(E) for every fit: E a device tensor laid out like the outputs (plan dtype)."""
        torch = _torch()
        U = torch.empty_like(E)
        _lib.check(_lib.load().blml_matvec(self.plan, ctypes.c_void_p(E.data_ptr()), ctypes.c_void_p(U.data_ptr()),
                                           self.stream()))
        return U

    def certificate(self, max_iter=20000, tol=1e-6):
        """lambda_max(D^1/2 S_0 D^1/2) / n of every fit (NaN without a shared kernel) and the power iterations used
        (negative: max_iter reached before the eigen-residual fell below tol * lambda)."""
        torch = _torch()
        lam = torch.zeros(len(self.fits), dtype=torch.float64, device='cuda')
        its = torch.zeros(len(self.fits), dtype=torch.int32, device='cuda')
        _lib.check(_lib.load().blml_certificate(self.plan, ctypes.c_void_p(self.d_e.data_ptr()),
                                                ctypes.c_void_p(self.d_counts.data_ptr()), int(max_iter), float(tol),
                                                ctypes.c_void_p(lam.data_ptr()), ctypes.c_void_p(its.data_ptr()),
                                                self.stream()))
        return lam.cpu().numpy(), its.cpu().numpy()

    def scalars(self):
        """The per-fit blml_result_t rows (loglik, residual, sweeps, stop reason, ...) as a numpy structured array,
        without the vectors (a small copy: what a judge of the results needs)."""
        torch = _torch()
        torch.cuda.synchronize()
        res = np.frombuffer(self.d_res.cpu().numpy().tobytes(), dtype=_lib.RESULT_DTYPE)
        if np.any(res['reserved'] != 0):
            raise RuntimeError("a grid barrier of the spread fit timed out: its results are void")
        return res

    def results(self):
        """Per-fit host outputs: list of dicts with e_hat, Se, density and the scalar attributes."""
        torch = _torch()
        torch.cuda.synchronize()
        e = self.d_e.cpu().numpy()
        se = self.d_se.cpu().numpy()
        dens = self.d_dens.cpu().numpy()
        res = np.frombuffer(self.d_res.cpu().numpy().tobytes(), dtype=_lib.RESULT_DTYPE)
        if np.any(res['reserved'] != 0):
            raise RuntimeError("a grid barrier of the spread fit timed out: its results are void")
        mh = self.d_mh.cpu().numpy()
        rh = self.d_rh.cpu().numpy()
        out = []
        for i, fs in enumerate(self.fits):
            F = self._Fits[i]
            V = fs.K * fs.B * fs.nc
            ev = e[F.out_off:F.out_off + V].reshape(fs.K, fs.B, fs.nc)
            sv = se[F.out_off:F.out_off + V].reshape(fs.K, fs.B, fs.nc)
            if fs.complex_g:
                cdt = np.complex64 if self.code == _lib.F32 else np.complex128
                ev = (ev[..., 0] + 1j * ev[..., 1]).astype(cdt)
                sv = (sv[..., 0] + 1j * sv[..., 1]).astype(cdt)
            else:
                ev, sv = ev[..., 0], sv[..., 0]
            r = res[i]
            nh = min(int(r['n_hist']), self.hist_cap)
            out.append(dict(e_hat=ev.reshape((fs.K,) + fs.shape), Se=sv.reshape((fs.K,) + fs.shape),
                            density=dens[F.dens_off:F.dens_off + fs.B].reshape(fs.shape),
                            loglik=float(r['loglik']), residual=float(r['residual']), n_sweeps=int(r['n_sweeps']),
                            stop_reason=_lib.STOP_REASONS[int(r['stop_reason'])], converged=bool(r['converged']),
                            n_anderson=int(r['n_anderson']), n_damped=int(r['n_damped']),
                            n_gradient=int(r['n_gradient']), start_sweeps=int(r['start_sweeps']),
                            n_matvec=int(r['n_matvec']), device_seconds=dict(zip(
                                ('fit', 'start', 'matvec', 'direction', 'line_search'), map(float, r['seconds']))),
                            merit_history=mh[F.hist_off:F.hist_off + nh].tolist(),
                            rho_history=rh[F.hist_off:F.hist_off + nh].tolist()))
        return out

    def close(self):
        if self.plan is not None:
            _lib.load().blml_plan_destroy(self.plan)
            self.plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Polish:
    """The plain FP64 method started at another batch's results, on the device: the reference's polish test
    (results/pball_parity.py:92-113 -- a result is a local maximum of the FP64 method when the method started there
    moves loglik by less than 1e-6 relative).  The FP64 plan (every fit of `src` warm-started with all its modes,
    init= semantics, blml_binned.py:403-408) is built once; run() copies src's current coefficients into its init
    on the device (no host round trip) and fits; rel() gives (polished - src) / |src| per fit."""

    def __init__(self, src, params, max_sweeps=200):
        torch = _torch()
        from types import SimpleNamespace
        specs = []
        for fs in src.fits:
            cdt = np.complex128 if fs.complex_g else np.float64
            init = SimpleNamespace(e_hat_=np.zeros((fs.K,) + tuple(fs.shape), cdt), shape_=fs.shape)
            p = _FitSpec.__new__(_FitSpec)
            p.__dict__.update(fs.__dict__)
            p.set_start(init, 0)
            specs.append(p)
        self.src = src
        self.batch = FitBatch(specs, dict(params, max_sweeps=max_sweeps), np.float64, hist_cap=0).build()
        # where each coefficient of src's output goes in the polish init: (K, B, 2) doubles per fit
        srcs, dsts, imag0 = [], [], []
        for F, G, fs in zip(src._Fits, self.batch._Fits, src.fits):
            V = fs.K * fs.B * fs.nc
            j = np.arange(V, dtype=np.int64)
            srcs.append(F.out_off + j)
            if fs.nc == 2:
                dsts.append(G.init_off + j)
            else:
                dsts.append(G.init_off + 2 * j)
        self._src_idx = torch.from_numpy(np.concatenate(srcs)).to('cuda')
        self._dst_idx = torch.from_numpy(np.concatenate(dsts)).to('cuda')

    def run(self):
        """Copies src's coefficients into the polish init and launches the FP64 fits (asynchronous)."""
        d = self.batch.d_init
        d.zero_()
        d.index_copy_(0, self._dst_idx, self.src.d_e.index_select(0, self._src_idx).double())
        self.batch.run()
        return self

    def rel(self, src_scalars=None):
        """(polished loglik - src loglik) / |src loglik| per fit, and the polish's blml_result_t rows."""
        a = self.src.scalars() if src_scalars is None else src_scalars
        b = self.batch.scalars()
        return (b['loglik'] - a['loglik']) / np.abs(a['loglik']), b

    def close(self):
        self.batch.close()


class BLMLBinned:
    """BLML on counts on a full regular lattice, fitted on the MI355X (blml_binned.py:104-518).  engine: 'hip'
    (alias 'gpu'); dtype np.float64 keeps the plain FP64 method, np.float32 runs FP32 transforms with FP64
    reductions.  workers and verbose are accepted for compatibility (the device has no host threads to set)."""

    def __init__(self, tol=1e-5, max_sweeps=None, anderson=5, engine='hip', dtype=np.float64, phase_noise=0.1,
                 seed=0, workers=-1, verbose=False, stall=1000, ll_tol=None, ll_window=100, solver='cd', lbfgs_m=20):
        if solver not in ('cd', 'lbfgs'):
            raise ValueError("solver: 'cd' (coordinate / Jacobi updates with Anderson mixing) or 'lbfgs'")
        if engine not in ('hip', 'gpu'):
            raise ValueError("engine: 'hip' (alias 'gpu'): this estimator runs on the MI355X engine only")
        self.solver, self.lbfgs_m = solver, int(lbfgs_m)
        self.tol, self.max_sweeps, self.anderson = tol, max_sweeps, anderson
        self.engine, self.dtype, self.phase_noise, self.seed = engine, np.dtype(dtype), phase_noise, seed
        self.workers, self.verbose, self.stall = workers, verbose, stall
        self.ll_tol, self.ll_window = ll_tol, ll_window

    def _params(self):
        return dict(tol=self.tol, max_sweeps=self.max_sweeps, anderson=self.anderson, stall=self.stall,
                    phase_noise=self.phase_noise, ll_tol=self.ll_tol, ll_window=self.ll_window, solver=self.solver,
                    lbfgs_m=self.lbfgs_m)

    def fit(self, counts, modes, h=1.0, origin=0.0, complex_g=False, signs=None, init=None):
        fs = _FitSpec(counts, modes, h, origin, complex_g, signs, init, self.seed)
        batch = FitBatch([fs], self._params(), self.dtype).build().run()
        self._adopt(fs, batch, batch.results()[0], 0)
        return self

    def _adopt(self, fs, batch, r, index):
        self._batch, self._index = batch, index
        self._gplans = self._iplan = None
        self.h_, self.origin_, self.pad_, self.shape_, self.d_ = fs.h, fs.origin, fs.pad, fs.shape, fs.d
        self.specs_, self.modes_, self.K_, self.shared_ = fs.specs, fs.modes, fs.K, fs.shared
        self.s_ = fs.s
        self.rdt_ = np.float32 if self.dtype == np.float32 else np.float64
        self.cdt_ = np.complex64 if self.dtype == np.float32 else np.complex128
        self.w_ = fs.counts.astype(self.rdt_)
        self.n_ = fs.n
        self.c_ = self.n_ * self.w_
        self.occ_ = fs.counts > 0
        self.complex_g_, self.psd_ = fs.complex_g, fs.psd
        self.signs_ = fs.signs
        self.e_hat_, self.Se_, self.density_ = r['e_hat'], r['Se'], r['density']
        self.loglik_ = r['loglik']
        self.n_samples_ = self.n_
        self.n_sweeps_, self.residual_, self.converged_ = r['n_sweeps'], r['residual'], r['converged']
        self.stop_reason_ = r['stop_reason']
        self.merit_history_, self.rho_history_ = r['merit_history'], r['rho_history']
        self.n_anderson_, self.n_damped_, self.n_gradient_ = r['n_anderson'], r['n_damped'], r['n_gradient']
        self.start_sweeps_ = r['start_sweeps']
        return self

    def centres(self):
        g = np.meshgrid(*[self.origin_[j] + self.h_[j] * np.arange(m) for j, m in enumerate(self.shape_)],
                        indexing='ij')
        return np.stack([x.ravel() for x in g], 1)

    def sqrt_density(self, X, chunk=None):
        """g_k(x) = (1/n) sum_b e_bk K_k(x - z_b) at points X (Nq, d) over the occupied bins (blml_binned.py:452-466):
        (Nq,) for rank 1, else (Nq, K); complex for complex_g fits.  The sums run on the device (blml_point_rows,
        the units' closed forms, FP64); chunk is accepted for compatibility."""
        return self._g_at(X, np.asarray(self.occ_).ravel())

    def _g_at(self, X, bins):
        """g_k at points X from the coefficients of the bins selected by the mask `bins`."""
        from .points import point_rows, point_units
        X = np.atleast_2d(np.asarray(X, float))
        if X.shape[1] != self.d_:
            raise ValueError(f"X: points with {self.d_} coordinates")
        Zb = self.centres()[bins]
        E = np.asarray(self.e_hat_).reshape(self.K_, -1)[:, bins]
        out = np.zeros((len(X), self.K_), dtype=complex if self.complex_g_ else float)
        for k, sp in enumerate(self.specs_):
            units, harm = point_units(sp)
            W = np.stack([E[k].real, E[k].imag], 1) if self.complex_g_ else np.asarray(E[k].real, float)
            S = point_rows(units, harm, X, Zb, W)
            out[:, k] = (S[:, 0] + 1j * S[:, 1] if self.complex_g_ else S[:, 0]) / self.n_
        return out[:, 0] if self.K_ == 1 else out

    def predict(self, X):
        """The density f(x) = sum_k |g_k(x)|^2 at points X (blml_binned.py:468-470)."""
        G = self.sqrt_density(X)
        return np.abs(G) ** 2 if G.ndim == 1 else (np.abs(G) ** 2).sum(1)

    def predict_grid(self, r):
        """f on the lattice refined r times per dim (origin, steps h / r; the last fine point is the last bin):
        shape ((m_j - 1) r + 1, ...) (blml_binned.py:472-499).  Evaluated at the fine points by the device's direct
        sums: the reference's shifted-table FFT pairs give the same values (the shift never reaches the table's
        mask, and circular columns wrap the periodic kernel)."""
        r = int(r)
        if r < 1:
            raise ValueError("predict_grid: r >= 1")
        fine = tuple((m - 1) * r + 1 for m in self.shape_)
        g = np.meshgrid(*[self.origin_[j] + self.h_[j] * np.arange(s) / r for j, s in enumerate(fine)], indexing='ij')
        X = np.stack([x.ravel() for x in g], 1)
        return self.predict(X).reshape(fine)

    # ---------------------------------------------------------------- fc gradients (blml_binned.py:520-563, 685-693)
    def _grad_plans(self):
        """The derivative mat-vecs of this fit, built once: [(FitBatch, entry modes)] (at most 8 entries a plan, each
        mode k's chain with unit u's table replaced by dK/dlog fc_u) and per mode the entry of each unit (None where
        fc moves in whole harmonics)."""
        if getattr(self, '_gplans', None) is not None:
            return self._gplans
        fs = self._batch.fits[self._index]
        circ = circular_dims(fs.specs, fs.shape, fs.h)
        chains, emode, where = [], [], []
        for k, sp in enumerate(fs.specs):
            _, facs = mode_factors(sp, fs.shape, fs.h, circ)
            row = []
            for u, g in enumerate(mode_factor_dlogs(sp, fs.shape, fs.h, circ)):
                if g is None:
                    row.append(None)
                    continue
                ch = list(facs)
                ch[u] = g
                row.append(len(chains))
                chains.append(ch)
                emode.append(k)
            where.append(row)
        plans = []
        for c0 in range(0, len(chains), _lib.MAXK):
            ds = _FitSpec.__new__(_FitSpec)
            ds.__dict__.update(fs.__dict__)
            part = chains[c0:c0 + _lib.MAXK]
            ds.chains, ds.K, ds.mode_chain = part, len(part), list(range(len(part)))
            ds.s, ds.shared, ds.equal_s = np.ones(len(part)), False, True
            ds.init, ds.init_K, ds.signs, ds.draws = None, 0, None, np.zeros(1)
            plans.append((FitBatch([ds], self._params(), self.dtype, hist_cap=0).build(), emode[c0:c0 + _lib.MAXK]))
        self._gplans = (plans, where)
        return self._gplans

    def _dots(self, counts, eps, plan=None, emode=()):
        """blml_grad_dots over this fit (and the derivative mat-vecs of one plan): host array of 4 + 2 nE doubles."""
        torch = _torch()
        b, i = self._batch, self._index
        F = b._Fits[i]
        fs = b.fits[i]
        V = fs.K * fs.B * fs.nc
        E = b.d_e[F.out_off:F.out_off + V]
        U = b.d_se[F.out_off:F.out_off + V]
        m = np.asarray(counts, float).reshape(-1)
        if m.size != fs.B:
            raise ValueError("counts: on the fit's lattice")
        dm = torch.from_numpy(np.ascontiguousarray(m)).to('cuda')
        nE = len(emode)
        dU = None
        if nE:
            step_ = fs.B * fs.nc
            Ed = torch.cat([b.d_e[F.out_off + k * step_:F.out_off + (k + 1) * step_] for k in emode])
            dU = plan.matvec(Ed)
        out = torch.zeros(4 + 2 * nE, dtype=torch.float64, device='cuda')
        em = (ctypes.c_int32 * max(1, nE))(*emode)
        ptr = lambda t: ctypes.c_void_p(t.data_ptr())
        _lib.check(_lib.load().blml_grad_dots(b.code, fs.K, fs.nc, fs.B, ptr(E), ptr(U), None if dU is None else ptr(dU),
                                              nE, em, ptr(dm), float(eps), ptr(out), b.stream()))
        return out.cpu().numpy()

    def param_grad(self, counts, eps=0.0):
        """d/dlog fc_u of sum_b m_b log(F_b / Z + eps) / sum m at the current coefficients, for counts m on the
        lattice (blml_binned.py:543-563): a list per mode of arrays, one entry per unit (nan where fc moves in whole
        harmonics).  The derivative mat-vecs and the reductions run on the device (blml_matvec over a plan of the
        derivative kernels, blml_grad_dots)."""
        plans, where = self._grad_plans()
        vals = []
        for plan, emode in plans:
            o = self._dots(counts, eps, plan, emode)
            Z, M, fz = o[0], o[1], o[2] / o[1]
            for j in range(len(emode)):
                vals.append(2.0 * o[4 + 2 * j] / M - fz * o[5 + 2 * j] / Z)
        return [np.array([np.nan if e is None else vals[e] for e in row]) for row in where]

    # ------------------------------------------------------- the refit response (blml_binned.py:566-628, 1485-1496)
    def _ift_ops(self):
        """The fit's FP64 device state and operators for implicit_grad: (E, U, w, S, dot, comb, elem, mask, K, B, nc),
        S applying _Sall through a one-fit plan of this fit's kernels (built once)."""
        torch = _torch()
        lib = _lib.load()
        b, i = self._batch, self._index
        F = b._Fits[i]
        fs = b.fits[i]
        K, B, nc = fs.K, fs.B, fs.nc
        V = K * B * nc
        if getattr(self, '_iplan', None) is None:
            ds = _FitSpec.__new__(_FitSpec)
            ds.__dict__.update(fs.__dict__)
            ds.init, ds.init_K, ds.signs, ds.draws = None, 0, None, np.zeros(1)
            self._iplan = FitBatch([ds], self._params(), self.dtype, hist_cap=0).build()
        plan = self._iplan
        tdt = b.d_e.dtype
        E = b.d_e[F.out_off:F.out_off + V].to(torch.float64)
        U = b.d_se[F.out_off:F.out_off + V].to(torch.float64)
        w = torch.from_numpy(np.ascontiguousarray(fs.counts.reshape(-1), np.float64)).to('cuda')
        st = b.stream()
        P = lambda t: None if t is None else ctypes.c_void_p(t.data_ptr())

        def S(x):
            return plan.matvec(x.to(tdt).contiguous()).to(torch.float64)

        def dot(x, y):
            out = torch.empty(4, dtype=torch.float64, device='cuda')
            A = np.eye(2)
            X = (ctypes.c_void_p * 2)(x.data_ptr(), y.data_ptr())
            _lib.check(lib.blml_vec_gram(2, X, 2, A.ctypes.data_as(ctypes.c_void_p), x.numel(), P(out), st))
            return float(out[1].item())

        def comb(xs, c):
            out = torch.empty_like(xs[0])
            X = (ctypes.c_void_p * len(xs))(*[x.data_ptr() for x in xs])
            _lib.check(lib.blml_vec_lincomb(len(xs), X, (ctypes.c_double * len(c))(*c), 0, xs[0].numel(), P(out), st))
            return out

        def elem(op, Kk=K, k=0, n=0.0, eps=0.0, Q=0.0, M=1.0, U_=None, dU=None, X=None, m=None, two=False):
            out = torch.empty(Kk * B * nc, dtype=torch.float64, device='cuda')
            out2 = torch.empty(B * nc, dtype=torch.float64, device='cuda') if two else None
            _lib.check(lib.blml_ift_elem(op, Kk, nc, B, k, n, eps, Q, M, P(U_), P(dU), P(X), P(w), P(m), P(out),
                                         P(out2), st))
            return (out, out2) if two else out

        mask = lambda x: elem(4, X=x)
        return E, U, w, S, dot, comb, elem, mask, K, B, nc

    def implicit_grad(self, counts, eps=0.0, tol=1e-6, maxit=300):
        """d/dlog fc_u of param_grad's score with the coefficients' refit response (blml_binned.py:566-581):
        G_u - <v, d grad / dlog fc_u>, H v = dP/de by CG (blml_binned.py:1485-1496) over the occupied bins, H the
        Hessian of the fit's objective at its stationary point.  Every vector is FP64 on the device; the
        Hessian-vector products are two mat-vecs each through the fit's own kernels (blml_matvec) plus their
        elementwise terms (blml_ift_elem); ift_iters_ and ift_res_ hold the CG's iterations and relative residual."""
        torch = _torch()
        G = self.param_grad(counts, eps)
        E, U, w, S, dot, comb, elem, mask, K, B, nc = self._ift_ops()
        n = float(self.n_)
        m = np.asarray(counts, float).reshape(-1)
        if m.size != B:
            raise ValueError("counts: on the fit's lattice")
        dm = torch.from_numpy(np.ascontiguousarray(m)).to('cuda')
        M = float(m.sum())
        Q = dot(E, U)
        src = elem(3, eps=eps, Q=Q, M=M, U_=U, m=dm)
        fz = dot(src, U)
        b = mask(comb([S(src), U], [2.0, -2.0 * fz / Q]))
        psd = bool(self.psd_)

        def A(x):
            dU = S(x)
            if not psd:
                return mask(elem(1, n=n, dU=dU, X=E, m=x))
            inner = elem(0, U_=U, dU=dU)
            dQ = 2.0 * dot(U, x)
            return mask(comb([S(inner), dU, U], [-2.0, 2.0 * n / Q, -2.0 * n * dQ / Q ** 2]))

        # CG for the symmetric positive semidefinite H on the real view of the coefficients (_icg)
        x = torch.zeros_like(b)
        r = b.clone()
        p_ = r.clone()
        rs = dot(r, r)
        bn = np.sqrt(rs) or 1.0
        its, res = maxit, np.sqrt(rs) / bn
        for it in range(1, maxit + 1):
            Ap = A(p_)
            pAp = dot(p_, Ap)
            if not pAp > 0:
                its, res = it, np.sqrt(rs) / bn
                break
            a = rs / pAp
            x = comb([x, p_], [1.0, a])
            r = comb([r, Ap], [1.0, -a])
            rs_new = dot(r, r)
            if np.sqrt(rs_new) / bn < tol:
                its, res = it, np.sqrt(rs_new) / bn
                break
            p_ = comb([r, p_], [1.0, rs_new / rs])
            rs = rs_new
            its, res = it, np.sqrt(rs) / bn
        self.ift_iters_, self.ift_res_ = its, res
        v = x
        plans, where = self._grad_plans()
        step = B * nc
        mode = lambda X, k: X[k * step:(k + 1) * step]
        tdt = self._batch.d_e.dtype
        mixed = {}
        for pi, (plan, emode) in enumerate(plans):
            Ed = torch.cat([mode(E, k) for k in emode]).to(tdt)
            dUe = plan.matvec(Ed).to(torch.float64)
            if not psd:
                for j, k in enumerate(emode):
                    mixed[pi * _lib.MAXK + j] = dot(v, mask(mode(dUe, j)))
                continue
            parts = [elem(2, k=k, U_=U, dU=mode(dUe, j), two=True) for j, k in enumerate(emode)]
            wsU = torch.cat([q[1] for q in parts]).to(tdt)
            app = plan.matvec(wsU).to(torch.float64)             # apply(dKf_e, ws U_k(e)) for every entry
            for j, k in enumerate(emode):
                dUk = mode(dUe, j)
                dQ = dot(mode(E, k), dUk)
                g = comb([S(parts[j][0]), U], [-2.0, -2.0 * n * dQ / Q ** 2])
                gk = comb([mode(g, k), mode(app, j), dUk], [1.0, -2.0, 2.0 * n / Q])
                g = torch.cat([gk if q == k else mode(g, q) for q in range(K)])
                mixed[pi * _lib.MAXK + j] = dot(v, mask(g))
        return [np.array([np.nan if e is None else G[k][u] - mixed[e] for u, e in enumerate(row)])
                for k, row in enumerate(where)]

    def fc_grad(self):
        """d loglik_ / d log fc of every (mode, unit) at the fitted coefficients (blml_binned.py:520-530): the
        training counts' param_grad."""
        return self.param_grad(self.w_)

    def log_density(self, counts, eps=0.0):
        """Mean over counts m on the lattice of log(f_b + eps), f = F / Z (blml_binned.py:685-693; loglik_ for the
        training counts, a new batch's prequential score)."""
        o = self._dots(counts, eps)
        return float(o[3] / o[1])

    def certificate(self, k=1, tol=1e-6):
        """lambda_max(D^1/2 S D^1/2) / n (shared kernel only, blml_binned.py:501-518), by power iteration on the
        device with FP64 transforms; tol is eigsh's: the iteration stops when the eigen-residual of the unit iterate
        is at most tol * lambda.  certificate_iters_ holds the iterations (negative if 100000 were not enough)."""
        if not self.shared_:
            raise ValueError("certificate: shared kernel only (per-mode kernels: no convex lift)")
        if k != 1:
            raise ValueError("certificate: k = 1 (the largest eigenvalue) on this engine")
        lam, its = self._batch.certificate(max_iter=100000, tol=tol)
        self.certificate_iters_ = int(its[self._index])
        return float(lam[self._index])


def fit_many(jobs, tol=1e-5, max_sweeps=None, anderson=5, dtype=np.float64, phase_noise=0.1, seed=0, stall=1000,
             ll_tol=None, ll_window=100, solver='cd', lbfgs_m=20, hist_cap=4096, warm=None, judge=True):
    """Fits many independent lattices in one persistent launch.  jobs: iterable of (counts, modes, fit kwargs)
    with the keywords of BLMLBinned.fit.  Returns one fitted BLMLBinned per job (same attributes).

    warm: None (cold starts), True (the packaged warm-start network) or a blml_amd.warm.WarmNet.  The network
    predicts the start of every 2-D job without init= (blml_amd.warm.apply); each such fit still runs to the
    solver's tolerance and is then judged: with judge=True the plain FP64 method is started at its result (one more
    launch, Polish) and a fit it moves by 1e-6 relative or more -- not a local maximum of the plain method to that
    tolerance -- or one that ended without converging (stop reason 'stall' or 'max_sweeps') is fitted again from
    the cold start in a second launch, whose result it keeps.  Fitted estimators carry warm_start_ (True when the
    kept result started from the network) and, for judged fits, polish_rel_ (the FP64 polish's relative change)."""
    est = BLMLBinned(tol=tol, max_sweeps=max_sweeps, anderson=anderson, dtype=dtype, phase_noise=phase_noise,
                     seed=seed, stall=stall, ll_tol=ll_tol, ll_window=ll_window, solver=solver, lbfgs_m=lbfgs_m)
    specs = []
    jobs_list = [(counts, modes, dict(kw)) for counts, modes, kw in jobs]
    for counts, modes, kw in jobs_list:
        kw = dict(kw)
        specs.append(_FitSpec(counts, modes, kw.pop('h', 1.0), kw.pop('origin', 0.0), kw.pop('complex_g', False),
                              kw.pop('signs', None), kw.pop('init', None), seed))
        if kw:
            raise TypeError(f"unknown fit keywords {sorted(kw)}")
    if warm is not None and warm is not False:
        from . import warm as _warm
        _warm.apply(_warm.WarmNet.load() if warm is True else warm, specs)
    batch = FitBatch(specs, est._params(), dtype, hist_cap=hist_cap).build().run()
    res = batch.results()
    where = [(batch, i) for i in range(len(specs))]
    prel = [None] * len(specs)
    if judge and any(fs.warm for fs in specs):
        pol = Polish(batch, est._params())
        rel, _ = pol.run().rel()
        pol.close()
        prel = [float(r) if fs.warm else None for r, fs in zip(rel, specs)]
    redo = [i for i, (fs, r) in enumerate(zip(specs, res)) if fs.warm and (
        r['stop_reason'] in ('stall', 'max_sweeps') or (prel[i] is not None and abs(prel[i]) >= 1e-6))]
    if redo:
        cold = []
        for i in redo:
            counts, modes, kw = jobs_list[i]
            cold.append(_FitSpec(counts, modes, kw.get('h', 1.0), kw.get('origin', 0.0), kw.get('complex_g', False),
                                 kw.get('signs', None), None, seed))
        b2 = FitBatch(cold, est._params(), dtype, hist_cap=hist_cap).build().run()
        for j, (i, r) in enumerate(zip(redo, b2.results())):
            specs[i], res[i], where[i] = cold[j], r, (b2, j)
    out = []
    for fs, r, (b, i) in zip(specs, res, where):
        m = BLMLBinned(tol=tol, max_sweeps=max_sweeps, anderson=anderson, dtype=dtype, phase_noise=phase_noise,
                       seed=seed, stall=stall, ll_tol=ll_tol, ll_window=ll_window, solver=solver, lbfgs_m=lbfgs_m)
        out.append(m._adopt(fs, b, r, i))
        m.warm_start_ = fs.warm
    for m, p in zip(out, prel):
        if m.warm_start_:
            m.polish_rel_ = p
    return out
