"""Learned warm start for the binned fit: a small convolutional network maps a lattice's counts to the start of the
fit, trained with the engine's own kernels (libblml_hip.so: blml_net_*, BF16 matrix cores, FP32 accumulation).

What the network predicts.  A cold complex / rank-K fit starts from the real rank-1 solution under mode 0 (a
coordinate-descent solve, blml_binned.py:415-427), then draws phases and the other modes (:429-436).  That solve is
convex with a unique optimum, so it is a well-posed regression target: the network predicts
log(e_b / sqrt(c_b / s_0)) on a 64 x 64 grid (c_b the counts, s_0 = K_0(0)), from log(1 + counts), the occupancy and
four planes of the kernel's parameters.  The warm start is the cold start with that solve replaced by the prediction
-- the same phase and mode draws from default_rng(seed), applied by the engine (a fit's init_K = -1).  The fit then
runs to the plain method's tolerance as usual (fit_many(..., warm=True) applies it and fits again
from the cold start any warm fit that ends without converging).  The C3-like fits have two modes and are
non-convex: from another start they end at another local maximum, as an FP32 and an FP64 run from the same start
do too, so a warm fit's likelihood is judged by the FP64 polish test (a local maximum) rather than by equality
with the cold fit's.

Network: 4 convolutions 3 x 3 (6 -> 32 -> 32 -> 32 -> 1 channels), GELU between, BF16 weights and activations, FP32
accumulation and FP32 master weights; Adam.  Inputs are resampled to 64 x 64 and the output back to the lattice
bilinearly (corners aligned).

    python -m blml_amd.warm train --fits 50000 --out warm_net.npz     # data from the engine, then training
    python -m blml_amd.warm eval                                      # cold vs warm on 2048 held-out fits
"""
import argparse
import ctypes
import functools
import math
import os
import sys
import time

import numpy as np

from . import _lib

GRID = 64
CHANNELS = (6, 32, 32, 32, 1)
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'warm_net.npz')


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("the warm-start network needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch


# ------------------------------------------------------------------------------------------------ grids
@functools.lru_cache(maxsize=512)
def _interp_matrix(n_out, n_in):
    """Linear interpolation from n_in points to n_out points, both spanning [0, 1] (corners aligned)."""
    A = np.zeros((n_out, n_in))
    if n_in == 1:
        A[:, 0] = 1.0
        return A
    x = np.linspace(0.0, n_in - 1.0, n_out)
    i0 = np.minimum(np.floor(x).astype(int), n_in - 2)
    f = x - i0
    A[np.arange(n_out), i0] = 1.0 - f
    A[np.arange(n_out), i0 + 1] += f
    return A


def resample(a, shape):
    """Bilinear resampling of a 2-D array to shape."""
    return _interp_matrix(shape[0], a.shape[0]) @ a @ _interp_matrix(shape[1], a.shape[1]).T


def describe(modes, spec=None):
    """(is_pball, fc, fill) of the signal mode: the first of a make_modes / modes list, or spec (a parsed mode)."""
    from .modes import parse_modes
    sp = spec if spec is not None else parse_modes(modes, 2)[0][0]
    kind = str(sp['kinds'][0])
    name, _, tag = kind.partition(':')
    fill = float(tag) if tag and name in ('slits', 'dirichlet') else 0.0
    return float(name == 'pball'), float(np.mean(sp['fc'])), fill


def features(counts, modes, spec=None):
    """(6, 64, 64) float32 network input of one 2-D lattice (spec: its first mode parsed already, else modes')."""
    c = np.asarray(counts, float)
    occ = (c > 0).astype(float)
    lc = np.log1p(c)
    is_pb, fc, fill = describe(modes, spec)
    X = np.empty((6, GRID, GRID), np.float32)
    X[0] = resample(lc / max(lc.max(), 1e-12), (GRID, GRID))
    X[1] = resample(occ, (GRID, GRID))
    X[2] = is_pb
    X[3] = fc
    X[4] = fill
    X[5] = math.log10(max(c.sum(), 1.0)) / 6.0
    return X


def start_target(counts, er, s0):
    """(Y, M) on the 64 x 64 grid: the log ratio of the start solution er to sqrt(c / s0) on occupied bins, and the
    resampled occupancy as the loss weight."""
    c = np.asarray(counts, float)
    occ = c > 0
    Y = np.zeros_like(c)
    Y[occ] = np.log(np.maximum(np.abs(er[occ]), 1e-300) / np.sqrt(c[occ] / s0))
    m64 = resample(occ.astype(float), (GRID, GRID))
    y64 = resample(Y, (GRID, GRID)) / np.maximum(m64, 1e-6)
    return y64.astype(np.float32), m64.astype(np.float32)


def start_from_prediction(counts, pred64, s0):
    """The start solution on the lattice from the network's 64 x 64 output."""
    c = np.asarray(counts, float)
    r = resample(np.asarray(pred64, float), c.shape)
    return np.where(c > 0, np.sqrt(c / s0) * np.exp(r), 0.0)


# ------------------------------------------------------------------------------------------------ network
class WarmNet:
    """The network's FP32 master weights, their BF16 copies and Adam moments on the device."""

    def __init__(self, seed=0):
        torch = _torch()
        rng = np.random.default_rng(seed)
        self.W, self.b = [], []
        for l in range(4):
            ci, co = CHANNELS[l], CHANNELS[l + 1]
            std = math.sqrt(2.0 / (ci * 9)) * (0.1 if l == 3 else 1.0)
            self.W.append(torch.from_numpy(rng.standard_normal((co, ci, 3, 3)).astype(np.float32) * std).cuda())
            self.b.append(torch.zeros(co, dtype=torch.float32, device='cuda'))
        self._setup()

    def _setup(self):
        torch = _torch()
        self.Wb = [torch.empty(w.shape, dtype=torch.int16, device='cuda') for w in self.W]
        self.Wf = [torch.empty((w.shape[1], w.shape[0], 3, 3), dtype=torch.int16, device='cuda') for w in self.W]
        self.mW = [torch.zeros_like(w) for w in self.W]
        self.vW = [torch.zeros_like(w) for w in self.W]
        self.mb = [torch.zeros_like(b) for b in self.b]
        self.vb = [torch.zeros_like(b) for b in self.b]
        self.gW = [torch.zeros_like(w) for w in self.W]
        self.gb = [torch.zeros_like(b) for b in self.b]
        self.t = 0
        for w, wb in zip(self.W, self.Wb):
            self._c(_lib.load().blml_net_to_bf16(_p(w), _p(wb), w.numel(), self._st()))

    @staticmethod
    def _st():
        return ctypes.c_void_p(_torch().cuda.current_stream().cuda_stream)

    @staticmethod
    def _c(rc):
        _lib.check(rc)

    def forward(self, Xb):
        """Xb: (N, 6, 64, 64) bf16 (int16 storage).  Returns the pre-activations Z[1..4] and activations A[0..3]."""
        torch = _torch()
        lib = _lib.load()
        N = Xb.shape[0]
        A, Z = [Xb], [None]
        for l in range(4):
            co = CHANNELS[l + 1]
            z = torch.empty((N, co, GRID, GRID), dtype=torch.float32, device='cuda')
            a = torch.empty((N, co, GRID, GRID), dtype=torch.int16, device='cuda') if l < 3 else None
            self._c(lib.blml_net_conv(_p(A[l]), _p(self.Wb[l]), _p(self.b[l]), _p(z), _p(a) if a is not None else None,
                                      N, CHANNELS[l], co, GRID, GRID, int(l < 3), self._st()))
            Z.append(z)
            A.append(a)
        return Z, A

    def step(self, Xb, Y, M, msum, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        """One Adam step on a batch; returns the batch loss (masked mean squared error)."""
        torch = _torch()
        lib = _lib.load()
        N = Xb.shape[0]
        Z, A = self.forward(Xb)
        n = Z[4].numel()
        dZ = torch.empty(Z[4].shape, dtype=torch.int16, device='cuda')
        lp = torch.empty((n + 255) // 256, dtype=torch.float32, device='cuda')
        self._c(lib.blml_net_mse(_p(Z[4]), _p(Y), _p(M), _p(dZ), _p(lp), n, 1.0 / msum, self._st()))
        for l in range(3, -1, -1):
            ci, co = CHANNELS[l], CHANNELS[l + 1]
            self._c(lib.blml_net_conv_grad_w(_p(A[l]), _p(dZ), _p(self.gW[l]), _p(self.gb[l]), N, ci, co, GRID, GRID,
                                             self._st()))
            if l > 0:
                self._c(lib.blml_net_flip(_p(self.Wb[l]), _p(self.Wf[l]), ci, co, self._st()))
                dA = torch.empty((N, ci, GRID, GRID), dtype=torch.float32, device='cuda')
                self._c(lib.blml_net_conv(_p(dZ), _p(self.Wf[l]), None, _p(dA), None, N, co, ci, GRID, GRID, 0,
                                          self._st()))
                dZ = torch.empty((N, ci, GRID, GRID), dtype=torch.int16, device='cuda')
                self._c(lib.blml_net_gelu_grad(_p(dA), _p(Z[l]), _p(dZ), dA.numel(), self._st()))
        self.t += 1
        for l in range(4):
            self._c(lib.blml_net_adam(_p(self.W[l]), _p(self.gW[l]), _p(self.mW[l]), _p(self.vW[l]), _p(self.Wb[l]),
                                      self.W[l].numel(), lr, b1, b2, eps, self.t, self._st()))
            self._c(lib.blml_net_adam(_p(self.b[l]), _p(self.gb[l]), _p(self.mb[l]), _p(self.vb[l]), None,
                                      self.b[l].numel(), lr, b1, b2, eps, self.t, self._st()))
        return float(lp.sum().item())

    def predict(self, X, batch=256):
        """X: (N, 6, 64, 64) float32 host array -> (N, 64, 64) float32 predictions."""
        torch = _torch()
        out = np.empty((len(X), GRID, GRID), np.float32)
        for i in range(0, len(X), batch):
            xb = to_bf16(torch.from_numpy(np.ascontiguousarray(X[i:i + batch])).cuda())
            Z, _ = self.forward(xb)
            out[i:i + batch] = Z[4][:, 0].cpu().numpy()
        return out

    def save(self, path):
        np.savez(path, **{f'W{l}': self.W[l].cpu().numpy() for l in range(4)},
                 **{f'b{l}': self.b[l].cpu().numpy() for l in range(4)})

    @classmethod
    def load(cls, path=WEIGHTS):
        torch = _torch()
        z = np.load(path, allow_pickle=False)
        net = cls.__new__(cls)
        net.W = [torch.from_numpy(z[f'W{l}']).cuda() for l in range(4)]
        net.b = [torch.from_numpy(z[f'b{l}']).cuda() for l in range(4)]
        net._setup()
        return net


def _p(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def to_bf16(x):
    torch = _torch()
    y = torch.empty(x.shape, dtype=torch.int16, device='cuda')
    _lib.check(_lib.load().blml_net_to_bf16(_p(x.contiguous()), _p(y), x.numel(),
                                            ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
    return y


def train(X, Y, M, epochs=20, batch=256, lr=1e-3, seed=0, log=None):
    """Adam on (X, Y, M) host arrays ((N, 6, 64, 64), (N, 64, 64), (N, 64, 64)); returns (net, epoch losses)."""
    torch = _torch()
    net = WarmNet(seed)
    Xd = to_bf16(torch.from_numpy(np.ascontiguousarray(X, np.float32)).cuda())
    Yd = torch.from_numpy(np.ascontiguousarray(Y, np.float32)).cuda()
    Md = torch.from_numpy(np.ascontiguousarray(M, np.float32)).cuda()
    msum = M.reshape(len(M), -1).sum(1)
    rng = np.random.default_rng(seed + 1)
    hist = []
    for ep in range(epochs):
        perm = rng.permutation(len(X))
        tot, nb = 0.0, 0
        for i in range(0, len(perm) - batch + 1, batch):
            idx = np.sort(perm[i:i + batch])
            it = torch.from_numpy(idx).cuda()
            tot += net.step(Xd[it].contiguous(), Yd[it].contiguous(), Md[it].contiguous(), float(msum[idx].sum()), lr)
            nb += 1
        hist.append(tot / max(nb, 1))
        if log:
            log(f'epoch {ep + 1}/{epochs}: loss {hist[-1]:.5f}')
    return net, hist


# ------------------------------------------------------------------------------------------------ data
def start_solutions(jobs, dtype=np.float32, max_sweeps=3000):
    """The real rank-1 solution under mode 0 of every job (counts, modes) -- the cold start's first solve -- from
    one batched launch of the engine (CD, the reference's defaults: tol 1e-5, anderson 5)."""
    from .binned import _FitSpec, FitBatch
    specs = []
    for counts, modes in jobs:
        specs.append(_FitSpec(counts, [modes[0]], 1.0, 0.0, False, None, None, 0))
    params = dict(tol=1e-5, max_sweeps=max_sweeps, anderson=5, stall=1000, phase_noise=0.1, ll_tol=None,
                  ll_window=100, solver='cd', lbfgs_m=20)
    b = FitBatch(specs, params, dtype, hist_cap=0).build().run()
    res = b.results()
    b.close()
    return [np.asarray(r['e_hat'][0], float) for r in res], [float(fs.s[0]) for fs in specs]


def apply(net, specs, solve=False):
    """Sets the network's warm start on every 2-D fit spec (blml_amd.binned._FitSpec) that has no init of the
    caller's (a warm start of an earlier call is replaced): the predicted real rank-1 solution under mode 0 (times
    the signs, when given) in place of the start solve; the engine draws the phases and the other modes from the
    cold start's draws (blml_binned.py:429-436).  solve=True: the prediction is the initial point of the start
    solve instead, which still runs to its tolerance (the same convex optimum, from nearer).  Returns the number of
    specs set."""
    idx = [i for i, fs in enumerate(specs) if fs.d == 2 and (fs.init is None or fs.warm)]
    if not idx:
        return 0
    X = np.stack([features(specs[i].counts, None, specs[i].specs[0]) for i in idx])
    pred = net.predict(X)
    for i, p in zip(idx, pred):
        fs = specs[i]
        er = start_from_prediction(fs.counts, p, fs.s[0])
        if fs.signs is not None:
            er = er * fs.signs
        fs.set_start_solution(er, solve)
        fs.warm = True
    return len(idx)


def _run_batch(specs, params, dtype):
    """(results, seconds) of one batched launch, the plan built before the clock starts."""
    import torch
    from .binned import FitBatch
    b = FitBatch(specs, params, dtype, hist_cap=0).build()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = b.run().results()
    dt = time.perf_counter() - t0
    b.close()
    return r, dt


def evaluate(net, jobs, params, dtype=np.float32, seed=0):
    """Cold and warm fits of jobs [(counts, modes, complex_g)], each as one batched launch; the warm time includes
    the features, the network and the start draws (apply).  Returns a dict: times, iterations, stop reasons and
    the warm fits' likelihoods against the cold ones' (better / worse by more than 1e-6 relative)."""
    from .binned import _FitSpec
    mk = lambda: [_FitSpec(c, m, 1.0, 0.0, cg, None, None, seed) for c, m, cg in jobs]
    rc, t_cold = _run_batch(mk(), params, dtype)
    specs = mk()
    t0 = time.perf_counter()
    apply(net, specs)
    t_pred = time.perf_counter() - t0
    rw, t_warm = _run_batch(specs, params, dtype)
    llc = np.array([r['loglik'] for r in rc])
    llw = np.array([r['loglik'] for r in rw])
    itc = np.array([r['n_sweeps'] + r['start_sweeps'] for r in rc])
    itw = np.array([r['n_sweeps'] + r['start_sweeps'] for r in rw])
    rel = (llw - llc) / np.abs(llc)
    stops = lambda rr: {k: int(sum(r['stop_reason'] == k for r in rr)) for k in ('tol', 'stall', 'll_tol', 'max_sweeps')}
    return dict(fits=len(jobs), cold_seconds=t_cold, warm_seconds=t_warm + t_pred, predict_seconds=t_pred,
                speedup=t_cold / (t_warm + t_pred), cold_iterations=float(itc.mean()),
                warm_iterations=float(itw.mean()), warm_worse=int((rel < -1e-6).sum()),
                warm_better=int((rel > 1e-6).sum()), median_rel_ll=float(np.median(rel)),
                rel_ll_quantiles=[float(q) for q in np.quantile(rel, [0.01, 0.1, 0.5, 0.9, 0.99])],
                cold_stops=stops(rc), warm_stops=stops(rw),
                cold_device_fit_s=float(sum(r['device_seconds']['fit'] for r in rc)),
                warm_device_fit_s=float(sum(r['device_seconds']['fit'] for r in rw)))


def main(argv=None):
    ap = argparse.ArgumentParser(description='train the warm-start network on fits the engine makes')
    ap.add_argument('cmd', choices=['train', 'eval'])
    ap.add_argument('--weights', default=WEIGHTS)
    ap.add_argument('--eval-fits', type=int, default=2048)
    ap.add_argument('--fits', type=int, default=16384)
    ap.add_argument('--seed0', type=int, default=1000)
    ap.add_argument('--epochs', type=int, default=20)
    ap.add_argument('--batch', type=int, default=256)
    ap.add_argument('--out', default=WEIGHTS)
    args = ap.parse_args(argv)
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))
    import bench
    t0 = time.time()
    if args.cmd == 'eval':
        import json
        net = WarmNet.load(args.weights)
        jobs = [(c, bench.modes_of(g, __import__('blml_amd')), kw['complex_g'])
                for c, g, kw, _ in bench.uneven_fits(args.eval_fits, seed=4)]
        print(json.dumps(evaluate(net, jobs, bench.fit_params())), flush=True)
        return
    X, Y, M = [], [], []
    for c0 in range(0, args.fits, 2048):
        seeds = range(args.seed0 + c0, args.seed0 + min(args.fits, c0 + 2048))
        jobs = []
        for s in seeds:
            counts, groups, kw, _ = bench.uneven_fits(1, seed=s)[0]
            jobs.append((counts, bench.modes_of(groups, __import__('blml_amd'))))
        er, s0 = start_solutions(jobs)
        for (counts, modes), e, s in zip(jobs, er, s0):
            X.append(features(counts, modes))
            y, m = start_target(counts, e, s)
            Y.append(y)
            M.append(m)
        print(f'data: {len(X)} fits ({time.time() - t0:.0f} s)', flush=True)
    X, Y, M = np.stack(X), np.stack(Y), np.stack(M)
    net, hist = train(X, Y, M, args.epochs, args.batch, log=lambda s: print(s, flush=True))
    net.save(args.out)
    print(f'saved {args.out}; {time.time() - t0:.0f} s', flush=True)


if __name__ == '__main__':
    main()
