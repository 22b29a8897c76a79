"""fc gradients on the GPU (BLMLBinned.param_grad / fc_grad / log_density: derivative mat-vecs through blml_matvec over
a plan of the derivative kernels, reductions through blml_grad_dots) against the oracle on the same coefficients
(its derivative tables by central differences in log fc; 1e-7 relative, FP64 plans) and against the derivative of
the refitted loglik (the envelope theorem, the reference's own check binned_gpu_job.py:84-86; 1e-4), for sinc,
slits, ball, dirichlet (nan: whole harmonics) and a spread 3-D fit."""
import sys

import numpy as np
import pytest

from oracle import blml_oracle as O

pytestmark = pytest.mark.gpu


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


CASES = [
    ('1d sinc', (40,), [dict(fc=[0.3], kinds=['sinc'])], False, -1),
    ('2d ball complex', (18, 16), [dict(fc=[0.4, 0.4], kinds=['ball', 'ball'])], True, -1),
    ('2d slits x dirichlet rank 2', (20, 16), [dict(fc=[0.45, 0.3], kinds=['slits:0.5', 'dirichlet'],
                                                    periods=[None, 16.0], rank=2)], False, -1),
    ('2d sinc + ball bg complex', (16, 18), [dict(fc=[0.4, 0.35], kinds=['sinc', 'sinc']),
                                             dict(fc=[0.15, 0.15], kinds=['ball', 'ball'])], True, -1),
    ('3d sinc spread', (12, 10, 9), [dict(fc=[0.35] * 3, kinds=['sinc'] * 3)], False, 1),
]


@pytest.mark.parametrize('name,shape,mm,cplx,spread', CASES, ids=[c[0] for c in CASES])
def test_param_grad_matches_oracle(name, shape, mm, cplx, spread):
    import blml_amd
    from blml_amd.binned import BLMLBinned, FitBatch, _FitSpec
    rng = np.random.default_rng(6)
    g = np.meshgrid(*[np.arange(s, dtype=float) for s in shape], indexing='ij')
    p = np.exp(-0.5 * sum(((x - s / 2.3) / (s / 6)) ** 2 for x, s in zip(g, shape))) + 0.05
    N = rng.multinomial(8000, (p / p.sum()).ravel()).reshape(shape).astype(float)
    modes = [md for kw in mm for md in blml_amd.make_modes(**kw)]
    kw = dict(tol=1e-9, max_sweeps=5000, solver='lbfgs' if (cplx or len(modes) > 1) else 'cd')
    m = BLMLBinned(**kw)
    fs = _FitSpec(N, modes, 1.0, 0.0, cplx, None, None, 0)
    b = FitBatch([fs], m._params(), np.float64, spread=spread).build().run()
    m._adopt(fs, b, b.results()[0], 0)
    _note(f'param_grad {name} ({m.stop_reason_}, {m.n_sweeps_} sweeps)')
    o = O.BLMLBinnedOracle()
    o._setup(N, [md for kw_ in mm for md in O.make_modes(**kw_)], 1.0, 0.0)
    o.e_hat_, o.Se_ = m.e_hat_, m.Se_
    Nb = rng.poisson(N * 0.3 + 0.2)                      # another batch's counts: the prequential gradient
    for counts, eps in ((N, 0.0), (Nb, 1e-3)):
        G, Go = m.param_grad(counts, eps), o.param_grad(counts, eps)
        for k, (a, c) in enumerate(zip(G, Go)):
            _note(f'   mode {k} eps {eps}: {np.round(a, 8)} oracle {np.round(c, 8)}')
            assert np.array_equal(np.isnan(a), np.isnan(c))
            ok = ~np.isnan(c)
            assert np.all(np.abs(a[ok] - c[ok]) <= 1e-7 * np.maximum(1.0, np.abs(c[ok])))
        assert abs(m.log_density(counts, eps) - o.log_density(counts, eps)) < 1e-12 * abs(o.log_density(counts, eps))
    if not cplx and len(modes) == 1 and spread < 0:
        # the envelope theorem: the refitted loglik's derivative in log fc of every column at once
        d = 1e-4
        scale = lambda s: [md for kw_ in mm for md in blml_amd.make_modes(**dict(kw_, fc=list(np.asarray(kw_['fc'])
                                                                                           * np.exp(s * d))))]
        ll = [BLMLBinned(**dict(kw, tol=1e-11)).fit(N, scale(s), init=m).loglik_ for s in (1, -1)]
        fd = (ll[0] - ll[1]) / (2 * d)
        tot = float(np.nansum(m.fc_grad()[0]))
        _note(f'   fc_grad sum {tot:.8f} refit difference {fd:.8f}')
        assert abs(tot - fd) < 1e-4 * max(1.0, abs(fd))


def _truth1(x):
    return np.exp(-0.5 * ((x - 20) / 4) ** 2) + 0.6 * np.exp(-0.5 * ((x - 42) / 3) ** 2) + 0.02


def _truth2(x, y):
    return np.exp(-0.5 * (((x - 12) / 3) ** 2 + ((y - 14) / 4) ** 2)) + 0.5 * np.exp(-0.5 * (((x - 21) / 2.5) ** 2 +
                                                                                           ((y - 9) / 2) ** 2)) + 0.02


# the reference's implicit-gradient cells (results/binned_igrad_checks.py:46-51): fits converged to tol 1e-11
IGRAD = [('1d sinc real', (64,), _truth1, [dict(fc=[0.3], kinds=['sinc'])], False),
         ('1d sinc complex', (64,), _truth1, [dict(fc=[0.3], kinds=['sinc'])], True),
         ('1d rank2 per-mode real', (64,), _truth1, [dict(fc=[[0.3], [0.15]], kinds=['sinc'])], False),
         ('2d ball complex', (32, 32), _truth2, [dict(fc=[0.35, 0.35], kinds=['ball', 'ball'])], True)]


@pytest.mark.parametrize('name,shape,f,mm,cplx', IGRAD, ids=[c[0] for c in IGRAD])
def test_implicit_grad_matches_oracle(name, shape, f, mm, cplx):
    """implicit_grad (CG on the device's Hessian-vector products) against the oracle's on the same coefficients
    (both CGs to 1e-10: 1e-5 relative, or 1e-8 absolute) and against the central difference of the held-out score
    of warm-started engine refits at fc exp(+-2e-3) (the reference's own check, results/binned_igrad_checks.py:20-35:
    5e-3 relative where the CG converged, 5e-2 where it stopped at maxit), on the reference's cells, each fit
    converged to tol 1e-11 (the Hessian's gauge directions are null only at a stationary point)."""
    import blml_amd
    from blml_amd.binned import BLMLBinned
    rng = np.random.default_rng(1)
    g = np.meshgrid(*[np.arange(s, dtype=float) for s in shape], indexing='ij')
    p = f(*g)
    p = (p / p.sum()).ravel()
    N = rng.multinomial(20000, p).reshape(shape).astype(float)
    Nb = rng.multinomial(5000, p).reshape(shape).astype(float)
    modes = [md for kw in mm for md in blml_amd.make_modes(**kw)]
    kw = dict(tol=1e-11, max_sweeps=200000)
    m = BLMLBinned(**kw).fit(N, modes, complex_g=cplx)
    assert m.converged_, (name, m.stop_reason_, m.residual_)
    eps = 1e-3 * float(np.median(np.asarray(m.density_)[Nb > 0]))
    _note(f'implicit_grad {name} ({m.stop_reason_}, {m.n_sweeps_} sweeps)')
    o = O.BLMLBinnedOracle()
    omodes = [md for kw_ in mm for md in O.make_modes(**kw_)]
    o._setup(N, omodes, 1.0, 0.0)
    o.e_hat_, o.Se_ = m.e_hat_, m.Se_
    o.w_, o.n_, o.occ_ = N, float(N.sum()), N > 0
    o.complex_g_, o.psd_ = cplx, cplx or len(omodes) > 1
    G = m.implicit_grad(Nb, eps, tol=1e-10, maxit=3000)
    Go = o.implicit_grad(Nb, eps, tol=1e-10, maxit=3000)
    Gf = m.param_grad(Nb, eps)
    F0 = np.array([sp['fc'] for sp in m.specs_], float)
    for k, (a, c, fx) in enumerate(zip(G, Go, Gf)):
        _note(f'   mode {k}: implicit {np.round(a, 9)} oracle {np.round(c, 9)} fixed {np.round(fx, 9)} '
              f'(CG {m.ift_iters_} / {o.ift_iters_} iterations)')
        assert np.array_equal(np.isnan(a), np.isnan(c))
        ok = ~np.isnan(c)
        assert np.all(np.abs(a[ok] - c[ok]) <= np.maximum(1e-5 * np.abs(c[ok]), 1e-8))
        d = 2e-3
        ll = []
        for sg in (1, -1):
            F = F0.copy()
            F[k] *= np.exp(sg * d)
            fk = [list(r) for r in F] if len(F0) > 1 else list(F[0])
            mk = blml_amd.make_modes(fk, mm[0]['kinds'])
            ll.append(BLMLBinned(**kw).fit(N, mk, complex_g=cplx, init=m).log_density(Nb, eps))
        fd = (ll[0] - ll[1]) / (2 * d)
        _note(f'   refit central difference {fd:.9f} (CG relative residual {m.ift_res_:.1e})')
        # a CG stopped by maxit short of its tolerance (a 2-D complex fit's Hessian is nearly singular along its
        # phases) gives the derivative to a few percent
        rel = 5e-3 if m.ift_res_ < 1e-8 else 5e-2
        assert abs(fd - a[0]) < rel * max(abs(fd), 1e-3), (fd, a[0])
