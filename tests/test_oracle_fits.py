"""Pins the oracle's solvers: golden vectors of the reference's own fits (complex / rank K, written by
tests/golden/make_golden.py), a dense sequential coordinate-descent solve for real rank 1
(results/binned_parity.py:78-86), the certificate against dense eigvalsh (results/binned_checks.py:42-47), and the
secular root against its defining equation (blml_binned.py:1556-1573)."""
import json
import os

import numpy as np
import pytest

from oracle import blml_oracle as O
from cases import golden_cases

GOLD = os.path.join(os.path.dirname(__file__), 'golden', 'binned_fits.npz')


def _golden():
    z = np.load(GOLD, allow_pickle=False)
    meta = json.loads(str(z['meta']))
    return z, meta


@pytest.mark.parametrize('i', range(len(golden_cases())), ids=[c[0] for c in golden_cases()])
def test_oracle_matches_reference_golden(i):
    z, meta = _golden()
    name, counts, mm, est_kw, fit_kw = golden_cases()[i]
    assert meta[i]['name'] == name
    assert np.array_equal(z[f'c{i}_counts'], counts)
    modes = [md for kw in mm for md in O.make_modes(**kw)]
    m = O.BLMLBinnedOracle(**est_kw).fit(counts, modes, **fit_kw)
    assert list(m.pad_) == meta[i]['pad'] and m.shared_ == meta[i]['shared']
    ref_ll = float(z[f'c{i}_loglik'])
    assert abs(m.loglik_ - ref_ll) <= 1e-9 * abs(ref_ll)
    d_ref = z[f'c{i}_density']; occ = counts > 0
    assert np.max(np.abs(m.density_ - d_ref)[occ] / d_ref[occ]) < 1e-6
    assert m.stop_reason_ == meta[i]['stop_reason']
    assert m.n_sweeps_ == int(z[f'c{i}_n_sweeps'])


def _hist(rng, shape, n, f):
    g = np.meshgrid(*[np.arange(mm, dtype=float) for mm in shape], indexing='ij')
    p = f(*g); p = p / p.sum()
    return rng.multinomial(n, p.ravel()).reshape(shape).astype(float)


REAL_CASES = [
    ('1d sinc', (40,), dict(fc=[0.35], kinds=['sinc']), 1.0),
    ('1d dirichlet', (32,), dict(fc=[0.3], kinds=['dirichlet'], periods=[32.0]), 1.0),
    ('2d sinc x sinc', (16, 18), dict(fc=[0.4, 0.3], kinds=['sinc', 'sinc']), 1.0),
    ('2d ball h=0.5', (14, 14), dict(fc=[0.8, 0.8], kinds=['ball', 'ball']), 0.5),
    ('2d pball', (15, 17), dict(fc=[0.45, 0.45], kinds=['pball', 'pball'], periods=[15.0, 17.0]), 1.0),
]


@pytest.mark.parametrize('name,shape,mm,h', REAL_CASES, ids=[c[0] for c in REAL_CASES])
def test_real_rank1_matches_dense_sequential_cd(name, shape, mm, h):
    rng = np.random.default_rng(5)
    f = lambda *x: np.exp(-0.5 * sum(((xi - s / 2.3) / (s / 5)) ** 2 for xi, s in zip(x, shape))) + 0.05
    N = _hist(rng, shape, 3000, f)
    modes = O.make_modes(**mm)
    m = O.BLMLBinnedOracle(tol=1e-9).fit(N, modes, h=h)
    occ = N.ravel() > 0; w = N.ravel()[occ]; n = w.sum()
    S = O.dense_gram(m.specs_[0], shape, m.h_)[np.ix_(occ, occ)]
    s0 = S[0, 0]; e = np.sqrt(n * w / s0); u = S @ e
    for _ in range(20000):
        for b in range(len(e)):
            t = O.cd_root(u[b] - s0 * e[b], n * w[b], s0); u += S[:, b] * (t - e[b]); e[b] = t
        if np.max(np.abs(e * u / (n * w) - 1)) < 1e-11:
            break
    ll_ref = float((w * np.log(u ** 2 / n ** 2)).sum() / n)
    assert abs(m.loglik_ - ll_ref) < 1e-9
    fr = u ** 2 / n ** 2
    assert np.max(np.abs(m.density_.ravel()[occ] - fr) / fr) < 1e-6


@pytest.mark.parametrize('cplx', [False, True])
def test_certificate_equals_dense_eigvalsh(cplx):
    rng = np.random.default_rng(2)
    c2 = lambda x, y: np.exp(-0.5 * (((x - 11) / 2.5) ** 2 + ((y - 13) / 3) ** 2)) + 0.01
    N = _hist(rng, (16, 16), 4000, c2)
    m = O.BLMLBinnedOracle(tol=1e-8).fit(N, O.make_modes([0.4, 0.4], ['ball', 'ball']), complex_g=cplx)
    occ = N.ravel() > 0
    S = O.dense_gram(m.specs_[0], N.shape, m.h_)[np.ix_(occ, occ)]
    E = m.e_hat_.reshape(m.K_, -1)[:, occ]; D = (np.abs(E) ** 2).sum(0) / N.ravel()[occ]
    lam = np.linalg.eigvalsh(np.sqrt(D)[:, None] * S * np.sqrt(D)[None])[-1] / m.n_
    assert abs(m.certificate() - lam) < 1e-6 * lam


def test_certificate_rejects_per_mode_kernels():
    N = np.ones((31, 31)) * 3
    m = O.BLMLBinnedOracle(solver='lbfgs', max_sweeps=5).fit(
        N, O.make_modes([[0.33, 0.33], [3 / 31, 3 / 31]], ['pball', 'pball'], periods=[31.0, 31.0]), complex_g=True)
    with pytest.raises(ValueError):
        m.certificate()


def test_secular_root_solves_equation():
    rng = np.random.default_rng(0)
    K, B = 3, 500
    s = np.array([0.3, 0.12, 0.05]).reshape(K, 1)
    R2 = rng.exponential(1.0, (K, B)) ** 2
    c = rng.exponential(5.0, B)
    lam = O.secular(R2, s, c)
    assert np.all(lam > s.max())
    phi = lam * (R2 / (lam[None] - s) ** 2).sum(0)
    assert np.max(np.abs(phi - c) / c) < 1e-10
    # equal s: the closed form t = 2c / (a + sqrt(a^2 + 4 s c)), lam = s + a / t
    se = np.full((K, 1), 0.2); lam_e = O.secular(R2, se, c)
    a = np.sqrt(R2.sum(0)); t = 2 * c / (a + np.sqrt(a * a + 4 * 0.2 * c))
    assert np.max(np.abs(lam_e - (0.2 + a / t)) / lam_e) < 1e-10


def test_validation_errors():
    m = O.BLMLBinnedOracle()
    with pytest.raises(ValueError):
        m.fit(np.array([1.0, -1.0, 2.0]), O.make_modes([0.3], ['sinc']))
    with pytest.raises(ValueError):
        m.fit(np.zeros(8), O.make_modes([0.3], ['sinc']))
    with pytest.raises(ValueError):
        O.BLMLBinnedOracle(solver='newton')


@pytest.mark.parametrize('kinds,fc,periods,shape,cplx', [
    (['ball', 'ball'], [0.4, 0.4], None, (14, 12), True),
    (['pball', 'pball'], [0.45, 0.45], [14.0, 12.0], (14, 12), False),
    (['sinc', 'dirichlet'], [0.4, 0.3], [None, 12.0], (14, 12), True),
    (['slits:0.5'], [0.35], None, (30,), False),
], ids=['ball complex', 'pball', 'sinc x dirichlet complex', 'slits 1d'])
def test_predict_grid_is_the_direct_sum_at_the_fine_points(kinds, fc, periods, shape, cplx):
    """The oracle's predict_grid (one FFT pair per sub-lattice shift, blml_binned.py:468-499) equals its predict
    (direct sums of the closed forms, :452-470) at the fine lattice's points -- the reference's own check
    (binned_gpu_job.py:79-82) -- so each pins the other."""
    rng = np.random.default_rng(4)
    N = rng.poisson(6.0, shape).astype(float)
    m = O.BLMLBinnedOracle(tol=1e-6, max_sweeps=300).fit(N, O.make_modes(fc, kinds, periods), complex_g=cplx,
                                                         h=0.7, origin=-1.0)
    r = 3
    G = m.predict_grid(r)
    fine = np.meshgrid(*[m.origin_[j] + m.h_[j] * np.arange(s) / r for j, s in enumerate(G.shape)], indexing='ij')
    X = np.stack([x.ravel() for x in fine], 1)
    P = m.predict(X)
    assert np.max(np.abs(G.ravel() - P)) < 1e-10 * np.max(P)
    # the bins themselves: the fit's density_
    assert np.allclose(G[tuple(slice(0, None, r) for _ in shape)], m.density_, rtol=1e-9, atol=1e-12 * m.density_.max())


@pytest.mark.parametrize('kinds,fc,cplx', [(['sinc'], [0.3], False), (['ball', 'ball'], [0.4, 0.4], False),
                                           (['slits:0.5', 'sinc'], [0.45, 0.35], True)],
                         ids=['sinc', 'ball', 'slits x sinc complex'])
def test_param_grad_is_the_refit_derivative(kinds, fc, cplx):
    """fc_grad (param_grad of the training counts at the fitted coefficients, blml_binned.py:520-563) is the
    derivative of the refitted loglik (the envelope theorem: the coefficients are stationary), the reference's own
    check (binned_gpu_job.py:84-86): the sum over units against a central difference of refits at fc e^{+-d}."""
    rng = np.random.default_rng(8)
    shape = (40,) if len(kinds) == 1 else (18, 16)
    g = np.meshgrid(*[np.arange(s_, dtype=float) for s_ in shape], indexing='ij')
    p = np.exp(-0.5 * sum(((x - s_ / 2.3) / (s_ / 6)) ** 2 for x, s_ in zip(g, shape))) + 0.05
    N = rng.multinomial(6000, (p / p.sum()).ravel()).reshape(shape).astype(float)
    mk = lambda f: O.make_modes(list(f), kinds)
    kw = dict(tol=1e-11, max_sweeps=20000)
    m = O.BLMLBinnedOracle(**kw).fit(N, mk(fc), complex_g=cplx)
    G = m.param_grad(N)[0]
    d = 1e-4
    ll = [O.BLMLBinnedOracle(**kw).fit(N, mk(np.asarray(fc) * np.exp(s * d)), complex_g=cplx, init=m).loglik_
          for s in (1, -1)]
    fd = (ll[0] - ll[1]) / (2 * d)
    assert abs(np.nansum(G) - fd) < 1e-5 * max(1.0, abs(fd)), (G, fd)
    assert abs(m.log_density(N) - m.loglik_) < 1e-12 * abs(m.loglik_)


IGRAD_CASES = [('1d sinc real', [0.3], 'sinc', 1, False), ('1d sinc complex', [0.3], 'sinc', 1, True),
               ('1d rank2 per-mode real', [[0.3], [0.15]], 'sinc', 1, False)]


@pytest.mark.parametrize('name,fc,kind,rank,cplx', IGRAD_CASES, ids=[c[0] for c in IGRAD_CASES])
def test_implicit_grad_is_the_refit_derivative(name, fc, kind, rank, cplx):
    """implicit_grad (the held-out score's derivative with the coefficients' refit response) against the central
    difference of the held-out score of warm-started refits at fc exp(+-delta) (the reference's own check,
    results/binned_igrad_checks.py:20-35), and away from the fixed-coefficient param_grad."""
    rng = np.random.default_rng(1)
    x = np.arange(64.0)
    p = np.exp(-0.5 * ((x - 20) / 4) ** 2) + 0.6 * np.exp(-0.5 * ((x - 42) / 3) ** 2) + 0.02
    p /= p.sum()
    tr, te = rng.multinomial(20000, p).astype(float), rng.multinomial(5000, p).astype(float)
    m = O.BLMLBinnedOracle(tol=1e-11).fit(tr, O.make_modes(fc, [kind]), complex_g=cplx)
    eps = 1e-3 * float(np.median(np.asarray(m.density_)[te > 0]))
    Gi, Gf = m.implicit_grad(te, eps, tol=1e-10, maxit=2000), m.param_grad(te, eps)
    F0 = np.array([sp['fc'] for sp in m.specs_], float)
    d = 2e-3
    for k in range(len(F0)):
        ll = []
        for sg in (1, -1):
            F = F0.copy()
            F[k] *= np.exp(sg * d)
            ll.append(O.BLMLBinnedOracle(tol=1e-11).fit(tr, O.make_modes([list(f) for f in F] if len(F0) > 1 else
                                                                         list(F[0]), [kind]),
                                                        complex_g=cplx, init=m).log_density(te, eps))
        fd = (ll[0] - ll[1]) / (2 * d)
        assert abs(fd - Gi[k][0]) < 5e-3 * max(abs(fd), 1e-3), (name, k, fd, Gi[k][0])
        assert abs(Gf[k][0] - Gi[k][0]) > 10 * abs(fd - Gi[k][0]), (name, k, Gf[k][0], Gi[k][0], fd)
