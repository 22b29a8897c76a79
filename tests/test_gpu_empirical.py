"""BLMLEmpirical on the GPU (blml_empirical: CG / FISTA, the hidden-phase alternation with Anderson mixing and the
normalisation in one launch) through the C-ABI, against the oracle (oracle/blml_empirical.py, pinned by dense
linear algebra and the reference's own outputs) and the reference's golden fits:
  * linear fits (ridge, l2, l1; 1-D and 2-D; window and plane normalisation; Anscombe targets): the oracle's
    iteration counts (within 2%) and the fit to 1e-8 (FP64 transforms) / 1e-4 (FP32) of the density's scale;
  * retrieve fits: the reference's fixed point (tests/golden/empirical_fits.npz) -- loglik 1e-6, density 1e-3 of its
    maximum (the alternation stops at outer_tol 1e-5 along a slow approach, and the trajectory's length depends on
    the summation order of the dots) -- with J_history_ never rising;
  * warm starts (init=, select along the lam grid), score, predict_grid over every bin against the oracle's FFT
    grid."""
import os
import sys

import numpy as np
import pytest

from oracle import blml_empirical as OE
from oracle import blml_oracle as O
from cases_empirical import empirical_cases

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), 'golden', 'empirical_fits.npz')


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _counts1(bins=64, n=5000, seed=0):
    rng = np.random.default_rng(seed)
    x = np.concatenate([rng.normal(-1.5, 0.5, int(0.6 * n)), rng.normal(1.8, 0.8, n - int(0.6 * n))])
    return np.histogram(x, bins=bins, range=(-6, 6))[0].astype(float)


def _counts2(shape=(24, 20), seed=1):
    rng = np.random.default_rng(seed)
    g = np.meshgrid(*[np.arange(m, dtype=float) for m in shape], indexing='ij')
    p = np.exp(-0.5 * sum(((x - m / 2.3) / (m / 5)) ** 2 for x, m in zip(g, shape))) + 0.05
    return rng.multinomial(6000, (p / p.sum()).ravel()).reshape(shape).astype(float)


LINEAR = [
    ('1d ridge', 1, [dict(fc=[0.2], kinds=['sinc'])], dict(reg='ridge', lam=1e-2), {}),
    ('1d l2 plane', 1, [dict(fc=[0.2], kinds=['sinc'])], dict(reg='l2', lam=1e-3, normalise='plane'), {}),
    ('1d l1', 1, [dict(fc=[0.2], kinds=['sinc'])], dict(reg='l1', lam=1e-2, max_iter=3000), {}),
    ('1d ridge anscombe h', 1, [dict(fc=[0.08], kinds=['sinc'])], dict(reg='ridge', lam=1e-2, anscombe=True),
     dict(h=0.4)),
    ('2d sinc x slits ridge', 2, [dict(fc=[0.4, 0.35], kinds=['sinc', 'slits:0.5'])], dict(reg='ridge', lam=1e-2), {}),
    ('2d pball l2', 2, [dict(fc=[0.45, 0.45], kinds=['pball', 'pball'], periods=[24.0, 20.0])],
     dict(reg='l2', lam=1e-3), {}),
]


@pytest.mark.parametrize('dtype', [np.float64, np.float32], ids=['f64', 'f32'])
@pytest.mark.parametrize('name,d,mm,kw,fk', LINEAR, ids=[c[0] for c in LINEAR])
def test_linear_fits_match_oracle(name, d, mm, kw, fk, dtype):
    import blml_amd
    counts = _counts1() if d == 1 else _counts2()
    modes = [md for k in mm for md in blml_amd.make_modes(**k)]
    omodes = [md for k in mm for md in O.make_modes(**k)]
    tol = 1e-10 if dtype == np.float64 else 1e-6
    _note(f'empirical {name} {np.dtype(dtype).name}')
    m = blml_amd.BLMLEmpirical(tol=tol, dtype=dtype, **kw).fit(counts, modes, **fk)
    ref = OE.BLMLEmpiricalOracle(tol=tol, **kw).fit(counts, omodes, **fk)
    # CG / FISTA stop at tol on the relative residual: two FP64 runs whose transforms round differently (power-of-
    # two FFTs here, 5-smooth pocketfft in the oracle) stop within cond(A) tol of each other; FP32 transforms carry
    # ~6e-8 relative error into a system of condition ~1 + max|Kf| / lam (about 100 at lam = 1e-2)
    dens_tol = 1e-8 if dtype == np.float64 else 1e-4
    assert np.max(np.abs(m.density_ - ref.density_)) < dens_tol * ref.density_.max()
    assert abs(m.loglik_ - ref.loglik_) < dens_tol * abs(ref.loglik_)
    if dtype == np.float64:
        assert abs(m.n_iter_ - ref.n_iter_) <= max(2, 0.02 * ref.n_iter_)
        assert m.window_norm_ == pytest.approx(ref.window_norm_, rel=1e-10)
        assert m.plane_norm_ == pytest.approx(ref.plane_norm_, rel=1e-10)
    assert m.c_raw_.shape == ref.c_raw_.shape and m.e_hat_.shape == ref.e_hat_.shape
    if kw['reg'] == 'l1':
        assert 0 < m.sparsity_ <= 1


@pytest.mark.parametrize('i', range(len(empirical_cases())), ids=[c[0] for c in empirical_cases()])
def test_retrieve_matches_reference_golden(i):
    import blml_amd
    z = np.load(GOLD)
    name, counts, mm, est_kw, fit_kw = empirical_cases()[i]
    modes = [md for k in mm for md in blml_amd.make_modes(**k)]
    _note(f'empirical golden {name}')
    m = blml_amd.BLMLEmpirical(**est_kw).fit(counts, modes, **fit_kw)
    p = f'c{i}_'
    ref = float(z[p + 'loglik'])
    assert abs(m.loglik_ - ref) < 1e-6 * abs(ref), (m.loglik_, ref)
    assert np.max(np.abs(m.density_ - z[p + 'density'])) < 1e-3 * z[p + 'density'].max()
    J = np.asarray(m.J_history_)
    assert len(J) == m.n_outer_ + 1 and np.all(np.diff(J) <= 1e-12 * np.abs(J[:-1]))
    assert J[-1] <= float(z[p + 'J_history'][-1]) * (1 + 1e-4)


def test_warm_start_select_and_outputs():
    """select() along a lam grid (each fit warm-started from the previous one's c_raw_) picks the oracle's lam with
    the same held-out scores; a warm refit at the same lam starts at the solution (one CG iteration); predict_grid
    sums every bin's coefficient, as the oracle's FFT grid."""
    import blml_amd
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.normal(-1.5, 0.5, 6000), rng.normal(1.8, 0.8, 4000)])
    full = np.histogram(x, bins=64, range=(-6, 6))[0]
    train = rng.binomial(full, 0.7).astype(float)
    val = (full - train).astype(float)
    lams = [1e-1, 1e-2, 1e-3]
    _note('empirical select')
    best = blml_amd.BLMLEmpirical.select(train, val, blml_amd.make_modes([0.2], ['sinc']), lams, tol=1e-10)
    rows = []
    prev = None
    for lam in sorted(lams, reverse=True):
        mo = OE.BLMLEmpiricalOracle(lam=lam, tol=1e-10).fit(train, O.make_modes([0.2], ['sinc']), init=prev)
        rows.append((lam, mo.score(val)))
        prev = mo
    want = max(rows, key=lambda r: r[1])
    assert best.lam == want[0]
    assert np.allclose([s for _, s in best.path_], [s for _, s in sorted(rows)], rtol=1e-9, atol=0)
    again = blml_amd.BLMLEmpirical(lam=best.lam, tol=1e-8).fit(train, blml_amd.make_modes([0.2], ['sinc']), init=best)
    assert again.n_iter_ <= 2
    og = OE.BLMLEmpiricalOracle(lam=best.lam, tol=1e-10).fit(train, O.make_modes([0.2], ['sinc']))
    g_hip, g_ref = best.predict_grid(3), og.predict_grid(3)
    assert np.max(np.abs(g_hip - g_ref)) < 1e-8 * g_ref.max()


def test_many_fits_in_one_launch_equal_single_fits():
    """empirical_many: linear and retrieve fits of uneven lattices in one launch each equal the same fits made
    alone (the same device code per fit; FP64 sums in a fixed order)."""
    import blml_amd
    jobs = []
    for i, (bins, cplx, kind) in enumerate([(48, False, 'sinc'), (64, True, 'sinc'), (40, False, 'slits:0.5'),
                                            (56, True, 'slits:0.4')]):
        jobs.append((_counts1(bins=bins, seed=10 + i), blml_amd.make_modes([0.25], [kind]), dict(complex_g=cplx)))
    jobs.append((_counts2((18, 16), seed=4), blml_amd.make_modes([0.4, 0.35], ['sinc', 'sinc']), {}))
    kw = dict(tol=1e-9, outer_tol=1e-5, max_outer=200)
    _note('empirical_many')
    many = blml_amd.empirical_many(jobs, **kw)
    for (c, modes, fk), m in zip(jobs, many):
        one = blml_amd.BLMLEmpirical(**kw).fit(c, modes, **fk)
        assert m.retrieve_ == one.retrieve_
        assert m.loglik_ == one.loglik_ and np.array_equal(m.density_, one.density_)
        assert m.n_iter_ == one.n_iter_ and m.n_outer_ == one.n_outer_
