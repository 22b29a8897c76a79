"""Pins the BLMLEmpirical oracle (oracle/blml_empirical.py): the linear fits against dense linear algebra (ridge /
l2 normal equations solved by numpy, l1 by its optimality conditions), the FISTA bound against dense eigvalsh, the
host's max|Kf| (blml_amd.units.mode_kf_max, the engine's FISTA bound) against the oracle's spectrum, the retrieve
alternation against the reference's own outputs (tests/golden/make_golden_empirical.py) and its invariants (J never
rises; the stop is a stationary point), and the output normalisations."""
import json
import os

import numpy as np
import pytest

from oracle import blml_empirical as OE
from oracle import blml_oracle as O
from cases_empirical import empirical_cases

GOLD = os.path.join(os.path.dirname(__file__), 'golden', 'empirical_fits.npz')


def _counts(n=4000, bins=48, seed=0):
    rng = np.random.default_rng(seed)
    x = np.concatenate([rng.normal(-1.5, 0.5, int(0.6 * n)), rng.normal(1.8, 0.8, n - int(0.6 * n))])
    return np.histogram(x, bins=bins, range=(-6, 6))[0].astype(float)


def _target(m, counts, anscombe=False):
    w = np.asarray(counts, float)
    n = w.sum()
    r = np.sqrt(w + 0.375) if anscombe else np.sqrt(w)
    r = r * np.sqrt(n / (r ** 2).sum())
    return r / np.sqrt(n * np.prod(m.h_)), n


def _gram(m):
    return O.dense_gram(m.specs_[0], m.shape_, m.h_)


@pytest.mark.parametrize('shape,fc,kinds', [((48,), [0.3], ['sinc']), ((14, 12), [0.4, 0.35], ['sinc', 'slits:0.5'])])
@pytest.mark.parametrize('reg', ['ridge', 'l2'])
def test_linear_fit_solves_the_normal_equations(shape, fc, kinds, reg):
    """_lin's CG solution of (S + lam I) c = n t or (S^2 + lam I) c = n S t equals numpy's dense solve."""
    rng = np.random.default_rng(1)
    counts = rng.poisson(30.0, shape).astype(float)
    lam = 1e-2 if reg == 'ridge' else 1e-3
    m = OE.BLMLEmpiricalOracle(reg=reg, lam=lam, tol=1e-13, max_iter=20000).fit(counts, O.make_modes(fc, kinds))
    t, n = _target(m, counts)
    G = _gram(m)
    A = G + lam * np.eye(len(G)) if reg == 'ridge' else G @ G + lam * np.eye(len(G))
    b = n * t.ravel() if reg == 'ridge' else n * (G @ t.ravel())
    c = np.linalg.solve(A, b)
    assert np.max(np.abs(m.c_raw_.ravel() - c)) < 1e-7 * np.max(np.abs(c))
    # a planted mistake (the regulariser dropped) fails the same check
    c_bad = np.linalg.lstsq(A - lam * np.eye(len(G)), b, rcond=None)[0]
    assert np.max(np.abs(c_bad - c)) > 1e-4 * np.max(np.abs(c))


def test_l1_fit_meets_its_optimality_conditions():
    """FISTA's a = c / n minimises ||S a - t||^2 + lam ||a||_1: the gradient 2 S (S a - t) equals -lam sign(a_i)
    where a_i != 0 and lies in [-lam, lam] where a_i = 0."""
    counts = _counts(bins=40)
    lam = 2e-2
    m = OE.BLMLEmpiricalOracle(reg='l1', lam=lam, tol=1e-13, max_iter=200000).fit(counts, O.make_modes([0.3], ['sinc']))
    t, n = _target(m, counts)
    G = _gram(m)
    a = m.c_raw_.ravel() / n
    g = 2 * G @ (G @ a - t.ravel())
    nz = np.abs(a) > 1e-9 * np.abs(a).max()
    assert nz.sum() > 0 and (~nz).sum() > 0
    assert np.max(np.abs(g[nz] + lam * np.sign(a[nz]))) < 1e-5 * lam
    assert np.max(np.abs(g[~nz])) <= lam * (1 + 1e-5)
    assert m.sparsity_ == pytest.approx(np.mean(np.abs(m.c_raw_) > 0))


@pytest.mark.parametrize('shape,fc,kinds,periods', [
    ((48,), [0.3], ['sinc'], None), ((20, 18), [0.4, 0.35], ['sinc', 'slits:0.5'], None),
    ((16, 16), [0.4, 0.4], ['ball', 'ball'], None), ((20, 18), [0.45, 0.45], ['pball', 'pball'], [20.0, 18.0]),
    ((24, 30), [0.3, 0.2], ['dirichlet', 'sinc'], [24.0, None]), ((31,), [0.3], ['dirichlet:0.5'], [31.0])])
def test_kf_max_bounds_the_gram_and_matches_the_host(shape, fc, kinds, periods):
    """max|Kf| >= lambda_max(S) (S is a section of the circulant of Kf), so FISTA's step 1 / (2 max|Kf|^2) is below
    the inverse Lipschitz constant of its gradient; the host's product-of-units maximum (mode_kf_max, which the
    engine receives) equals the oracle's max over the whole padded spectrum."""
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(__file__)), 'band-limited-ml_amd'))
    from blml_amd import modes as M, units
    m = OE.BLMLEmpiricalOracle()
    md = O.make_modes(fc, kinds, periods=periods)
    m._setup(np.ones(shape), md, 1.0, 0.0)
    kf = OE.kf_max(m.kern_[0])
    lam = np.linalg.eigvalsh(_gram(m))
    assert kf >= lam[-1] * (1 - 1e-12) and kf < 1.5 * lam[-1]
    sp = M.parse_modes(md, len(shape))[0][0]
    assert units.mode_kf_max(sp, shape, m.h_, m.pad_) == pytest.approx(kf, rel=1e-12)


def _golden():
    z = np.load(GOLD, allow_pickle=False)
    return z, json.loads(str(z['meta']))


FAST = [i for i, c in enumerate(empirical_cases()) if c[1].ndim == 1]


@pytest.mark.parametrize('i', FAST, ids=[empirical_cases()[i][0] for i in FAST])
def test_oracle_matches_reference_golden(i):
    z, meta = _golden()
    name, counts, mm, est_kw, fit_kw = empirical_cases()[i]
    assert meta[i]['name'] == name and np.array_equal(z[f'c{i}_counts'], counts)
    modes = [md for kw in mm for md in O.make_modes(**kw)]
    m = OE.BLMLEmpiricalOracle(**est_kw).fit(counts, modes, **fit_kw)
    p = f'c{i}_'
    assert abs(m.loglik_ - float(z[p + 'loglik'])) <= 1e-12 * abs(float(z[p + 'loglik']))
    assert np.max(np.abs(m.density_ - z[p + 'density'])) <= 1e-10 * np.max(z[p + 'density'])
    assert m.n_outer_ == int(z[p + 'n_outer']) and m.n_iter_ == int(z[p + 'n_iter'])
    assert m.n_anderson_ == int(z[p + 'n_anderson'])
    assert np.allclose(m.J_history_, z[p + 'J_history'], rtol=1e-10, atol=0)


def test_retrieve_never_raises_J_and_stops_stationary():
    """Each half-step of the alternation minimises J over its own variables and an Anderson step is kept only when
    it lowers J further, so J_history_ never rises; the stop rule is |grad J_red| / its first value < outer_tol."""
    counts = _counts(bins=48, seed=3)
    m = OE.BLMLEmpiricalOracle(tol=1e-12, outer_tol=1e-5, max_outer=2000).fit(
        counts, O.make_modes([0.3], ['sinc']), complex_g=True)
    J = np.asarray(m.J_history_)
    assert m.n_outer_ < 2000 and m.outer_crit_ < 1e-5 and m.n_anderson_ > 0
    assert np.all(np.diff(J) <= 1e-12 * np.abs(J[:-1]))


@pytest.mark.parametrize('normalise', ['window', 'plane'])
def test_outputs_are_normalised(normalise):
    """window: sum_b g_b^2 Delta = sum_b density_b Delta = 1 (f a density on the box); plane: c'Sc / n^2 = 1."""
    counts = _counts(bins=40)
    h = 0.3
    m = OE.BLMLEmpiricalOracle(normalise=normalise, tol=1e-12).fit(counts, O.make_modes([0.1], ['sinc']), h=h)
    if normalise == 'window':
        assert np.sum(m.density_) * h == pytest.approx(1.0, rel=1e-12)
    else:
        assert O._rdot(m.e_hat_, m.Se_) / m.n_ ** 2 == pytest.approx(1.0, rel=1e-12)


def test_estimator_errors():
    with pytest.raises(ValueError, match='reg'):
        OE.BLMLEmpiricalOracle(reg='lasso')
    with pytest.raises(ValueError, match='retrieve=True'):
        OE.BLMLEmpiricalOracle(retrieve=False).fit(_counts(), O.make_modes([0.3], ['sinc']), complex_g=True)
    with pytest.raises(ValueError, match='no observations'):
        OE.BLMLEmpiricalOracle().fit(np.zeros(16), O.make_modes([0.3], ['sinc']))
