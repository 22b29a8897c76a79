"""BLMLQuick's complex / rank-K counts fits on the GPU (blml_amd/quick.py, libblml_hip.so blml_quick_psd_* and
blml_vec_* / blml_lbfgs_twoloop) against the oracle's restatement of _psd_solve / lbfgs_psd (oracle/blml_quick.py)
with dense Gram matrices of the same kernels (oracle/blml_pairs.py):
- FP64 tables, one start (n_starts = (0, 0), as the reference's own parity check, results/binned_parity.py:52),
  for a shared and a per-mode kernel, real rank 2 and complex, 1-D sinc, 2-D sinc x dirichlet and a 2-D periodic
  ball, CD and L-BFGS: the same draws from default_rng(seed), so the same path -- the residual histories agree to
  1e-6 (or 1e-11 absolute) entry by entry; where they agree to the end, the same count and stop, the likelihood to 1e-10 and E to 1e-7
  of its scale.  Where rounding moves the stop (a secular root's bisection or an L-BFGS direction turning on the
  last bits, then the stall rule or the 1e-11 likelihood stop firing a few iterations apart), the paths agree for
  at least the first quarter (up to 100 entries), the stop kind is the same, the likelihood agrees to 1e-10
  (converged CD) or 1e-8, and the density at the bins to 1e-3 (E's components in the near-null space of S stay
  where each path left them, as lbfgs_psd's docstring says);
- FP32 tables: the likelihood within 1e-6 of the FP64 oracle's;
- sqrt_density at the bins is Se / n (the exact kernel at lattice points);
- the start search (blml_complex_starts) step for step against the oracle on short runs, and a default complex fit
  with its five starts against the oracle's fixed point;
- a real fit warm-started from a fit on another lattice (re-binned through the exact kernel) reaches the cold
  fit's fixed point;
- a complex conditional fit (dep column 1): the joint and the marginal (the counts summed over column 1) as two
  separate fits, each against the oracle's psd_fit on its own Gram matrix (likelihood to 1e-8)."""
import os
import sys

import numpy as np
import pytest

from oracle import blml_pairs as OP
from oracle import blml_quick as OQ

pytestmark = pytest.mark.gpu


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _hist(shape, total, f, seed=1):
    rng = np.random.default_rng(seed)
    g = np.meshgrid(*[np.arange(m, dtype=float) for m in shape], indexing='ij')
    p = f(*g)
    N = rng.multinomial(total, (p / p.sum()).ravel()).reshape(shape).astype(float)
    Z = np.stack([x.ravel() for x in np.meshgrid(*[np.arange(m) for m in shape], indexing='ij')], 1)
    return Z, N.ravel()


c1 = lambda x: np.exp(-0.5 * ((x - 14) / 3) ** 2) + 0.5 * np.exp(-0.5 * ((x - 27) / 2) ** 2) + 0.02
c2 = lambda x, y: (np.exp(-0.5 * (((x - 11) / 2.5) ** 2 + ((y - 13) / 3) ** 2))
                   + 0.4 * np.exp(-0.5 * (((x - 17) / 2) ** 2 + ((y - 6) / 2) ** 2)) + 0.01)
c3 = lambda x, y: np.exp(np.cos(2 * np.pi * x / 16) + 0.6 * np.sin(2 * np.pi * y / 16)) + 0.1

# name, shape, density, product columns per mode (without h), oracle units per mode, complex
CASES = [
    ('1d sinc complex', (40,), c1, [[dict(fc=0.35)]], [[('offset', 'sinc', [0.35], [0.0], [0])]], True),
    ('1d sinc rank2 shared complex', (40,), c1, [[dict(fc=0.35)]] * 2, [[('offset', 'sinc', [0.35], [0.0], [0])]],
     True),
    ('1d sinc rank2 shared real', (40,), c1, [[dict(fc=0.35)]] * 2, [[('offset', 'sinc', [0.35], [0.0], [0])]],
     False),
    ('1d sinc rank2 per-mode complex', (40,), c1, [[dict(fc=0.35)], [dict(fc=0.12)]],
     [[('offset', 'sinc', [0.35], [0.0], [0])], [('offset', 'sinc', [0.12], [0.0], [0])]], True),
    ('2d sinc x dirichlet complex', (24, 24), c2, [[dict(fc=0.4), dict(kind='dirichlet', fc=0.3, period=24.0)]],
     [[('offset', 'sinc', [0.4], [0.0], [0]), ('circular', 'dirichlet', [0.3], [24.0], [1])]], True),
    ('2d periodic ball rank2 per-mode complex', (16, 16), c3,
     [[{(0, 1): dict(kind='pball', fc=0.4, period=16.0)}], [{(0, 1): dict(kind='pball', fc=0.2, period=16.0)}]],
     [[('torus', 'pball', [0.4, 0.4], [16.0, 16.0], [0, 1])], [('torus', 'pball', [0.2, 0.2], [16.0, 16.0], [0, 1])]],
     True),
]


def _modes(cols_per_mode):
    out = []
    for cols in cols_per_mode:
        if len(cols) == 1 and isinstance(cols[0], dict) and any(isinstance(k, tuple) for k in cols[0]):
            out.append({key: dict(c, h=1.0) for key, c in cols[0].items()})
        else:
            out.append({j: dict(c, h=1.0) for j, c in enumerate(cols)})
    return out


def _fit_both(case, solver, tol, dtype):
    from blml_amd.quick import BLMLQuick
    name, shape, f, cols, units, cplx = case
    Z, w = _hist(shape, 3000 if len(shape) == 1 else 8000, f)
    keep = w > 0
    K = len(cols)
    _note(f'psd quick: {name} ({solver}, {np.dtype(dtype).name}), {int(keep.sum())} bins, K = {K}')
    q = BLMLQuick(tol=tol, dtype=dtype, solver=solver, seed=0)
    q.n_starts = (0, 0)
    q.fit(Z.astype(float), _modes(cols), sample_weight=w, complex_g=cplx)
    Gs = [OP.dense_pair_gram(u, np.ones(len(shape)), Z[keep], Z[keep]) for u in units]
    o = OQ.psd_fit(Gs, w[keep], tol=tol, seed=0, solver=solver, K=K, complex_g=cplx)
    return q, o, Gs, w[keep]


def _split(a, b, rtol):
    """The first index where two residual histories differ by more than rtol (their common length if none); an
    absolute 1e-11 as well: near convergence the residual is a cancellation (S e |e|^2 / (n w) - e at 1e-7 of e),
    whose relative rounding is about 1e-5."""
    m = min(len(a), len(b))
    bad = np.flatnonzero(~np.isclose(np.asarray(a[:m]), np.asarray(b[:m]), rtol=rtol, atol=1e-11))
    return int(bad[0]) if len(bad) else m


@pytest.mark.parametrize('solver', ['cd', 'lbfgs'])
@pytest.mark.parametrize('case', CASES, ids=[c[0] for c in CASES])
def test_psd_fit_matches_oracle_fp64(case, solver):
    q, o, Gs, w = _fit_both(case, solver, 1e-7, np.float64)
    n = w.sum()
    Eo, SEo, its, conv = o[0], o[1], o[2], o[3]
    E = q.e_hat_ if q.e_hat_.ndim == 2 else q.e_hat_[:, None]
    SEx = OQ.psd_apply(Gs, E)
    l_q, l_o = OQ.psd_loglik(E, SEx, w, n), OQ.psd_loglik(Eo, OQ.psd_apply(Gs, Eo), w, n)
    fq, fo = (np.abs(SEx) ** 2).sum(1), (np.abs(OQ.psd_apply(Gs, Eo)) ** 2).sum(1)
    split = _split(q.rho_history_, o[5], 1e-6)
    _note(f'   engine: {q.n_sweeps_} {"sweeps" if solver == "cd" else "iterations"}, converged {q.converged_}, '
          f'residual {q.residual_:.3e}, loglik {l_q:.12f}; oracle {its}, {conv}, {o[4]:.3e}, {l_o:.12f}; '
          f'max |dE| / max |E| {np.abs(E - Eo).max() / np.abs(Eo).max():.2e}, density max rel '
          f'{np.abs(fq / fo - 1).max():.2e}; residual histories equal to 1e-6 for the first {split} entries')
    assert q.converged_ == conv
    if q.n_sweeps_ == its:            # the same path to the end
        assert split == len(o[5]) and abs(l_q - l_o) < 1e-10 * abs(l_o)
        assert np.abs(E - Eo).max() < 1e-7 * np.abs(Eo).max()
    else:                             # rounding moved the stop: the same path for a long stretch, the same answer
        assert split >= min(100, its // 4), split
        assert abs(l_q - l_o) < (1e-10 if conv and solver == 'cd' else 1e-8) * abs(l_o)
        assert np.abs(fq / fo - 1).max() < 1e-3
    if q.Se_ is not None:
        Se = q.Se_ if q.Se_.ndim == 2 else q.Se_[:, None]
        assert np.abs(Se - SEx).max() < 1e-9 * np.abs(SEx).max()


@pytest.mark.parametrize('case', [CASES[0], CASES[3], CASES[4]], ids=[CASES[i][0] for i in (0, 3, 4)])
def test_psd_fit_fp32_tables(case):
    q, o, Gs, w = _fit_both(case, 'cd', 1e-6, np.float32)
    n = w.sum()
    E = q.e_hat_ if q.e_hat_.ndim == 2 else q.e_hat_[:, None]
    l_q = OQ.psd_loglik(E, OQ.psd_apply(Gs, E), w, n)
    l_o = OQ.psd_loglik(o[0], OQ.psd_apply(Gs, o[0]), w, n)
    _note(f'   FP32 tables: loglik {l_q:.10f} against the FP64 oracle {l_o:.10f} ({q.n_sweeps_} / {o[2]} sweeps)')
    assert q.converged_
    assert abs(l_q - l_o) < 1e-6 * abs(l_o)


def test_psd_queries_and_start_search():
    from blml_amd.quick import BLMLQuick
    case = CASES[1]
    Z, w = _hist((40,), 3000, c1)
    q = BLMLQuick(tol=1e-7, dtype=np.float64)
    q.n_starts = (0, 0)
    q.fit(Z.astype(float), _modes(case[3]), sample_weight=w, complex_g=True)
    g = q.sqrt_density(q.Z_)
    assert g.shape == (len(q.Z_), 2) and np.iscomplexobj(g)
    assert np.abs(g - q.Se_ / q.n_samples_).max() < 1e-10 * np.abs(g).max()
    f = q.predict(q.Z_ + 0.37)
    assert f.shape == (len(q.Z_),) and np.all(f >= 0)


@pytest.mark.parametrize('i', range(4), ids=['1d', '2d', '1d coarse q=5', '2d ball'])
def test_start_search_short_runs_match_oracle(i):
    """blml_complex_starts through the product's complex_starts, 4 random runs and every symmetry-breaking run of
    3 HIO + 2 ER steps (the iteration is chaotic over hundreds of steps, so two FFT implementations agree over a
    few only), against the oracle's search with the same draws: the same families in the same order (the same
    runs selected), scores to 1e-9 relative, starts e0 to 1e-7 of their scale (the coarse case through
    blml_band_eval's exact band sum)."""
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), 'golden'))
    from cases_starts import starts_cases
    from oracle import blml_starts as OS
    from blml_amd.quick import complex_starts
    name, Z, w, h, fc, kinds = starts_cases()[i]
    kw = dict(hio_pool=4, n_hio_iter=3, n_er_iter=2, seed=0)
    _note(f'start search, short runs: {name}')
    got = complex_starts(Z, w, h, fc, kinds, **kw)
    want = OS.complex_starts(Z, w, h, fc, kinds, **kw)
    assert [s[0] for s in got] == [s[0] for s in want]
    for (_, sg, eg), (_, sw, ew) in zip(got, want):
        assert sg == pytest.approx(sw, rel=1e-9)
        assert np.abs(eg - ew).max() < 1e-7 * np.abs(ew).max()


def test_complex_fit_with_the_start_search():
    """The default complex counts fit (n_starts = (3, 2): 30 random runs of 400 HIO + 100 ER steps and the
    symmetry-breaking runs, on the device): five starts, each fitted; the best start's fit equals the oracle's
    fixed point on this case (every start of the reference's own search reaches it too) to 1e-9; the search's best
    random score is within 5e-4 of the reference's (tests/golden/complex_starts.npz: the same draws, a chaotic
    iteration, so the same distribution rather than the same runs)."""
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), 'golden'))
    from cases_starts import starts_cases
    from blml_amd.quick import BLMLQuick, complex_starts
    name, Z, w, h, fc, kinds = starts_cases()[0]
    _note('complex fit with the start search: 1d sinc 40 bins')
    q = BLMLQuick(tol=1e-7, dtype=np.float64, seed=0).fit(Z.astype(float), [{0: dict(fc=0.35, h=1.0)}],
                                                           sample_weight=w, complex_g=True)
    assert len(q.starts_) == 5 and [s['family'] for s in q.starts_] == ['hio'] * 3 + ['symbreak'] * 2
    assert all(s['converged'] for s in q.starts_) and q.converged_
    G0 = [OP.dense_pair_gram([('offset', 'sinc', [0.35], [0.0], [0])], np.ones(1), Z, Z)]
    one = OQ.psd_fit(G0, w, tol=1e-7, seed=0, K=1, complex_g=True)
    n = w.sum()
    l_o = OQ.psd_loglik(one[0], OQ.psd_apply(G0, one[0]), w, n)
    E = q.e_hat_[:, None]
    l_q = OQ.psd_loglik(E, OQ.psd_apply(G0, E), w, n)
    assert abs(l_q - l_o) < 1e-9 * abs(l_o)
    assert max(s['loglik'] for s in q.starts_) == pytest.approx(l_q * n, rel=1e-9)
    z = np.load(os.path.join(os.path.dirname(__file__), 'golden', 'complex_starts.npz'))
    ref_best = float(z['c0_score'][0])
    got_best = max(s['start_score'] for s in q.starts_ if s['family'] == 'hio')
    assert abs(got_best - ref_best) < 5e-4 * abs(ref_best), (got_best, ref_best)
    s2 = complex_starts(Z, w, np.ones(1), np.array([0.35]), ['sinc'])
    assert [s[1] for s in s2] == [s['start_score'] for s in q.starts_]


def test_rebinned_warm_start_reaches_the_cold_fit():
    """A fit on a 2-unit lattice warm-starts one on the 1-unit lattice (re-binned: e = w / g_prev at the new bins);
    the real rank-1 problem is convex, so it ends at the cold fit's fixed point."""
    from blml_amd.quick import BLMLQuick
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.normal(30, 5, 4000), rng.normal(55, 3, 2000)])
    mk = lambda h: [{0: dict(fc=0.15, h=h)}]

    def binned(h):
        zi, cnt = np.unique(np.round(x / h).astype(int), return_counts=True)
        return (zi * h).astype(float)[:, None], cnt.astype(float)

    Z2, w2 = binned(2.0)
    Z1, w1 = binned(1.0)
    coarse = BLMLQuick(tol=1e-9, dtype=np.float64).fit(Z2, mk(2.0), sample_weight=w2)
    warm = BLMLQuick(tol=1e-9, dtype=np.float64).fit(Z1, mk(1.0), sample_weight=w1, init=coarse)
    cold = BLMLQuick(tol=1e-9, dtype=np.float64).fit(Z1, mk(1.0), sample_weight=w1)
    _note(f'rebinned warm start: {warm.n_sweeps_} sweeps ({warm.n_new_bins_} bins started at their 1-D root) '
          f'against {cold.n_sweeps_} cold')
    assert warm.rebinned_ and warm.converged_ and cold.converged_
    assert np.abs(warm.e_hat_ - cold.e_hat_).max() < 1e-6 * np.abs(cold.e_hat_).max()


def test_complex_conditional_fit_matches_oracle():
    from blml_amd import make_modes
    from blml_amd.quick import BLMLQuick
    Z, w = _hist((24, 20), 8000, c2)
    modes = [{j: dict(c, h=1.0) for j, c in m.items()}
             for m in make_modes([0.4, 0.3], ['sinc', 'dirichlet'], periods=[None, 20.0], dep=[1])]
    _note('psd quick: complex conditional fit, dep column 1')
    q = BLMLQuick(tol=1e-7, dtype=np.float64, seed=0)
    q.n_starts = (0, 0)
    q.fit(Z.astype(float), modes, sample_weight=w, complex_g=True)
    keep = w > 0
    Gj = [OP.dense_pair_gram([('offset', 'sinc', [0.4], [0.0], [0]), ('circular', 'dirichlet', [0.3], [20.0], [1])],
                             np.ones(2), Z[keep], Z[keep])]
    Zm, inv = np.unique(Z[:, [0]], axis=0, return_inverse=True)
    wm = np.bincount(np.ravel(inv), weights=w)
    km = wm > 0
    Gm = [OP.dense_pair_gram([('offset', 'sinc', [0.4], [0.0], [0])], np.ones(1), Zm[km], Zm[km])]
    oj = OQ.psd_fit(Gj, w[keep], tol=1e-7, seed=0, K=1, complex_g=True)
    om = OQ.psd_fit(Gm, wm[km], tol=1e-7, seed=0, K=1, complex_g=True)
    assert q.dep_ == [1] and q.converged_ and q._marg.converged_
    for est, o, Gs, ww in ((q._joint, oj, Gj, w[keep]), (q._marg, om, Gm, wm[km])):
        n = ww.sum()
        E = est.e_hat_[:, None]
        l_q = OQ.psd_loglik(E, OQ.psd_apply(Gs, E), ww, n)
        l_o = OQ.psd_loglik(o[0], OQ.psd_apply(Gs, o[0]), ww, n)
        assert abs(l_q - l_o) < 1e-8 * abs(l_o), (l_q, l_o)
    r = q.predict_ratio(Z[keep][:5].astype(float) + 0.25)
    assert r.shape == (5,) and np.all(r > 0)
