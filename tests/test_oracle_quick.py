"""The CPU restatement of the BLMLQuick counts path (oracle/blml_quick.py), pinned without the reference:
- the one-coordinate root zeroes its quadratic and the 1-D objective's derivative, on both branches;
- quick_fit's converged e equals the unique minimiser of the convex objective found by an independent damped Newton
  solve, and satisfies (G e)_b e_b = n w_b;
- every safeguarded block step does not raise the objective (checked by evaluating it directly);
- the returned Se is G e; a capped fit is scaled to e'Se = n^2;
- the conditional fit's marginal is the merged-bin minimiser split by the counts (the reference's own statement of
  what its fused fit computes, blml_gpu.py:67-72)."""
import numpy as np
import pytest

from oracle import blml_pairs as OP
from oracle import blml_quick as OQ


def problem(shape, units, h, total, seed):
    """(Z integer coordinates of the occupied bins, counts, G) for a smooth density on a box lattice."""
    rng = np.random.default_rng(seed)
    grids = np.meshgrid(*[np.arange(m) for m in shape], indexing='ij')
    Z = np.stack([g.ravel() for g in grids], 1)
    c = np.array(shape, float) / 2
    r2 = sum(((Z[:, k] - c[k]) / (0.3 * shape[k])) ** 2 for k in range(len(shape)))
    p = np.exp(-0.5 * r2) + 0.05
    w = rng.poisson(total * p / p.sum()).astype(float)
    keep = w > 0
    Z, w = Z[keep], w[keep]
    G = OP.dense_pair_gram(units, np.asarray(h, float), Z, Z)
    return Z, w, G


CASES = [
    ('1d sinc', (120,), [('offset', 'sinc', [0.3], [0.0], [0])], (1.0,), 4000),
    ('2d sinc', (14, 12), [('offset', 'sinc', [0.35], [0.0], [0]), ('offset', 'sinc', [0.3], [0.0], [1])],
     (1.0, 1.0), 6000),
    ('2d sinc x dirichlet circular', (12, 16), [('offset', 'sinc', [0.4], [0.0], [0]),
                                                ('circular', 'dirichlet', [0.3], [16.0], [1])], (1.0, 1.0), 5000),
]


def test_conditional_fit_marginal_is_the_merged_fit_split_by_counts():
    """quick_fit_dep (the joint and the marginal of dep column 1 in one loop): the joint is the convex minimiser of
    the joint problem (newton_solve); the marginal, fitted on the joint's bins with the kernel of the units off the
    dep side, is the minimiser on the merged bins (bins that differ only in column 1 merged, counts summed) split
    back in proportion to the counts.  Leaving out the split (every bin of a merged bin given its whole
    coefficient) fails the same check."""
    units = [('offset', 'sinc', [0.35], [0.0], [0]), ('offset', 'sinc', [0.3], [0.0], [1])]
    off = [u for u in units if 1 not in u[4]]
    Z, w, G = problem((14, 12), units, (1.0, 1.0), 6000, 3)
    Gm = OP.dense_pair_gram(off, np.ones(2), Z, Z)
    (ej, Sej, nsj, cj, rj), (em, Sem, nsm, cm, rm) = OQ.quick_fit_dep(G, Gm, w, tol=1e-10, block_size=37, seed=2)
    assert cj and cm, (nsj, rj, nsm, rm)
    en = OQ.newton_solve(G, w)
    assert np.abs(ej - en).max() < 1e-7 * np.abs(en).max()
    Zr, first, inv = np.unique(Z[:, [0]], axis=0, return_index=True, return_inverse=True)
    inv = np.ravel(inv)
    W = np.bincount(inv, weights=w)
    Gmm = OP.dense_pair_gram(off, np.ones(2), Z[first], Z[first])
    emm = OQ.newton_solve(Gmm, W)
    want = emm[inv] * w / W[inv]
    assert np.abs(em - want).max() < 1e-7 * np.abs(want).max()
    assert np.abs(emm[inv] - want).max() > 1e-2 * np.abs(want).max()
    n = w.sum()
    assert np.abs(Sem * em / (n * w) - 1).max() < 1e-10


def test_root_solves_its_quadratic():
    rng = np.random.default_rng(0)
    b = rng.normal(size=200) * 10
    cc = rng.uniform(0.1, 50, 200)
    sbb = 0.37
    e = OQ.root(b, cc, sbb)
    assert np.all(e > 0)
    assert np.abs(sbb * e * e + b * e - cc).max() < 1e-10 * np.max(cc + np.abs(b * e))
    assert np.abs(sbb * e + b - cc / e).max() < 1e-9 * np.max(np.abs(b) + cc / e)


@pytest.mark.parametrize('name,shape,units,h,total', CASES, ids=[c[0] for c in CASES])
def test_fixed_point_is_the_convex_minimiser(name, shape, units, h, total):
    Z, w, G = problem(shape, units, h, total, 1)
    e, Se, ns, conv, rho, rh, oh, safe = OQ.quick_fit(G, w, tol=1e-10, block_size=37, seed=2)
    assert conv, (name, ns, rho)
    n = w.sum()
    assert np.abs(Se - G @ e).max() < 1e-9 * np.abs(Se).max()
    assert np.abs(Se * e / (n * w) - 1).max() < 1e-10
    en = OQ.newton_solve(G, w)
    assert np.abs(e - en).max() < 1e-7 * np.abs(en).max(), name
    assert abs(e @ Se - n * n) < 1e-8 * n * n


def test_safeguarded_steps_do_not_raise_the_objective():
    # one block holding every bin of a 1-D lattice: the fully parallel update overshoots and the safeguard shrinks
    # it (blml_gpu.py:316-319); every chosen step is checked against the objective evaluated directly
    Z, w, G = problem((200,), [('offset', 'sinc', [0.02], [0.0], [0])], (1.0,), 20000, 4)
    n = w.sum()
    e = np.ones(len(w))
    Se = G @ e
    log = []
    rng = np.random.default_rng(0)
    for _ in range(30):
        OQ.sweep(G, w, n, float(G[0, 0]), e, Se, rng.permutation(len(w)), len(w), True, log)
    ts = [t for _, _, t, _ in log]
    assert min(ts) < 1.0 and max(ts) > 0.0
    for e0, dl, t, I in log:
        e1 = e0.copy()
        e1[I] = e0[I] + t * dl
        assert OQ.objective(G, e1, w, n) <= OQ.objective(G, e0, w, n) + 1e-12 * abs(OQ.objective(G, e0, w, n))
    assert np.abs(Se - G @ e).max() < 1e-9 * np.abs(Se).max()


def test_capped_fit_is_normalised():
    Z, w, G = problem((14, 12), CASES[1][2], (1.0, 1.0), 6000, 1)
    e, Se, ns, conv, rho, rh, oh, safe = OQ.quick_fit(G, w, tol=1e-12, seed=0, max_sweeps=2)
    assert ns == 2 and Se is not None
    n = w.sum()
    assert abs(e @ (G @ e) - n * n) < 1e-9 * n * n
    assert np.abs(Se - G @ e).max() < 1e-9 * np.abs(Se).max()


def test_warm_start_reaches_the_convex_minimiser():
    """init= on the same lattice: a fit's e carried to a second problem (some bins gone, some new, counts moved):
    each new bin starts where the 1-D objective's derivative in it vanishes (the others fixed), and the warm fit
    reaches the second problem's Newton fixed point, in fewer sweeps than the cold fit."""
    name, shape, units, h, total = CASES[1]
    Z, w, G = problem(shape, units, h, total, 1)
    rng = np.random.default_rng(5)
    m1, m2 = rng.uniform(size=len(w)) < 0.85, rng.uniform(size=len(w)) < 0.85
    Z1, w1, G1 = Z[m1], w[m1], G[np.ix_(m1, m1)]
    Z2, w2, G2 = Z[m2], w[m2] + rng.poisson(3.0, int(m2.sum())), G[np.ix_(m2, m2)]
    e1 = OQ.quick_fit(G1, w1, tol=1e-10, block_size=37, seed=2)[0]
    key = lambda Z: [z.tobytes() for z in np.ascontiguousarray(Z)]
    prev = dict(zip(key(Z1), e1))
    e_prev = np.array([prev.get(k, np.nan) for k in key(Z2)])
    new = np.isnan(e_prev)
    assert new.any() and (~new).any(), (new.sum(), len(new))
    e0 = OQ.warm_start(e_prev, new, G2, w2)
    n2 = w2.sum()
    Se0 = G2 @ np.where(new, 0.0, e0)                     # the others' part at the new bins
    sbb = G2[0, 0]
    assert np.abs(sbb * e0[new] + Se0[new] - n2 * w2[new] / e0[new]).max() < 1e-9 * np.abs(Se0[new]).max()
    assert np.array_equal(e0[~new], e_prev[~new])
    warm = OQ.quick_fit(G2, w2, tol=1e-10, block_size=37, seed=2, e0=e0)
    cold = OQ.quick_fit(G2, w2, tol=1e-10, block_size=37, seed=2)
    en = OQ.newton_solve(G2, w2)
    assert warm[3] and np.abs(warm[0] - en).max() < 1e-7 * np.abs(en).max()
    assert warm[2] < cold[2], (warm[2], cold[2])


# ------------------------------------------------------------------------------------------ complex / rank K
def _hist1d(total, seed):
    """The 40-bin two-bump histogram of the reference's parity cells (results/binned_parity.py:21-26)."""
    rng = np.random.default_rng(seed)
    x = np.arange(40, dtype=float)
    p = np.exp(-0.5 * ((x - 14) / 3) ** 2) + 0.5 * np.exp(-0.5 * ((x - 27) / 2) ** 2) + 0.02
    return rng.multinomial(total, p / p.sum()).astype(float)


def _grams(fcs, Z):
    return [OP.dense_pair_gram([('offset', 'sinc', [f], [0.0], [0])], np.ones(1), Z, Z) for f in fcs]


def _stat_resid(Gs, E, w):
    """max_b |(S e)_b |e_b|^2 / (n w_b) - e_b| / |e_b| under the dense Grams, written from the stationarity
    condition (S_k e_k)_b = n w_b e_bk / |e_b|^2 (results/binned_parity.py:64-70), not from the oracle's code."""
    n = w.sum()
    U = np.stack([(Gs[0] if len(Gs) == 1 else Gs[k]) @ E[:, k] for k in range(E.shape[1])], 1)
    nr = (np.abs(E) ** 2).sum(1)
    return float(np.max(np.sqrt((np.abs(U * (nr / (n * w))[:, None] - E) ** 2).sum(1)) / np.sqrt(nr)))


@pytest.mark.parametrize('shared', [True, False])
def test_psd_row_update_is_row_stationary(shared):
    """The new row solves its own stationarity condition R_bk + s_k e_bk = n w_b e_bk / |e_b|^2 (R = S e - s e
    with the row's own term removed): the closed form for one kernel, the secular root for per-mode kernels."""
    rng = np.random.default_rng(4)
    m, K = 300, 3
    R = rng.normal(size=(m, K)) + 1j * rng.normal(size=(m, K))
    EI = rng.normal(size=(m, K)) + 1j * rng.normal(size=(m, K))
    cc = rng.uniform(0.5, 80.0, m)
    s = np.full(K, 0.7) if shared else np.array([0.7, 0.31, 0.52])
    e = OQ.psd_row_update(R, EI, cc, s, shared)
    lam = cc / (np.abs(e) ** 2).sum(1)
    assert np.abs(R + s[None] * e - lam[:, None] * e).max() < 1e-9 * np.abs(R).max()
    # a row with R = 0 keeps its direction (shared kernel)
    if shared:
        e0 = OQ.psd_row_update(np.zeros((1, K), complex), EI[:1], cc[:1], s, True)
        assert np.allclose(e0 / np.linalg.norm(e0), EI[:1] / np.linalg.norm(EI[:1]))
        assert abs(0.7 * np.linalg.norm(e0) ** 2 - cc[0]) < 1e-9 * cc[0]


PSD_CASES = [
    ('d1 sinc complex', [0.35], 1, True),
    ('d1 sinc rank2 shared complex', [0.35], 2, True),
    ('d1 sinc rank2 shared real', [0.35], 2, False),
    ('d1 sinc rank2 per-mode complex', [0.35, 0.12], 2, True),
]


@pytest.mark.parametrize('name,fcs,K,cplx', PSD_CASES, ids=[c[0] for c in PSD_CASES])
@pytest.mark.parametrize('solver', ['cd', 'lbfgs'])
def test_psd_fit_is_stationary(name, fcs, K, cplx, solver):
    w = _hist1d(3000, 1)
    keep = w > 0
    Z = np.flatnonzero(keep)[:, None]
    w = w[keep]
    Gs = _grams(fcs, Z)
    out = OQ.psd_fit(Gs, w, tol=1e-6, block_size=4096, seed=0, solver=solver, K=K, complex_g=cplx)
    E, SE, conv = out[0], out[1], out[3]
    n = w.sum()
    assert conv, (name, solver, out[2], out[4])
    assert E.shape == (len(w), K) and np.iscomplexobj(E) == cplx
    SEx = OQ.psd_apply(Gs, E)
    assert np.abs(SE - SEx).max() < 1e-9 * np.abs(SEx).max()
    assert abs(np.real(np.vdot(E, SEx)) - n * n) < 1e-6 * n * n     # stationarity: e'Se = sum n w = n^2
    if solver == 'cd':
        assert _stat_resid(Gs, E, w) < 1e-5
    else:            # the likelihood stop: the near-null directions of S can keep the residual above tol
        assert _stat_resid(Gs, E, w) < 1e-2


@pytest.mark.parametrize('name,fcs,K,cplx', PSD_CASES, ids=[c[0] for c in PSD_CASES])
def test_psd_cd_and_lbfgs_reach_the_same_likelihood(name, fcs, K, cplx):
    w = _hist1d(3000, 1)
    keep = w > 0
    Z = np.flatnonzero(keep)[:, None]
    w = w[keep]
    Gs = _grams(fcs, Z)
    a = OQ.psd_fit(Gs, w, tol=1e-8, seed=0, solver='cd', K=K, complex_g=cplx)
    b = OQ.psd_fit(Gs, w, tol=1e-8, seed=0, solver='lbfgs', K=K, complex_g=cplx)
    la = OQ.psd_loglik(a[0], a[1], w, w.sum())
    lb = OQ.psd_loglik(b[0], b[1], w, w.sum())
    assert abs(la - lb) < 1e-7, (name, la, lb)


def test_complex_fit_not_below_the_real_fit():
    """f = |g|^2 over complex g contains the real fits: the complex optimum's likelihood is at least the real
    rank-1 fit's."""
    w = _hist1d(3000, 1)
    keep = w > 0
    Z = np.flatnonzero(keep)[:, None]
    w = w[keep]
    Gs = _grams([0.35], Z)
    e, Se = OQ.quick_fit(Gs[0], w, tol=1e-10)[:2]
    n = w.sum()
    l_real = float(w @ np.log(Se ** 2) - n * np.log(e @ Se)) / n
    E, SE = OQ.psd_fit(Gs, w, tol=1e-9, complex_g=True)[:2]
    l_c = OQ.psd_loglik(E, SE, w, n)
    assert l_c >= l_real - 1e-9


@pytest.mark.parametrize('name,fcs,K,cplx', [c for c in PSD_CASES if c[3]], ids=[c[0] for c in PSD_CASES if c[3]])
def test_psd_counts_fit_matches_the_binned_oracle(name, fcs, K, cplx):
    """The reference's own parity check (results/binned_parity.py:52-57): BLMLBinned on the full lattice and the
    counts fit on its occupied bins (sinc: the lattice tables are exact, so both have the same Gram) reach the same
    likelihood and density at the bins."""
    from oracle import blml_oracle as O
    N = _hist1d(3000, 1)
    keep = N > 0
    Z = np.flatnonzero(keep)[:, None]
    w = N[keep]
    modes = O.make_modes([[f] for f in fcs] if len(fcs) > 1 else fcs, ['sinc'], rank=K if len(fcs) == 1 else 1)
    mb = O.BLMLBinnedOracle(tol=1e-7).fit(N, modes, complex_g=cplx)
    E, SE = OQ.psd_fit(_grams(fcs, Z), w, tol=1e-7, K=K, complex_g=cplx)[:2]
    n = w.sum()
    fq = (np.abs(SE) ** 2).sum(1) / n ** 2
    assert abs(OQ.psd_loglik(E, SE, w, n) - mb.loglik_) < 1e-6        # both sum w log(F / Q) / n
    assert np.abs(fq / np.asarray(mb.density_)[keep] - 1).max() < 1e-4
