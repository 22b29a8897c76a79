"""The CPU restatement of the BLMLQuick counts path (oracle/blml_quick.py), pinned without the reference:
- the one-coordinate root zeroes its quadratic and the 1-D objective's derivative, on both branches;
- quick_fit's converged e equals the unique minimiser of the convex objective found by an independent damped Newton
  solve, and satisfies (G e)_b e_b = n w_b;
- every safeguarded block step does not raise the objective (checked by evaluating it directly);
- the returned Se is G e; a capped fit is scaled to e'Se = n^2."""
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
