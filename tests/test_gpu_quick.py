"""BLMLQuick(engine='hip') on counts on a fixed lattice (blml_amd/quick.py, libblml_hip.so blml_quick_*) against
the CPU restatement (oracle/blml_quick.py) with the dense Gram matrix of the same kernel (oracle/blml_pairs.py):
- FP64 tables: the same sweep count, safeguard and convergence, e and Se to 1e-9 of their scale, the residual
  history to 1e-6 relative (same permutations from default_rng(seed), so the same trajectory);
- a safeguarded sweep with one block holding every bin, through the C-ABI, equals the oracle's sweep to 1e-10;
- FP32 tables (the reference's default policy): e to 1e-4 of the FP64 oracle's converged e;
- a capped fit is scaled to e'Se = n^2; sqrt_density at the bins is Se / n;
- a conditional fit (dep column 1): the joint's e and the marginal's e on the joint's bins (e_marg_) against the
  oracle's fused fit (quick_fit_dep, the reference's loop) to 1e-7 of their scale; predict_marginal and
  predict_ratio at the bins against the oracle's S e / n; a warm refit from it keeps both fixed points."""
import ctypes
import sys

import numpy as np
import pytest

from oracle import blml_pairs as OP
from oracle import blml_quick as OQ

pytestmark = pytest.mark.gpu


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _problem(shape, total, seed, zero_frac=0.1):
    rng = np.random.default_rng(seed)
    grids = np.meshgrid(*[np.arange(m) for m in shape], indexing='ij')
    Z = np.stack([g.ravel() for g in grids], 1)
    c = np.array(shape, float) / 2
    r2 = sum(((Z[:, k] - c[k]) / (0.3 * shape[k])) ** 2 for k in range(len(shape)))
    p = np.exp(-0.5 * r2) + 0.05
    w = rng.poisson(total * p / p.sum()).astype(float)
    w[rng.uniform(size=len(w)) < zero_frac] = 0.0          # empty bins are dropped by the fit
    return Z, w


# name, shape, h, column dicts (without h), oracle units, block_size, seed, signs
CASES = [
    ('1d sinc', (300,), (0.5,), [dict(fc=0.6)], [('offset', 'sinc', [0.6], [0.0], [0])], 4096, 0, False),
    ('2d sinc blocks of 37', (24, 20), (1.0, 1.0), [dict(fc=0.35), dict(fc=0.3)],
     [('offset', 'sinc', [0.35], [0.0], [0]), ('offset', 'sinc', [0.3], [0.0], [1])], 37, 3, False),
    ('2d sinc x dirichlet circular', (18, 16), (1.0, 1.0), [dict(fc=0.4), dict(kind='dirichlet', fc=0.3, period=16.0)],
     [('offset', 'sinc', [0.4], [0.0], [0]), ('circular', 'dirichlet', [0.3], [16.0], [1])], 64, 1, False),
    ('3d sinc with signs', (10, 9, 8), (1.0, 1.0, 1.0), [dict(fc=0.4)] * 3,
     [('offset', 'sinc', [0.4], [0.0], [k]) for k in range(3)], 100, 2, True),
]


def _modes(cols, h):
    return [{j: dict(c, h=float(h[j])) for j, c in enumerate(cols)}]


@pytest.mark.parametrize('name,shape,h,cols,units,bs,seed,signed', CASES, ids=[c[0] for c in CASES])
def test_fit_matches_oracle_fp64(name, shape, h, cols, units, bs, seed, signed):
    from blml_amd.quick import BLMLQuick
    Z, w = _problem(shape, 8000, seed)
    hv = np.asarray(h, float)
    sg = None
    if signed:
        sg = np.where(Z[:, 0] < 3, -1.0, 1.0)              # g changes sign across a plane
    _note(f'quick: {name}, {int((w > 0).sum())} bins')
    q = BLMLQuick(tol=1e-8, block_size=bs, seed=seed, dtype=np.float64).fit(Z * hv, _modes(cols, hv), sample_weight=w,
                                                                           signs=sg)
    keep = w > 0
    G = OP.dense_pair_gram(units, hv, Z[keep], Z[keep])
    if signed:
        G = G * sg[keep][:, None] * sg[keep][None, :]
    e, Se, ns, conv, rho, rh, oh, safe = OQ.quick_fit(G, w[keep], tol=1e-8, block_size=bs, seed=seed)
    assert conv and q.converged_, (name, q.n_sweeps_, q.residual_)
    assert q.n_sweeps_ == ns and q.safeguard_used_ == safe
    es = e * (sg[keep] if signed else 1.0)
    assert np.abs(q.e_hat_ - es).max() < 1e-9 * np.abs(es).max(), name
    Ses = Se * (sg[keep] if signed else 1.0)
    assert np.abs(q.Se_ - Ses).max() < 1e-9 * np.abs(Ses).max(), name
    assert np.allclose(q.rho_history_, rh, rtol=1e-6, atol=1e-12)
    assert np.allclose(q.obj_history_, oh, rtol=1e-10)


def test_safeguarded_sweep_matches_oracle():
    import torch
    from blml_amd import _lib, local
    Z, w = _problem((200,), 20000, 4, zero_frac=0.0)
    keep = w > 0
    Z, w = Z[keep], w[keep]
    hv = np.array([1.0])
    tab = local.lattice_tables(['sinc'], [0.02], None, hv, Z)
    t, dZ, _, hold = local._prepare(Z, Z, tab, False, False)
    G = OP.dense_pair_gram([('offset', 'sinc', [0.02], [0.0], [0])], hv, Z, Z)
    B, n = len(w), float(w.sum())
    lib = _lib.load()
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    P = lambda x: None if x is None else ctypes.c_void_p(x.data_ptr())
    dw = torch.from_numpy(w).cuda()
    e = torch.ones(B, dtype=torch.float64, device='cuda')
    Se = torch.empty_like(e)
    _lib.check(lib.blml_quick_exact(ctypes.byref(t), P(dZ), B, P(e), None, P(Se), st))
    eo, So = np.ones(B), G @ np.ones(B)
    rng = np.random.default_rng(0)
    log = []
    _note('quick: five safeguarded sweeps, one block of every bin')
    for _ in range(5):
        perm = rng.permutation(B)
        _lib.check(lib.blml_quick_sweep(ctypes.byref(t), P(dZ), B, P(dw), None,
                                        P(torch.from_numpy(perm.astype(np.int32)).cuda()), B, n, float(G[0, 0]), 1,
                                        P(e), P(Se), st))
        OQ.sweep(G, w, n, float(G[0, 0]), eo, So, perm, B, True, log)
    assert min(tt for _, _, tt, _ in log) < 1.0
    ed, Sd = e.cpu().numpy(), Se.cpu().numpy()
    assert np.abs(ed - eo).max() < 1e-10 * np.abs(eo).max()
    assert np.abs(Sd - So).max() < 1e-10 * np.abs(So).max()


def test_fp32_tables_reach_the_fp64_solution():
    from blml_amd.quick import BLMLQuick
    name, shape, h, cols, units, bs, seed, _ = CASES[1]
    Z, w = _problem(shape, 8000, seed)
    hv = np.asarray(h, float)
    q = BLMLQuick(tol=1e-6, block_size=bs, seed=seed, dtype=np.float32).fit(Z * hv, _modes(cols, hv), sample_weight=w)
    keep = w > 0
    G = OP.dense_pair_gram(units, hv, Z[keep], Z[keep])
    en = OQ.newton_solve(G, w[keep])
    assert q.converged_
    assert np.abs(q.e_hat_ - en).max() < 1e-4 * np.abs(en).max()


def test_capped_fit_and_sqrt_density():
    from blml_amd.quick import BLMLQuick
    name, shape, h, cols, units, bs, seed, _ = CASES[2]
    Z, w = _problem(shape, 8000, seed)
    hv = np.asarray(h, float)
    q = BLMLQuick(tol=1e-12, block_size=bs, seed=seed, dtype=np.float64, max_sweeps=3).fit(
        Z * hv, _modes(cols, hv), sample_weight=w)
    n = q.n_samples_
    assert q.n_sweeps_ == 3 and q.Se_ is not None
    assert abs(q.e_hat_ @ q.Se_ - n * n) < 1e-9 * n * n
    g = q.sqrt_density(q.Z_)
    assert np.abs(g - q.Se_ / n).max() < 1e-9 * np.abs(q.Se_ / n).max()


def test_warm_start_matches_oracle():
    """init= on the same lattice (FP64 tables): a second fit on other bins (some gone, some new, counts moved) warm-
    started from the first equals the oracle's quick_fit from warm_start(the first fit's e) -- the same sweeps, e to
    1e-9 -- and reaches the cold fit's solution (warm starts from another lattice: test_gpu_quick_psd.py)."""
    from blml_amd.quick import BLMLQuick
    name, shape, h, cols, units, bs, seed, signed = CASES[1]
    Z, w = _problem(shape, 8000, seed, zero_frac=0.0)
    rng = np.random.default_rng(11)
    m1, m2 = rng.uniform(size=len(w)) < 0.85, rng.uniform(size=len(w)) < 0.85
    hv = np.asarray(h, float)
    Z1, w1 = Z[m1 & (w > 0)], w[m1 & (w > 0)]
    Z2, w2 = Z[m2 & (w > 0)], w[m2 & (w > 0)] + rng.poisson(3.0, int((m2 & (w > 0)).sum()))
    _note(f'quick warm start: {len(w1)} -> {len(w2)} bins')
    kw = dict(tol=1e-8, block_size=bs, seed=seed, dtype=np.float64)
    q1 = BLMLQuick(**kw).fit(Z1 * hv, _modes(cols, hv), sample_weight=w1)
    q2 = BLMLQuick(**kw).fit(Z2 * hv, _modes(cols, hv), sample_weight=w2, init=q1)
    key = lambda Z: [z.tobytes() for z in np.ascontiguousarray(Z, np.int64)]
    prev = dict(zip(key(Z1), q1.e_hat_))
    e_prev = np.array([prev.get(k, np.nan) for k in key(Z2)])
    new = np.isnan(e_prev)
    assert q2.n_new_bins_ == int(new.sum()) > 0 and (~new).any()
    G = OP.dense_pair_gram(units, hv, Z2, Z2)
    e0 = OQ.warm_start(np.where(new, 0.0, e_prev), new, G, w2)
    e, Se, ns, conv, rho, rh, oh, safe = OQ.quick_fit(G, w2, tol=1e-8, block_size=bs, seed=seed, e0=e0)
    assert conv and q2.converged_ and q2.n_sweeps_ == ns, (q2.n_sweeps_, ns)
    assert np.abs(q2.e_hat_ - e).max() < 1e-9 * np.abs(e).max()
    cold = BLMLQuick(**kw).fit(Z2 * hv, _modes(cols, hv), sample_weight=w2)
    assert np.abs(q2.e_hat_ - cold.e_hat_).max() < 1e-6 * np.abs(cold.e_hat_).max()


def test_conditional_fit_matches_oracle():
    from blml_amd import make_modes
    from blml_amd.quick import BLMLQuick
    units = [('offset', 'sinc', [0.35], [0.0], [0]), ('offset', 'sinc', [0.3], [0.0], [1])]
    off = [u for u in units if 1 not in u[4]]
    Z, w = _problem((14, 12), 6000, 3)
    modes = [{j: dict(c, h=1.0) for j, c in m.items()} for m in make_modes([0.35, 0.3], ['sinc', 'sinc'], dep=[1])]
    _note('quick: conditional fit, dep column 1')
    q = BLMLQuick(tol=1e-10, block_size=37, seed=2, dtype=np.float64).fit(Z.astype(float), modes, sample_weight=w)
    keep = w > 0
    Zk, wk = Z[keep], w[keep]
    G = OP.dense_pair_gram(units, np.ones(2), Zk, Zk)
    Gm = OP.dense_pair_gram(off, np.ones(2), Zk, Zk)
    (ej, Sej, _, cj, _), (em, Sem, _, cm, _) = OQ.quick_fit_dep(G, Gm, wk, tol=1e-10, block_size=37, seed=2)
    assert q.dep_ == [1] and q.converged_ and q.marg_converged_ and cj and cm
    assert np.abs(q.e_hat_ - ej).max() < 1e-7 * np.abs(ej).max()
    assert np.abs(q.e_marg_ - em).max() < 1e-7 * np.abs(em).max()
    n = wk.sum()
    pm = q.predict_marginal(Zk[:, [0]].astype(float))
    assert np.abs(pm - (Sem / n) ** 2).max() < 1e-7 * ((Sem / n) ** 2).max()
    r = q.predict_ratio(Zk.astype(float))
    want = (Sej / n) ** 2 / (Sem / n) ** 2
    assert np.abs(r - want).max() < 1e-6 * want.max()
    again = BLMLQuick(tol=1e-10, block_size=37, seed=2, dtype=np.float64).fit(Z.astype(float), modes,
                                                                             sample_weight=w, init=q)
    assert again.converged_ and again.n_sweeps_ <= 2 and again.marg_sweeps_ <= 2
    assert np.abs(again.e_marg_ - em).max() < 1e-7 * np.abs(em).max()
    with pytest.raises(ValueError, match='signs'):
        BLMLQuick(dtype=np.float64).fit(Z.astype(float), modes, sample_weight=w, signs=np.ones(len(w)))
    with pytest.raises(RuntimeError, match='dep'):
        BLMLQuick(tol=1e-6, dtype=np.float64).fit(Z.astype(float), [{j: dict(c, h=1.0) for j, c in m.items()}
                                                                    for m in make_modes([0.35, 0.3])],
                                                  sample_weight=w).predict_ratio(Zk[:3].astype(float))


@pytest.mark.parametrize('case', [1, 4], ids=['2d sinc x dirichlet', '2d periodic ball'])
def test_points_fit_matches_oracle_on_its_bins(case):
    """BLMLQuick on points (no h in the modes): binned on the fit's own lattice as the reference bins them (pinned
    on the CPU by test_host_units), then the counts fit.  Its e and loglik's stationarity against the oracle's
    quick_fit / Newton solve with the dense Gram of the same bins (FP64 tables, tol 1e-8): e to 1e-6 of its scale;
    a refit with 10% more points keeps the lattice, starts from the previous e and ends at the cold fit's solution."""
    import os
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), 'golden'))
    from cases_binning import binning_cases
    from blml_amd import make_modes
    from blml_amd.modes import parse_modes
    from blml_amd.quick import BLMLQuick, bin_points
    name, X, modes, bs = binning_cases(make_modes)[case]
    _note(f'quick on points: {name}')
    q = BLMLQuick(tol=1e-8, bin_scale=bs, dtype=np.float64).fit(X, modes)
    specs, _ = parse_modes(modes, 2)
    bp = bin_points(X, specs, bin_scale=bs)
    assert np.array_equal(q.Z_, bp['Z']) and np.array_equal(q.w_, bp['w']) and np.allclose(q._h_vec, bp['h'])
    Zi = np.round(bp['Z'] / bp['h']).astype(np.int64)
    sp = specs[0]
    if case == 1:
        units = [('offset', 'sinc', [sp['fc'][0]], [0.0], [0]), ('circular', 'dirichlet', [sp['fc'][1]], [10.0], [1])]
    else:
        units = [('torus', 'pball', [sp['fc'][0], sp['fc'][1]], [12.0, 9.0], [0, 1])]
    G = OP.dense_pair_gram(units, bp['h'], Zi, Zi)
    en = OQ.newton_solve(G, bp['w'])
    assert q.converged_
    assert np.abs(q.e_hat_ - en).max() < 1e-6 * np.abs(en).max()
    more = np.concatenate([X, X[:len(X) // 10] + 0.3 * bp['h']])
    warm = BLMLQuick(tol=1e-8, bin_scale=bs, dtype=np.float64).fit(more, modes, init=q)
    # the cold fit of the same counts on the kept lattice (a fresh fit of the points would pick its own lattice)
    cold = BLMLQuick(tol=1e-8, bin_scale=bs, dtype=np.float64)._fit_counts(warm.Z_, warm.w_, warm._h_vec, modes,
                                                                            None, None, False)
    _note(f'   refit with 10% more points: {warm.n_sweeps_} sweeps warm, {cold.n_sweeps_} cold')
    assert not warm.rebinned_ and np.allclose(warm._h_vec, q._h_vec) and warm.converged_
    assert np.abs(warm.e_hat_ - cold.e_hat_).max() < 1e-6 * np.abs(cold.e_hat_).max()
