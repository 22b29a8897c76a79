"""The packaged warm-start network (blml_amd/warm_net.npz, trained by `python -m blml_amd.warm train` on fits of
seeds 1000-50999) on held-out fits of the batched uneven-lattice workload (seed 4):
- its prediction of the real rank-1 start solution (the engine's solve, log of e / sqrt(c / s_0)) is closer than
  the cold start's own first guess sqrt(c / s_0) (a prediction of 0), in the masked squared error it was trained on;
- warm-started fits converge with at most half the cold start's iterations and less device time, and their
  likelihoods are no worse than the cold fits' in distribution (these two-mode fits are non-convex: each start ends
  at a local maximum of its own, better or worse by about 1e-3 relative either way);
- fit_many(..., warm=True) sets the start on 2-D jobs only and returns converged fits."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _jobs(n, seed=4):
    import bench
    import blml_amd
    return [(c, bench.modes_of(g, blml_amd), kw['complex_g']) for c, g, kw, _ in bench.uneven_fits(n, seed=seed)]


def test_prediction_beats_the_cold_first_guess():
    from blml_amd import warm
    net = warm.WarmNet.load()
    jobs = _jobs(64)
    _note('warm: start solutions of 64 held-out fits')
    er, s0 = warm.start_solutions([(c, m) for c, m, _ in jobs])
    X = np.stack([warm.features(c, m) for c, m, _ in jobs])
    YM = [warm.start_target(c, e, s) for (c, m, _), e, s in zip(jobs, er, s0)]
    Y = np.stack([y for y, _ in YM])
    M = np.stack([m for _, m in YM])
    pred = net.predict(X)
    loss_net = float(np.sum(M * (pred - Y) ** 2) / M.sum())
    loss_zero = float(np.sum(M * Y ** 2) / M.sum())
    _note(f'warm: masked MSE network {loss_net:.4f}, cold first guess {loss_zero:.4f}')
    assert loss_net < 0.5 * loss_zero


def test_warm_start_halves_iterations_and_keeps_the_likelihoods():
    import bench
    from blml_amd import warm
    net = warm.WarmNet.load()
    _note('warm: 256 held-out fits, cold then warm')
    out = warm.evaluate(net, _jobs(256), bench.fit_params())
    _note(f'warm: {out}')
    assert out['warm_iterations'] <= 0.5 * out['cold_iterations']
    assert out['warm_device_fit_s'] < out['cold_device_fit_s']
    assert out['warm_stops']['stall'] == 0 and out['warm_stops']['max_sweeps'] == 0
    assert out['warm_worse'] <= 1.5 * out['warm_better'] + 8


def test_fit_many_warm_sets_2d_starts_and_converges():
    import bench
    import blml_amd
    jobs = [(c, m, dict(complex_g=cg)) for c, m, cg in _jobs(12, seed=5)]
    x = np.random.default_rng(0).normal(size=4000)
    c1, _ = np.histogram(x, bins=256, range=(-4, 4))
    jobs.append((c1.astype(float), blml_amd.make_modes([0.05], ['sinc']), {}))
    P = bench.fit_params()
    _note('warm: fit_many(warm=True) over 12 2-D fits and one 1-D fit')
    fits = blml_amd.fit_many(jobs, tol=P['tol'], max_sweeps=P['max_sweeps'], dtype=np.float32, ll_tol=P['ll_tol'],
                             solver=P['solver'], lbfgs_m=P['lbfgs_m'], hist_cap=0, warm=True)
    assert [f.warm_start_ for f in fits] == [True] * 12 + [False]
    for f in fits[:12]:
        assert f.stop_reason_ in ('tol', 'll_tol') and np.isfinite(f.loglik_)
        assert f.density_.shape == f.shape_
