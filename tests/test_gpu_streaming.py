"""BLMLBinnedStreaming on the GPU (blml_amd/streaming.py), against the reference's own streaming checks
(results/binned_stream_checks.py) on the same constructions:
- loo: a held batch's score and implicit gradient inside the leave-batch-out context equal those of an exact refit
  on the held counts less the batch (real and complex; :37-53);
- epochs: a sinc cut-off started 1.6x above and 0.6x below the truth's moves over epochs toward the batch selector's
  choice on the same counts (real fits: the sign-changing truth is not band-limited as a positive root, so both go to
  the Nyquist cap; complex fits: near the truth), with the residual test met (converged_, epochs_, residual_; :56-76);
- burn-in: an fc left out of the modes (NaN) is the burn-in's selection, equal to select_fc's on the same binomial
  split; then the batches' prequential scores are finite and the history records every batch."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def truth(shape, fc0, n_src=8, seed=3, floor=0.0):
    """A sign-changing sum of sinc bumps, squared (the reference's streaming truth, results/binned_stream_checks.py)."""
    r = np.random.default_rng(seed)
    G = np.meshgrid(*[np.arange(m, dtype=float) for m in shape], indexing='ij')
    g = 0.0
    for _ in range(n_src):
        x0 = [r.uniform(0.3 * m, 0.7 * m) for m in shape]
        a = r.choice([-1, 1]) * r.uniform(0.5, 1.0)
        K = 1.0
        for j in range(len(shape)):
            K = K * fc0 * np.sinc(fc0 * (G[j] - x0[j]))
        g = g + a * K
    p = g ** 2
    p = p / p.sum()
    return (1 - floor) * p + floor / p.size


def draws(p, n, seed):
    return np.random.default_rng(seed).multinomial(n, p.ravel()).reshape(p.shape).astype(float)


@pytest.mark.parametrize('cg', [False, True], ids=['real', 'complex'])
def test_leave_batch_out_equals_exact_refit(cg):
    import blml_amd
    from blml_amd import BLMLBinned, BLMLBinnedStreaming
    p = truth((96,), 0.2)
    batches = [draws(p, 600, 10 + i) for i in range(10)]
    tot = sum(batches)
    modes = blml_amd.make_modes([0.2], ['sinc'])
    s = BLMLBinnedStreaming(tol=1e-10, fc_update=None).fit(batches, modes, epochs=1, complex_g=cg)
    s.max_sweeps = 20000
    s._warm_fit(s.w_, np.asarray(s.e_hat_), None)                # the held fit to convergence
    es, eg = [], []
    for j in (0, 4, 8):
        b = batches[j]
        eps = 1e-3
        with s._left_out(b):
            sl, gl = s.log_density(b), s.implicit_grad(b, eps)[0][0]
        ex = BLMLBinned(tol=1e-10, max_sweeps=20000).fit(tot - b, modes, complex_g=cg, init=s)
        se, ge = ex.log_density(b), ex.implicit_grad(b, eps)[0][0]
        es.append(abs(sl - se))
        eg.append(abs(gl - ge) / max(abs(ge), 1e-3))
        _note(f'loo {"complex" if cg else "real"} batch {j}: score {sl:.10f} vs exact {se:.10f}, implicit grad '
              f'{gl:.6f} vs {ge:.6f}')
    assert max(es) < 1e-6 and max(eg) < 1e-2


@pytest.mark.parametrize('cg', [False, True], ids=['real', 'complex'])
def test_epochs_move_fc_toward_the_selector(cg):
    import blml_amd
    from blml_amd import BLMLBinnedStreaming
    p = truth((128,), 0.2)
    batches = [draws(p, 400, 100 + i) for i in range(15)]
    tr, va = sum(batches[:12]), sum(batches[12:])
    from blml_amd.select import FcSearch, val_score

    def score_many(vs):
        ok = [i for i, v in enumerate(vs) if v[0] <= 0.5]
        ms = blml_amd.fit_many([(tr, blml_amd.make_modes([vs[i][0]], ['sinc']), dict(complex_g=cg)) for i in ok],
                               tol=1e-8) if ok else []
        out = [-np.inf] * len(vs)
        for i, m in zip(ok, ms):
            out[i] = val_score(m.density_, va)
        return out
    ref = float(FcSearch(1, ['sinc'], [None], fc0=0.02, patience=2, max_sweeps=6).run(score_many)[0][0])
    _note(f'batch selector fc {ref:.4f}')
    for start in (1.6, 0.6):
        m0 = blml_amd.make_modes([0.2 * start], ['sinc'])
        ep = BLMLBinnedStreaming(tol=1e-8, lr=0.05, param_every=1, batch_sweeps=50).fit(
            batches, m0, epochs=12, complex_g=cg, tol=1e-3, tol_check='exact')
        fe = ep._fc_now()[0][0]
        _note(f'epochs start x{start}: fc {0.2 * start:.3f} -> {fe:.4f} ({ep.epochs_} epochs, converged '
              f'{ep.converged_}, residual {ep.residual_:.1e}); selector {ref:.4f}')
        assert abs(np.log(fe / ref)) < max(0.5 * abs(np.log(0.2 * start / ref)), 0.1)
        assert np.all(np.isfinite([h['preq'] for h in ep.history_ if 'preq' in h][1:]))


def test_burn_in_selects_the_left_out_fc():
    import blml_amd
    from blml_amd import BLMLBinnedStreaming
    p = truth((96,), 0.2)
    batches = [draws(p, 600, 40 + i) for i in range(6)]
    kw = dict(tol=1e-8, dtype=np.float64)
    s = BLMLBinnedStreaming(burn_in=1500, fc_update='adam', implicit_grad=False, lr=0.05, **kw)
    for b in batches:
        s.partial_fit(b, [{0: dict(fc=np.nan)}])
    held = batches[0] + batches[1] + batches[2]                  # 1800 >= burn_in after the third batch
    fc_sel = blml_amd.select_fc(held, ['sinc'], search='shared', seed=0, **kw)[0][0]
    burn = [h for h in s.history_ if h.get('burn_in')][0]
    _note(f'burn-in: selected fc {burn["fc"][0][0]:.5f} (select_fc on the same split {fc_sel:.5f}); '
          f'{len(s.preq_)} history entries, fc now {s._fc_now()[0][0]:.5f}')
    assert abs(burn['fc'][0][0] - fc_sel) < 1e-12 * fc_sel
    assert len(s.preq_) == len(batches) - 2 and np.all(np.isfinite(s.preq_[1:]))
