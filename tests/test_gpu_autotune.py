"""The plan's autotuner on the GPU (blml_plan_autotune through the C-ABI): on a batch of uneven 2-D fits (padded
FFT lines and band DFTs) it times every candidate -- the FFT radix-stage order and LDS-resident intermediates --
and keeps the fastest; the mat-vec after tuning equals the untuned plan's to the transforms' rounding (FP64: 1e-12
of its scale, FP32: 2e-6) and convex fits end at the same likelihood (1e-10 / 1e-6), whatever it picks.  A separable spread plan has no FFT passes:
its one configuration is timed."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize('dtype', [np.float64, np.float32], ids=['f64', 'f32'])
def test_autotune_keeps_the_operator(dtype):
    import torch
    import bench
    import blml_amd
    from blml_amd.binned import _FitSpec, FitBatch
    fits = bench.uneven_fits(48, seed=7)
    specs = [_FitSpec(c, bench.modes_of(g, blml_amd), 1.0, 0.0, kw['complex_g'], None, None, 0) for c, g, kw, _ in fits]
    params = dict(bench.fit_params(), max_sweeps=60)
    _note(f'autotune {np.dtype(dtype).name}: {len(specs)} fits')
    b = FitBatch(specs, params, dtype, hist_cap=0).build()
    g = torch.Generator(device='cuda').manual_seed(1)
    tdt = torch.float32 if dtype == np.float32 else torch.float64
    E = torch.randn(b.sizes['out'], device='cuda', dtype=tdt, generator=g)
    U0 = b.matvec(E).double().cpu().numpy()
    b.autotune(reps=2)
    t = b.tuned
    assert len(t['seconds']) == 4 and all(v > 0 for v in t['seconds'].values())
    U1 = b.matvec(E).double().cpu().numpy()
    scale = np.abs(U0).max()
    assert np.max(np.abs(U1 - U0)) < (1e-12 if dtype == np.float64 else 2e-6) * scale
    b.close()
    # convex fits (real rank 1: a unique optimum) on the same lattices end at the same likelihood under any choice
    cs = [_FitSpec(c, blml_amd.make_modes([0.3, 0.3], ['sinc', 'sinc']), 1.0, 0.0, False, None, None, 0)
          for c, *_ in fits]
    cp = dict(params, solver='cd', max_sweeps=None, tol=1e-8)
    a = FitBatch(cs, cp, dtype, hist_cap=0).build()
    ra = a.run().scalars().copy()
    a.autotune(reps=2)
    rb = a.run().scalars()
    tol = 1e-10 if dtype == np.float64 else 1e-6
    assert np.all(np.abs(rb['loglik'] - ra['loglik']) < tol * np.abs(ra['loglik']))
    a.close()


def test_autotune_separable_spread_plan():
    import torch
    import blml_amd
    from blml_amd.binned import _FitSpec, FitBatch
    rng = np.random.default_rng(2)
    counts = rng.poisson(3.0, (32, 32, 32)).astype(float)
    fs = _FitSpec(counts, blml_amd.make_modes([0.12] * 3, ['sinc'] * 3), 1.0, 0.0, False, None, None, 0)
    b = FitBatch([fs], dict(tol=1e-5, max_sweeps=5, anderson=5, stall=1000, phase_noise=0.1, ll_tol=None,
                            ll_window=100, solver='cd', lbfgs_m=20), np.float32, hist_cap=0, spread=1).build()
    assert b.info()['matvec_kernel'] in ('matvec_grid_kernel', 'matvec_pipe_kernel')
    b.autotune(reps=2)
    assert list(b.tuned['seconds']) == ['sched0_global'] and b.tuned['seconds']['sched0_global'] > 0
    b.close()
