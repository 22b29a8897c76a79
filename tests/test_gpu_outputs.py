"""BLMLBinned's outputs at points on the GPU (blml_point_rows: the units' closed forms, FP64 sums) against the
oracle's restatement of the reference (blml_binned.py:452-499) on the same coefficients: sqrt_density and predict
at random points inside and outside the lattice to 1e-10 of their scale, predict_grid(3) against the oracle's
shifted-table FFT grid to 1e-10 of its maximum, and equal to density_ at the bins (1e-9); every kernel kind."""
import sys

import numpy as np
import pytest

from oracle import blml_oracle as O

pytestmark = pytest.mark.gpu


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


CASES = [
    ('1d sinc complex', (40,), [dict(fc=[0.35], kinds=['sinc'])], True, 0.5),
    ('1d slits', (48,), [dict(fc=[0.4], kinds=['slits:0.5'])], False, 1.0),
    ('2d ball complex', (16, 14), [dict(fc=[0.4, 0.4], kinds=['ball', 'ball'])], True, 1.0),
    ('2d sinc x dirichlet rank 2', (14, 12), [dict(fc=[0.4, 0.3], kinds=['sinc', 'dirichlet'], periods=[None, 12.0],
                                                   rank=2)], False, 1.0),
    ('2d pball + bg complex', (15, 15), [dict(fc=[[0.45, 0.45], [0.2, 0.2]], kinds=['pball', 'pball'],
                                              periods=[15.0, 15.0])], True, 1.0),
    ('3d ball', (8, 9, 7), [dict(fc=[0.5] * 3, kinds=['ball'] * 3)], False, 1.0),
]


@pytest.mark.parametrize('name,shape,mm,cplx,h', CASES, ids=[c[0] for c in CASES])
def test_outputs_match_oracle(name, shape, mm, cplx, h):
    import blml_amd
    rng = np.random.default_rng(3)
    N = rng.poisson(5.0, shape).astype(float)
    modes = [md for kw in mm for md in blml_amd.make_modes(**kw)]
    omodes = [md for kw in mm for md in O.make_modes(**kw)]
    _note(f'outputs {name}')
    m = blml_amd.BLMLBinned(tol=1e-6, max_sweeps=400, solver='lbfgs').fit(N, modes, h=h, origin=-2.0,
                                                                         complex_g=cplx)
    o = O.BLMLBinnedOracle()
    o._setup(N, omodes, h, -2.0)
    o.e_hat_, o.n_, o.occ_, o.complex_g_ = m.e_hat_, m.n_, N > 0, cplx
    lo, hi = -2.0 - 2 * h, -2.0 + h * (np.array(shape) + 1)
    X = rng.uniform(lo, hi, (300, len(shape)))
    g, go = m.sqrt_density(X), o.sqrt_density(X)
    err = float(np.abs(g - go).max())
    _note(f'   sqrt_density max err {err:.2e} of {np.abs(go).max():.2e}')
    assert err < 1e-10 * np.abs(go).max()
    assert np.allclose(m.predict(X), o.predict(X), rtol=0, atol=1e-10 * o.predict(X).max())
    G, Go = m.predict_grid(3), o.predict_grid(3)
    gerr = float(np.abs(G - Go).max())
    _note(f'   predict_grid(3) {G.shape} max err {gerr:.2e} of {Go.max():.2e}')
    assert G.shape == Go.shape and gerr < 1e-10 * Go.max()
    at_bins = G[tuple(slice(0, None, 3) for _ in shape)]
    assert np.allclose(at_bins, m.density_, rtol=1e-9, atol=1e-12 * m.density_.max())
