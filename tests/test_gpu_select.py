"""The fc selection and sweep on the GPU (blml_amd.select, blml_amd.frames):
- select_fc (every round of candidates one batched launch of the FP64 engine) picks the same fc with the same
  held-out score (1e-9 relative) as the oracle's sequential search (oracle/blml_select.py) over the oracle's fits of
  the same split, on a 1-D sinc lattice and a 2-D sinc x dirichlet lattice (harmonic-aware steps), in fewer
  launches than the sequential search's fits;
- a bead frame (a synthetic 2-D .npy) through load_frame / split_frame and an fc sweep of a complex ball kernel
  over six multiples of the band in one FP32 launch: every fit converged, each equal to the same fit run alone, and
  the held-out score peaks inside the sweep."""
import sys

import numpy as np
import pytest

from oracle import blml_oracle as O
from oracle import blml_select as OS

pytestmark = pytest.mark.gpu


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


CASES = [
    ('1d sinc 256', (256,), ['sinc'], [None], 'refine'),
    ('2d sinc x dirichlet 40x48', (40, 48), ['sinc', 'dirichlet'], [None, 48.0], 'refine'),
]


@pytest.mark.parametrize('name,shape,kinds,periods,search', CASES, ids=[c[0] for c in CASES])
def test_select_matches_sequential_search(name, shape, kinds, periods, search):
    import blml_amd
    from blml_amd import select as S
    rng = np.random.default_rng(17)
    g = np.meshgrid(*[np.arange(m, dtype=float) for m in shape], indexing='ij')
    p = sum(np.exp(-0.5 * sum(((x - c * m) / (0.07 * m)) ** 2 for x, m in zip(g, shape))) for c in (0.3, 0.62)) + 0.05
    counts = rng.multinomial(20000, (p / p.sum()).ravel()).reshape(shape).astype(float)
    kw = dict(tol=1e-9, max_sweeps=20000)
    _note(f'select {name}')
    fc, best, info = blml_amd.select_fc(counts, kinds, periods, search=search, dtype=np.float64, **kw)
    tr, va = S.split_counts(counts, 0)
    d = len(shape)
    per_s = [None if q is None else float(q) for q in periods]

    def obj(v):
        if np.any(np.asarray(v) > 0.5 + 1e-12):
            return -np.inf
        m = O.BLMLBinnedOracle(**kw).fit(tr, O.make_modes(list(v), kinds, periods))
        return S.val_score(m.density_, va)
    fc_o, best_o, visited = OS.search(obj, d, kinds, per_s, fc0=2.0 / max(shape), patience=2,
                                      max_sweeps=0 if search == 'shared' else 6,
                                      factors=(np.sqrt(2.0), 2.0 ** 0.25, 2.0 ** 0.125))
    _note(f'   engine fc {np.round(fc, 5)} score {best:.10f} ({info["evaluations"]} fits in {info["launches"]} '
          f'launches); oracle fc {np.round(fc_o, 5)} score {best_o:.10f} ({len(visited)} fits one by one)')
    assert np.allclose(fc, fc_o, rtol=1e-12)
    assert abs(best - best_o) < 1e-9 * abs(best_o)
    assert info['launches'] < len(visited)


def test_frame_fc_sweep(tmp_path):
    import blml_amd
    from blml_amd import frames as FR
    rng = np.random.default_rng(5)
    shape = (96, 96)
    y, x = np.meshgrid(np.arange(shape[0]), np.arange(shape[1]), indexing='ij')
    lam = 0.5 + sum(80 * np.exp(-0.5 * ((y - a) ** 2 + (x - b) ** 2) / 1.6 ** 2)
                    for a, b in [(20, 30), (60, 25), (45, 70), (80, 80)])
    path = tmp_path / 'frame.npy'
    np.save(path, rng.poisson(lam).astype(np.uint16))
    N = FR.load_frame(path)
    tr, va, te = FR.split_frame(N, seed=0)
    fc0 = 0.2
    mults = [0.5, 0.7, 1.0, 1.4, 2.0, 2.4]
    fcs = [[k * fc0, k * fc0] for k in mults]
    kw = dict(tol=1e-4, max_sweeps=4000, ll_tol=1e-9, solver='lbfgs')
    _note(f'frame fc sweep: {shape} frame, {int(N.sum())} counts, {len(fcs)} fc in one launch')
    rows = blml_amd.fc_sweep(tr, va, te, ['ball', 'ball'], fcs, complex_g=True, dtype=np.float32, **kw)
    for r in rows:
        _note(f'   fc {r["fc"][0]:.3f}: train {r["train"]:.5f} val {r["val"]:.5f} test {r["test"]:.5f} '
              f'({r["sweeps"]} sweeps, {r["stop"]})')
    assert all(r['stop'] in ('tol', 'll_tol') for r in rows)
    one = blml_amd.BLMLBinned(dtype=np.float32, **kw).fit(tr, blml_amd.make_modes(fcs[2], ['ball', 'ball']),
                                                          complex_g=True)
    assert one.loglik_ == rows[2]['loglik']
    val = [r['val'] for r in rows]
    assert 0 < int(np.argmax(val)) < len(val) - 1
