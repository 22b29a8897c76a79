"""Pins the start-search oracle (oracle/blml_starts.py) without a GPU:
- the reference's own outputs (tests/golden/complex_starts.npz, written by make_golden_starts.py from the
  reference's complex_starts with its default settings): the same families in the same order, scores to 1e-9
  relative and the starts e0 to 1e-6 of their scale;
- error reduction is a projection method: the data error ||X - proj(X)|| never rises over its steps;
- the runs end band-limited, and each score is the Poisson log-likelihood of |g|^2 normalised on the box."""
import json
import os

import numpy as np
import pytest

from oracle import blml_starts as OS
from cases_starts import starts_cases

GOLD = os.path.join(os.path.dirname(__file__), 'golden', 'complex_starts.npz')


@pytest.mark.parametrize('i', range(len(starts_cases())), ids=[c[0] for c in starts_cases()])
def test_starts_match_reference_golden(i):
    z = np.load(GOLD, allow_pickle=False)
    meta = json.loads(str(z['meta']))[i]
    name, Z, w, h, fc, kinds = starts_cases()[i]
    assert meta['name'] == name
    out = OS.complex_starts(Z, w, h, fc, kinds, seed=0)
    assert [s[0] for s in out] == meta['family']
    sc = np.array([s[1] for s in out])
    ref = z[f'c{i}_score']
    assert np.allclose(sc, ref, rtol=1e-9, atol=0), (sc, ref)
    e0 = np.stack([s[2] for s in out])
    ge = z[f'c{i}_e0']
    assert np.abs(e0 - ge).max() < 1e-6 * np.abs(ge).max()


def test_error_reduction_never_raises_the_data_error():
    name, Z, w, h, fc, kinds = starts_cases()[1]
    G = OS.search_grid(Z, w, h, fc)
    U, band = OS.band_grid(G['shape'], G['hc'], fc, kinds)
    rng = np.random.default_rng(7)
    P0 = band * np.exp(2j * np.pi * rng.random(G['shape']))
    errs = []
    P = OS.run(P0, G, band, 0, 60, errors=errs)
    e = np.array(errs)
    assert np.all(np.diff(e) <= 1e-10 * e[:-1]), e
    assert e[-1] < e[0]
    assert np.all(P[~band] == 0)
    f = np.abs(np.fft.ifftn(P)) ** 2
    direct = sum(wb * np.log(f[tuple(zb)] / f[G['box']].sum()) for zb, wb in zip(Z - Z.min(0), w))
    assert OS.score(P, G) == pytest.approx(direct, rel=1e-12)
