"""The batched fc search (blml_amd.select.FcSearch: candidates scored a launch at a time, decisions replayed on the
scores) against the oracle's sequential restatement of the reference's search (oracle/blml_select.py, after
selector.py:490-553) on synthetic objectives of log fc: the same final fc and score, every point the sequential
search scored among the batched search's, and fewer batches than sequential calls.  Objectives: concave bumps with
plateaus and ripples (near-ties at the 1e-6 step threshold), a dirichlet column (harmonic-aware steps), a radial
group moving as one block, fixed and started columns."""
import numpy as np
import pytest

from oracle import blml_select as OS
from blml_amd import select as S


def _objective(rng, d, ripple):
    """A concave bump in log fc; with ripple, a flat top (exact ties, which the bracket's strict comparison keeps)
    over a quantised slope in steps of 5e-7 (gains below the refinement's 1e-6 threshold)."""
    c = rng.uniform(-2.5, -0.8, d)
    w = rng.uniform(0.3, 1.2, d)
    top = rng.uniform(0.3, 0.9)

    def f(fc):
        x = np.log(np.asarray(fc, float))
        r2 = float(np.sum(((x - c) / w) ** 2))
        if ripple:
            v = -max(r2 - top, 0.0) * 1e-5
            v = np.round(v / 5e-7) * 5e-7
        else:
            v = -r2
        return v if np.all(np.asarray(fc) <= 0.5) else -np.inf
    return f


CASES = [
    ('1 sinc column', ['sinc'], [None], {}),
    ('2 sinc columns', ['sinc', 'sinc'], [None, None], {}),
    ('sinc x dirichlet', ['sinc', 'dirichlet'], [None, 24.0], {}),
    ('ball group + sinc', ['ball', 'ball', 'sinc'], [None, None, None], {}),
    ('fixed and started columns', ['sinc', 'sinc', 'sinc'], [None] * 3,
     dict(start=np.array([np.nan, 0.1, np.nan]), free=np.array([True, True, False]))),
]


@pytest.mark.parametrize('name,kinds,periods,kw', CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize('ripple', [False, True])
def test_batched_search_makes_the_sequential_decisions(name, kinds, periods, kw, ripple):
    d = len(kinds)
    for seed in range(12):
        rng = np.random.default_rng(seed)
        f = _objective(rng, d, ripple)
        kw2 = dict(kw)
        if 'free' in kw2 and not kw2['free'][2]:
            kw2['start'] = np.array([np.nan, 0.1, 0.2])       # a fixed column needs a value
        args = dict(fc0=0.02, patience=2, max_sweeps=6, factors=(np.sqrt(2.0), 2 ** 0.25), **kw2)
        fc_o, best_o, visited = OS.search(f, d, kinds, periods, **args)
        calls = []
        s = S.FcSearch(d, kinds, periods, **args)

        def many(pts):
            calls.append(len(pts))
            return [f(p) for p in pts]
        fc_b, best_b = s.run(many)
        assert np.allclose(fc_b, fc_o, rtol=1e-12) and best_b == best_o, (name, seed, fc_b, fc_o)
        keys = set(s.scores)
        assert all(S._key(v) in keys for v in visited)
        assert len(calls) < len(visited) or len(visited) <= 2
