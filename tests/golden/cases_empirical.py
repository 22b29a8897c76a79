"""The golden BLMLEmpirical cases: fits whose result depends on the alternation's trajectory (the hidden-phase
retrieve path, blml_binned.py:1229-1289: non-convex, started from the rank-1 fit and the seeded phase draws).  The
linear fits (reg ridge / l2 / l1 without retrieve) are pinned by dense linear algebra instead and are not here.
Shared by make_golden_empirical.py (writes the vectors) and the tests (which rebuild the inputs and compare)."""
import numpy as np

from cases import _hist


def empirical_cases():
    """[(name, counts, [make_modes kwargs per group], estimator kwargs, fit kwargs)]."""
    rng = np.random.default_rng(21)
    c1 = lambda x: np.exp(-0.5 * ((x - 14) / 3) ** 2) + 0.5 * np.exp(-0.5 * ((x - 27) / 2) ** 2) + 0.02
    c2 = lambda x, y: (np.exp(-0.5 * (((x - 9) / 2.5) ** 2 + ((y - 10) / 3) ** 2))
                       + 0.4 * np.exp(-0.5 * (((x - 14) / 2) ** 2 + ((y - 5) / 2) ** 2)) + 0.01)
    N1 = _hist(rng, (48,), 4000, c1)
    N2 = _hist(rng, (20, 18), 6000, c2)
    kw = dict(tol=1e-10, outer_tol=1e-5, max_outer=400)
    return [
        ('d1 sinc complex ridge', N1, [dict(fc=[0.3], kinds=['sinc'])], dict(kw), dict(complex_g=True)),
        ('d1 sinc complex l2', N1, [dict(fc=[0.3], kinds=['sinc'])], dict(kw, reg='l2', lam=1e-3),
         dict(complex_g=True)),
        ('d1 sinc real sign retrieve ridge', N1, [dict(fc=[0.3], kinds=['sinc'])], dict(kw, retrieve=True), {}),
        ('d1 sinc rank2 real ridge plane', N1, [dict(fc=[0.3], kinds=['sinc'], rank=2)],
         dict(kw, normalise='plane'), {}),
        ('d1 sinc complex l1', N1, [dict(fc=[0.3], kinds=['sinc'])],
         dict(kw, reg='l1', lam=1e-2, tol=1e-8, max_iter=4000), dict(complex_g=True)),
        ('d2 sinc x sinc complex anscombe', N2, [dict(fc=[0.4, 0.35], kinds=['sinc', 'sinc'])],
         dict(kw, anscombe=True), dict(complex_g=True)),
        ('d2 pball complex ridge', N2, [dict(fc=[0.45, 0.45], kinds=['pball', 'pball'], periods=[20.0, 18.0])],
         dict(kw), dict(complex_g=True)),
    ]
