"""Inputs of the binning golden vectors (tests/golden/make_golden_binning.py): points and modes of counts fits on
the fit's own lattice.  make_modes: the reference's or this package's (the same dictionaries)."""
import numpy as np


def binning_cases(make_modes):
    """[(name, X (n, d), modes, bin_scale)]"""
    rng = np.random.default_rng(21)
    out = []
    X = np.concatenate([rng.normal(-1.0, 0.6, (3000, 1)), rng.normal(1.5, 0.4, (2000, 1))])
    out.append(('1d sinc', X, make_modes([0.8], ['sinc']), 0.1))
    X = np.stack([rng.normal(0, 1.0, 4000), rng.uniform(0, 10, 4000)], 1)
    out.append(('2d sinc x dirichlet P=10', X, make_modes([0.7, 0.6], ['sinc', 'dirichlet'], periods=[None, 10.0]),
                0.1))
    X = rng.normal(0, 1.0, (5000, 2))
    out.append(('2d ball', X, make_modes([0.9, 0.9], ['ball', 'ball']), 0.2))
    X = np.stack([rng.normal(0, 1.0, 3000), rng.normal(0, 2.0, 3000)], 1)
    out.append(('2d two modes', X, make_modes([[0.5, 0.4], [1.1, 0.3]], ['sinc', 'slits:0.5']), 0.1))
    X = np.stack([rng.uniform(0, 12, 3000), rng.uniform(0, 9, 3000)], 1)
    out.append(('2d periodic ball', X, make_modes([0.6, 0.6], ['pball', 'pball'], periods=[12.0, 9.0]), 0.1))
    return out
