"""Inputs of the start-search golden vectors (tests/golden/make_golden_starts.py): lattice bins, counts, h, fc and
kinds of small complex counts fits."""
import numpy as np


def _hist(shape, total, f, seed):
    rng = np.random.default_rng(seed)
    g = np.meshgrid(*[np.arange(m, dtype=float) for m in shape], indexing='ij')
    p = f(*g)
    N = rng.multinomial(total, (p / p.sum()).ravel()).astype(float)
    Z = np.stack([x.ravel() for x in np.meshgrid(*[np.arange(m) for m in shape], indexing='ij')], 1)
    keep = N > 0
    return Z[keep].astype(np.int64), N[keep]


def starts_cases():
    """[(name, Zi (B, d) int64, w (B,), h (d,), fc (d,), kinds)]"""
    c1 = lambda x: np.exp(-0.5 * ((x - 14) / 3) ** 2) + 0.5 * np.exp(-0.5 * ((x - 27) / 2) ** 2) + 0.02
    c2 = lambda x, y: (np.exp(-0.5 * (((x - 11) / 2.5) ** 2 + ((y - 13) / 3) ** 2))
                       + 0.4 * np.exp(-0.5 * (((x - 17) / 2) ** 2 + ((y - 6) / 2) ** 2)) + 0.01)
    c3 = lambda x: np.exp(-0.5 * ((x - 25) / 6) ** 2) + 0.7 * np.exp(-0.5 * ((x - 40) / 4) ** 2) + 0.01
    c4 = lambda x, y: np.exp(-0.5 * (((x - 7) / 2) ** 2 + ((y - 8) / 2.5) ** 2)) + 0.05
    out = []
    Z, w = _hist((40,), 3000, c1, 1)
    out.append(('1d sinc 40 bins', Z, w, np.array([1.0]), np.array([0.35]), ['sinc']))
    Z, w = _hist((24, 20), 8000, c2, 2)
    out.append(('2d sinc 24x20', Z, w, np.array([1.0, 1.0]), np.array([0.4, 0.3]), ['sinc', 'sinc']))
    Z, w = _hist((60,), 4000, c3, 3)
    out.append(('1d sinc coarse q=5', Z, w, np.array([1.0]), np.array([0.1]), ['sinc']))
    Z, w = _hist((16, 16), 5000, c4, 4)
    out.append(('2d ball 16x16', Z, w, np.array([1.0, 1.0]), np.array([0.45, 0.45]), ['ball:a', 'ball:a']))
    return out
