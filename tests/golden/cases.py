"""The golden fit cases: small complex and rank-K binned fits (seeded synthetic histograms, the shapes of the
reference's own parity cells, results/binned_parity.py:26-47, results/pball_parity.py, and of the 31x31 pball
window with a background mode, README.md:53-59).  Shared by make_golden.py (writes the vectors) and the tests
(which rebuild the inputs and compare)."""
import numpy as np


def _hist(rng, shape, n, f):
    g = np.meshgrid(*[np.arange(m, dtype=float) for m in shape], indexing='ij')
    p = f(*g); p = p / p.sum()
    return rng.multinomial(n, p.ravel()).reshape(shape).astype(float)


def pball_field_counts(shape, fc, n_total, seed, bg=0.05):
    """Counts of a complex harmonic field inside the pball band on a periodic window (the README form): coefficients
    CN(0,1) on the harmonics, density |g|^2 normalised, a uniform background share, Poisson counts."""
    rng = np.random.default_rng(seed)
    m1, m2 = shape
    R = fc / 2.0
    k1 = np.arange(-int(np.floor(R * m1 + 1e-9)), int(np.floor(R * m1 + 1e-9)) + 1)
    k2 = np.arange(-int(np.floor(R * m2 + 1e-9)), int(np.floor(R * m2 + 1e-9)) + 1)
    K1, K2 = np.meshgrid(k1, k2, indexing='ij')
    keep = (K1 / m1) ** 2 + (K2 / m2) ** 2 <= R * R * (1 + 1e-9) + 1e-12
    ks = np.stack([K1[keep], K2[keep]], 1)
    a = (rng.standard_normal(len(ks)) + 1j * rng.standard_normal(len(ks))) / np.sqrt(2)
    x1, x2 = np.meshgrid(np.arange(m1), np.arange(m2), indexing='ij')
    ph = 2 * np.pi * (ks[:, 0, None, None] * x1[None] / m1 + ks[:, 1, None, None] * x2[None] / m2)
    g = (a[:, None, None] * np.exp(1j * ph)).sum(0)
    p = np.abs(g) ** 2; p = p / p.sum()
    p = (1 - bg) * p + bg / p.size
    return rng.poisson(n_total * p).astype(float)


def golden_cases():
    """[(name, counts, [make_modes kwargs per group of modes], estimator kwargs, fit kwargs)]; the modes are the
    concatenation of make_modes(**kw) over the groups."""
    rng = np.random.default_rng(11)
    c1 = lambda x: np.exp(-0.5 * ((x - 14) / 3) ** 2) + 0.5 * np.exp(-0.5 * ((x - 27) / 2) ** 2) + 0.02
    cp = lambda x: np.exp(np.cos(2 * np.pi * x / 32)) + 0.3 * np.exp(-0.5 * ((x - 5) / 1.5) ** 2)
    c2 = lambda x, y: (np.exp(-0.5 * (((x - 11) / 2.5) ** 2 + ((y - 13) / 3) ** 2))
                       + 0.4 * np.exp(-0.5 * (((x - 17) / 2) ** 2 + ((y - 6) / 2) ** 2)) + 0.01)
    N1 = _hist(rng, (40,), 3000, c1)
    Np = _hist(rng, (32,), 3000, cp)
    N2 = _hist(rng, (24, 24), 8000, c2)
    N3 = _hist(rng, (20, 27), 6000, c2)
    out = [
        ('d1 sinc complex cd', N1, [dict(fc=[0.35], kinds=['sinc'])], dict(tol=1e-7), dict(complex_g=True)),
        ('d1 sinc complex lbfgs', N1, [dict(fc=[0.35], kinds=['sinc'])], dict(tol=1e-7, solver='lbfgs'),
         dict(complex_g=True)),
        ('d1 sinc rank2 shared real cd', N1, [dict(fc=[0.35], kinds=['sinc'], rank=2)], dict(tol=1e-7), {}),
        ('d1 sinc rank2 per-mode complex cd', N1, [dict(fc=[[0.35], [0.12]], kinds=['sinc'])], dict(tol=1e-7),
         dict(complex_g=True)),
        ('d1 dirichlet exact period complex cd', Np, [dict(fc=[0.3], kinds=['dirichlet'], periods=[32.0])],
         dict(tol=1e-7), dict(complex_g=True)),
        ('d2 sinc x dirichlet complex cd', N2, [dict(fc=[0.4, 0.3], kinds=['sinc', 'dirichlet'], periods=[None, 24.0])],
         dict(tol=1e-7), dict(complex_g=True)),
        ('d2 ball complex cd', N2, [dict(fc=[0.4, 0.4], kinds=['ball', 'ball'])], dict(tol=1e-7),
         dict(complex_g=True)),
        ('d2 slits x slits + dirichlet bg complex lbfgs', N3,
         [dict(fc=[0.35, 0.3], kinds=['slits:0.5', 'slits:0.5']),
          dict(fc=[3.0 / 20, 3.0 / 27], kinds=['dirichlet', 'dirichlet'], periods=[20.0, 27.0])],
         dict(tol=1e-7, solver='lbfgs', max_sweeps=1500), dict(complex_g=True)),
        ('d2 pball 31x31 + bg complex lbfgs', pball_field_counts((31, 31), 0.33, 2e4, 1),
         [dict(fc=[[0.33, 0.33], [3 / 31, 3 / 31]], kinds=['pball', 'pball'], periods=[31.0, 31.0])],
         dict(tol=1e-7, solver='lbfgs', max_sweeps=1500), dict(complex_g=True)),
        ('d2 pball 17x23 + bg complex cd', pball_field_counts((17, 23), 0.4, 5e3, 5),
         [dict(fc=[[0.4, 0.4], [0.15, 0.15]], kinds=['pball', 'pball'], periods=[17.0, 23.0])],
         dict(tol=1e-7, max_sweeps=3000), dict(complex_g=True)),
    ]
    return out
