"""Pins the oracle's kernel math and mat-vec with closed forms and brute force (no reference import):
harmonic sums (results/pball_parity.py:28-39, results/pslits_parity.py:24-35), the dense Gram
(blml_kernels.py:243-251 style), the 5-smooth pad rule (blml_binned.py:130-135, 1533-1546)."""
import numpy as np
import pytest

from oracle import blml_oracle as O


def test_pball_is_harmonic_sum():
    P, fc = 31.0, 0.33
    Kv, sc = O.pball_harmonics(fc, [P, P])
    assert len(Kv) == 89                              # the disk of the 31x31 window at fc 0.33
    t = np.random.default_rng(0).uniform(-40, 40, (2, 500))
    ref = sc * np.real(np.exp(2j * np.pi * (Kv @ t) / P).sum(0))
    got = O.kernel_eval('pball:0-1', [fc, fc], [P, P], [t[0], t[1]])
    assert np.abs(got - ref).max() < 1e-12
    Kb, _ = O.pball_harmonics(3 / 31, [P, P])
    assert len(Kb) == 9                               # the background band of the same window


def test_pball_one_column_is_dirichlet():
    x = np.random.default_rng(1).uniform(-500, 500, 2000)
    a = O.kernel_eval('pball:0', [0.0214], [392.0], [x])
    b = O.kernel_eval('dirichlet', [0.0214], [392.0], [x])
    assert np.abs(a - b).max() < 1e-12


def test_dirichlet_fill_is_harmonic_sum():
    P, fc, w = 392.0, 0.0214, 0.47
    t = np.random.default_rng(0).uniform(-600, 600, 4000)
    N = int(np.floor(fc * P / 2 + 1e-9)); lo = int(np.ceil((1 - w) * fc * P / 2 - 1e-9))
    ks = np.array([k for k in range(-N, N + 1) if abs(k) >= lo])
    ref = np.real(np.exp(2j * np.pi * np.outer(t, ks) / P).sum(1)) / P
    assert np.abs(O.kernel_eval(f'dirichlet:{w!r}', [fc], [P], [t]) - ref).max() < 1e-10
    one = O.kernel_eval('dirichlet:1.0', [fc], [P], [t]) - O.kernel_eval('dirichlet', [fc], [P], [t])
    assert np.abs(one).max() < 1e-12


def test_ball_in_one_dim_is_sinc_and_slits_band():
    t = np.linspace(-30, 30, 1001)
    assert np.abs(O.kernel_eval('ball:0', [0.37], [None], [t]) - 0.37 * np.sinc(0.37 * t)).max() < 1e-13
    # slits: the spectrum is the band (1 - fill) fc / 2 <= |nu| <= fc / 2; K(0) = fill fc
    assert abs(O.kernel_eval('slits:0.3', [0.5], [None], [np.zeros(1)])[0] - 0.15) < 1e-15
    # a band-limited kernel reproduces: int K(t) e^{-2 pi i nu t} dt = 1 inside the band, 0 outside
    tt = np.linspace(-4000, 4000, 800001); dt = tt[1] - tt[0]
    K = O.kernel_eval('slits:0.5', [0.4], [None], [tt])
    for nu, want in ((0.05, 0.0), (0.15, 1.0), (0.3, 0.0)):
        got = float(np.sum(K * np.cos(2 * np.pi * nu * tt)) * dt)
        assert abs(got - want) < 2e-3


def test_ball_2d_profile_value_at_zero():
    # the ball |nu| <= 1/2 in 2-D has area pi / 4
    assert abs(O.kernel_eval('ball:0-1', [1.0, 1.0], [None, None], [np.zeros(1), np.zeros(1)])[0] - np.pi / 4) < 1e-14


def _smooth5(n):
    best = None
    for a in range(0, 20):
        for b in range(0, 13):
            for c in range(0, 9):
                v = 2 ** a * 3 ** b * 5 ** c
                if v >= n and (best is None or v < best):
                    best = v
    return best


@pytest.mark.parametrize('n', [1, 7, 31, 59, 63, 127, 255, 8191, 20000])
def test_next_fast_len_is_5_smooth(n):
    assert O.next_fast_len5(n) == _smooth5(n)


CASES = [
    ('1d sinc', (40,), [dict(fc=[0.35], kinds=['sinc'])], 1.0),
    ('1d dirichlet circular', (32,), [dict(fc=[0.3], kinds=['dirichlet'], periods=[32.0])], 1.0),
    ('1d dirichlet fill circular', (32,), [dict(fc=[0.6], kinds=['dirichlet:0.5'], periods=[32.0])], 1.0),
    ('1d dirichlet long period', (32,), [dict(fc=[0.3], kinds=['dirichlet'], periods=[50.0])], 1.0),
    ('2d sinc x dirichlet', (12, 16), [dict(fc=[0.4, 0.3], kinds=['sinc', 'dirichlet'], periods=[None, 16.0])], 1.0),
    ('2d ball h=0.5', (11, 13), [dict(fc=[0.8, 0.8], kinds=['ball', 'ball'])], 0.5),
    ('2d pball circular', (17, 23), [dict(fc=[0.4, 0.4], kinds=['pball', 'pball'], periods=[17.0, 23.0])], 1.0),
    ('2d slits + dirichlet bg', (14, 19), [dict(fc=[0.35, 0.3], kinds=['slits:0.5', 'slits:0.4']),
                                           dict(fc=[0.15, 0.15], kinds=['dirichlet', 'dirichlet'],
                                                periods=[14.0, 19.0])], 1.0),
    ('3d sinc', (6, 7, 5), [dict(fc=[0.5, 0.4, 0.3], kinds=['sinc'] * 3)], 1.0),
]


@pytest.mark.parametrize('name,shape,mm,h', CASES, ids=[c[0] for c in CASES])
def test_matvec_equals_dense_gram(name, shape, mm, h):
    modes = [md for kw in mm for md in O.make_modes(**kw)]
    rng = np.random.default_rng(3)
    counts = rng.poisson(5.0, shape).astype(float)
    m = O.BLMLBinnedOracle()
    m._setup(counts, modes, h, 0.0)
    for k, sp in enumerate(m.specs_):
        G = O.dense_gram(sp, shape, m.h_)
        e = rng.standard_normal(shape)
        got = m._S(k, e).ravel()
        assert np.abs(got - G @ e.ravel()).max() < 1e-12 * max(1.0, np.abs(G).sum(1).max())
        assert abs(m.kern_[k].s - G[0, 0]) < 1e-14


def test_pad_rule():
    m = O.BLMLBinnedOracle()
    c = np.ones((31, 31))
    m._setup(c, O.make_modes([[0.33, 0.33], [3 / 31, 3 / 31]], ['pball', 'pball'], periods=[31.0, 31.0]), 1.0, 0)
    assert m.pad_ == (31, 31) and not m.shared_
    m._setup(c, O.make_modes([0.33, 0.33], ['pball', 'pball'], periods=[40.0, 31.0]), 1.0, 0)
    assert m.pad_ == (64, 31)
    m._setup(np.ones(4096), O.make_modes([0.05], ['sinc']), 1.0, 0)
    assert m.pad_ == (8192,)
    m._setup(np.ones((32,) * 3), O.make_modes([0.12] * 3, ['sinc'] * 3), 0.25, 0)
    assert m.pad_ == (64, 64, 64)
