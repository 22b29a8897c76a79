"""The frame loader and thinning (blml_amd.frames): the split conserves every count and is a binomial thinning
(each set's total within 6 standard deviations of its binomial mean, the validation share of the rest likewise),
the same seed gives the same split and another seed another; the crop is the window around the centre; the loader
refuses object (pickled) arrays, negative and non-finite counts."""
import numpy as np
import pytest

from blml_amd import frames as FR


def _frame(rng, shape=(300, 320)):
    y, x = np.meshgrid(np.arange(shape[0]), np.arange(shape[1]), indexing='ij')
    lam = 2.0 + sum(400 * np.exp(-0.5 * ((y - a) ** 2 + (x - b) ** 2) / 2.0 ** 2)
                    for a, b in [(50, 60), (150, 200), (230, 90)])
    return rng.poisson(lam)


def test_split_is_a_conserving_binomial_thinning():
    N = _frame(np.random.default_rng(0))
    tr, va, te = FR.split_frame(N, seed=0)
    assert np.array_equal(tr + va + te, N) and tr.min() >= 0 and va.min() >= 0 and te.min() >= 0
    n = N.sum()
    assert abs(tr.sum() - 0.7 * n) < 6 * np.sqrt(n * 0.7 * 0.3)
    rest = n - tr.sum()
    assert abs(va.sum() - rest / 3) < 6 * np.sqrt(rest * (1 / 3) * (2 / 3))
    again = FR.split_frame(N, seed=0)
    other = FR.split_frame(N, seed=1)
    assert all(np.array_equal(a, b) for a, b in zip((tr, va, te), again))
    assert not np.array_equal(tr, other[0])
    with pytest.raises(ValueError):
        FR.split_frame(N + 0.5)


def test_loader_crop_and_refusals(tmp_path):
    N = _frame(np.random.default_rng(1), (900, 1300))
    p = tmp_path / 'frame.npy'
    np.save(p, N.astype(np.uint16))
    full = FR.load_frame(p)
    assert full.dtype == np.int64 and np.array_equal(full, N)
    c = FR.load_frame(p, crop=257)
    assert c.shape == (257, 257) and np.array_equal(c, N[661 - 128:661 + 129, 1068 - 128:1068 + 129])
    bad = tmp_path / 'obj.npy'
    np.save(bad, np.array([{'a': 1}], dtype=object), allow_pickle=True)
    with pytest.raises(ValueError):
        FR.load_frame(bad)
    neg = tmp_path / 'neg.npy'
    np.save(neg, -np.ones((4, 4)))
    with pytest.raises(ValueError):
        FR.load_frame(neg)
    with pytest.raises(ValueError):
        FR.load_frame(p, crop=2001)
    P = np.full(N.shape, 1.0 / N.size)
    assert abs(FR.score(P, N) - np.log(1.0 / N.size)) < 1e-12
