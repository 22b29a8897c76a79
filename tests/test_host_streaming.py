"""Host pieces of the streaming learner (blml_amd/streaming.py), without a GPU:
- adam_step: the first step from zero moments is lr in log value in the gradient's direction (bias-corrected Adam);
- _row_update: the real root zeroes its quadratic; a shared-kernel rank-K row is row-stationary
  (R + s e = c e / |e|^2); per-mode rows solve the secular equation (lam sum |R|^2 / (lam - s)^2 = c);
- _batch_order keeps groups together and is a permutation; modes_fc / _modes_with_fc round-trip a template."""
import numpy as np

from blml_amd import streaming as S


def test_adam_first_step_is_lr_in_log():
    for g in (0.37, -2.5):
        st = [0.0, 0.0, 0]
        v = S.adam_step(0.2, g, st, 0.1, (0.9, 0.999))
        assert abs(np.log(v / 0.2) - 0.1 * np.sign(g)) < 1e-9
        assert st[2] == 1


def test_row_update_stationarity():
    rng = np.random.default_rng(0)
    B = 400
    c = rng.uniform(0.5, 60.0, B)
    occ = np.ones(B, bool)
    # real rank 1: s e^2 + b e - c = 0 with b = U - s e
    E = rng.uniform(0.1, 2.0, (1, B))
    U = rng.normal(size=(1, B)) * 5
    s = np.array([0.8])
    t = S._row_update(E, U, s, c, occ, False)[0]
    b = (U - s[:, None] * E)[0]
    assert np.all(t > 0) and np.abs(0.8 * t * t + b * t - c).max() < 1e-9 * np.max(c + np.abs(b * t))
    # shared kernel, complex rank 3
    K = 3
    E = rng.normal(size=(K, B)) + 1j * rng.normal(size=(K, B))
    U = rng.normal(size=(K, B)) + 1j * rng.normal(size=(K, B))
    s = np.full(K, 0.6)
    e = S._row_update(E, U, s, c, occ, True)
    R = U - s[:, None] * E
    lam = c / (np.abs(e) ** 2).sum(0)
    assert np.abs(R + s[:, None] * e - lam[None] * e).max() < 1e-9 * np.abs(R).max()
    # per-mode kernels: the secular root
    s = np.array([0.6, 0.25, 0.4])
    e = S._row_update(E, U, s, c, occ, True)
    R = U - s[:, None] * E
    lam = c / (np.abs(e) ** 2).sum(0)
    assert np.abs(R + s[:, None] * e - lam[None] * e).max() < 1e-8 * np.abs(R).max()
    assert np.abs(lam * ((np.abs(R) ** 2) / (lam[None] - s[:, None]) ** 2).sum(0) / c - 1).max() < 1e-10


def test_batch_order_and_modes():
    rng = np.random.default_rng(1)
    groups = ['a', 'b', 'a', 'c', 'b', 'c', 'c']
    order = S._batch_order(len(groups), groups, rng)
    assert sorted(order) == list(range(len(groups)))
    seq = [groups[i] for i in order]
    assert all(seq[i] == seq[i + 1] or seq[i + 1] not in seq[:i + 1] for i in range(len(seq) - 1))
    modes = [{0: dict(fc=np.nan), 1: dict(kind='dirichlet', fc={'value': 0.3, 'trainable': False}, period=16.0)}]
    F = S.modes_fc(modes, 2)
    assert np.isnan(F[0, 0]) and F[0, 1] == 0.3
    assert S.modes_trainable(modes, 2).tolist() == [[True, False]]
    m2 = S._modes_with_fc(modes, np.array([[0.25, 0.3]]))
    assert m2[0][0]['fc'] == 0.25 and m2[0][1]['fc'] == {'value': 0.3, 'trainable': False}
