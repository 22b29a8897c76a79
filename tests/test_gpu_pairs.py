"""The HIP pair kernels (blml_pair_rows / blml_pair_cols through blml_amd.local) against the oracle's restatement of
the reference's rows / cols (oracle/blml_pairs.py), on the same seeded inputs: FP64 tables to 1e-12 of the rows'
absolute scale (sum_j |S_rj w_j|), FP32 tables to 1e-5 of it (FP32 products, FP64 sums); cols elementwise 1e-13
(FP64) / 1e-6 (FP32) of max |S|."""
import sys

import numpy as np
import pytest

from oracle import blml_pairs as OP
from test_oracle_pairs import CASES, _bins

pytestmark = pytest.mark.gpu


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _radial_tab(rng):
    LR = 80
    return dict(T=rng.standard_normal(30), base=np.array([0, 0, 0]), Jk=np.array([0, 0, 14]), mk=np.zeros(3, int),
                vals=[None, None, None], kinds=np.array([2, 2, 0]), nbins=np.zeros(3, int), grp=np.array([0, 0, -1]),
                R=rng.standard_normal(LR)[None], dim_unit=np.array([-1, -1, 1]), grp_unit=np.array([0]), U=2, nW=0)


def _all_cases(rng):
    out = []
    for name, shape, units, h, J in CASES:
        out.append((name, shape, OP.build_tables(units, np.asarray(h, float), J), 60, 90))
    Z = _bins(rng, (40, 6), 25)
    out.append(('rank x offset', (40, 6), OP.build_tables(
        [('rank', 'sinc', [0.2], [0.0], [0]), ('offset', 'sinc', [0.5], [0.0], [1])], np.ones(2), (40, 6),
        vals=[np.unique(Z[:, 0]), None]), None, None))
    out.append(('radial random rows', (8, 8, 8), _radial_tab(rng), 40, 50))
    from blml_amd import local
    Zb = _bins(rng, (20, 18, 9), 400)
    out.append(('ball x sinc (radial group from lattice_tables)', (20, 18, 9),
                local.lattice_tables(['ball', 'ball', 'sinc'], [0.3, 0.3, 0.4], None, [1.0, 1.0, 1.0], Zb), 70, 120))
    Z4 = _bins(rng, (9, 8, 10, 7), 400)
    out.append(('two ball groups (lattice_tables)', (9, 8, 10, 7),
                local.lattice_tables(['ball:a', 'ball:a', 'ball:b', 'ball:b'], [0.3, 0.3, 0.35, 0.35], None,
                                     [1.0] * 4, Z4), 90, 150))
    # enough rows and sources for several slices, chunk tails and a table too big for LDS (float64, > 32 KB)
    out.append(('3d slits large', (64, 64, 64), OP.build_tables(
        [('offset', 'slits:0.5', [0.2], [0.0], [k]) for k in range(3)], np.ones(3), (1500, 1500, 1500)), 3000,
        20000))
    return out, Z


@pytest.mark.parametrize('f32', [False, True], ids=['f64', 'f32'])
def test_rows_and_cols_match_oracle(f32):
    from blml_amd import local
    rng = np.random.default_rng(21)
    cases, Zrank = _all_cases(rng)
    for name, shape, tab, nA, nB in cases:
        _note(f'pairs {name} f32={f32}')
        if nA is None:
            ZA = ZB = Zrank
        else:
            ZA, ZB = _bins(rng, shape, nA), _bins(rng, shape, nB)
        w = rng.standard_normal(len(ZB))
        w[::5] = 0.0
        want = OP.pair_rows(ZA, ZB, w, tab)
        out_o, _ = OP.pair_cols(ZA[:50], ZB[:70], tab)
        scale = (np.abs(out_o).max() if out_o.size else 1.0)
        s, ds = local.rows(ZA, ZB, w, tab, f32=f32)
        absrow = np.abs(OP.pair_cols(ZA[:20], ZB, tab)[0]) @ np.abs(w)
        tol = (1e-5 if f32 else 1e-12) * max(float(np.max(absrow, initial=0.0)), 1e-300)
        assert s.shape == (len(ZA),) and ds.shape[0] == len(ZA)
        assert np.abs(s - want).max() <= tol * 10, (name, np.abs(s - want).max(), tol)
        out = local.cols(ZA[:50], ZB[:70], tab, f32=f32)
        assert np.abs(out - out_o).max() <= (1e-6 if f32 else 1e-13) * scale, name


def test_cols_marginal_and_errors():
    from blml_amd import local
    rng = np.random.default_rng(22)
    name, shape, units, h, J = CASES[4]
    tab = OP.build_tables(units, np.asarray(h, float), J)
    ZA, ZB = _bins(rng, shape, 40), _bins(rng, shape, 50)
    out, marg = local.cols(ZA, ZB, tab, depmask=0b010)
    o_out, o_marg = OP.pair_cols(ZA, ZB, tab, depmask=0b010)
    assert np.abs(out - o_out).max() < 1e-13 and np.abs(marg - o_marg).max() < 1e-13
    bad = ZA.copy()
    bad[0, 0] = 100                        # beyond the offset row's half width: the tables predate this bin
    with pytest.raises(RuntimeError):
        local.rows(bad, ZB, np.ones(len(ZB)), tab)
    s, ds = local.rows(ZA, ZB, np.ones(len(ZB)), tab, grad=True)      # no derivative entries: ds zeros
    assert np.array_equal(s, local.rows(ZA, ZB, np.ones(len(ZB)), tab)[0]) and ds.shape == (len(ZA), 1) and not ds.any()


@pytest.mark.parametrize('f32', [False, True], ids=['f64', 'f32'])
def test_derivative_rows_match_oracle(f32):
    """rows(grad=True) (blml_pair_rows_grad) against the oracle's pair_rows_grad on the same tables: column entries
    (raw-offset rows, a rank table, a dirichlet period row; the full 16 entries), a ball group's fc and shift rows,
    a periodic ball's torus shift rows, and a large case with many slices and a table beyond LDS.  s to the value
    rows' bar; ds[:, q] to 1e-11 (FP64) / 1e-4 (FP32) of max |ds[:, q]|."""
    from blml_amd import local
    from test_oracle_pairs import _column_case
    rng = np.random.default_rng(41)
    Z, kinds, fc, periods, want, _, _ = _column_case(rng)
    cases = [('columns', local.lattice_tables(kinds, fc, periods, np.ones(4), Z, want=want), Z, Z),
             ('columns, 16 entries', local.lattice_tables(kinds, fc, periods, np.ones(4), Z, want=want + want[::-1]),
              Z, Z)]
    Zb = _bins(rng, (20, 18, 9), 300)
    cases.append(('ball x sinc', local.lattice_tables(
        ['ball', 'ball', 'sinc'], [0.3, 0.3, 0.4], None, np.ones(3), Zb,
        want=[(1, 'fc'), (1, ('shift', 0)), (1, ('shift', 1)), (0, 'fc'), (0, 'shift')]), Zb[:120], Zb))
    Zp = _bins(rng, (10, 8), 60)
    cases.append(('pball', local.lattice_tables(['pball', 'pball'], [0.45, 0.45], [10.0, 8.0], [1.0, 1.0], Zp,
                                                want=[(0, 'shift'), (0, ('shift', 1))]), Zp, Zp))
    Z4 = _bins(rng, (9, 8, 10, 7), 300)
    cases.append(('two ball groups', local.lattice_tables(
        ['ball:a', 'ball:a', 'ball:b', 'ball:b'], [0.3, 0.3, 0.35, 0.35], None, [1.0] * 4, Z4,
        want=[(0, 'fc'), (1, ('shift', 1)), (1, 'fc')]), Z4[:100], Z4))
    ZL = np.concatenate([_bins(rng, (1500, 1500, 1500), 12000), np.array([[0, 0, 0], [1499, 1499, 1499]])])
    cases.append(('3d slits large', local.lattice_tables(
        ['slits:0.5'] * 3, [0.2] * 3, None, np.ones(3), ZL,
        want=[(k, p) for k in range(3) for p in ('fc', 'shift')]), ZL[:1000], ZL))
    cases.append(('3d slits large, 2 entries', local.lattice_tables(
        ['slits:0.5'] * 3, [0.2] * 3, None, np.ones(3), ZL, want=[(0, 'fc'), (2, 'shift')]), ZL[:500], ZL))
    for name, tab, ZA, ZB in cases:
        _note(f'derivative rows {name} f32={f32} nA={len(ZA)} nB={len(ZB)} nW={tab["nW"]}')
        w = rng.standard_normal(len(ZB))
        w[::5] = 0.0
        s, ds = local.rows(ZA, ZB, w, tab, grad=True, f32=f32)
        s0, ds0 = OP.pair_rows_grad(ZA, ZB, w, tab)
        scale = np.abs(OP.pair_rows(ZA, ZB, np.abs(w), dict(tab, T=np.abs(tab['T']), R=np.abs(tab['R'])))).max()
        assert np.abs(s - s0).max() <= (1e-5 if f32 else 1e-12) * scale, name
        assert ds.shape == ds0.shape
        for q in range(int(tab['nW'])):
            ref = np.abs(ds0[:, q]).max()
            assert np.abs(ds[:, q] - ds0[:, q]).max() <= (1e-4 if f32 else 1e-11) * ref, (name, q)
    big = local.lattice_tables(kinds, fc, periods, np.ones(4), Z, want=want * 3)
    with pytest.raises(ValueError, match='max 16'):
        local.rows(Z, Z, np.ones(len(Z)), big, grad=True)
    s, ds = local.rows(Z, Z, np.ones(len(Z)), local.lattice_tables(kinds, fc, periods, np.ones(4), Z), grad=True)
    assert ds.shape == (len(Z), 1) and not ds.any()


@pytest.mark.parametrize('f32', [False, True], ids=['f64', 'f32'])
def test_fused_densities_match_oracle(f32):
    """rows_multi (blml_pair_rows_multi) against the oracle's pair_rows_multi: the joint and two marginals with five
    derivative entries (15 accumulators), four densities without derivatives over a ball group and a sinc column,
    and the many-slice case; bars as the derivative rows."""
    from blml_amd import local
    from test_oracle_pairs import _column_case
    rng = np.random.default_rng(43)
    Z, kinds, fc, periods, want, _, _ = _column_case(rng)
    Zb = _bins(rng, (20, 18, 9), 300)
    ZL = np.concatenate([_bins(rng, (1500, 1500, 1500), 9000), np.array([[0, 0, 0], [1499, 1499, 1499]])])
    cases = [('columns, 3 densities x 5 entries', local.lattice_tables(kinds, fc, periods, np.ones(4), Z, want=want[:5]),
              Z, Z, np.array([[1, 1, 1, 1], [1, 0, 1, 0], [0, 1, 0, 1]]), True),
             ('ball x sinc, 4 densities', local.lattice_tables(['ball', 'ball', 'sinc'], [0.3, 0.3, 0.4], None,
                                                              np.ones(3), Zb), Zb[:150], Zb,
              np.array([[1, 1], [1, 0], [0, 1], [0, 0]]), False),
             ('3d slits large, 2 densities x 6 entries', local.lattice_tables(
                 ['slits:0.5'] * 3, [0.2] * 3, None, np.ones(3), ZL,
                 want=[(k, p) for k in range(3) for p in ('fc', 'shift')]), ZL[:800], ZL,
              np.array([[1, 1, 1], [1, 0, 1]]), True)]
    for name, tab, ZA, ZB, umask, grad in cases:
        _note(f'fused rows {name} f32={f32} nA={len(ZA)} nB={len(ZB)}')
        Wt = rng.standard_normal((len(ZB), umask.shape[0]))
        Wt[::5] = 0.0
        Wt[1::4, 0] = 0.0
        s, ds = local.rows_multi(ZA, ZB, Wt, tab, umask, grad=grad, f32=f32)
        s0, ds0 = OP.pair_rows_multi(ZA, ZB, Wt, tab, umask, grad=grad)
        assert s.shape == s0.shape and ds.shape == ds0.shape, name
        for m in range(umask.shape[0]):
            ref = np.abs(s0[:, m]).max()
            assert np.abs(s[:, m] - s0[:, m]).max() <= (1e-4 if f32 else 1e-11) * ref, (name, m)
            for q in range(ds0.shape[2]):
                ref = np.abs(ds0[:, m, q]).max()
                assert np.abs(ds[:, m, q] - ds0[:, m, q]).max() <= (1e-4 if f32 else 1e-11) * ref, (name, m, q)
    tab = cases[0][1]
    with pytest.raises(ValueError, match='max 16'):
        local.rows_multi(Z, Z, np.ones((len(Z), 4)), tab, np.ones((4, 4)), grad=True)
    with pytest.raises(ValueError, match='umask'):
        local.rows_multi(Z, Z, np.ones((len(Z), 2)), tab, np.ones((2, 3)))
