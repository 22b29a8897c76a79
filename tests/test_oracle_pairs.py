"""The CPU restatement of the reference's pair accumulation (oracle/blml_pairs.py), pinned without the reference:
rows and cols over tables built from the kernels' closed forms equal the dense Gram gram_int @ w
(blml_kernels.py:243-251; results/pball_parity.py:66-76) to 1e-12; radial rows follow the interpolation rule of
blml_local.py:250-252 (checked against a per-pair loop); rank columns reject values their table lacks."""
import numpy as np
import pytest

from oracle import blml_oracle as O
from oracle import blml_pairs as OP

# (name, shape of the lattice box, units, h, J) -- units as blml_pairs.build_tables takes them
CASES = [
    ('2d sinc x slits offset rows', (14, 11),
     [('offset', 'sinc', [0.4], [0.0], [0]), ('offset', 'slits:0.5', [0.3], [0.0], [1])], (1.0, 1.0), (14, 11)),
    ('1d dirichlet circular row', (16,), [('circular', 'dirichlet', [0.3], [16.0], [0])], (1.0,), (16,)),
    ('2d dirichlet unfolded x sinc h', (9, 12),
     [('offset', 'dirichlet', [0.5], [20.0], [0]), ('offset', 'sinc', [1.2], [0.0], [1])], (1.0, 0.25), (9, 12)),
    ('2d pball torus', (10, 8), [('torus', 'pball', [0.45], [10.0, 8.0], [0, 1])], (1.0, 1.0), (10, 8)),
    ('3d slits x dirichlet circular x sinc', (7, 10, 6),
     [('offset', 'slits:0.4', [0.5], [0.0], [0]), ('circular', 'dirichlet:0.5', [0.6], [10.0], [1]),
      ('offset', 'sinc', [0.3], [0.0], [2])], (1.0, 1.0, 1.0), (7, 10, 6)),
]


def _bins(rng, shape, n):
    Z = np.stack([rng.integers(0, m, n) for m in shape], 1)
    return np.unique(Z, axis=0)


@pytest.mark.parametrize('name,shape,units,h,J', CASES, ids=[c[0] for c in CASES])
def test_rows_and_cols_equal_dense_gram(name, shape, units, h, J):
    rng = np.random.default_rng(3)
    tab = OP.build_tables(units, np.asarray(h, float), J)
    ZA, ZB = _bins(rng, shape, 60), _bins(rng, shape, 90)
    w = rng.standard_normal(len(ZB))
    w[::7] = 0.0
    G = OP.dense_pair_gram(units, np.asarray(h, float), ZA, ZB)
    s = OP.pair_rows(ZA, ZB, w, tab)
    assert np.abs(s - G @ w).max() < 1e-12 * max(1.0, np.abs(G).sum(1).max()), name
    out, _ = OP.pair_cols(ZA, ZB, tab)
    assert np.abs(out - G).max() < 1e-13 * max(1.0, np.abs(G).max()), name


def test_cols_marginal_is_the_other_columns():
    units = CASES[4][2]
    h = np.ones(3)
    tab = OP.build_tables(units, h, CASES[4][4])
    rng = np.random.default_rng(4)
    ZA, ZB = _bins(rng, (7, 10, 6), 30), _bins(rng, (7, 10, 6), 40)
    out, marg = OP.pair_cols(ZA, ZB, tab, depmask=0b010)
    want = OP.dense_pair_gram([units[0], units[2]], h, ZA, ZB)
    assert np.abs(marg - want).max() < 1e-13 and np.abs(out - OP.dense_pair_gram(units, h, ZA, ZB)).max() < 1e-13


def test_rank_table_equals_offset_row():
    """A rank column (the occupied values' pairs) gives the same kernel as the offset row it replaces."""
    rng = np.random.default_rng(5)
    Z = _bins(rng, (40, 6), 25)
    vals = [np.unique(Z[:, 0]), None]
    units = [('rank', 'sinc', [0.2], [0.0], [0]), ('offset', 'sinc', [0.5], [0.0], [1])]
    tab = OP.build_tables(units, np.ones(2), (40, 6), vals=vals)
    w = rng.standard_normal(len(Z))
    G = OP.dense_pair_gram(units, np.ones(2), Z, Z)
    assert np.abs(OP.pair_rows(Z, Z, w, tab) - G @ w).max() < 1e-12
    with pytest.raises(RuntimeError):
        OP.rank_coords(np.array([[41, 0]]), tab)


def test_radial_rows_interpolate_per_pair():
    """Radial rows: the value at sqrt(sum off^2) * 4 on the row, linear between grid points (random table)."""
    rng = np.random.default_rng(6)
    LR = 80
    row = rng.standard_normal(LR)
    tab = dict(T=rng.standard_normal(30), base=np.array([0, 0, 0]), Jk=np.array([0, 0, 14]), mk=np.zeros(3, int),
               vals=[None, None, None], kinds=np.array([2, 2, 0]), nbins=np.zeros(3, int), grp=np.array([0, 0, -1]),
               R=row[None], dim_unit=np.array([-1, -1, 1]), grp_unit=np.array([0]), U=2, nW=0)
    ZA, ZB = _bins(rng, (8, 8, 8), 12), _bins(rng, (8, 8, 8), 15)
    w = rng.standard_normal(len(ZB))
    want = np.zeros(len(ZA))
    for r in range(len(ZA)):
        for j in range(len(ZB)):
            o = ZA[r] - ZB[j]
            t = np.sqrt(float(o[0] ** 2 + o[1] ** 2)) * 4
            i0 = int(t)
            f = t - i0
            want[r] += (row[i0] * (1 - f) + row[i0 + 1] * f) * tab['T'][o[2] + 14] * w[j]
    assert np.abs(OP.pair_rows(ZA, ZB, w, tab) - want).max() < 1e-12


def test_product_tables_match_oracle_tables():
    """blml_amd.local.lattice_tables (the product's builder) gives the rows the oracle's closed-form builder gives,
    so rows over either equal the dense Gram."""
    from blml_amd import local
    rng = np.random.default_rng(8)
    Z = _bins(rng, (30, 12, 16), 200)
    kinds = ['slits:0.5', 'sinc', 'dirichlet']
    fc, per, h = [0.3, 0.45, 0.4], [None, None, 16.0], [1.0, 1.0, 1.0]
    tab = local.lattice_tables(kinds, fc, per, h, Z)
    units = [('offset', 'slits:0.5', [0.3], [0.0], [0]), ('offset', 'sinc', [0.45], [0.0], [1]),
             ('circular', 'dirichlet', [0.4], [16.0], [2])]
    w = rng.standard_normal(len(Z))
    G = OP.dense_pair_gram(units, np.asarray(h), Z, Z)
    assert np.abs(OP.pair_rows(Z, Z, w, tab) - G @ w).max() < 1e-12 * np.abs(G).sum(1).max()
    Zp = _bins(rng, (10, 8), 50)
    tabp = local.lattice_tables(['pball', 'pball'], [0.45, 0.45], [10.0, 8.0], [1.0, 1.0], Zp)
    Gp = OP.dense_pair_gram([('torus', 'pball', [0.45], [10.0, 8.0], [0, 1])], np.ones(2), Zp, Zp)
    wp = rng.standard_normal(len(Zp))
    assert np.abs(OP.pair_rows(Zp, Zp, wp, tabp) - Gp @ wp).max() < 1e-12


def test_c_rows_equal_the_dense_gram():
    """The C restatement (oracle/pairs_cpu.c, the pair path's CPU baseline) over offset, circular and rank columns
    equals the dense Gram (FP64 tables) and the numpy restatement; with FP32 tables its FP32 products stay within
    FP32 rounding of them; a planted shift of the offset rows (J + 1) fails the check."""
    from blml_amd import local
    rng = np.random.default_rng(9)
    Z = _bins(rng, (30, 12, 16), 300)
    kinds = ['slits:0.5', 'sinc', 'dirichlet']
    fc, per, h = [0.3, 0.45, 0.4], [None, None, 16.0], [1.0, 1.0, 1.0]
    tab = local.lattice_tables(kinds, fc, per, h, Z)
    units = [('offset', 'slits:0.5', [0.3], [0.0], [0]), ('offset', 'sinc', [0.45], [0.0], [1]),
             ('circular', 'dirichlet', [0.4], [16.0], [2])]
    w = rng.standard_normal(len(Z))
    w[::7] = 0.0
    G = OP.dense_pair_gram(units, np.asarray(h), Z, Z)
    scale = np.abs(G).sum(1).max()
    s = OP.pair_rows_c(Z, Z, w, tab, threads=4)
    assert np.abs(s - G @ w).max() < 1e-12 * scale
    assert np.abs(s - OP.pair_rows(Z, Z, w, tab)).max() < 1e-13 * scale
    s32 = OP.pair_rows_c(Z, Z, w, tab, f32=True, threads=4)
    assert np.abs(s32 - G @ w).max() < 1e-6 * scale
    bad = dict(tab, Jk=np.asarray(tab['Jk']) + 1)
    assert np.abs(OP.pair_rows_c(Z, Z, w, bad, threads=4) - G @ w).max() > 1e-4 * scale


def test_ball_tables_radial_rows():
    """A ball over columns (one radial group) next to a sinc column: lattice_tables' radial row is the oracle's ball
    (kernel_eval, the reference's closed form) on the grid r = j / 4 lattice steps to 1e-14, so rows through it
    equal the exact Gram at pairs whose distance lies on the grid (a shift along one column) and stay within the
    linear interpolation's error bound elsewhere; fc h must be equal across the group."""
    from blml_amd import local
    rng = np.random.default_rng(9)
    Z = _bins(rng, (14, 12, 10), 120)
    fc, h = [0.35, 0.35, 0.4], [1.0, 1.0, 1.0]
    tab = local.lattice_tables(['ball', 'ball', 'sinc'], fc, None, h, Z)
    assert list(tab['kinds']) == [2, 2, 0] and int(tab['U']) == 2 and list(tab['grp']) == [0, 0, -1]
    row = np.asarray(tab['R'])[0]
    r = np.arange(len(row)) / OP.RADIAL_Q
    want = O.kernel_eval('ball', [0.35, 0.35], [0.0, 0.0], [r, 0.0 * r])
    assert np.abs(row - want).max() < 1e-14 * np.abs(want).max()
    span = Z.max(0) - Z.min(0)
    assert (len(row) - 2) / OP.RADIAL_Q >= np.sqrt((span[:2] ** 2).sum())   # every pair's distance is on the row
    units = [('offset', 'ball', [0.35, 0.35], [0.0, 0.0], [0, 1]), ('offset', 'sinc', [0.4], [0.0], [2])]
    ZA = Z.copy()
    ZB = Z.copy()
    ZB[:, 0] += 1                                      # distances on the grid: |off| = integer when column 1 matches
    ZB[:, 1] = ZA[:, 1]
    w = rng.standard_normal(len(Z))
    G = OP.dense_pair_gram(units, np.asarray(h), ZA, ZB)
    got = OP.pair_rows(ZA, ZB, w, local.lattice_tables(['ball', 'ball', 'sinc'], fc, None, h, np.vstack([ZA, ZB])))
    # pairs with both column offsets nonzero fall between grid points: |K'' dr^2 / 8| with dr = 1/4, |K''| <= fc^2 pi^2 K0
    bound = (0.35 ** 2 * np.pi ** 2 * 0.35 ** 2 * (1.0 / OP.RADIAL_Q) ** 2 / 8) * np.abs(
        OP.dense_pair_gram(units[1:], np.asarray(h), ZA, ZB)).sum(1).max() * np.abs(w).max()
    assert np.abs(got - G @ w).max() <= bound
    with pytest.raises(ValueError, match='equal across the group'):
        local.lattice_tables(['ball', 'ball'], [0.35, 0.3], None, [1.0, 1.0], Z[:, :2])


def test_multi_column_groups_by_tag_in_order():
    """Two periodic balls with their own tags get a torus table each (their own fc), and a ball listed before a
    periodic ball is group 0 / unit 0 as in the reference's radial_groups and kernel_units (blml_trivial.py:251-261,
    330-335); rows over either table equal the dense Gram of the units."""
    from blml_amd import local
    rng = np.random.default_rng(21)
    Z = _bins(rng, (10, 8, 12, 6), 150)
    kinds = ['pball:a', 'pball:a', 'pball:b', 'pball:b']
    fc, per, h = [0.3, 0.3, 0.45, 0.45], [10.0, 8.0, 12.0, 6.0], [1.0] * 4
    tab = local.lattice_tables(kinds, fc, per, h, Z)
    assert list(tab['grp']) == [0, 0, 1, 1] and int(tab['U']) == 2 and list(tab['grp_unit']) == [0, 1]
    units = [('torus', 'pball', [0.3], [10.0, 8.0], [0, 1]), ('torus', 'pball', [0.45], [12.0, 6.0], [2, 3])]
    w = rng.standard_normal(len(Z))
    G = OP.dense_pair_gram(units, np.asarray(h), Z, Z)
    assert np.abs(OP.pair_rows(Z, Z, w, tab) - G @ w).max() < 1e-12 * max(1.0, np.abs(G).sum(1).max())
    # a ball before a periodic ball, then a sinc column: units sinc (0), ball (1), pball (2); groups ball 0, pball 1
    Zb = np.stack([rng.integers(0, 9, 140), rng.integers(0, 9, 140), rng.integers(0, 10, 140),
                   rng.integers(0, 8, 140), rng.integers(0, 7, 140)], 1)
    Zb = np.unique(Zb, axis=0)
    kinds_b = ['ball:r', 'ball:r', 'pball:p', 'pball:p', 'sinc']
    fc_b, per_b, h_b = [0.35, 0.35, 0.4, 0.4, 0.3], [None, None, 10.0, 8.0, None], [1.0] * 5
    tb = local.lattice_tables(kinds_b, fc_b, per_b, h_b, Zb)
    assert list(tb['grp']) == [0, 0, 1, 1, -1] and list(tb['grp_unit']) == [1, 2]
    assert list(tb['kinds']) == [2, 2, 4, 4, 0] and int(tb['U']) == 3 and list(tb['dim_unit']) == [-1, -1, -1, -1, 0]
    # every pair's ball distance on the radial grid (column 1 the same for every bin): the ball's row is exact there
    ZA = Zb.copy()
    ZA[:, 1] = 4
    ZB = ZA.copy()
    ZB[:, 0] = (ZA[:, 0] + 3) % 9
    tb = local.lattice_tables(kinds_b, fc_b, per_b, h_b, np.vstack([ZA, ZB]))
    ub = [('offset', 'ball', [0.35, 0.35], [0.0, 0.0], [0, 1]), ('torus', 'pball', [0.4], [10.0, 8.0], [2, 3]),
          ('offset', 'sinc', [0.3], [0.0], [4])]
    wb = rng.standard_normal(len(ZB))
    Gb = OP.dense_pair_gram(ub, np.asarray(h_b), ZA, ZB)
    assert np.abs(OP.pair_rows(ZA, ZB, wb, tb) - Gb @ wb).max() < 1e-12 * max(1.0, np.abs(Gb).sum(1).max())
    with pytest.raises(ValueError, match='mixes kinds'):
        local.lattice_tables(['ball:x', 'pball:x'], [0.3, 0.3], [None, 8.0], [1.0, 1.0], Z[:, :2])


# ------------------------------------------------------------------------------------------------ derivative rows
def _fd_rows(units, h, ZA, ZB, w, ui, how, eps=1e-5):
    """Central finite difference of the dense Gram's rows (the kernels' closed forms, oracle.kernel_eval) in one
    unit's log fc ('logfc', a group's fc scaled together), log period ('logP') or coordinate offset ('t', i)."""
    def gram(delta):
        G = np.ones((len(ZA), len(ZB)))
        for u, (kind, fcs, pers, dims) in enumerate(units):
            fcs, pers = np.asarray(fcs, float), np.asarray(pers, float)
            diffs = [(ZA[:, k, None] - ZB[None, :, k]) * h[k] for k in dims]
            if u == ui:
                if how == 'logfc':
                    fcs = fcs * np.exp(delta)
                elif how == 'logP':
                    pers = pers * np.exp(delta)
                else:
                    diffs[how[1]] = diffs[how[1]] + delta
            G = G * O.kernel_eval(kind, fcs, pers, diffs)
        return G
    return (gram(eps) - gram(-eps)) @ w / (2 * eps)


def _column_case(rng):
    """sinc, slits, dirichlet (folded on P = 16) and a rank column (three occupied values) with every derivative
    entry the reference allows for them."""
    n = 70
    Z = np.stack([rng.integers(0, 14, n), rng.integers(0, 12, n), rng.integers(0, 16, n),
                  rng.choice([0, 7, 19], n)], 1)
    kinds, fc, periods = ['sinc', 'slits:0.5', 'dirichlet', 'sinc'], [0.4, 0.45, 0.4, 0.3], [None, None, 16.0, None]
    want = [(0, 'fc'), (0, 'shift'), (1, 'fc'), (1, 'shift'), (2, 'period'), (2, 'shift'), (3, 'fc'), (3, 'shift')]
    units = [('sinc', [0.4], [0.0], [0]), ('slits:0.5', [0.45], [0.0], [1]), ('dirichlet', [0.4], [16.0], [2]),
             ('sinc', [0.3], [0.0], [3])]
    how = {'fc': 'logfc', 'period': 'logP', 'shift': ('t', 0)}
    return Z, kinds, fc, periods, want, units, [(u, how[p]) for u, p in want]


def test_derivative_rows_match_finite_differences():
    """rows(grad=True) through the derivative tables of lattice_tables (raw-offset rows, a rank table, a dirichlet
    period row) equals central differences of the dense Gram in each unit's log fc / log period / offset to 1e-7 of
    the rows' scale; the value rows are unchanged.  A wrong sign, a dropped term or a transposed offset fails it."""
    from blml_amd import local
    rng = np.random.default_rng(31)
    Z, kinds, fc, periods, want, units, fd = _column_case(rng)
    h = np.ones(4)
    tab = local.lattice_tables(kinds, fc, periods, h, Z, want=want)
    assert int(tab['nW']) == len(want) and int(tab['kinds'][3]) == 3 and int(tab['kinds'][2]) == 1
    w = rng.standard_normal(len(Z))
    w[::4] = 0.0
    s, ds = OP.pair_rows_grad(Z, Z, w, tab)
    assert np.abs(s - OP.pair_rows(Z, Z, w, tab)).max() < 1e-14 * np.abs(s).max()
    for q, (ui, how) in enumerate(fd):
        ref = _fd_rows(units, h, Z, Z, w, ui, how)
        assert np.abs(ds[:, q] - ref).max() < 1e-7 * np.abs(ref).max(), (q, want[q])


def test_derivative_rows_ball_and_pball():
    """A ball group's fc and shift rows (on the r grid): exact at pairs on the grid (one column offset 0) to 1e-7, and
    within the linear interpolation's error elsewhere (3e-2 of the scale; a wrong term is O(1)); its table entries
    are the closed forms' differences at the grid points.  A periodic ball's shift rows (torus tables) to 1e-7."""
    from blml_amd import local
    rng = np.random.default_rng(32)
    kinds, fc, h = ['ball', 'ball', 'sinc'], [0.35, 0.35, 0.4], np.ones(3)
    want = [(1, 'fc'), (1, ('shift', 0)), (1, ('shift', 1)), (0, 'fc'), (0, 'shift')]
    units = [('sinc', [0.4], [0.0], [2]), ('ball', [0.35, 0.35], [0.0, 0.0], [0, 1])]
    fd = [(1, 'logfc'), (1, ('t', 0)), (1, ('t', 1)), (0, 'logfc'), (0, ('t', 0))]
    Z = _bins(rng, (14, 12, 10), 90)
    tab = local.lattice_tables(kinds, fc, None, h, Z, want=want)
    r = np.arange(np.asarray(tab['dR']).shape[1]) / OP.RADIAL_Q
    drow = np.asarray(tab['dR'])
    for q, how in ((0, 'logfc'), (1, ('t', 0))):
        e = 1e-5
        if how == 'logfc':
            ref = (O.kernel_eval('ball', [0.35 * np.exp(e)] * 2, [0, 0], [r, 0 * r]) -
                   O.kernel_eval('ball', [0.35 * np.exp(-e)] * 2, [0, 0], [r, 0 * r])) / (2 * e)
            got = drow[q]
        else:                                      # the shift row times the offset r along column 0
            ref = (O.kernel_eval('ball', [0.35] * 2, [0, 0], [r + e, 0 * r]) -
                   O.kernel_eval('ball', [0.35] * 2, [0, 0], [r - e, 0 * r])) / (2 * e)
            got = drow[q] * r
        assert np.abs(got - ref).max() < 1e-8 * np.abs(ref).max(), q
    w = rng.standard_normal(len(Z))
    s, ds = OP.pair_rows_grad(Z, Z, w, tab)
    for q, (ui, how) in enumerate(fd):
        ref = _fd_rows(units, h, Z, Z, w, ui, how)
        assert np.abs(ds[:, q] - ref).max() < 3e-2 * np.abs(ref).max(), (q, want[q])
    Zg = Z.copy()
    Zg[:, 1] = 5                                   # every pair's distance |off_0| on the grid
    tabg = local.lattice_tables(kinds, fc, None, h, Zg, want=want)
    s, ds = OP.pair_rows_grad(Zg, Zg, w, tabg)
    for q, (ui, how) in enumerate(fd):
        if q == 2:
            assert np.abs(ds[:, q]).max() == 0.0   # d/dt_1 at offset 0 along column 1
            continue
        ref = _fd_rows(units, h, Zg, Zg, w, ui, how)
        assert np.abs(ds[:, q] - ref).max() < 1e-7 * np.abs(ref).max(), (q, want[q])
    Zp = _bins(rng, (10, 8), 50)
    tabp = local.lattice_tables(['pball', 'pball'], [0.45, 0.45], [10.0, 8.0], [1.0, 1.0], Zp,
                                want=[(0, 'shift'), (0, ('shift', 1))])
    wp = rng.standard_normal(len(Zp))
    s, ds = OP.pair_rows_grad(Zp, Zp, wp, tabp)
    unitp = [('pball', [0.45, 0.45], [10.0, 8.0], [0, 1])]
    for q in range(2):
        ref = _fd_rows(unitp, np.ones(2), Zp, Zp, wp, 0, ('t', q))
        assert np.abs(ds[:, q] - ref).max() < 1e-7 * np.abs(ref).max(), q
    with pytest.raises(ValueError, match='shifts only'):
        local.lattice_tables(['pball', 'pball'], [0.45, 0.45], [10.0, 8.0], [1.0, 1.0], Zp, want=[(0, 'fc')])
    with pytest.raises(ValueError, match='whole harmonics'):
        local.lattice_tables(['dirichlet'], [0.4], [16.0], [1.0], Zp[:, :1], want=[(0, 'fc')])
    with pytest.raises(ValueError, match='fixed shape'):
        local.lattice_tables(['slits:0.5'], [0.4], None, [1.0], Zp[:, :1], want=[(0, 'fill')])


def test_fused_densities_are_the_marginal_grams():
    """pair_rows_multi: the joint mask gives pair_rows_grad; a marginal mask gives the dense Gram of its units alone
    (closed forms, 1e-12) and that Gram's finite differences for the entries on its units (1e-7), zero for the
    others; a source whose weights are all zero is skipped."""
    from blml_amd import local
    rng = np.random.default_rng(33)
    Z, kinds, fc, periods, want, units, fd = _column_case(rng)
    h = np.ones(4)
    tab = local.lattice_tables(kinds, fc, periods, h, Z, want=want)
    umask = np.array([[1, 1, 1, 1], [1, 0, 1, 0], [0, 1, 0, 0]])
    Wt = rng.standard_normal((len(Z), 3))
    Wt[::6] = 0.0
    Wt[1::7, 1] = 0.0
    s, ds = OP.pair_rows_multi(Z, Z, Wt, tab, umask, grad=True)
    s0, ds0 = OP.pair_rows_grad(Z, Z, Wt[:, 0], tab)
    assert np.abs(s[:, 0] - s0).max() < 1e-14 * np.abs(s0).max() and np.abs(ds[:, 0] - ds0).max() < 1e-14 * np.abs(ds0).max()
    for m in (1, 2):
        keep = [u for u in range(4) if umask[m, u]]
        sub = [units[u] for u in keep]
        G = np.ones((len(Z), len(Z)))
        for kind, fcs, pers, dims in sub:
            G = G * O.kernel_eval(kind, np.asarray(fcs), np.asarray(pers), [(Z[:, k, None] - Z[None, :, k]) * h[k] for k in dims])
        assert np.abs(s[:, m] - G @ Wt[:, m]).max() < 1e-12 * np.abs(G).sum(1).max()
        for q, (ui, how) in enumerate(fd):
            if not umask[m, ui]:
                assert not ds[:, m, q].any()
                continue
            ref = _fd_rows(sub, h, Z, Z, Wt[:, m], keep.index(ui), how)
            assert np.abs(ds[:, m, q] - ref).max() < 1e-7 * np.abs(ref).max(), (m, q)
