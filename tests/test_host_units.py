"""Host side of plan creation, checked without a GPU: the factor tables blml_amd.units builds, applied with the
engine's conventions (a numpy model of the device passes: zero-padded power-of-two convolution for padded axes,
band-restricted DFT for circular axes, dense spectrum for multi-column units), reproduce the oracle's dense Gram;
K(0) matches the oracle's table; the pad rule and the library's exported symbols."""
import ctypes
import os
import re

import numpy as np
import pytest

from oracle import blml_oracle as O
from blml_amd import units as U
from blml_amd import modes as M
from blml_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _apply_axis_line(X, axis, fa):
    X = np.moveaxis(X, axis, -1)
    m = X.shape[-1]
    if fa['circ'][0]:
        F = fa['freqs'][0]
        W = np.exp(-2j * np.pi * np.outer(F, np.arange(m)) / m)
        Y = (X @ W.T) * fa['data']
        out = Y @ np.conj(W)
    else:
        L = fa['tlen'][0]
        j = np.arange(L)
        row = np.zeros(L)
        row[:m] = fa['data']
        row[L - m + 1:] = fa['data'][1:][::-1]
        Kf = np.real(np.fft.fft(row))
        out = np.fft.ifft(np.fft.fft(X, n=L, axis=-1) * Kf, axis=-1)[..., :m]
        assert np.allclose(row[(L - j) % L], row)                       # even row
    return np.moveaxis(out, -1, axis)


def _apply_group(X, fa, shape):
    axes = fa['axes']
    if all(fa['circ']):
        Y = X
        for a, ax in reversed(list(enumerate(axes))):
            F = fa['freqs'][a]; m = shape[ax]
            W = np.exp(-2j * np.pi * np.outer(F, np.arange(m)) / m)
            Y = np.moveaxis(np.moveaxis(Y, ax, -1) @ W.T, -1, ax)
        spec = fa['data'].reshape(fa['tlen'])
        sh = [1] * X.ndim
        for a, ax in enumerate(axes):
            sh[ax] = fa['tlen'][a]
        Y = Y * spec.reshape(sh)
        for a, ax in enumerate(axes):
            F = fa['freqs'][a]; m = shape[ax]
            W = np.exp(-2j * np.pi * np.outer(F, np.arange(m)) / m)
            Y = np.moveaxis(np.moveaxis(Y, ax, -1) @ np.conj(W), -1, ax)
        return Y
    T = fa['data'].reshape(fa['tlen'])
    Kf = np.real(np.fft.fftn(T))
    Ls = fa['tlen']
    Y = np.fft.fftn(X, s=Ls, axes=axes)
    sh = [1] * X.ndim
    for a, ax in enumerate(axes):
        sh[ax] = Ls[a]
    Y = np.fft.ifftn(Y * Kf.reshape(sh), axes=axes)
    sl = tuple(slice(0, shape[j]) for j in range(X.ndim))
    return Y[sl]


def apply_mode(facs, e, shape):
    X = e.astype(complex)
    for fa in facs:
        X = _apply_axis_line(X, fa['axes'][0], fa) if fa['kind'] == 'line' else _apply_group(X, fa, shape)
    return X


CASES = [
    ('1d sinc', (40,), [dict(fc=[0.35], kinds=['sinc'])], 1.0),
    ('1d dirichlet circular', (32,), [dict(fc=[0.3], kinds=['dirichlet'], periods=[32.0])], 1.0),
    ('1d dirichlet fill circular', (32,), [dict(fc=[0.6], kinds=['dirichlet:0.5'], periods=[32.0])], 1.0),
    ('1d dirichlet wraps', (8,), [dict(fc=[1.3], kinds=['dirichlet'], periods=[8.0])], 1.0),
    ('1d dirichlet long period', (32,), [dict(fc=[0.3], kinds=['dirichlet'], periods=[50.0])], 1.0),
    ('2d sinc x dirichlet', (12, 16), [dict(fc=[0.4, 0.3], kinds=['sinc', 'dirichlet'], periods=[None, 16.0])], 1.0),
    ('2d ball h=0.5', (11, 13), [dict(fc=[0.8, 0.8], kinds=['ball', 'ball'])], 0.5),
    ('2d pball circular', (17, 23), [dict(fc=[0.4, 0.4], kinds=['pball', 'pball'], periods=[17.0, 23.0])], 1.0),
    ('2d pball padded', (9, 10), [dict(fc=[0.4, 0.4], kinds=['pball', 'pball'], periods=[12.0, 15.0])], 1.0),
    ('2d slits + dirichlet bg', (14, 19), [dict(fc=[0.35, 0.3], kinds=['slits:0.5', 'slits:0.4']),
                                           dict(fc=[0.15, 0.15], kinds=['dirichlet', 'dirichlet'],
                                                periods=[14.0, 19.0])], 1.0),
    ('3d sinc', (6, 7, 5), [dict(fc=[0.5, 0.4, 0.3], kinds=['sinc'] * 3)], 1.0),
    ('3d pball x sinc', (7, 9, 5), [dict(fc=[0.5, 0.5, 0.3], kinds=['pball', 'pball', 'sinc'],
                                         periods=[7.0, 9.0, None])], 1.0),
]


@pytest.mark.parametrize('name,shape,mm,h', CASES, ids=[c[0] for c in CASES])
def test_factors_reproduce_dense_gram(name, shape, mm, h):
    modes = [md for kw in mm for md in M.make_modes(**kw)]
    specs, _ = M.parse_modes(modes, len(shape))
    ospecs = O.parse_modes([md for kw in mm for md in O.make_modes(**kw)], len(shape))
    hv = np.full(len(shape), h)
    circ = U.circular_dims(specs, shape, hv)
    om = O.BLMLBinnedOracle()
    om._setup(np.ones(shape), [md for kw in mm for md in O.make_modes(**kw)], h, 0.0)
    assert tuple(m if c else U.next_fast_len5(2 * m - 1) for m, c in zip(shape, circ)) == om.pad_
    rng = np.random.default_rng(7)
    for k, (sp, osp) in enumerate(zip(specs, ospecs)):
        s, facs = U.mode_factors(sp, shape, hv, circ)
        assert s == om.s_[k]                                          # K(0) bit-identical to the oracle's T[0]
        G = O.dense_gram(osp, shape, hv)
        e = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
        got = apply_mode(facs, e, shape).ravel()
        want = G @ e.ravel()
        assert np.abs(got - want).max() < 1e-12 * max(1.0, np.abs(G).sum(1).max()), name


def test_parse_modes_errors_mirror_reference():
    with pytest.raises(TypeError):
        M.parse_modes({0: dict(kind='sinc', fc=0.3)}, 1)
    with pytest.raises(ValueError):
        M.parse_modes([{0: dict(kind='sinc', fc=-0.3)}], 1)
    with pytest.raises(ValueError):
        M.parse_modes([{0: dict(kind='slits', fc=0.3)}], 1)              # fill missing
    with pytest.raises(ValueError):
        M.parse_modes([{0: dict(kind='ball', fc=0.3)}], 1)               # multi-column kind on an int key
    with pytest.raises(ValueError):
        M.parse_modes([{0: dict(kind='sinc', fc=0.3)}], 2)               # column 1 missing
    with pytest.raises(ValueError):
        M.parse_modes([{0: dict(kind='dirichlet', fc=0.3)}], 1)          # period missing
    with pytest.raises(ValueError):
        M.parse_modes([{0: dict(kind='sinc', fc=0.3, bogus=1)}], 1)


def test_make_modes_matches_oracle_tokens():
    for kw in [dict(fc=[0.3, 0.2], kinds=['pball', 'pball'], periods=[31.0, 31.0]),
               dict(fc=[[0.3, 0.2], [0.1, 0.1]], kinds=['slits:0.5', 'dirichlet:0.7'], periods=[None, 20.0]),
               dict(fc=[0.4, 0.4, 0.2], kinds=['ball', 'ball', 'sinc'])]:
        a, _ = M.parse_modes(M.make_modes(**kw), len(kw['kinds']))
        b = O.parse_modes(O.make_modes(**kw), len(kw['kinds']))
        for x, y in zip(a, b):
            assert x['kinds'] == y['kinds'] and np.array_equal(x['fc'], y['fc']) and x['periods'] == y['periods']


def _declared_functions():
    src = open(os.path.join(ROOT, 'include', 'blml_hip.h')).read()
    return sorted(set(re.findall(r'^\s*(?:int|const char \*)\s*(blml_\w+)\s*\(', src, re.M)))


def test_library_exports_every_declared_symbol():
    path = _lib.LIB_PATH
    if not os.path.exists(path):
        pytest.skip('libblml_hip.so not built')
    lib = ctypes.CDLL(path)
    declared = _declared_functions()
    assert set(declared) == set(_lib.EXPORTS)
    for name in declared:
        assert hasattr(lib, name), name


def test_struct_layouts_match_header(tmp_path):
    """ctypes mirrors of the C structs agree with the header's layout (sizes and field offsets, compiled by gcc)."""
    import shutil
    import subprocess
    if shutil.which('gcc') is None:
        pytest.skip('gcc not available')
    from blml_amd import local, points
    fields = {'blml_factor_t': _lib.Factor, 'blml_fit_t': _lib.Fit, 'blml_params_t': _lib.Params,
              'blml_result_t': _lib.Result, 'blml_pairtab_t': local.PairTab, 'blml_pairgrad_t': local.PairGrad,
              'blml_pointunit_t': points.PointUnit, 'blml_emp_params_t': _lib.EmpParams,
              'blml_hio_params_t': _lib.HioParams}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "blml_hip.h"', 'int main(void) {']
    for cname, py in fields.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'  printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ['  return 0;', '}']
    src = tmp_path / 'layout.c'
    src.write_text('\n'.join(lines))
    exe = tmp_path / 'layout'
    subprocess.run(['gcc', '-I', os.path.join(ROOT, 'include'), str(src), '-o', str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split('\n')
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for cname, py in fields.items():
        assert got[(cname, 'size')] == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)


def test_fit_input_errors_mirror_reference():
    """The fit's input checks run on the host before any device work and raise what blml_binned.py:119-124,
    383-400 raise (ValueError), plus this engine's own limits (at most 5 axes, 8 modes)."""
    from blml_amd.binned import _FitSpec
    modes = M.make_modes([0.3], ['sinc'])
    ok = np.ones(12)
    spec = lambda c, **kw: _FitSpec(c, kw.pop('modes', modes), 1.0, 0.0, kw.pop('complex_g', False),
                                    kw.pop('signs', None), None, 0)
    assert spec(ok).B == 12
    for bad in (np.array(3.0), -ok, np.where(np.arange(12) == 3, np.nan, 1.0), np.where(np.arange(12) == 5, np.inf, 1.0)):
        with pytest.raises(ValueError, match='non-negative finite'):
            spec(bad)
    with pytest.raises(ValueError, match='no observations'):
        spec(np.zeros(12))
    with pytest.raises(ValueError, match='phases'):
        spec(ok, complex_g=True, signs=np.ones(12))
    with pytest.raises(ValueError, match='nonzero on every occupied bin'):
        spec(ok, signs=np.where(np.arange(12) == 2, 0.0, 1.0))
    assert spec(np.where(np.arange(12) == 2, 0.0, 1.0), signs=np.where(np.arange(12) == 2, 0.0, -1.0)).signs is not None
    with pytest.raises(ValueError, match='dimensions'):
        spec(np.ones((2,) * 6), modes=M.make_modes([0.3] * 6, ['sinc'] * 6))
    with pytest.raises(ValueError, match='modes'):
        spec(ok, modes=M.make_modes([[0.3]] * 9, ['sinc']))
    with pytest.raises(ValueError, match='dep'):
        M.parse_modes(M.make_modes([0.3, 0.2], ['sinc', 'sinc'], dep=[1]), 2)


def test_conditional_modes_parse_and_restrict():
    """dep columns (blml_trivial.py:1247-1323, 1467-1505, 1585-1602): make_modes marks them (a radial unit through
    its first column), parse_modes returns them for the counts path and rejects modes that disagree; the joint's
    modes drop the flags and the marginal's keep the other columns, renumbered, a unit whole or not at all."""
    from blml_amd import modes as M
    md = M.make_modes([0.3, 0.2, 0.25], ['ball:a', 'ball:a', 'sinc'], dep=[2])
    specs, dep = M.parse_modes(md, 3, allow_dep=True)
    assert dep == [2] and specs[0]['kinds'] == ['ball:0-1', 'ball:0-1', 'sinc']
    md2 = M.make_modes([0.3, 0.3, 0.25], ['ball:a', 'ball:a', 'sinc'], dep=[0])
    assert M.parse_modes(md2, 3, allow_dep=True)[1] == [0, 1]
    assert M.parse_modes(M.strip_dep(md), 3)[1] is None
    mg = M.marginal_modes(md, [0, 1])
    assert list(mg[0]) == [(0, 1)] and 'dep' not in mg[0][(0, 1)]
    assert M.marginal_modes(M.make_modes([0.3, 0.2, 0.25], dep=[0]), [1, 2]) == [{0: dict(kind='sinc', fc=0.2),
                                                                               1: dict(kind='sinc', fc=0.25)}]
    with pytest.raises(ValueError, match='straddles'):
        M.marginal_modes(md, [0, 2])
    bad = [dict(md[0]), {k: {a: v for a, v in c.items() if a != 'dep'} for k, c in md[0].items()}]
    with pytest.raises(ValueError, match='same columns'):
        M.parse_modes(bad, 3, allow_dep=True)


def test_pair_spans_reject_ranks_outside_the_table():
    """A ranked=True caller's rank column must hold ranks 0 .. mk - 1: a negative rank would be clamped into another
    column's rows on the device, so the host check raises (the reference's stale-table RuntimeError,
    blml_kernels.py:214-215) for ranks above and below the table."""
    from blml_amd import local
    Z = np.stack([np.array([0, 3, 7, 40]), np.arange(4)], 1)          # column 0: 4 values over a span of 40
    tab = local.lattice_tables(['sinc', 'sinc'], [0.3, 0.3], None, [1.0, 1.0], Z)
    assert int(tab['kinds'][0]) == 3 and int(tab['mk'][0]) == 4
    A = local.rank_coords(Z, tab)
    local._check_spans(A, A, tab)
    for bad in (-1, 4):
        B = A.copy()
        B[0, 0] = bad
        with pytest.raises(RuntimeError, match='rank outside'):
            local._check_spans(A, B, tab)


def test_points_are_binned_as_the_reference_bins_them():
    """BLMLQuick on points bins them on the fit's own lattice first (blml_amd.quick.bin_points, blml_gpu.py:128-176):
    the target widths, the steps (a folded period tiled by whole bins), the bin centres and their counts equal the
    reference's own binning (tests/golden/binning.npz, written by make_golden_binning.py from the reference's
    lattice_target, lattice_fold_specs and ProductKernel.quantize), every point lands in its nearest bin (circular
    distance on a folded column), and a refit whose targets moved less than (1 + rebin_tol)^(1/4) keeps the lattice."""
    import json
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), 'golden'))
    from cases_binning import binning_cases
    from blml_amd import make_modes
    from blml_amd.modes import parse_modes
    from blml_amd.quick import bin_points, quantize, lattice_fold
    z = np.load(os.path.join(os.path.dirname(__file__), 'golden', 'binning.npz'))
    meta = json.loads(str(z['meta']))
    for i, (name, X, modes, bs) in enumerate(binning_cases(make_modes)):
        assert meta[i]['name'] == name
        specs, _ = parse_modes(modes, X.shape[1])
        bp = bin_points(X, specs, bin_scale=bs)
        p = f'c{i}_'
        assert np.allclose(bp['h'], z[p + 'h'], rtol=1e-14, atol=0), name
        assert np.allclose(bp['h_target'], z[p + 'h_target'], rtol=1e-14, atol=0), name
        assert np.array_equal(bp['Z'], z[p + 'Z']) and np.array_equal(bp['w'], z[p + 'w']), name
        assert bp['w'].sum() == len(X)
        Xq = quantize(X, specs[0]['kinds'], specs[0]['periods'], lattice_fold(specs), bp['h'])
        for k in range(X.shape[1]):
            dist = np.abs(X[:, k] - Xq[:, k])
            if str(specs[0]['kinds'][k]).partition(':')[0] in ('dirichlet', 'pball'):
                P = specs[0]['periods'][k]
                assert abs(P / bp['h'][k] - round(P / bp['h'][k])) < 1e-9 and P / bp['h'][k] >= 8
                dist = np.minimum(np.mod(dist, P), P - np.mod(dist, P))
            assert np.all(dist <= 0.5 * bp['h'][k] * (1 + 1e-9)), (name, k)

    class Prev:
        pass
    name, X, modes, bs = binning_cases(make_modes)[1]
    specs, _ = parse_modes(modes, 2)
    first = bin_points(X, specs, bin_scale=bs)
    prev = Prev()
    prev.h_target_, prev.n_lattice_, prev._h_vec = first['h_target'], first['n_lattice'], first['h']
    more = np.concatenate([X, X[:600] + 0.01])                       # 15% more points: targets move 3.6%
    kept = bin_points(more, specs, init=prev, bin_scale=bs)
    assert not kept['rebinned'] and np.array_equal(kept['h'], first['h'])
    many = np.concatenate([X] * 3)                                    # 3x the points: 24% finer targets
    moved = bin_points(many, specs, init=prev, bin_scale=bs)
    assert moved['rebinned'] and np.all(moved['h'] < first['h'])
    signs = np.where(X[:, 0] < 0, -1.0, 1.0)
    sg = bin_points(X, specs, signs=signs, bin_scale=bs)['signs']
    assert set(np.unique(sg)) <= {-1.0, 1.0}
    with pytest.raises(ValueError, match='sample_weight'):
        bin_points(X, specs, sample_weight=np.zeros(len(X)), bin_scale=bs)
