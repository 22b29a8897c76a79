"""Pair accumulation over lattice bins on the MI355X: the reference's blml_local_gpu.rows / .cols
(blml_local_gpu.py:453-475) over the same table dicts (blml_kernels.ProductKernel.tables, blml_kernels.py:133-200),
computed by libblml_hip.so (blml_pair_rows / blml_pair_cols).

    s, ds = rows(ZA, ZB, w, tab)           s_r = sum_j S(ZA_r, ZB_j) w_j   (ds: zeros)
    s, ds = rows(ZA, ZB, w, tab, grad=True)   with ds_rq = sum_j dS_q(ZA_r, ZB_j) w_j for tab's derivative entries
    s, ds = rows_multi(ZA, ZB, Wt, tab, umask, grad)   several densities (unit masks) in one pass
    out = cols(ZA, ZB, tab)                the block S(ZA_r, ZB_j) (R, M)
    out, out_m = cols(ZA, ZB, tab, depmask=mask)   with the marginal over the columns not in mask

Inputs are host arrays (integer lattice coordinates, int64 as the reference's); results are numpy.  f32=True reads
float32 tables and multiplies in float32, with FP64 sums (the reference's f32 policy, blml_local_gpu.py:5-8).  The
coordinates must lie inside the tables' spans: a bin outside them means the tables predate it (the reference's
stale-table case) and raises RuntimeError before anything runs."""
import ctypes

import numpy as np

from . import _lib

__all__ = ['rows', 'rows_multi', 'cols', 'rank_coords', 'lattice_tables']

MAXD, MAXG = 8, 8
RADIAL_Q = 4


class PairTab(ctypes.Structure):
    _fields_ = [('d', ctypes.c_int32), ('G', ctypes.c_int32), ('f32', ctypes.c_int32), ('reserved', ctypes.c_int32),
                ('kind', ctypes.c_int32 * MAXD), ('grp', ctypes.c_int32 * MAXD), ('base', ctypes.c_int64 * MAXD),
                ('J', ctypes.c_int64 * MAXD), ('mk', ctypes.c_int64 * MAXD), ('nbins', ctypes.c_int64 * MAXD),
                ('nT', ctypes.c_int64), ('LR', ctypes.c_int64), ('T', ctypes.c_void_p), ('RT', ctypes.c_void_p)]


MAXW, MAXM = 16, 4


class PairGrad(ctypes.Structure):
    _fields_ = [('nW', ctypes.c_int32), ('U', ctypes.c_int32), ('reserved', ctypes.c_int32 * 2),
                ('wunit', ctypes.c_int32 * MAXW), ('wcol', ctypes.c_int32 * MAXW), ('wgrp', ctypes.c_int32 * MAXW),
                ('dim_unit', ctypes.c_int32 * MAXD), ('grp_unit', ctypes.c_int32 * MAXG),
                ('dbase', ctypes.c_int64 * MAXW), ('ndT', ctypes.c_int64), ('dT', ctypes.c_void_p),
                ('dR', ctypes.c_void_p)]


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("the pair kernels need a ROCm GPU (torch.cuda.is_available() is False)")
    return torch


def rank_coords(Z, tab):
    """Coordinates as the row kernels read them: a rank column's value replaced by its rank among the table's values
    (blml_kernels.py:203-217); a value the table lacks raises RuntimeError (stale tables)."""
    Z = np.array(Z, np.int64, copy=True, ndmin=2)
    for k, v in enumerate(tab['vals']):
        if v is None:
            continue
        v = np.asarray(v, np.int64)
        r = np.minimum(np.searchsorted(v, Z[:, k]), len(v) - 1)
        if len(Z) and not np.array_equal(v[r], Z[:, k]):
            raise RuntimeError(f"rank column {k}: a bin value is not in the kernel table (stale tables)")
        Z[:, k] = r
    return Z


def lattice_tables(kinds, fc, periods, h, Zi, want=()):
    """Value tables in the layout of the reference's ProductKernel.tables (blml_kernels.py:133-200) for the bins Zi
    (N, d) and per-column kinds: sinc / slits columns get an offset row K(o h), o = -J..J, J = ceil(1.25 span)
    (the reference's 25% headroom), or a rank table over their occupied values when that is smaller
    (len(vals)^2 < 2 J + 1, blml_trivial.py:405-413); dirichlet columns (folded on their period) a circular row
    K(m h), m = 0..P/h - 1; a periodic ball over its columns (pball) its torus table over the columns' circular
    offsets (row-major); a ball over its columns (one radial group per kind tag, blml_trivial.py radial_groups) its
    row on the grid r = j / RADIAL_Q lattice steps, j < floor(sqrt(sum span^2) RADIAL_Q) + 2, with fc h equal across
    the group (blml_trivial.py:284-303; the row kernels interpolate it linearly in r).

    Units: the one-column units in column order, then the multi-column units (balls and periodic balls) by kind tag
    in order of first appearance, group g being the g-th of them (the reference's kernel_units / radial_groups).
    want: derivative entries (unit, p) for rows(grad=True), p a parameter (dK/dlog p: 'fc' of sinc / slits / ball,
    'period' of dirichlet) or a coordinate shift ('shift' or ('shift', i), dK/dt_i), built as the reference's
    ProductKernel._assemble (blml_kernels.py:165-182): a column's row over raw offsets -J..J (its rank table for a
    rank column) at dT[dbase[q]:]; a ball's row on its r grid in dR[q] (a shift row divided by the offset of its
    column, which the row kernels multiply back, blml_trivial.py:317-327); a periodic ball's shift as its torus
    table of dK/dt_i."""
    from .modes import split_kind
    from .units import (one_column_values, multi_values, one_column_dvalues, shift_index, ball_profile,
                        ball_profile_dlog, pball_dshift)
    kinds = list(kinds)
    d = len(kinds)
    fc = np.broadcast_to(np.asarray(fc, float), (d,))
    periods = list(periods) if periods is not None else [None] * d
    h = np.broadcast_to(np.asarray(h, float), (d,))
    Zi = np.asarray(Zi, np.int64)
    span = np.maximum(Zi.max(0) - Zi.min(0), 1) if len(Zi) else np.ones(d, np.int64)
    alloc = np.ceil(span * 1.25).astype(np.int64)
    kd_arr = np.zeros(d, np.int64)
    base = np.zeros(d, np.int64)
    Jk = np.zeros(d, np.int64)
    mk = np.zeros(d, np.int64)
    nbins = np.zeros(d, np.int64)
    grp = np.full(d, -1, np.int64)
    dim_unit = np.full(d, -1, np.int64)
    vals = [None] * d
    rows, rrows, grp_unit, uinfo = [], [], [], []
    single = [k for k in range(d) if split_kind(kinds[k])[0] not in ('pball', 'ball')]
    # the multi-column units (balls and periodic balls) by kind tag, in order of first appearance; group g is the
    # g-th of them (blml_trivial.py:251-261, kernel_units :330-335)
    rgroups = {}
    for k in range(d):
        name, tag = split_kind(kinds[k])
        if name in ('ball', 'pball'):
            g = rgroups.setdefault(tag, (name, []))
            if g[0] != name:
                raise ValueError(f"radial group {tag!r} mixes kinds {g[0]!r} and {name!r}")
            g[1].append(k)
    at = 0
    for ui, k in enumerate(single):
        name = split_kind(kinds[k])[0]
        dim_unit[k] = ui
        uinfo.append(('col', k))
        Jk[k] = int(alloc[k])                     # every column's raw-offset half width (its derivative rows')
        if name in ('sinc', 'slits'):
            J = int(alloc[k])
            occ = np.unique(Zi[:, k])
            if len(occ) ** 2 < 2 * J + 1:
                row = one_column_values(kinds[k], fc[k], None, (occ[:, None] - occ[None, :]).ravel() * h[k])
                kd_arr[k], mk[k], vals[k] = 3, len(occ), occ
            else:
                row = one_column_values(kinds[k], fc[k], None, np.arange(-J, J + 1) * h[k])
                kd_arr[k] = 0
        elif name == 'dirichlet':
            nb = int(round(float(periods[k]) / h[k]))
            if abs(nb * h[k] - float(periods[k])) > 1e-9 * float(periods[k]):
                raise ValueError(f"column {k}: the period is not a whole number of bins")
            row = one_column_values(kinds[k], fc[k], periods[k], np.arange(nb) * h[k])
            kd_arr[k], nbins[k] = 1, nb
        else:
            raise NotImplementedError(f"lattice_tables: kind {kinds[k]!r} has no lattice row on this engine")
        base[k] = at
        rows.append(np.asarray(row, float))
        at += len(row)
    for name, dims in rgroups.values():
        g = len(rrows)
        grp[dims] = g
        grp_unit.append(len(single) + g)
        if name == 'pball':
            P = [float(periods[k]) for k in dims]
            ms = [int(round(p / h[k])) for p, k in zip(P, dims)]
            offs = np.stack([x.ravel() for x in np.meshgrid(*[np.arange(m) for m in ms], indexing='ij')], 1)
            rrows.append(multi_values('pball', [fc[k] for k in dims], P,
                                      [offs[:, i] * h[k] for i, k in enumerate(dims)]))
            kd_arr[dims] = 4
            nbins[dims] = ms
            uinfo.append(('pball', dims, P, ms, offs))
            continue
        c = fc[dims] * h[dims]
        if not np.allclose(c, c[0], rtol=1e-9, atol=0):
            raise ValueError(f"radial group dims {dims}: fc*h must be equal across the group, got {c}")
        r2max = int((alloc[dims] ** 2).sum())
        r = np.arange(int(np.floor(np.sqrt(r2max) * RADIAL_Q)) + 2, dtype=float) / RADIAL_Q
        diffs = [r * h[dims[0]]] + [np.zeros_like(r)] * (len(dims) - 1)     # along the first column: K is radial
        kd_arr[dims] = 2
        rrows.append(multi_values('ball', [fc[k] for k in dims], None, diffs))
        uinfo.append(('ball', dims, r, float(c[0])))
    nW = len(want)
    wunit = np.zeros(max(nW, 1), np.int64)
    wcol = np.full(max(nW, 1), -1, np.int64)
    wgrp = np.full(max(nW, 1), -1, np.int64)
    dbase = np.zeros(max(nW, 1), np.int64)
    drows, drr, dat = [], {}, 0
    for q, (ui, p) in enumerate(want):
        ui = int(ui)
        if not 0 <= ui < len(uinfo):
            raise ValueError(f"derivative entry {q}: unit {ui} of {len(uinfo)}")
        info = uinfo[ui]
        wunit[q] = ui
        i = shift_index(p)
        if info[0] == 'col':
            k = info[1]
            wcol[q] = k
            if kd_arr[k] == 3:
                t = (vals[k][:, None] - vals[k][None, :]).ravel() * h[k]
            else:
                t = np.arange(-Jk[k], Jk[k] + 1) * h[k]
            row = np.asarray(one_column_dvalues(kinds[k], fc[k], periods[k], p, t), float)
            dbase[q] = dat
            drows.append(row)
            dat += len(row)
        elif info[0] == 'pball':
            pc, P, ms, offs = info[1:]
            if i is None:
                raise ValueError("pball: no parameter gradient (fc in whole harmonics, fixed periods); shifts only")
            wgrp[q], wcol[q] = grp[pc[0]], pc[i]
            drr[q] = pball_dshift([fc[k] for k in pc], P, i, [offs[:, j] * h[k] for j, k in enumerate(pc)])
        else:
            dims, r, cg = info[1:]
            D = len(dims)
            sg = cg * r
            pf = float(np.prod(fc[dims]))
            wgrp[q] = grp[dims[0]]
            if i is None:
                if p != 'fc':
                    raise ValueError(f"ball has no parameter {p!r}")
                drr[q] = pf * (D * ball_profile(sg, D) + ball_profile_dlog(sg, D))
            else:
                wcol[q] = dims[i]
                out = pf * fc[dims[i]] * ball_profile_dlog(sg, D) / (cg * np.where(r > 0, r * r, 1.0))
                out[0] = out[1]                   # offset 0 there: the entry only bounds the interpolation
                drr[q] = out
    T = np.concatenate(rows) if rows else np.zeros(1)
    LR = max([len(r) for r in rrows + list(drr.values())] or [2])
    R = np.zeros((max(len(rrows), 1), LR))
    for g, r in enumerate(rrows):
        R[g, :len(r)] = r
    dR = np.zeros((max(nW, 1), LR))
    for q, r in drr.items():
        dR[q, :len(r)] = r
    return dict(T=T, base=base, Jk=Jk, mk=mk, vals=vals, kinds=kd_arr, nbins=nbins, grp=grp, R=R,
                dim_unit=dim_unit, grp_unit=np.asarray(grp_unit or [0], np.int64), U=len(single) + len(grp_unit),
                dT=np.concatenate(drows) if drows else np.zeros(1), dbase=dbase, dR=dR, wunit=wunit, wcol=wcol,
                wgrp=wgrp, nW=nW)


def _groups(tab):
    grp = np.asarray(tab['grp'])
    return int(np.atleast_2d(tab['R']).shape[0]) if (grp >= 0).any() else 0


def _check_spans(A, B, tab):
    """Every pair's table index inside its table (the kernels clamp; a clamp would be a stale table)."""
    if not len(A) or not len(B):
        return
    kinds, Jk, mk, nbins, grp = (np.asarray(tab[k]) for k in ('kinds', 'Jk', 'mk', 'nbins', 'grp'))
    lo_a, hi_a, lo_b, hi_b = A.min(0), A.max(0), B.min(0), B.max(0)
    far = np.maximum(hi_a - lo_b, hi_b - lo_a)                  # the largest |za - zb| per column
    G = _groups(tab)
    LR = np.atleast_2d(tab['R']).shape[1]
    r2max = np.zeros(max(G, 1), np.int64)
    for k in range(A.shape[1]):
        kd = int(kinds[k])
        if kd == 0 and far[k] > Jk[k]:
            raise RuntimeError(f"column {k}: offsets up to {far[k]} beyond the table's half width {Jk[k]} "
                               "(stale tables)")
        if kd == 3 and (max(hi_a[k], hi_b[k]) >= mk[k] or min(lo_a[k], lo_b[k]) < 0):
            raise RuntimeError(f"rank column {k}: rank outside the table's 0..{int(mk[k]) - 1} (stale tables)")
        if kd in (1, 4) and far[k] >= nbins[k] + (kd == 1):
            raise RuntimeError(f"column {k}: offsets beyond one period (stale tables)")
        if kd == 2:
            r2max[grp[k]] += int(far[k]) ** 2
    for g in range(G):
        tor = bool(np.any((kinds == 4) & (grp == g)))
        if not tor and np.sqrt(float(r2max[g])) * RADIAL_Q + 1 >= LR:
            raise RuntimeError(f"radial group {g}: distances beyond its row (stale tables)")


def _prepare(ZA, ZB, tab, f32, ranked):
    torch = _torch()
    A = np.ascontiguousarray(ZA if ranked else rank_coords(ZA, tab), np.int64)
    B = np.ascontiguousarray(ZB if ranked else rank_coords(ZB, tab), np.int64)
    A, B = np.atleast_2d(A), np.atleast_2d(B)
    d = A.shape[1]
    if d > MAXD or B.shape[1] != d:
        raise ValueError(f"coordinates: the same number of columns for ZA and ZB, at most {MAXD}")
    G = _groups(tab)
    if G > MAXG:
        raise ValueError(f"at most {MAXG} radial groups on this engine")
    for Z in (A, B):
        if len(Z) and (Z.min() < -(1 << 31) or Z.max() >= (1 << 31)):
            raise ValueError("lattice coordinates beyond int32")
    _check_spans(A, B, tab)
    rt = torch.float32 if f32 else torch.float64
    dT = torch.as_tensor(np.asarray(tab['T'], np.float64)).to('cuda', rt).contiguous()
    dR = torch.as_tensor(np.atleast_2d(np.asarray(tab['R'], np.float64))[:max(G, 1)]).to('cuda', rt).contiguous()
    t = PairTab()
    t.d, t.G, t.f32 = d, G, int(bool(f32))
    for k in range(d):
        t.kind[k] = int(tab['kinds'][k])
        t.grp[k] = int(tab['grp'][k])
        t.base[k] = int(tab['base'][k])
        t.J[k] = int(tab['Jk'][k])
        t.mk[k] = int(tab['mk'][k])
        t.nbins[k] = int(tab['nbins'][k])
    t.nT = dT.numel()
    t.LR = dR.shape[1]
    t.T = dT.data_ptr()
    t.RT = dR.data_ptr()
    dA = torch.from_numpy(A.astype(np.int32)).to('cuda')
    dB = torch.from_numpy(B.astype(np.int32)).to('cuda')
    return t, dA, dB, (dT, dR)


def _check_dspans(A, B, tab):
    """A column entry's derivative row is over raw offsets -J..J (or its rank table): every pair inside it."""
    if not len(A) or not len(B):
        return
    kinds, Jk = np.asarray(tab['kinds']), np.asarray(tab['Jk'])
    far = np.maximum(A.max(0) - B.min(0), B.max(0) - A.min(0))
    for q in range(int(tab['nW'])):
        k = int(tab['wcol'][q])
        if int(tab['wgrp'][q]) < 0 and kinds[k] != 3 and far[k] > Jk[k]:
            raise RuntimeError(f"derivative entry {q}: offsets up to {far[k]} in column {k} beyond its row's half "
                               f"width {Jk[k]} (stale tables)")


def _rows_fused(ZA, ZB, Wt, tab, umask, grad, f32, ranked):
    """s (R, Mm), ds (R, Mm, nW) through blml_pair_rows_grad (umask None: one density of every unit) or
    blml_pair_rows_multi."""
    launch, result = fused_rows_launcher(ZA, ZB, Wt, tab, umask, grad, f32, ranked)
    launch()
    return result()


def fused_rows_launcher(ZA, ZB, Wt, tab, umask=None, grad=True, f32=False, ranked=False):
    """(launch, result): the device inputs of _rows_fused uploaded once; launch() runs the kernels on the current
    stream (no host synchronisation), result() returns (s, ds) as _rows_fused does."""
    torch = _torch()
    nW, U = (int(tab['nW']) if grad else 0), int(tab['U'])
    Wt = np.ascontiguousarray(np.asarray(Wt, np.float64).reshape(len(np.atleast_2d(ZB)), -1))
    Mm = Wt.shape[1]
    if Mm > MAXM or Mm * nW > MAXW or U > MAXD:
        raise ValueError(f"pair rows: {U} kernel units (max {MAXD}), {nW} derivative entries x {Mm} densities "
                         f"(max {MAXW}), {Mm} densities (max {MAXM})")
    A = np.atleast_2d(np.asarray(ZA if ranked else rank_coords(ZA, tab), np.int64))
    B = np.atleast_2d(np.asarray(ZB if ranked else rank_coords(ZB, tab), np.int64))
    if A.shape[1] == B.shape[1] and nW:
        _check_dspans(A, B, tab)
    t, dA, dB, keep = _prepare(A, B, tab, f32, True)
    R, M = dA.shape[0], dB.shape[0]
    rt = torch.float32 if f32 else torch.float64
    ddT = torch.as_tensor(np.asarray(tab.get('dT', np.zeros(1)), np.float64).ravel()).to('cuda', rt).contiguous()
    dRh = np.atleast_2d(np.asarray(tab.get('dR', np.zeros((1, 2))), np.float64))
    if dRh.shape[1] != t.LR:                      # derivative rows on the value rows' length
        pad = np.zeros((dRh.shape[0], t.LR))
        n = min(t.LR, dRh.shape[1])
        pad[:, :n] = dRh[:, :n]
        dRh = pad
    ddR = torch.as_tensor(dRh).to('cuda', rt).contiguous()
    g = PairGrad()
    g.nW, g.U = nW, U
    for q in range(nW):
        g.wunit[q] = int(tab['wunit'][q])
        g.wcol[q] = int(tab['wcol'][q])
        g.wgrp[q] = int(tab['wgrp'][q])
        g.dbase[q] = int(tab['dbase'][q])
    for k in range(A.shape[1]):
        g.dim_unit[k] = int(tab['dim_unit'][k])
    for c in range(_groups(tab)):
        g.grp_unit[c] = int(tab['grp_unit'][c])
    g.ndT = ddT.numel()
    g.dT = ddT.data_ptr()
    g.dR = ddR.data_ptr()
    dw = torch.from_numpy(Wt.ravel().copy()).to('cuda')
    s = torch.zeros(max(R * Mm, 1), dtype=torch.float64, device='cuda')
    ds = torch.zeros(max(R * Mm * nW, 1), dtype=torch.float64, device='cuda')
    lib = _lib.load()
    P = ctypes.c_void_p
    um = None if umask is None else np.ascontiguousarray(np.asarray(umask, np.int64).reshape(Mm, U) != 0, np.int32)
    held = (keep, ddT, ddR)                       # device tables the kernels read: alive with either closure

    def launch():
        assert held is not None
        st = P(torch.cuda.current_stream().cuda_stream)
        if um is None:
            rc = lib.blml_pair_rows_grad(ctypes.byref(t), ctypes.byref(g), P(dA.data_ptr()), P(dB.data_ptr()),
                                         P(dw.data_ptr()), R, M, P(s.data_ptr()), P(ds.data_ptr()), st)
        else:
            rc = lib.blml_pair_rows_multi(ctypes.byref(t), ctypes.byref(g), P(dA.data_ptr()), P(dB.data_ptr()),
                                          P(dw.data_ptr()), um.ctypes.data_as(P), Mm, R, M, P(s.data_ptr()),
                                          P(ds.data_ptr()), st)
        _lib.check(rc)

    def result():
        assert held is not None
        out = s.cpu().numpy()[:R * Mm].reshape(R, Mm)
        dout = ds.cpu().numpy()[:R * Mm * nW].reshape(R, Mm, nW)
        if nW == 0:
            dout = np.zeros((R, Mm, max(int(tab.get('nW', 0)), 1)))
        return out, dout

    return launch, result


def rows_multi(ZA, ZB, Wt, tab, umask, grad=False, f32=False, ranked=False):
    """Several densities in one pass (blml_local_gpu.py:107-180, 432-439): s (R, Mm) with s[r, m] = sum_j
    prod_{u: umask[m, u]} fac_u(ZA_r, ZB_j) Wt[j, m] (the joint and marginals: units left out of a density count
    as 1), and ds (R, Mm, max(nW, 1)): with grad, each density's derivative rows (0 where its mask leaves the
    entry's unit out), else zeros.  At most 4 densities and 16 densities x derivative entries."""
    Wt = np.asarray(Wt, np.float64)
    if Wt.ndim != 2 or Wt.shape[0] != len(np.atleast_2d(ZB)):
        raise ValueError("Wt: (M, densities), one row per row of ZB")
    if np.shape(umask) != (Wt.shape[1], int(tab['U'])):
        raise ValueError(f"umask: (densities, units) = ({Wt.shape[1]}, {int(tab['U'])})")
    return _rows_fused(ZA, ZB, Wt, tab, umask, grad, f32, ranked)


def rows(ZA, ZB, w, tab, grad=False, f32=False, ranked=False):
    """(s, ds): s (R,) = sum over j with w_j != 0 of S(ZA_r, ZB_j) w_j (FP64); ds (R, max(nW, 1)): with grad and
    derivative entries in the tables (lattice_tables(want=...)), ds[r, q] = sum_j dS_q(ZA_r, ZB_j) w_j, dS_q entry q's
    derivative of its unit's factor times the other units' factors (blml_local_gpu.py:94-103, 423-429); else zeros."""
    if grad and int(tab.get('nW', 0)) > 0:
        wv = np.ascontiguousarray(w, np.float64).ravel()
        if len(wv) != len(np.atleast_2d(ZB)):
            raise ValueError("w: one weight per row of ZB")
        s, ds = _rows_fused(ZA, ZB, wv[:, None], tab, None, True, f32, ranked)
        return s[:, 0], ds[:, 0, :]
    torch = _torch()
    t, dA, dB, keep = _prepare(ZA, ZB, tab, f32, ranked)
    R, M = dA.shape[0], dB.shape[0]
    wv = np.ascontiguousarray(w, np.float64).ravel()
    if len(wv) != M:
        raise ValueError("w: one weight per row of ZB")
    dw = torch.from_numpy(wv).to('cuda')
    s = torch.zeros(max(R, 1), dtype=torch.float64, device='cuda')
    lib = _lib.load()
    _lib.check(lib.blml_pair_rows(ctypes.byref(t), ctypes.c_void_p(dA.data_ptr()), ctypes.c_void_p(dB.data_ptr()),
                                  ctypes.c_void_p(dw.data_ptr()), R, M, ctypes.c_void_p(s.data_ptr()),
                                  ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
    out = s.cpu().numpy()[:R]
    del keep
    return out, np.zeros((R, max(int(tab.get('nW', 0)), 1)))


def cols(ZA, ZB, tab, out=None, out_m=None, depmask=0, f32=False, ranked=False):
    """The block S(ZA_r, ZB_j) (R, M) in the tables' precision; with depmask also the marginal over the columns not
    in depmask (returned as (out, out_m)).  out / out_m: optional host arrays to fill."""
    torch = _torch()
    t, dA, dB, keep = _prepare(ZA, ZB, tab, f32, ranked)
    R, M = dA.shape[0], dB.shape[0]
    rt = torch.float32 if f32 else torch.float64
    dO = torch.empty(max(R * M, 1), dtype=rt, device='cuda')
    dM = torch.empty(max(R * M, 1), dtype=rt, device='cuda') if depmask else None
    lib = _lib.load()
    _lib.check(lib.blml_pair_cols(ctypes.byref(t), ctypes.c_void_p(dA.data_ptr()), ctypes.c_void_p(dB.data_ptr()),
                                  R, M, int(depmask), ctypes.c_void_p(dO.data_ptr()),
                                  ctypes.c_void_p(dM.data_ptr() if dM is not None else 0),
                                  ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
    res = dO.cpu().numpy()[:R * M].reshape(R, M)
    if out is not None:
        out[...] = res
        res = out
    del keep
    if not depmask:
        return res
    resm = dM.cpu().numpy()[:R * M].reshape(R, M)
    if out_m is not None:
        out_m[...] = resm
        resm = out_m
    return res, resm
