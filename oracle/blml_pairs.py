"""CPU FP64 restatement of the reference's pair accumulation -- TEST INFRASTRUCTURE ONLY (the checker of the HIP pair
kernels; imported by tests and the benchmark's CPU baseline, never by the product).

Restated (numpy, float64):
  rank_coords(Z, tab)        blml_kernels.py:203-217 (rank columns -> ranks among the table's values; stale -> error)
  pair_rows(ZA, ZB, w, tab)  blml_local.py:203-208 with _rows_numba's pair loop (blml_local.py:230-253, no
                             derivatives): s_r = sum_{j: w_j != 0} prod_u fac_u(r, j) w_j, the unit factors from
                             _col_index (blml_local.py:86-96), radial rows by linear interpolation on the r grid
                             (RADIAL_Q = 4, blml_trivial.py:218), torus tables at the row-major circular offset
  pair_cols(ZA, ZB, tab, depmask)   blml_local.py:271-282 / blml_local_gpu.py:331-361 (joint and marginal)
  pair_rows_grad(ZA, ZB, w, tab)   with grad (blml_local.py:99-116 _dval, 146-191; blml_local_gpu.py:21-40, 94-103):
                             ds[r, q] = sum_j dval_q(r, j) prod_{u != wunit[q]} fac_u(r, j) w_j
  pair_rows_multi(ZA, ZB, Wt, tab, umask, grad)   blml_local.py:212-268 / blml_local_gpu.py:107-180: one density per
                             column of Wt, its product over the units its mask row keeps
Table builders for the tests (the rows ProductKernel.tables holds, blml_kernels.py:133-200, for the kinds with
closed forms): offset rows K(o h), o = -J..J (sinc, slits, unfolded dirichlet); circular rows K(m h), m = 0..nb-1
(folded dirichlet); rank tables K((v_a - v_b) h) over the occupied values; a periodic ball's torus table over its
columns' circular offsets.

Pinning (tests/test_oracle_pairs.py): pair_rows / pair_cols against the dense Gram gram_int @ w
(blml_kernels.py:243-251; the reference's own check, results/pball_parity.py:66-76) for tables built from the
kernels' closed forms; the lookup of radial rows against a brute-force per-pair loop of the same rule; pair_rows_grad
over the derivative tables of blml_amd.local.lattice_tables against central finite differences of the dense Gram
(the kernels' closed forms at fc e^+-eps, or at coordinate offsets +-eps), which pins the tables and the lookup;
pair_rows_multi's marginal densities against the dense Gram of their units alone (and its finite differences)."""
import numpy as np

from oracle.blml_oracle import kernel_eval

RADIAL_Q = 4
KIND = dict(offset=0, circular=1, radial=2, rank=3, torus=4)


def rank_coords(Z, tab):
    """Coordinates as the row kernels read them: a rank column's value replaced by its rank in tab['vals'][k]."""
    Z = np.array(Z, np.int64, copy=True)
    for k, v in enumerate(tab['vals']):
        if v is None:
            continue
        v = np.asarray(v, np.int64)
        r = np.minimum(np.searchsorted(v, Z[:, k]), len(v) - 1)
        if len(Z) and not np.array_equal(v[r], Z[:, k]):
            raise RuntimeError(f"rank column {k}: a bin value is not in the kernel table (stale tables)")
        Z[:, k] = r
    return Z


def _factors(A, B, tab):
    """Per-unit factor arrays (U, nA, nB) of every pair: column units from T, groups from R."""
    A = np.asarray(A, np.int64)
    B = np.asarray(B, np.int64)
    d = A.shape[1]
    kinds, base, Jk, mk, nbins, grp = (np.asarray(tab[k]) for k in ('kinds', 'base', 'Jk', 'mk', 'nbins', 'grp'))
    T = np.asarray(tab['T'], float)
    Rt = np.atleast_2d(np.asarray(tab['R'], float))
    U = int(tab['U'])
    fac = np.ones((U, len(A), len(B)))
    G = int((grp >= 0).any() and Rt.shape[0])
    r2 = np.zeros((max(G, 1), len(A), len(B)), np.int64)
    tor = np.zeros(max(G, 1), bool)
    for k in range(d):
        za, zb = A[:, k][:, None], B[:, k][None, :]
        off = za - zb
        kd = int(kinds[k])
        if kd == 2:
            r2[grp[k]] += off * off
        elif kd == 4:
            nb = int(nbins[k])
            r2[grp[k]] = r2[grp[k]] * nb + np.mod(off, nb)
            tor[grp[k]] = True
        else:
            if kd == 0:
                idx = base[k] + off + Jk[k]
            elif kd == 3:
                idx = base[k] + za * mk[k] + zb
            else:
                nb = int(nbins[k])
                m = np.mod(off, nb)
                idx = base[k] + np.minimum(m, nb - m)
            fac[int(tab['dim_unit'][k])] = T[idx]
    for g in range(G):
        u = int(tab['grp_unit'][g])
        if tor[g]:
            fac[u] = Rt[g][r2[g]]
        else:
            t = np.sqrt(r2[g] * 1.0) * RADIAL_Q
            i0 = t.astype(np.int64)
            f = t - i0
            fac[u] = Rt[g][i0] * (1.0 - f) + Rt[g][i0 + 1] * f
    return fac


def pair_rows(ZA, ZB, w, tab, chunk=512):
    """s (R,): sum over sources with w_j != 0 of the unit-order product times w_j (float64)."""
    A, B = rank_coords(ZA, tab), rank_coords(ZB, tab)
    w = np.asarray(w, float)
    keep = w != 0
    B, w = B[keep], w[keep]
    s = np.zeros(len(A))
    for a0 in range(0, len(A), chunk):
        fac = _factors(A[a0:a0 + chunk], B, tab)
        prod = np.ones(fac.shape[1:])
        for u in range(fac.shape[0]):
            prod = prod * fac[u]
        s[a0:a0 + chunk] = prod @ w
    return s


def pair_rows_c(ZA, ZB, w, tab, f32=False, threads=None):
    """pair_rows by the C restatement (oracle/pairs_cpu.c, OpenMP over rows) for tables of offset, circular and rank
    columns (no groups): the FP32 products of an f32 table (REAL = float, as the reference's GPU rows), FP64 sums.
    Builds oracle/libpairs_cpu.so with make when it is missing."""
    import ctypes
    import os
    import subprocess
    here = os.path.dirname(os.path.abspath(__file__))
    path = os.path.join(here, 'libpairs_cpu.so')
    if not os.path.exists(path):
        subprocess.run(['make', '-s', '-C', here], check=True)
    lib = ctypes.CDLL(path)
    A = np.ascontiguousarray(rank_coords(ZA, tab), np.int32)
    B = np.ascontiguousarray(rank_coords(ZB, tab), np.int32)
    d = A.shape[1]
    i64 = lambda k: np.ascontiguousarray(np.asarray(tab[k], np.int64))
    kinds = np.ascontiguousarray(np.asarray(tab['kinds'], np.int32))
    base, Jk, mk, nb = i64('base'), i64('Jk'), i64('mk'), i64('nbins')
    T = np.ascontiguousarray(np.asarray(tab['T'], np.float32 if f32 else np.float64))
    w = np.ascontiguousarray(np.asarray(w, np.float64))
    s = np.zeros(len(A))
    ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    nt = int(threads or len(os.sched_getaffinity(0)))
    rc = lib.pair_rows_cpu(ctypes.c_int(d), ptr(kinds), ptr(base), ptr(Jk), ptr(mk), ptr(nb),
                           ctypes.c_int64(len(T)), ptr(T), ctypes.c_int(int(f32)), ptr(A), ptr(B), ptr(w),
                           ctypes.c_int64(len(A)), ctypes.c_int64(len(B)), ptr(s), ctypes.c_int(nt))
    if rc:
        raise ValueError("pair_rows_c: radial / torus columns are not restated in C")
    return s


def _group_r2(A, B, tab, g):
    """The group's r^2 (radial) or row-major circular offset (torus) of every pair, and whether it is a torus."""
    kinds, nbins, grp = (np.asarray(tab[k]) for k in ('kinds', 'nbins', 'grp'))
    r2 = np.zeros((len(A), len(B)), np.int64)
    tor = False
    for k in range(A.shape[1]):
        if grp[k] != g:
            continue
        off = A[:, k][:, None] - B[:, k][None, :]
        if kinds[k] == 4:
            r2 = r2 * int(nbins[k]) + np.mod(off, int(nbins[k]))
            tor = True
        else:
            r2 = r2 + off * off
    return r2, tor


def _dval(q, A, B, tab):
    """Entry q's derivative value of every pair (blml_local.py:99-116 _dval): a column's row (raw offsets, or its rank
    table), a group's row on the r grid (times the offset of column wcol[q] for a shift entry), or a torus table
    read as is."""
    g, k = int(tab['wgrp'][q]), int(tab['wcol'][q])
    if g >= 0:
        row = np.atleast_2d(np.asarray(tab['dR'], float))[q]
        r2, tor = _group_r2(A, B, tab, g)
        if tor:
            return row[r2]
        t = np.sqrt(r2 * 1.0) * RADIAL_Q
        i0 = t.astype(np.int64)
        f = t - i0
        v = row[i0] * (1.0 - f) + row[i0 + 1] * f
        if k >= 0:
            v = v * (A[:, k][:, None] - B[:, k][None, :])
        return v
    dT = np.asarray(tab['dT'], float)
    za, zb = A[:, k][:, None], B[:, k][None, :]
    at = int(tab['dbase'][q])
    if int(tab['kinds'][k]) == 3:
        return dT[at + za * int(tab['mk'][k]) + zb]
    return dT[at + za - zb + int(tab['Jk'][k])]


def pair_rows_grad(ZA, ZB, w, tab, chunk=256):
    """(s, ds): pair_rows and ds[r, q] = sum over sources with w_j != 0 of entry q's derivative value times the other
    units' factors, times w_j (blml_local.py:146-191 _rows_numba with grad; blml_local_gpu.py:94-103)."""
    A, B = rank_coords(ZA, tab), rank_coords(ZB, tab)
    w = np.asarray(w, float)
    keep = w != 0
    B, w = B[keep], w[keep]
    nW = int(tab['nW'])
    s = np.zeros(len(A))
    ds = np.zeros((len(A), max(nW, 1)))
    for a0 in range(0, len(A), chunk):
        Ac = A[a0:a0 + chunk]
        fac = _factors(Ac, B, tab)
        s[a0:a0 + chunk] = np.prod(fac, axis=0) @ w
        for q in range(nW):
            u = int(tab['wunit'][q])
            ex = np.prod(np.delete(fac, u, axis=0), axis=0)
            ds[a0:a0 + chunk, q] = (_dval(q, Ac, B, tab) * ex) @ w
    return s, ds


def pair_rows_multi(ZA, ZB, Wt, tab, umask, grad=False, chunk=256):
    """(s (R, Mm), ds (R, Mm, max(nW, 1))): several densities at once (blml_local.py:212-268 _rows_numba_m,
    blml_local_gpu.py:107-180): density m multiplies the units with umask[m, u] != 0 (the others count as 1) with
    weights Wt[:, m]; a derivative entry whose unit the mask leaves out stays 0."""
    A, B = rank_coords(ZA, tab), rank_coords(ZB, tab)
    Wt = np.asarray(Wt, float)
    umask = np.asarray(umask) != 0
    keep = (Wt != 0).any(1)
    B, Wt = B[keep], Wt[keep]
    Mm, nW = Wt.shape[1], int(tab['nW'])
    s = np.zeros((len(A), Mm))
    ds = np.zeros((len(A), Mm, max(nW, 1)))
    for a0 in range(0, len(A), chunk):
        Ac = A[a0:a0 + chunk]
        fac = _factors(Ac, B, tab)
        units = np.arange(fac.shape[0])
        dv = [_dval(q, Ac, B, tab) for q in range(nW)] if grad else []
        for m in range(Mm):
            s[a0:a0 + chunk, m] = np.prod(fac[umask[m]], axis=0) @ Wt[:, m]
            for q in range(len(dv)):
                u = int(tab['wunit'][q])
                if umask[m, u]:
                    ex = np.prod(fac[umask[m] & (units != u)], axis=0)
                    ds[a0:a0 + chunk, m, q] = (dv[q] * ex) @ Wt[:, m]
    return s, ds


def pair_cols(ZA, ZB, tab, depmask=0):
    """(out, out_m): the block S(ZA_r, ZB_j) and, with depmask, the product over the columns not in depmask (a
    group on the dep side when any of its columns is) -- column order, then the groups (blml_local_gpu.py:341-357)."""
    A, B = rank_coords(ZA, tab), rank_coords(ZB, tab)
    fac = _factors(A, B, tab)
    d = A.shape[1]
    dep_unit = np.zeros(fac.shape[0], bool)
    grp = np.asarray(tab['grp'])
    for k in range(d):
        if (depmask >> k) & 1:
            dep_unit[int(tab['dim_unit'][k]) if grp[k] < 0 else int(tab['grp_unit'][grp[k]])] = True
    v = np.ones(fac.shape[1:])
    vd = np.ones(fac.shape[1:])
    for u in range(fac.shape[0]):
        if dep_unit[u]:
            vd = vd * fac[u]
        else:
            v = v * fac[u]
    return v * vd, v


# ------------------------------------------------------------------------------------------------ table builders
def build_tables(units, h, J, vals=None):
    """A table dict in ProductKernel.tables' layout for per-column units and periodic-ball groups.
    units: list of (code, kind, fc_cols, per_cols, dims): code 'offset' | 'circular' | 'rank' | 'torus', kind the
    kernel_eval kind string ('sinc', 'slits:0.5', 'dirichlet', 'pball' ...), fc_cols / per_cols its cut-offs and
    periods (0.0 where none), dims its columns; h: per-column step; J: per-column half width of offset rows; vals:
    per-column occupied values of rank columns (else None)."""
    d = len(h)
    kinds = np.zeros(d, np.int64)
    base = np.zeros(d, np.int64)
    Jk = np.zeros(d, np.int64)
    mk = np.zeros(d, np.int64)
    nbins = np.zeros(d, np.int64)
    grp = np.full(d, -1, np.int64)
    dim_unit = np.full(d, -1, np.int64)
    grp_unit = []
    rows, rrows = [], []
    rvals = [None] * d
    at = 0
    for ui, (code, kind, fc_cols, per_cols, dims) in enumerate(units):
        if code == 'torus':
            g = len(rrows)
            P = np.asarray(per_cols, float)
            ms = [int(round(p / h[k])) for p, k in zip(P, dims)]
            offs = np.stack([g_.ravel() for g_ in np.meshgrid(*[np.arange(m) for m in ms], indexing='ij')], 1)
            diffs = [offs[:, i] * h[k] for i, k in enumerate(dims)]
            rrows.append(kernel_eval(kind, np.asarray(fc_cols, float), P, diffs))
            grp[list(dims)] = g
            kinds[list(dims)] = KIND['torus']
            nbins[list(dims)] = ms
            grp_unit.append(ui)
            continue
        k = dims[0]
        dim_unit[k] = ui
        if code == 'offset':
            row = kernel_eval(kind, np.asarray(fc_cols, float), np.asarray(per_cols, float),
                              [np.arange(-J[k], J[k] + 1) * h[k]])
            Jk[k] = J[k]
        elif code == 'circular':
            nb = int(round(float(per_cols[0]) / h[k]))
            row = kernel_eval(kind, np.asarray(fc_cols, float), np.asarray(per_cols, float), [np.arange(nb) * h[k]])
            nbins[k] = nb
        else:
            v = np.asarray(vals[k], np.int64)
            row = kernel_eval(kind, np.asarray(fc_cols, float), np.asarray(per_cols, float),
                              [((v[:, None] - v[None, :]).ravel() * h[k]).astype(float)])
            mk[k] = len(v)
            rvals[k] = v
        kinds[k] = KIND[code]
        base[k] = at
        rows.append(np.asarray(row, float))
        at += len(row)
    T = np.concatenate(rows) if rows else np.zeros(1)
    LR = max([len(r) for r in rrows] or [2])
    R = np.zeros((max(len(rrows), 1), LR))
    for g, r in enumerate(rrows):
        R[g, :len(r)] = r
    return dict(T=T, base=base, Jk=Jk, mk=mk, vals=rvals, kinds=kinds, nbins=nbins, grp=grp, R=R,
                dim_unit=dim_unit, grp_unit=np.asarray(grp_unit or [0], np.int64), U=len(units), nW=0)


def dense_pair_gram(units, h, ZA, ZB):
    """gram_int (blml_kernels.py:243-251): prod over units of K(diffs) at lattice differences times h."""
    ZA, ZB = np.atleast_2d(ZA), np.atleast_2d(ZB)
    G = np.ones((len(ZA), len(ZB)))
    for code, kind, fc_cols, per_cols, dims in units:
        diffs = [(ZA[:, k, None] - ZB[None, :, k]) * h[k] for k in dims]
        G *= kernel_eval(kind, np.asarray(fc_cols, float), np.asarray(per_cols, float), diffs)
    return G
