"""Kernel units as lattice operators: the host side of plan creation.

Each mode's kernel is a product of units (blml_trivial.py:330-335).  On the fit's lattice every unit becomes one
factor of the engine (include/blml_hip.h, blml_factor_t):
  * a one-column unit on a padded axis: its even row K(o h), o = 0..m-1 (sinc blml_trivial.py:368-370, slits
    :442-444, dirichlet :509-515), convolved by a power-of-two FFT of length >= 2m - 1;
  * a periodic unit on circular axes (dirichlet :501-515, pball :642-665): its spectrum is a harmonic indicator,
    given as the band's frequencies per axis and (count of harmonics) / prod P on their product grid;
  * a multi-column unit on padded axes (ball :582-591, or pball whose period is not the box): its table on the
    power-of-two lattice, masked to |off| <= m - 1 as the reference's _offsets (blml_binned.py:63-72).
All values are computed here in FP64 from the kernel formulas; the device turns rows and tables into spectra."""
import numpy as np

from .modes import split_kind, units_of

PERIODIC = ('dirichlet', 'pball')


def pow2_at_least(n):
    p = 2
    while p < n:
        p <<= 1
    return p


def next_fast_len5(n):
    """Smallest 2^a 3^b 5^c >= n: the reference's pad per padded dim (blml_binned.py:134, 1533-1546)."""
    best = pow2_at_least(n)
    a = 1
    while a <= best:
        b = a
        while b <= best:
            c = b
            while c < n:
                c *= 5
            best = min(best, c)
            b *= 3
        a *= 2
    return best


def dirichlet_N(fc, P):
    return max(0, int(np.floor(fc * P / 2.0 + 1e-9)))


def dirichlet_Mi(fc, P, fill):
    if fill is None or fill >= 1.0:
        return 0
    return max(2 * (int(np.ceil((1.0 - fill) * fc * P / 2.0 - 1e-9)) - 1) + 1, 0)


def pball_harmonics(fc, periods):
    """(H, D) integer harmonics with sum (k_i / P_i)^2 <= (fc / 2)^2, and 1 / prod P (blml_trivial.py:632-639)."""
    P = np.asarray(periods, float)
    R = float(fc) / 2.0
    N = [max(0, int(np.floor(R * p + 1e-9))) for p in P]
    axes = [np.arange(-v, v + 1) for v in N]
    grid = np.stack([g.ravel() for g in np.meshgrid(*axes, indexing='ij')], 1)
    keep = ((grid / P) ** 2).sum(1) <= R * R * (1 + 1e-9) + 1e-12
    return grid[keep], 1.0 / float(np.prod(P))


def one_column_values(kind, fc, period, t):
    """K of a one-column unit at coordinate offsets t."""
    name, tag = split_kind(kind)
    t = np.asarray(t, float)
    if name == 'sinc':
        return fc * np.sinc(fc * t)
    if name == 'slits':
        fi = (1.0 - float(tag)) * fc
        return fc * np.sinc(fc * t) - fi * np.sinc(fi * t)
    if name == 'dirichlet':
        fill = float(tag) if tag else None
        P = float(period)
        M = 2 * dirichlet_N(fc, P) + 1
        Mi = dirichlet_Mi(fc, P, fill)
        u = np.mod(np.abs(t), P)
        u = np.minimum(u, P - u)
        with np.errstate(divide='ignore', invalid='ignore'):
            v = (np.sin(np.pi * M * u / P) - np.sin(np.pi * Mi * u / P)) / np.sin(np.pi * u / P)
        return (1.0 / P) * np.where(u < 1e-10, float(M - Mi), v)
    raise ValueError(f"not a one-column kind: {kind!r}")


def ball_profile(s, D):
    """FT of the ball |nu| <= 1/2 in D dims (blml_trivial.py:135-143)."""
    from scipy.special import gamma, jv
    s = np.asarray(s, float)
    out = np.full(s.shape, np.pi ** (D / 2) * 0.5 ** D / gamma(D / 2 + 1))
    nz = s > 0
    out[nz] = 0.5 ** (D / 2) * jv(D / 2, np.pi * s[nz]) / s[nz] ** (D / 2)
    return out


def multi_values(name, fcs, periods, diffs):
    """K of a multi-column unit (ball / pball) at per-column offsets."""
    if name == 'ball':
        s = np.sqrt(sum((f * dk) ** 2 for f, dk in zip(fcs, diffs)))
        return float(np.prod(fcs)) * ball_profile(s, len(fcs))
    H, sc = pball_harmonics(fcs[0], periods)
    acc = 0.0 * sum(diffs)
    for k in H:
        ph = 0.0
        for i, P in enumerate(periods):
            if k[i]:
                ph = ph + (2 * np.pi * k[i] / P) * diffs[i]
        acc = acc + np.cos(ph)
    return sc * acc


def shift_index(p):
    """The column (within its unit) of a coordinate derivative p = 'shift' (0) or ('shift', i); None for a parameter
    (blml_trivial.py:357-361)."""
    if isinstance(p, str):
        return 0 if p == 'shift' else None
    return int(p[1]) if isinstance(p, tuple) and len(p) == 2 and p[0] == 'shift' else None


def _sinc_dt(f, t):
    """dK/dt of fc sinc(fc t): (pi f t cos(pi f t) - sin(pi f t)) / (pi t^2), the series -f (pi f)^2 t / 3 where
    |pi f t| < 1e-4 (blml_trivial.py:379-387)."""
    pt = np.pi * f * t
    small = np.abs(pt) < 1e-4
    ts = np.where(small, 1.0, t)
    with np.errstate(divide='ignore', invalid='ignore'):
        full = (pt * np.cos(pt) - np.sin(pt)) / (np.pi * ts * ts)
    return np.where(small, -f * (np.pi * f) ** 2 * t / 3.0, full)


def _dirichlet_D(M, u):
    """D(u) = sin(M u / 2) / sin(u / 2) and D'(u), with D = M, D' = 0 where |sin(u / 2)| < 1e-9
    (blml_trivial.py:500-507)."""
    s2 = np.sin(u / 2)
    small = np.abs(s2) < 1e-9
    ss = np.where(small, 1.0, s2)
    Dv = np.where(small, float(M), np.sin(M * u / 2) / ss)
    dD = np.where(small, 0.0, (0.5 * M * np.cos(M * u / 2) * ss - 0.5 * np.sin(M * u / 2) * np.cos(u / 2)) / ss ** 2)
    return Dv, dD


def one_column_dvalues(kind, fc, period, p, t):
    """A derivative of a one-column unit at raw coordinate offsets t: dK/dlog p for a parameter p, dK/dt for
    p = 'shift' (sinc blml_trivial.py:372-387, 392-395; slits :445-460; dirichlet :517-543).  Parameters without a
    gradient (slits / dirichlet fill, dirichlet fc: whole harmonics) raise ValueError."""
    name, tag = split_kind(kind)
    t = np.asarray(t, float)
    shift = shift_index(p) is not None
    if name == 'sinc':
        if shift:
            return _sinc_dt(fc, t)
        if p != 'fc':
            raise ValueError(f"sinc has no parameter {p!r}")
        return fc * np.cos(np.pi * fc * t)
    if name == 'slits':
        fi = (1.0 - float(tag)) * fc
        if shift:
            return _sinc_dt(fc, t) - _sinc_dt(fi, t) if fi > 0 else _sinc_dt(fc, t)
        if p != 'fc':
            raise ValueError("slits fill is a fixed shape: no gradient" if p == 'fill' else
                             f"slits has no parameter {p!r}")
        return fc * np.cos(np.pi * fc * t) - fi * np.cos(np.pi * fi * t)
    if name == 'dirichlet':
        fill = float(tag) if tag else None
        P = float(period)
        M = 2 * dirichlet_N(fc, P) + 1
        Mi = dirichlet_Mi(fc, P, fill)
        u = 2 * np.pi * t / P
        Dv, dD = _dirichlet_D(M, u)
        if Mi:
            Di, dDi = _dirichlet_D(Mi, u)
            Dv, dD = Dv - Di, dD - dDi
        if shift:
            return (1.0 / P) * dD * (2 * np.pi / P)
        if p != 'period':
            raise ValueError("dirichlet fc moves in whole harmonics: no gradient" if p == 'fc' else
                             "dirichlet fill is a fixed shape: no gradient" if p == 'fill' else
                             f"dirichlet has no parameter {p!r}")
        return -(Dv + u * dD) / P
    raise ValueError(f"not a one-column kind: {kind!r}")


def ball_profile_dlog(s, D):
    """s p'(s) of the ball profile: -pi (1/2)^(D/2) s^(1 - D/2) J_{D/2 + 1}(pi s), 0 at s = 0
    (blml_trivial.py:146-155)."""
    from scipy.special import jv
    s = np.asarray(s, float)
    out = np.zeros(s.shape)
    nz = s > 0
    out[nz] = -np.pi * 0.5 ** (D / 2) * s[nz] ** (1 - D / 2) * jv(D / 2 + 1, np.pi * s[nz])
    return out


def pball_dshift(fcs, periods, i, diffs):
    """dK/dt_i of a periodic ball at per-column offsets: -(1 / prod P) sum_k (2 pi k_i / P_i) sin(2 pi sum_j k_j t_j
    / P_j) (blml_trivial.py:672-689)."""
    H, sc = pball_harmonics(fcs[0], periods)
    acc = 0.0 * sum(diffs)
    for k in H:
        if not k[i]:
            continue
        ph = 0.0
        for j, P in enumerate(periods):
            if k[j]:
                ph = ph + (2 * np.pi * k[j] / P) * diffs[j]
        acc = acc - (2 * np.pi * k[i] / periods[i]) * np.sin(ph)
    return sc * acc


def unit_at_zero(kind, dims, spec):
    name = split_kind(kind)[0]
    fc = np.asarray(spec['fc'], float)
    if name in ('ball', 'pball'):
        per = [spec['periods'][j] for j in dims]
        return float(multi_values(name, fc[dims], per, [np.zeros(1)] * len(dims))[0])
    return float(one_column_values(kind, float(fc[dims[0]]), spec['periods'][dims[0]], np.zeros(1))[0])


def circular_dims(specs, shape, h):
    """Per dim: circular (pad = m) iff every mode is periodic there with period m h (blml_binned.py:130-135)."""
    out = []
    for j, m in enumerate(shape):
        out.append(all(split_kind(sp['kinds'][j])[0] in PERIODIC and sp['periods'][j] is not None and
                       abs(m * h[j] - sp['periods'][j]) <= 1e-9 * sp['periods'][j] for sp in specs))
    return out


def mode_factors(spec, shape, h, circ):
    """(s = K(0), [factor dicts]) of one mode.  A factor dict: kind ('line' | 'group'), axes, circ per axis, tlen
    per axis, freqs per circular axis (int arrays), data (float64 array)."""
    d = len(shape)
    s = 1.0
    facs = []
    for kind, dims in units_of(spec['kinds']):
        name = split_kind(kind)[0]
        s = s * unit_at_zero(kind, dims, spec)
        cflags = [bool(circ[j]) for j in dims]
        if any(cflags) and not all(cflags):
            raise ValueError(f"unit {kind!r} over dims {dims}: some of its dims are circular (period = box) and some "
                             "are not; the engine needs all or none")
        fc = np.asarray(spec['fc'], float)
        if len(dims) == 1 and name not in ('ball', 'pball'):
            j = dims[0]
            m = shape[j]
            if cflags[0]:
                P = float(spec['periods'][j])
                N = dirichlet_N(fc[j], P)
                Mi = dirichlet_Mi(fc[j], P, float(split_kind(kind)[1]) if split_kind(kind)[1] else None)
                lo = (Mi + 1) // 2 if Mi else 0          # |k| >= lo kept (Mi = 2 Ni + 1 removed)
                ks = np.array([k for k in range(-N, N + 1) if abs(k) >= lo], dtype=np.int64)
                F, cnt = np.unique(np.mod(ks, m), return_counts=True)
                facs.append(dict(kind='line', axes=[j], circ=[1], tlen=[len(F)], freqs=[F.astype(np.int32)],
                                 data=cnt.astype(float) / P))
            else:
                L = pow2_at_least(2 * m - 1)
                row = one_column_values(kind, fc[j], spec['periods'][j], np.arange(m) * h[j])
                facs.append(dict(kind='line', axes=[j], circ=[0], tlen=[L], freqs=[None], data=row.astype(float)))
            continue
        per = [spec['periods'][j] for j in dims]
        if all(cflags):
            if name != 'pball':
                raise ValueError(f"unit {kind!r}: circular dims need a periodic kind")
            H, sc = pball_harmonics(fc[dims[0]], per)
            red = np.stack([np.mod(H[:, a], shape[j]) for a, j in enumerate(dims)], 1)
            Fs = [np.unique(red[:, a]) for a in range(len(dims))]
            idx = [np.searchsorted(Fs[a], red[:, a]) for a in range(len(dims))]
            dense = np.zeros([len(F) for F in Fs])
            np.add.at(dense, tuple(idx), 1.0)
            facs.append(dict(kind='group', axes=list(dims), circ=[1] * len(dims), tlen=[len(F) for F in Fs],
                             freqs=[F.astype(np.int32) for F in Fs], data=(dense * sc).ravel()))
        else:
            Ls = [pow2_at_least(2 * shape[j] - 1) for j in dims]
            offs = []
            for a, j in enumerate(dims):
                L = Ls[a]
                i = np.arange(L)
                off = np.where(i < (L + 1) // 2, i, i - L).astype(float)
                offs.append((off, np.abs(off) <= shape[j] - 1))
            grids = np.meshgrid(*[o[0] * h[j] for o, j in zip(offs, dims)], indexing='ij')
            masks = np.meshgrid(*[o[1] for o in offs], indexing='ij')
            T = multi_values(name, fc[dims], per, grids)
            for mk in masks:
                T = T * mk
            facs.append(dict(kind='group', axes=list(dims), circ=[0] * len(dims), tlen=Ls, freqs=[None] * len(dims),
                             data=np.ascontiguousarray(T, dtype=float).ravel()))
    return s, facs


def mode_kf_max(spec, shape, h, pad):
    """max |Kf| of one mode's kernel on the reference's padded lattice (the FISTA bound of BLMLEmpirical's l1 fit,
    blml_binned.py:1179, with Kf = rfftn of the masked table, :79-90): the table is a product of units over disjoint
    dims, so its spectrum is the outer product of the units' spectra and the maximum is the product of their
    maxima.  Each unit's table is evaluated on its dims of pad (offsets wrapped as _offsets does, :63-72; masked to
    |off| <= m - 1 on padded dims) and transformed in FP64."""
    d = len(shape)
    out = 1.0
    for kind, dims in units_of(spec['kinds']):
        name = split_kind(kind)[0]
        fc = np.asarray(spec['fc'], float)
        offs, masks = [], []
        for j in dims:
            L, m = int(pad[j]), int(shape[j])
            i = np.arange(L)
            off = np.where(i < (L + 1) // 2, i, i - L).astype(float)
            offs.append(off * h[j])
            masks.append(np.ones(L) if L == m else (np.abs(off) <= m - 1).astype(float))
        if len(dims) == 1 and name not in ('ball', 'pball'):
            T = one_column_values(kind, fc[dims[0]], spec['periods'][dims[0]], offs[0]) * masks[0]
        else:
            grids = np.meshgrid(*offs, indexing='ij')
            T = multi_values(name, fc[list(dims)], [spec['periods'][j] for j in dims], grids)
            for mk in np.meshgrid(*masks, indexing='ij'):
                T = T * mk
        out *= float(np.max(np.abs(np.fft.fftn(T))))
    return out


def mode_factor_dlogs(spec, shape, h, circ):
    """Per unit of mode_factors(spec, ...) (same order): the factor of dK/dlog fc_u (the unit's table replaced by its
    derivative: blml_binned.py:92-101's dKf, one table per unit), or None where fc moves in whole harmonics
    (dirichlet, periodic ball) or the unit is on circular dims."""
    _, facs = mode_factors(spec, shape, h, circ)
    out = []
    fc = np.asarray(spec['fc'], float)
    for (kind, dims), fa in zip(units_of(spec['kinds']), facs):
        name = split_kind(kind)[0]
        if name in ('dirichlet', 'pball') or any(fa['circ']):
            out.append(None)
            continue
        g = dict(fa)
        if fa['kind'] == 'line':
            j = dims[0]
            g['data'] = np.asarray(one_column_dvalues(kind, fc[j], spec['periods'][j], 'fc',
                                                      np.arange(shape[j]) * h[j]), float)
        else:                               # a ball over its columns: prod fc (D p(s) + s p'(s)), the same mask
            Ls = fa['tlen']
            offs = []
            for a, j in enumerate(dims):
                L = Ls[a]
                i = np.arange(L)
                off = np.where(i < (L + 1) // 2, i, i - L).astype(float)
                offs.append((off, np.abs(off) <= shape[j] - 1))
            grids = np.meshgrid(*[o[0] * h[j] for o, j in zip(offs, dims)], indexing='ij')
            masks = np.meshgrid(*[o[1] for o in offs], indexing='ij')
            sg = np.sqrt(sum((fc[j] * x) ** 2 for x, j in zip(grids, dims)))
            D = len(dims)
            T = float(np.prod(fc[dims])) * (D * ball_profile(sg, D) + ball_profile_dlog(sg, D))
            for mk in masks:
                T = T * mk
            g['data'] = np.ascontiguousarray(T, dtype=float).ravel()
        out.append(g)
    return out
