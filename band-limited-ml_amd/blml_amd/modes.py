"""The model specification of the binned fit: make_modes / parse_modes with the reference's arguments, keys and
errors (blml_trivial.py:1189-1323, 1467-1505), for the kinds the binned fit supports: 'sinc', 'slits:<fill>',
'dirichlet[:<fill>]' (one column each) and 'ball', 'pball' (multi-column units, tuple keys)."""
import numpy as np

KINDS = ('sinc', 'dirichlet', 'ball', 'slits', 'pball')
MULTI = ('ball', 'pball')
_PARAMS = {'sinc': ('fc',), 'slits': ('fc', 'fill'), 'dirichlet': ('fc', 'period', 'fill'), 'ball': ('fc',),
           'pball': ('fc', 'period')}
_STRUCT = {'kind', 'dep', 'h', '_unit'}


def split_kind(kind):
    name, _, tag = str(kind).partition(':')
    return name, tag


def is_multi(kind):
    name = split_kind(kind)[0]
    if name not in KINDS:
        raise ValueError(f"unknown kernel kind {kind!r}")
    return name in MULTI


def _value(v):
    """A parameter given bare or as {'value': v, 'trainable': t} (blml_trivial.py:1238-1244)."""
    if isinstance(v, dict):
        if set(v) - {'value', 'trainable'}:
            raise ValueError(f"parameter {v}: keys must be 'value' and/or 'trainable'")
        return v.get('value')
    return v


def units_of(kinds):
    """[(kind token, [columns])]: one-column units in column order, then multi-column groups in order of first
    appearance, grouped by tag (blml_trivial.py:252-263, 330-335)."""
    one = [(k, [j]) for j, k in enumerate(kinds) if not is_multi(k)]
    groups = {}
    for j, k in enumerate(kinds):
        if is_multi(k):
            name, tag = split_kind(k)
            g = groups.setdefault(tag, (name, []))
            if g[0] != name:
                raise ValueError(f"multi-column group {tag!r} mixes kinds {g[0]!r} and {name!r}")
            g[1].append(j)
    return one + [(name, dims) for name, dims in groups.values()]


def parse_modes(modes, d, allow_dep=False):
    """(specs, dep) as the reference returns them (blml_trivial.py:1247-1323): specs = K dicts {fc (d,), kinds (d,),
    periods (d,)}; dep, the sorted columns marked 'dep' (the same in every mode; a unit's flag covers its columns),
    or None.  The binned fit rejects dep (allow_dep False); the counts path fits conditionals."""
    if isinstance(modes, dict) or not isinstance(modes, (list, tuple)) or not modes:
        raise TypeError("modes: a non-empty list of mode dicts {column: {...}}; wrap one mode as [mode]")
    specs, deps = [], []
    for k, m in enumerate(modes):
        cols = {}
        for key, c in m.items():
            kind = c.get('kind', 'sinc')
            if kind not in KINDS:
                raise ValueError(f"mode {k} column {key}: kind {kind!r}, expected one of {KINDS}")
            bad = set(c) - _STRUCT - set(_PARAMS[kind])
            if bad:
                raise ValueError(f"mode {k} column {key}: unknown keys {sorted(bad)} (kind {kind!r} declares "
                                 f"{sorted(_PARAMS[kind])})")
            if isinstance(key, tuple):
                if not is_multi(kind):
                    raise ValueError(f"columns {key}: a tuple key is a multi-column unit (kind 'ball' or 'pball')")
                per = _value(c.get('period'))
                for i, j in enumerate(key):
                    cj = dict(c, _unit=key)
                    if isinstance(per, (list, tuple, np.ndarray)):
                        if len(per) != len(key):
                            raise ValueError(f"columns {key}: period needs one value per column, got {per}")
                        cj['period'] = float(per[i])
                    cols[j] = cj
            else:
                if is_multi(kind):
                    raise ValueError(f"column {key}: kind {kind!r} spans columns: give its unit as a tuple key")
                cols[key] = c
        if set(cols) != set(range(d)):
            raise ValueError(f"mode {k}: columns {sorted(cols)}, expected every column 0..{d - 1}")
        fc, kinds, periods, dep = [], [], [], []
        for j in range(d):
            c = cols[j]
            kind = c.get('kind', 'sinc')
            if c.get('dep') and not allow_dep:
                raise ValueError("BLMLBinned: conditional (dep) modes are not supported; fit the joint")
            dep.append(bool(c.get('dep', False)))
            v = _value(c.get('fc'))
            if v is None:
                raise ValueError(f"mode {k} column {j}: fc missing")
            if not float(v) > 0:
                raise ValueError(f"mode {k} column {j}: fc must be positive, got {v}")
            fc.append(float(v))
            per = _value(c.get('period'))
            if kind in ('dirichlet', 'pball'):
                if per is None:
                    raise ValueError(f"mode {k} column {j}: period missing (kind {kind!r} needs it)")
                if not float(per) > 0:
                    raise ValueError(f"mode {k} column {j}: period must be positive, got {per}")
                periods.append(float(per))
            else:
                periods.append(None)
            fill = _value(c.get('fill'))
            if kind == 'slits' and fill is None:
                raise ValueError(f"mode {k} column {j}: fill missing (kind 'slits' needs it)")
            if fill is not None and not 0 < float(fill) <= 1:
                raise ValueError(f"mode {k} column {j}: {kind} fill must be in (0, 1], got {fill}")
            if '_unit' in c:
                kinds.append(f"{kind}:{'-'.join(map(str, c['_unit']))}")
            elif kind in ('slits', 'dirichlet') and fill is not None:
                kinds.append(f"{kind}:{float(fill)!r}")
            else:
                kinds.append(kind)
        for name, dims in units_of(kinds):
            if name in MULTI and not np.allclose(np.array(fc)[dims], fc[dims[0]], rtol=1e-9, atol=0):
                raise ValueError(f"radial group {name} over dims {dims} uses one fc for every direction; "
                                 f"got {np.array(fc)[dims]}")
        specs.append(dict(fc=np.array(fc), kinds=kinds, periods=periods))
        deps.append(dep)
    if any(dp != deps[0] for dp in deps):
        raise ValueError("dep must mark the same columns in every mode")
    dep = [j for j in range(d) if deps[0][j]]
    return specs, (dep or None)


def make_modes(fc, kinds=None, periods=None, rank=1, dep=None, d=None):
    """Modes from per-dim settings (blml_trivial.py:1467-1505): fc scalar, (d,) or (K, d); kinds per dim
    ('sinc', 'dirichlet[:fill]', 'slits:fill', 'ball[:tag]', 'pball[:tag]'); periods per dim (dirichlet, pball);
    rank: K when fc has one row; dep: the dependent column(s) of a conditional counts fit (marked on a unit's
    first column, which carries the unit)."""
    deps = set() if dep is None else {int(j) for j in np.atleast_1d(dep)}
    F = np.asarray(fc, dtype=float)
    if d is None:
        d = len(kinds) if kinds is not None else (F.shape[-1] if F.ndim else None)
    if d is None:
        raise ValueError("make_modes: give kinds or d with a scalar fc")
    F = F if F.ndim == 2 else np.broadcast_to(F, (d,))[None]
    F = np.broadcast_to(F, (int(rank) if len(F) == 1 else len(F), d))
    kinds = list(kinds) if kinds is not None else ['sinc'] * d
    groups = {}
    for j, kd in enumerate(kinds):
        name, tag = split_kind(kd)
        if name in MULTI:
            groups.setdefault(tag, []).append(j)
    units = [tuple(v) for v in groups.values()]

    def col(k, j):
        kind, _, tag = str(kinds[j]).partition(':')
        c = dict(kind=kind)
        if kind in ('slits', 'dirichlet') and tag:
            c['fill'] = float(tag)
        if np.isfinite(F[k, j]):
            c['fc'] = float(F[k, j])
        if kind == 'dirichlet':
            c['period'] = periods[j]
        if kind == 'pball':
            key = next(u for u in units if j in u)
            c['period'] = [periods[i] for i in key]
        if j in deps:
            c['dep'] = True
        return c

    out = []
    for k in range(len(F)):
        m = {}
        for j in range(d):
            key = next((u for u in units if j in u), j)
            if key == j or j == key[0]:
                m[key] = col(k, j)
        out.append(m)
    return out


def strip_dep(modes):
    """The modes without their dep flags (the joint of a conditional fit)."""
    return [{key: {a: v for a, v in c.items() if a != 'dep'} for key, c in m.items()} for m in modes]


def marginal_modes(modes, keep):
    """The modes restricted to the columns in keep, renumbered 0.. in that order, dep flags dropped (the marginal of a
    conditional fit, blml_trivial.py:1585-1598); a multi-column unit is kept whole or not at all."""
    pos = {j: i for i, j in enumerate(keep)}
    out = []
    for m in modes:
        mm = {}
        for key, c in m.items():
            cols = key if isinstance(key, tuple) else (key,)
            inside = [j in pos for j in cols]
            if all(inside):
                nk = tuple(pos[j] for j in cols) if isinstance(key, tuple) else pos[key]
                mm[nk] = {a: v for a, v in c.items() if a != 'dep'}
            elif any(inside):
                raise ValueError(f"columns {key}: a unit straddles kept and dropped columns")
        out.append(dict(sorted(mm.items(), key=lambda t: t[0][0] if isinstance(t[0], tuple) else t[0])))
    return out
