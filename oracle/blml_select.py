"""Test infrastructure (the oracle): a plain sequential restatement of the reference's fc search, one objective
call at a time, to check the engine's batched search (blml_amd.select) against.

    search(obj, d, kinds, periods_scaled, fc0, patience, max_sweeps, factors, start, free) -> (fc, best, visited)

Follows selector.py: _blocks (:63-76), _free_blocks (:89-90), _harmonics / _move (:467-488), _search (:490-553):
stage 1a brackets the optimum by doubling every free fc without a start (continuing `patience` steps past a
non-improvement), stage 1b walks the volume grid centre * 2**(j / dg) outward from the best doubling (up first,
down only if no up-step improved), stage 2 refines block by block with steps of each factor until a sweep improves
nothing (a step must gain more than 1e-6).  visited: the fc vectors obj was called on, in order."""
import numpy as np


def radial_groups(kinds):
    """[(name, dims)] of the multi-column kinds (ball, pball) by tag, in order of first appearance
    (blml_trivial.py:251-261)."""
    groups = {}
    for k, kind in enumerate(kinds or []):
        name, _, tag = str(kind).partition(':')
        if name in ('ball', 'pball'):
            groups.setdefault(tag, (name, []))[1].append(k)
    return list(groups.values())


def blocks(kinds, d):
    groups = [dims for _, dims in radial_groups(kinds)]
    grouped = {k for dims in groups for k in dims}
    out, seen = [], set()
    for k in range(d):
        if k in grouped:
            g = next(dims for dims in groups if k in dims)
            if g[0] not in seen:
                out.append(list(g))
                seen.add(g[0])
        else:
            out.append([k])
    return out


def free_blocks(kinds, d, free):
    return [b for b in blocks(kinds, d) if all(free[b])]


def harmonics(fc_s, Ps):
    return int(np.floor(fc_s * Ps / 2.0))


def move(fc, block, factor, kinds, periods_scaled):
    new = fc.copy()
    new[block] *= factor
    if len(block) > 1:
        return new
    k = block[0]
    if kinds is not None and str(kinds[k]).partition(':')[0] == 'dirichlet':
        Ps = periods_scaled[k]
        n_old, n_new = harmonics(fc[k], Ps), harmonics(new[k], Ps)
        if n_new == n_old:
            n_new = n_old + 1 if factor > 1 else n_old - 1
            if n_new < 0:
                return None
            new[k] = (n_new + 0.5) * 2.0 / Ps
    return new


def search(obj, d, kinds, periods_scaled, fc0=0.2, patience=2, max_sweeps=6, factors=(np.sqrt(2.0),), start=None,
           free=None):
    cache, visited = {}, []

    def f(fc):
        key = tuple(np.round(np.log(fc), 6))
        if key not in cache:
            cache[key] = obj(fc)
            visited.append(np.array(fc, float))
        return cache[key]
    st = np.full(d, np.nan) if start is None else np.asarray(start, float)
    fr = np.ones(d, bool) if free is None else np.asarray(free, bool)
    grow = fr & ~np.isfinite(st)
    dg = int(grow.sum())
    fc = np.where(grow, float(fc0), st)
    best_fc, best = fc.copy(), f(fc)
    bad = 0
    while dg and bad <= patience:
        fc = np.where(grow, fc * 2.0, fc)
        ll = f(fc)
        if ll > best:
            best, best_fc, bad = ll, fc.copy(), 0
        else:
            bad += 1
    centre = best_fc.copy()
    for sign in (1, -1):
        improved = False
        for j in range(1, dg):
            cand = np.where(grow, centre * 2.0 ** (sign * j / dg), centre)
            ll = f(cand)
            if ll > best:
                best, best_fc, improved = ll, cand, True
            else:
                break
        if improved:
            break
    fc = best_fc
    for factor in factors:
        for sweep in range(max_sweeps):
            moved = False
            for blk in free_blocks(kinds, d, fr):
                for fac in (factor, 1.0 / factor):
                    stepped = False
                    while True:
                        cand = move(fc, blk, fac, kinds, periods_scaled)
                        if cand is None:
                            break
                        ll = f(cand)
                        if ll > best + 1e-6:
                            best, fc, moved, stepped = ll, cand, True, True
                        else:
                            break
                    if stepped:
                        break
            if not moved:
                break
    return fc, best, visited
