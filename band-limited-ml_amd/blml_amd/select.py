"""The fc search of the reference's selector (selector.py:490-553, used by the streaming classes' burn-in selection,
blml_binned.py:899-936) and the fc sweep of its frame jobs (frame_fc_gpu_job.py:26-45), with every round of
candidate fits as one batched launch of the engine.

    fc, score, info = select_fc(counts, kinds, periods, h, complex_g, rank, search='shared' | 'refine')
    table = fc_sweep(train, val, test, kinds, fcs, periods, h, complex_g, rank)

The search keeps the reference's decisions -- a doubling bracket of every free fc (continuing `patience` steps past
a non-improvement), the volume grid 2**(j / dg) around the best doubling (outward, up first), then block-wise
refinement with steps of each factor (harmonic-aware for dirichlet columns) until a sweep gains nothing above
1e-6 -- but evaluates ahead: when the search needs a score it does not have, it asks for the candidates its next
decisions can reach (the rest of the bracket, both halves of the volume grid, a chain of steps in both directions
from the current point), fits them all in one launch (blml_amd.fit_many), and replays its decisions on the scores.
The decisions only read scores, so the result is the sequential search's; the extra candidates cost launches'
width, not their number.  The objective is the held-out score of a fit on the training counts: the mean log
density at the validation counts (the streaming burn-in's _val_score, blml_binned.py:685-697), -inf for an fc
beyond the lattice's Nyquist limit (0.5 / h)."""
import numpy as np

__all__ = ['FcSearch', 'select_fc', 'fc_sweep', 'val_score', 'split_counts']


class _Need(Exception):
    def __init__(self, points):
        super().__init__('scores needed')
        self.points = points


def _key(fc):
    return tuple(np.round(np.log(np.asarray(fc, float)), 6))


def _multi_column_groups(kinds):
    """Column lists of the multi-column units (ball, pball), by kind tag in order of first appearance."""
    g = {}
    for k, kind in enumerate(kinds or []):
        name, _, tag = str(kind).partition(':')
        if name in ('ball', 'pball'):
            g.setdefault(tag, []).append(k)
    return list(g.values())


def search_blocks(kinds, d, free):
    """The search coordinates: a multi-column unit moves as one block (its band stays round), every other free column
    on its own; blocks with a fixed column are left out."""
    groups = _multi_column_groups(kinds)
    first = {dims[0]: dims for dims in groups}
    member = {k for dims in groups for k in dims}
    out = []
    for k in range(d):
        blk = first.get(k) if k in member else [k]
        if blk is not None and all(free[j] for j in blk):
            out.append(list(blk))
    return out


def step(fc, blk, factor, kinds, periods_scaled):
    """fc with the block's columns times factor; a dirichlet column moves to the next harmonic count
    N = floor(fc P / 2) in the step's direction when the factor alone leaves N unchanged (None below N = 0)."""
    out = np.array(fc, float)
    out[blk] = out[blk] * factor
    if len(blk) == 1 and kinds is not None and str(kinds[blk[0]]).partition(':')[0] == 'dirichlet':
        k = blk[0]
        P = float(periods_scaled[k])
        n0, n1 = int(np.floor(fc[k] * P / 2.0)), int(np.floor(out[k] * P / 2.0))
        if n1 == n0:
            n1 = n0 + (1 if factor > 1 else -1)
            if n1 < 0:
                return None
            out[k] = (n1 + 0.5) * 2.0 / P
    return out


class FcSearch:
    """The search as decisions replayed on a score table; run(score_many) feeds it batches."""

    def __init__(self, d, kinds=None, periods_scaled=None, fc0=0.2, patience=2, max_sweeps=6,
                 factors=(np.sqrt(2.0),), start=None, free=None, chain=3):
        self.d, self.kinds, self.periods_scaled = d, kinds, periods_scaled
        self.fc0, self.patience, self.max_sweeps, self.factors = float(fc0), int(patience), int(max_sweeps), factors
        self.start = np.full(d, np.nan) if start is None else np.asarray(start, float)
        self.free = np.ones(d, bool) if free is None else np.asarray(free, bool)
        self.chain = int(chain)
        self.scores = {}
        self.rounds = []

    def _score(self, fc, ahead):
        k = _key(fc)
        if k not in self.scores:
            pts = [np.asarray(fc, float)] + [np.asarray(p, float) for p in ahead() if p is not None]
            raise _Need(pts)
        return self.scores[k]

    def _decide(self):
        grow = self.free & ~np.isfinite(self.start)
        dg = int(grow.sum())
        fc = np.where(grow, self.fc0, self.start)
        best_fc, best = fc.copy(), self._score(fc, lambda: [])
        bad = 0
        # bracket: doublings of the growing columns
        while dg and bad <= self.patience:
            cur = fc.copy()
            fc = np.where(grow, fc * 2.0, fc)
            ll = self._score(fc, lambda c=cur: [np.where(grow, c * 2.0 ** t, c) for t in range(2, self.patience + 3)])
            if ll > best:
                best, best_fc, bad = ll, fc.copy(), 0
            else:
                bad += 1
        # the volume grid around the best doubling: up, then down if no up-step improved
        centre = best_fc.copy()
        grid = lambda sgn, j: np.where(grow, centre * 2.0 ** (sgn * j / dg), centre)
        every = lambda: [grid(s, j) for s in (1, -1) for j in range(1, dg)]
        for sgn in (1, -1):
            up = False
            for j in range(1, dg):
                ll = self._score(grid(sgn, j), every)
                if ll > best:
                    best, best_fc, up = ll, grid(sgn, j), True
                else:
                    break
            if up:
                break
        # refinement, block by block
        fc = best_fc
        for factor in self.factors:
            for _ in range(self.max_sweeps):
                moved = False
                for blk in search_blocks(self.kinds, self.d, self.free):
                    for fac in (factor, 1.0 / factor):
                        stepped = False
                        while True:
                            cand = step(fc, blk, fac, self.kinds, self.periods_scaled)
                            if cand is None:
                                break

                            def ahead(c=cand, b=blk, f=fac, base=fc):
                                pts, p = [], c
                                for _ in range(self.chain - 1):
                                    p = None if p is None else step(p, b, f, self.kinds, self.periods_scaled)
                                    pts.append(p)
                                pts.append(step(base, b, 1.0 / f, self.kinds, self.periods_scaled))
                                return pts
                            ll = self._score(cand, ahead)
                            if ll > best + 1e-6:
                                best, fc, moved, stepped = ll, cand, True, True
                            else:
                                break
                        if stepped:
                            break
                if not moved:
                    break
        return fc, best

    def run(self, score_many, max_rounds=10000):
        """score_many(list of fc vectors) -> list of scores, one batch per call.  Returns (fc, best score)."""
        for _ in range(max_rounds):
            try:
                return self._decide()
            except _Need as need:
                pts, seen = [], set()
                for p in need.points:
                    k = _key(p)
                    if k not in self.scores and k not in seen:
                        seen.add(k)
                        pts.append(p)
                vals = score_many(pts)
                for p, v in zip(pts, vals):
                    self.scores[_key(p)] = float(v)
                self.rounds.append(len(pts))
        raise RuntimeError("fc search: no decision after max_rounds batches")


def split_counts(counts, seed=0, share=0.8):
    """Train / validation counts: a binomial share of the (integer) counts, numpy default_rng(seed), as the
    streaming burn-in's split (blml_binned.py:914-916)."""
    counts = np.asarray(counts, float)
    if not np.allclose(counts, np.round(counts)):
        raise ValueError("selection: integer counts for the binomial split")
    rng = np.random.default_rng(seed)
    tr = rng.binomial(np.round(counts).astype(np.int64), share).astype(float)
    return tr, counts - tr


def val_score(density, counts):
    """Mean log density at counts on the lattice (log_density of a fit whose F / Q is density_,
    blml_binned.py:685-693)."""
    m = np.asarray(counts, float)
    occ = m > 0
    return float((m[occ] * np.log(np.maximum(np.asarray(density)[occ], 1e-300))).sum(dtype=np.float64) /
                 m.sum(dtype=np.float64))


def _fits(train, fcs, kinds, periods, h, complex_g, rank, fit_kw):
    from . import fit_many, make_modes
    jobs = [(train, make_modes(list(fc), kinds, periods, rank), dict(h=h, complex_g=complex_g)) for fc in fcs]
    return fit_many(jobs, **fit_kw) if jobs else []


def select_fc(counts, kinds, periods=None, h=1.0, complex_g=False, rank=1, search='shared', seed=0, fc0=None,
              patience=2, dtype=np.float32, **fit_kw):
    """The fc of every column of make_modes(fc, kinds, periods, rank) that maximises the held-out score of a fit on
    an 80/20 binomial split of counts (the streaming burn-in's selection, blml_binned.py:899-936): search='shared'
    is the bracket and volume grid only, 'refine' adds six refinement sweeps; factors sqrt(2), 2^(1/4), 2^(1/8).
    Returns (fc, score, dict(evaluations, launches))."""
    counts = np.asarray(counts, float)
    d = counts.ndim
    hh = np.broadcast_to(np.asarray(h, float), (d,))
    per = list(periods) if periods is not None else [None] * d
    per_s = [None if p is None else float(p) / hh[j] for j, p in enumerate(per)]
    tr, va = split_counts(counts, seed)
    nyq = 0.5 / hh
    if fc0 is None:
        fc0 = 2.0 / (max(counts.shape) * float(hh.max()))
    fit_kw = dict(fit_kw, dtype=dtype)

    def score_many(fcs):
        ok = [i for i, fc in enumerate(fcs) if not np.any(np.asarray(fc) > nyq + 1e-12)]
        out = [-np.inf] * len(fcs)
        for i, m in zip(ok, _fits(tr, [fcs[i] for i in ok], kinds, per, h, complex_g, rank, fit_kw)):
            out[i] = val_score(m.density_, va) if np.isfinite(m.loglik_) else -np.inf
        return out
    s = FcSearch(d, kinds, per_s, fc0=fc0, patience=patience, max_sweeps=0 if search == 'shared' else 6,
                 factors=(np.sqrt(2.0), 2.0 ** 0.25, 2.0 ** 0.125))
    fc, best = s.run(score_many)
    return fc, best, dict(evaluations=len(s.scores), launches=len(s.rounds), batch_sizes=list(s.rounds))


def fc_sweep(train, val, test, kinds, fcs, periods=None, h=1.0, complex_g=False, rank=1, dtype=np.float32,
             **fit_kw):
    """Fits of the training counts at every fc vector of fcs in one launch (the frame job's fc sweep,
    frame_fc_gpu_job.py:26-45, its fits cold-started side by side instead of warm-started one after another) and
    their train / validation / test scores (mean log of the normalised density at each count set)."""
    ms = _fits(np.asarray(train, float), [np.asarray(f, float) for f in fcs], kinds, periods, h, complex_g, rank,
               dict(fit_kw, dtype=dtype))
    out = []
    for fc, m in zip(fcs, ms):
        P = np.asarray(m.density_, float)
        P = P / P.sum()
        out.append(dict(fc=[float(x) for x in np.atleast_1d(fc)], train=val_score(P, train), val=val_score(P, val),
                        test=val_score(P, test), sweeps=int(m.n_sweeps_), stop=m.stop_reason_, loglik=m.loglik_))
    return out
