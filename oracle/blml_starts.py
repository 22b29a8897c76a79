"""CPU FP64 restatement of the start search of a complex counts fit (blml_trivial.py:1605-1676 complex_starts, with
_band_grid :1167-1178 and _even_kicks :1181-1186) -- TEST INFRASTRUCTURE ONLY (the checker of the HIP start search,
blml_complex_starts; imported by tests, never by the product).

Search grid: the bin lattice coarsened by q_j = max(1, floor(0.5 / (fc_j h_j))) (counts summed per coarse cell), the
box of the coarse bins, sides the next power of two of twice the box.  Band: |nu_j| <= fc_j / 2 per sinc-like
axis, ||nu_G|| <= fc / 2 per radial group, nu the grid's FFT frequencies.  A run alternates the data constraint
(|g| = sqrt(counts) s on the box, s = the field's own scale there) and the band constraint: hybrid input-output
(P <- P - beta Pn off the band) for n_hio steps, then error reduction (P <- 0 off the band) for n_er steps.  Random
runs start from band-limited random phases (default_rng(seed).random(shape) per run, hio_pool runs); symmetry-
breaking runs from the band-limited sqrt(counts) times exp(+-i delta p) for each even phase direction p, n_er ER
steps only.  Each family keeps its best runs by the Poisson score sum w log(|g|^2 / sum_box |g|^2); a start is
e0_b = w_b / conj(g_b) at the original bins (g by the exact band sum when q > 1).

Pinning (tests/test_oracle_starts.py): the reference's own outputs on small cases
(tests/golden/make_golden_starts.py writes them by importing complex_starts in this container), the error-reduction
steps' monotone data error (a projection method: ||X - proj_data X|| never rises), the band-limited output, and the
score's definition recomputed directly.  The iteration is chaotic: a relative change of 1e-15 in the start phases
moves a 500-step run's result at O(1), so two FFT implementations agree over a few steps only (the HIP search is
checked step for step on short runs and by these properties on long ones)."""
import numpy as np


def radial_groups(kinds):
    """[(dims)] of the multi-column units ('ball:<tag>', 'pball:<tag>': the columns sharing a tag)."""
    groups = {}
    for j, k in enumerate(kinds):
        name, _, tag = str(k).partition(':')
        if name in ('ball', 'pball'):
            groups.setdefault(tag, []).append(j)
    return list(groups.values())


def band_grid(shape, h, fc, kinds):
    """(U, band): the grid's FFT frequencies divided by fc / 2 per axis, and the band indicator."""
    U = np.stack(np.meshgrid(*[np.fft.fftfreq(m, hj) / (0.5 * f) for m, hj, f in zip(shape, h, fc)],
                             indexing='ij'), -1)
    band = np.ones(shape, bool)
    rad = set()
    for dims in radial_groups(kinds):
        band &= (U[..., dims] ** 2).sum(-1) <= 1
        rad |= set(dims)
    for j in range(len(shape)):
        if j not in rad:
            band &= np.abs(U[..., j]) <= 1
    return U, band


def even_kicks(U):
    """Even low-order phase directions on the unit band: 2 r^2 - 1, u_0^2 - u_j^2 (j > 0), 2 u_i u_j (i < j),
    6 r^4 - 6 r^2 + 1."""
    d = U.shape[-1]
    r2 = (U ** 2).sum(-1)
    out = [2 * r2 - 1]
    out += [U[..., 0] ** 2 - U[..., j] ** 2 for j in range(1, d)]
    out += [2 * U[..., i] * U[..., j] for i in range(d) for j in range(i + 1, d)]
    out.append(6 * r2 ** 2 - 6 * r2 + 1)
    return out


def search_grid(Zi, w, h, fc, max_cells=1 << 22):
    """The coarse grid of the search, or None when it has more than max_cells cells."""
    Zi0 = np.asarray(Zi, np.int64)
    w0 = np.asarray(w, np.float64)
    h0 = np.asarray(h, float)
    q = np.maximum(1, np.floor(0.5 / (np.asarray(fc, float) * h0))).astype(np.int64)
    Zc = np.floor_divide(Zi0 - Zi0.min(0), q)
    if np.all(q == 1):
        Zg, wg = Zc, w0
    else:
        Zg, inv = np.unique(Zc, axis=0, return_inverse=True)
        wg = np.bincount(np.ravel(inv), weights=w0)
    lo = Zg.min(0)
    L = Zg.max(0) - lo + 1
    shape = tuple(1 << int(np.ceil(np.log2(2 * int(l)))) for l in L)
    if np.prod(shape, dtype=float) > max_cells:
        return None
    idx = tuple((Zg - lo).T)
    A = np.zeros(shape)
    A[idx] = np.sqrt(wg)
    box = tuple(slice(0, int(l)) for l in L)
    return dict(q=q, h0=h0, hc=h0 * q, Zi0=Zi0, w0=w0, wg=wg, L=L, shape=shape, idx=idx, A=A, box=box,
                Ab=A[box], Ab2=float(A[box].ravel() @ A[box].ravel()))


def project(g, G):
    """The data constraint on the box: the field's phase kept, |g| = sqrt(counts) s."""
    gb = g[G['box']]
    s = np.sqrt((np.abs(gb) ** 2).sum() / G['Ab2'])
    g = g.copy()
    g[G['box']] = G['Ab'] * s * np.exp(1j * np.angle(gb))
    return g


def run(P, G, band, n_hio, n_er, beta=0.9, errors=None):
    """n_hio HIO steps then n_er ER steps from the spectrum P; the band-limited result.  errors: a list that
    receives, before each step, the data error ||X - proj(X)|| of the step's field X = ifftn(P)."""
    for it in range(n_hio + n_er):
        X = np.fft.ifftn(P)
        Xp = project(X, G)
        if errors is not None:
            errors.append(float(np.linalg.norm(X - Xp)))
        Pn = np.fft.fftn(Xp)
        P = np.where(band, Pn, (P - beta * Pn) if it < n_hio else 0)
    return P * band


def score(P, G):
    f = np.abs(np.fft.ifftn(P)) ** 2
    return float(G['wg'] @ np.log(np.maximum(f[G['idx']] / f[G['box']].sum(), 1e-300)))


def field_at_bins(P, G, band):
    """g at the original bins: ifftn(P) at the cells when q = 1, else the exact band sum."""
    if np.all(G['q'] == 1):
        return np.fft.ifftn(P)[G['idx']]
    nz = np.nonzero(band)
    nu = np.stack([np.fft.fftfreq(m, hj)[k] for m, hj, k in zip(G['shape'], G['hc'], nz)], 1)
    c = P[nz] / P.size
    x = (G['Zi0'] - G['Zi0'].min(0)) * G['h0']
    return np.exp(2j * np.pi * x @ nu.T) @ c


def start_runs(Zi, w, h, fc, kinds, hio_pool=30, n_hio_iter=400, n_er_iter=100, beta=0.9, sb_delta=1.0, seed=0,
               n_rand=None, n_sb=2, max_cells=1 << 22):
    """Every run of the search: (G, band, [(family, P_start, P_end, score)]) in the reference's order (random runs,
    then per even direction the + and - kick); n_rand: random runs (hio_pool, or 0 when no random starts are
    asked for); symmetry-breaking runs only when n_sb > 0."""
    G = search_grid(Zi, w, h, fc, max_cells)
    if G is None:
        return None, None, []
    U, band = band_grid(G['shape'], G['hc'], fc, kinds)
    rng = np.random.default_rng(seed)
    runs = []
    for _ in range(hio_pool if n_rand is None else n_rand):
        P0 = band * np.exp(2j * np.pi * rng.random(G['shape']))
        P = run(P0, G, band, n_hio_iter, n_er_iter, beta)
        runs.append(('hio', P0, P, score(P, G)))
    if n_sb > 0:
        F0 = np.fft.fftn(G['A']) * band
        for p in even_kicks(U):
            for sg in (1, -1):
                P0 = F0 * np.exp(1j * sg * sb_delta * p)
                P = run(P0, G, band, 0, n_er_iter, beta)
                runs.append(('symbreak', P0, P, score(P, G)))
    return G, band, runs


def best(runs, family, k):
    """The k best runs of a family by score (the first of equal scores first)."""
    cand = [r for r in runs if r[0] == family]
    order = sorted(range(len(cand)), key=lambda i: -cand[i][3])
    return [cand[i] for i in order[:k]]


def complex_starts(Zi, w, h, fc, kinds, n_hio=3, n_sb=2, hio_pool=30, sb_delta=1.0, n_hio_iter=400, n_er_iter=100,
                   beta=0.9, seed=0, max_cells=1 << 22):
    """[(family, score, e0)]: the top n_hio random runs and the top n_sb symmetry-breaking runs, e0 = w / conj(g)
    at the original bins; None when the grid has more than max_cells cells."""
    G, band, runs = start_runs(Zi, w, h, fc, kinds, hio_pool, n_hio_iter, n_er_iter, beta, sb_delta, seed,
                               n_rand=hio_pool if n_hio > 0 else 0, n_sb=n_sb, max_cells=max_cells)
    if G is None:
        return None
    out = []
    for family, k in (('hio', n_hio), ('symbreak', n_sb)):
        for fam, _, P, sc in best(runs, family, k):
            g = field_at_bins(P, G, band)
            out.append((fam, sc, G['w0'] / np.conj(np.where(np.abs(g) == 0, 1e-12, g))))
    return out
