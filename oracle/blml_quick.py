"""CPU FP64 restatement of the reference's BLMLQuick counts path, real rank 1 -- TEST INFRASTRUCTURE ONLY (the
checker of the HIP counts path; imported by tests, never by the product).

Restated (numpy, float64, a dense Gram matrix G over the occupied bins in place of the cols kernel's blocks):
  root(b, cc, sbb)        the one-coordinate minimiser of 1/2 sbb e^2 + b e - cc log e, e > 0 (blml_gpu.py:345-349)
  sweep(...)              one sweep of blocks in a given order, plain or safeguarded (blml_gpu.py:340-373)
  quick_fit(G, w, ...)    block-Jacobi CD (blml_gpu.py:320-441): sweeps over default_rng(seed).permutation(B) in
                          blocks of min(block_size, ceil(B / 8)); plain sweeps until one raises
                          Phi = (1/2 e'Ge - n sum w log e) / n, which is undone and redone with the block safeguard
                          (the largest t in 2^-k, k < 31, with dPhi(t) <= 0, blml_gpu.py:352-366) for good; Se = G e
                          afresh when the residual max|Se e / (n w) - 1| first drops under tol; the stall guard
                          (< 1% over 50 sweeps); max_sweeps, then e scaled to e'Se = n^2 (blml_gpu.py:432-439)
  warm_start(e_prev, new, G, w)   init= on the same lattice (blml_gpu.py:200-208, 304-320): the previous e at the
                          shared bins, each new bin at root(Se_b, n w_b, S_bb) from e with the new bins at 0
  newton_solve(G, w)      the unique minimiser of the same convex Phi by damped Newton (an independent solver)
  quick_fit_dep(G, Gm, w, ...)  the conditional fit (dep columns, blml_gpu.py:226-441 with nfit = 2): the joint
                          (Gram G) and the marginal (Gm: the kernel of the units off the dep side, on the joint's
                          bins) in one loop, each sweep's permutation shared by the fits still running, a fit's
                          redo sweep drawing its own, each fit with its own safeguard, stop and stall guard

Pinning (tests/test_oracle_quick.py): quick_fit_dep's joint by newton_solve, its marginal by newton_solve on the
merged bins (bins that differ only in dep columns merged, counts summed) split back in proportion to the counts --
the identity the reference's docstring states (blml_gpu.py:67-72); root by its quadratic's residual and the 1-D objective's derivative; the
fixed point of quick_fit by newton_solve (the problem is convex, so both must reach the same e); each safeguarded
block step against a direct evaluation of Phi; Se = G e at the end; a warm-started fit by the same Newton fixed
point, and each new bin's start by the 1-D objective's derivative.  The trajectory itself (sweep counts, when the
safeguard switches on) follows the reference's code and no reference test holds it: parity unpinned beyond the
fixed point."""
import numpy as np


def root(b, cc, sbb):
    """The positive root of sbb e^2 + b e - cc = 0, without cancellation."""
    disc = np.sqrt(b * b + 4.0 * sbb * cc)
    return np.where(b >= 0, 2.0 * cc / (b + disc), (-b + disc) / (2.0 * sbb))


def objective(G, e, w, n):
    return float(0.5 * e @ (G @ e) - n * np.sum(w * np.log(np.maximum(e, 1e-300)))) / n


def sweep(G, w, n, sbb, e, Se, perm, Bs, safe, log_steps=None):
    """One sweep in place over the bins in the order perm, blocks of Bs (blml_gpu.py:340-373)."""
    steps = 0.5 ** np.arange(31)
    B = len(w)
    for s0 in range(0, B, Bs):
        I = perm[s0:s0 + Bs]
        eo = e[I].copy()
        cc = n * w[I]
        es = root(Se[I] - sbb * eo, cc, sbb)
        dl = es - eo
        v = G[:, I] @ dl
        if safe:
            lin = dl @ Se[I]
            quad = dl @ v[I]
            with np.errstate(invalid='ignore', divide='ignore'):
                logs = np.log1p(steps[:, None] * (dl / eo)[None, :]) @ cc
                ok = steps * lin + 0.5 * steps ** 2 * quad - logs <= 0
            t = steps[np.argmax(ok)] if ok.any() else 0.0
            if log_steps is not None:
                log_steps.append((e.copy(), dl.copy(), t, I.copy()))
            e[I] = eo + t * dl
            Se += t * v
        else:
            e[I] = es
            Se += v


def warm_start(e_prev, new, G, w):
    """A warm start on the same lattice (blml_gpu.py:200-208, 304-320): the previous fit's e at the bins both fits
    share (new: the mask of bins it lacks, whose e starts at 0), then each new bin at the exact 1-D minimiser of
    the objective with the others fixed, root(b, n w_b, S_bb) with b = (S e)_b, all from the same S e."""
    G = np.asarray(G, float)
    w = np.asarray(w, float)
    e = np.where(new, 0.0, np.asarray(e_prev, float))
    if np.any(new):
        Se = G @ e
        e[new] = root(Se[new], float(w.sum()) * w[new], float(G[0, 0]))
    return e


def quick_fit(G, w, tol=1e-4, block_size=4096, seed=0, max_sweeps=None, log_steps=None, e0=None, rng=None):
    """(e, Se, n_sweeps, converged, max_rho, rho_hist, obj_hist, safeguard) of the real rank-1 counts fit with Gram
    matrix G (B x B, the signs already in it: D S D) and counts w (B,).  log_steps: a list that receives
    (e_before, dlt, t, I) for every safeguarded block step.  e0: the start (warm_start), else ones.  rng: the
    generator of the sweeps' permutations (default_rng(seed) when None; a rank-K fit shares one with its start)."""
    G = np.asarray(G, float)
    w = np.asarray(w, float)
    B = len(w)
    n = float(w.sum())
    sbb = float(G[0, 0])
    rng = np.random.default_rng(seed) if rng is None else rng
    e = np.ones(B) if e0 is None else np.array(e0, float)
    Se = G @ e
    Bs = max(1, min(int(block_size), -(-B // 8)))

    def resid(e, Se):
        rho = np.abs(Se * e / (n * w) - 1.0)
        return float(np.max(np.where((e > 0) & (w > 0), rho, 0.0)))

    def obj(e, Se):
        return float(0.5 * e @ Se - n * np.sum(w * np.log(np.maximum(e, 1e-300)))) / n

    def run(safe):
        sweep(G, w, n, sbb, e, Se, rng.permutation(B), Bs, safe, log_steps)

    safeguard, done, conv = False, False, False
    n_sweeps, max_rho = 0, np.inf
    rho_hist, obj_hist = [], []
    obj_prev = obj(e, Se)
    while not done:
        if max_sweeps is not None and n_sweeps >= max_sweeps:
            break
        n_sweeps += 1
        save = (e.copy(), Se.copy())
        run(safeguard)
        obj_now = obj(e, Se)
        if not safeguard and obj_now > obj_prev + 1e-10 * abs(obj_prev):
            e[:], Se[:] = save
            safeguard = True
            run(True)
            obj_now = obj(e, Se)
        obj_prev = obj_now
        if not np.all(np.isfinite(e)):
            max_rho = np.inf
            rho_hist.append(np.inf)
            break
        rho = resid(e, Se)
        if rho < tol:
            Se[:] = G @ e
            rho = resid(e, Se)
            if rho < tol:
                obj_now = obj(e, Se)
        max_rho = rho
        rho_hist.append(rho)
        obj_hist.append(obj_now)
        if rho < tol:
            conv = done = True
        elif len(rho_hist) > 50 and rho_hist[-1] > 0.99 * min(rho_hist[:-50]):
            done = True
    exact = conv
    if max_sweeps is not None and not conv and np.isfinite(max_rho):
        Se[:] = G @ e
        sc = n / np.sqrt(max(float(e @ Se), 1e-300))
        e *= sc
        Se *= sc
        max_rho = resid(e, Se)
        conv = max_rho < tol
        exact = True
    return e, (Se if exact else None), n_sweeps, conv, max_rho, rho_hist, obj_hist, safeguard


def newton_solve(G, w, tol=1e-13, max_iter=200):
    """argmin over e > 0 of 1/2 e'Ge - n sum w log e (G positive definite on the bins): Newton steps with
    backtracking that keeps e > 0 and does not raise the objective."""
    G = np.asarray(G, float)
    w = np.asarray(w, float)
    n = float(w.sum())
    e = np.sqrt(n * w / G[0, 0])
    f = lambda x: 0.5 * x @ (G @ x) - n * np.sum(w * np.log(x))
    fe = f(e)
    for _ in range(max_iter):
        g = G @ e - n * w / e
        H = G + np.diag(n * w / e ** 2)
        step = np.linalg.solve(H, g)
        t = 1.0
        while True:
            x = e - t * step
            if np.all(x > 0):
                fx = f(x)
                if fx <= fe + 1e-4 * t * (-(g @ step)):
                    break
            t *= 0.5
            if t < 1e-20:
                return e
        e, fe = x, fx
        if np.max(np.abs(t * step) / e) < tol:
            break
    return e


# ------------------------------------------------------------------------------------------ complex / rank K
def psd_start(e_real, K, complex_g, rng, phase_noise=0.1, ws=None):
    """The start of the rank-K / complex counts fit (blml_gpu.py:575-590): column 0 the real solution with phases
    exp(i pn N(0, 1)) (complex), or the warm start ws's columns; the missing columns pn |e0| times CN(0, 1) (real:
    N(0, 1)), drawn from rng in that order."""
    B = len(e_real) if ws is None else len(ws)
    e0 = np.zeros((B, K), dtype=complex if complex_g else np.float64)
    if ws is not None:
        ws = np.asarray(ws, dtype=complex if complex_g else np.float64)
        ws = ws[:, None] if ws.ndim == 1 else ws
        k0 = min(K, ws.shape[1])
        e0[:, :k0] = ws[:, :k0]
        er = np.abs(ws[:, 0])
    else:
        k0 = 1
        er = np.asarray(e_real, np.float64)
        e0[:, 0] = er * (np.exp(1j * phase_noise * rng.standard_normal(B)) if complex_g else 1.0)
    for k in range(k0, K):
        noise = ((rng.standard_normal(B) + 1j * rng.standard_normal(B)) / np.sqrt(2) if complex_g
                 else rng.standard_normal(B))
        e0[:, k] = phase_noise * er * noise
    return e0


def psd_row_update(R, EI, cc, sbbs, shared):
    """The block's new rows (blml_gpu.py:622-640): shared kernel e_b = t R_b / |R_b| with
    t = 2 c / (|R| + sqrt(|R|^2 + 4 s c)) (EI's direction where R_b = 0); per-mode kernels e_bk = R_bk / (lam - s_k),
    lam = smax + x, x by doubling from 1 (at most 200 times, while phi(x) >= c) then 60 bisections of
    phi(x) = (smax + x) sum_k |R_bk|^2 / (smax + x - s_k)^2 against c."""
    if shared:
        s0 = float(sbbs[0])
        ar = np.sqrt((np.abs(R) ** 2).sum(1))
        t = 2.0 * cc / (ar + np.sqrt(ar * ar + 4.0 * s0 * cc))
        en = np.sqrt((np.abs(EI) ** 2).sum(1))
        with np.errstate(invalid='ignore', divide='ignore'):
            dirn = np.where((ar > 0)[:, None], R / np.where(ar > 0, ar, 1.0)[:, None],
                            EI / np.maximum(en, 1e-30)[:, None])
        return t[:, None] * dirn
    a2 = np.abs(R) ** 2
    s = np.asarray(sbbs, float)
    smax = float(s.max())
    lo, hi = np.zeros(len(cc)), np.ones(len(cc))
    phi = lambda x: (smax + x) * (a2 / (smax + x[:, None] - s[None]) ** 2).sum(1)
    for _ in range(200):
        big = phi(hi) >= cc
        if not big.any():
            break
        hi = np.where(big, hi * 2, hi)
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        up = phi(mid) > cc
        lo = np.where(up, mid, lo)
        hi = np.where(up, hi, mid)
    lam = smax + 0.5 * (lo + hi)
    return R / (lam[:, None] - s[None])


def psd_loglik(E, SE, w, n):
    F = (np.abs(SE) ** 2).sum(1)
    tr = float(np.real(np.vdot(E, SE)))
    return float(w @ np.log(np.maximum(F, 1e-300)) - n * np.log(max(tr, 1e-300))) / n


def psd_residual(E, SE, w, n):
    """max_b |S e_b |e_b|^2 / (n w_b) - e_b| / |e_b| (blml_gpu.py:616-620)."""
    en2 = (np.abs(E) ** 2).sum(1)
    dev = SE * (en2 / (n * w))[:, None] - E
    return float(np.max(np.sqrt((np.abs(dev) ** 2).sum(1) / np.maximum(en2, 1e-300))))


def psd_apply(Gs, D, cols=None):
    """S restricted to the columns cols (all bins: None) times D (m, K): one Gram for every mode, or one per mode."""
    sl = slice(None) if cols is None else cols
    if len(Gs) == 1:
        return Gs[0][:, sl] @ D
    return np.stack([Gs[k][:, sl] @ D[:, k] for k in range(D.shape[1])], 1)


def psd_cd(Gs, w, e0, Bs, rng, tol, max_sweeps=None, anderson=5):
    """Block-Jacobi CD with Anderson mixing and the likelihood safeguard of the rank-K / complex counts fit
    (blml_gpu.py:591-744) with dense Grams Gs (one: shared kernel, else one per mode).  e0: the start (psd_start).
    Returns (E, SE, n_sweeps, converged, max_rho, rho_hist, obj_hist, safeguard, aa_accepted, SE_exact)."""
    w = np.asarray(w, float)
    n = float(w.sum())
    B, K = e0.shape
    shared = len(Gs) == 1
    sbbs = np.array([float(G[0, 0]) for G in Gs] * (K if shared else 1))
    E = np.array(e0)
    SE = psd_apply(Gs, E)
    sc = n / np.sqrt(max(float(np.real(np.vdot(E, SE))), 1e-300))
    E *= sc
    SE *= sc
    steps = 0.5 ** np.arange(0, 31)
    perm_fixed = rng.permutation(B) if anderson else None
    cplx = np.iscomplexobj(E)

    def sweep(safeguard):
        nonlocal E, SE
        perm = perm_fixed if perm_fixed is not None else rng.permutation(B)
        for start in range(0, B, Bs):
            I = perm[start:start + Bs]
            EI = E[I]
            R = SE[I] - sbbs[None] * EI
            d = psd_row_update(R, EI, n * w[I], sbbs, shared) - EI
            V = psd_apply(Gs, d, I)
            if safeguard:
                a0 = float(np.real(np.vdot(E, SE)))
                a1 = float(np.real(np.vdot(d, SE[I])) + np.real(np.vdot(E, V)))
                a2 = float(np.real(np.vdot(d, V[I])))
                F = (np.abs(SE[None] + steps[:, None, None] * V[None]) ** 2).sum(2)
                Ls = np.log(np.maximum(F, 1e-300)) @ w - n * np.log(np.maximum(a0 + a1 * steps + a2 * steps ** 2,
                                                                                1e-300))
                L0 = float(w @ np.log(np.maximum((np.abs(SE) ** 2).sum(1), 1e-300))) - n * np.log(max(a0, 1e-300))
                ok = Ls >= L0
                st = float(steps[int(np.argmax(ok))]) if ok.any() else 0.0
                E[I] = EI + st * d
                SE += st * V
            else:
                E[I] = EI + d
                SE += V

    flat = (lambda a: np.concatenate([a.real.ravel(), a.imag.ravel()])) if cplx else (lambda a: a.ravel())
    hG, hSG, hR = [], [], []
    aa_acc = 0
    safeguard, n_sweeps, converged = False, 0, False
    rho_hist, obj_hist = [], []
    L_prev = psd_loglik(E, SE, w, n)
    max_rho = psd_residual(E, SE, w, n)
    SE_end = None
    while True:
        if max_sweeps is not None and n_sweeps >= max_sweeps:
            break
        n_sweeps += 1
        E_save, SE_save = E.copy(), SE.copy()
        sweep(safeguard)
        L_now = psd_loglik(E, SE, w, n)
        if not safeguard and L_now < L_prev - 1e-10 * abs(L_prev):
            E, SE = E_save, SE_save
            safeguard = True
            sweep(True)
            L_now = psd_loglik(E, SE, w, n)
        if anderson:
            hG.append(E.copy())
            hSG.append(SE.copy())
            hR.append(flat(E - E_save).astype(np.float64))
            if len(hG) > anderson + 1:
                hG.pop(0)
                hSG.pop(0)
                hR.pop(0)
            if len(hG) >= 2:
                dR = np.stack([hR[j + 1] - hR[j] for j in range(len(hR) - 1)], 1)
                G = dR.T @ dR
                b = dR.T @ hR[-1]
                gam = np.linalg.lstsq(G + 1e-12 * np.trace(G) * np.eye(len(G)), b, rcond=None)[0]
                E_aa, SE_aa = hG[-1].copy(), hSG[-1].copy()
                for j, gj in enumerate(gam):
                    E_aa -= float(gj) * (hG[j + 1] - hG[j])
                    SE_aa -= float(gj) * (hSG[j + 1] - hSG[j])
                L_aa = psd_loglik(E_aa, SE_aa, w, n) if np.all(np.isfinite(E_aa)) else -np.inf
                if L_aa >= L_now:
                    E, SE, L_now = E_aa, SE_aa, L_aa
                    aa_acc += 1
                else:
                    hG, hSG, hR = hG[-1:], hSG[-1:], hR[-1:]
        L_prev = L_now
        if not np.all(np.isfinite(E)):
            max_rho = np.inf
            rho_hist.append(max_rho)
            break
        max_rho = psd_residual(E, SE, w, n)
        if max_rho < tol:
            SE = psd_apply(Gs, E)
            max_rho = psd_residual(E, SE, w, n)
        rho_hist.append(max_rho)
        obj_hist.append(-L_now)
        if max_rho < tol:
            converged = True
            break
        if (len(rho_hist) > 500 and rho_hist[-1] > 0.99 * min(rho_hist[:-500])
                and obj_hist[-501] - obj_hist[-1] < 1e-9 * abs(obj_hist[-1])):
            break
    if not converged and max_sweeps is not None and np.all(np.isfinite(E)):
        SE = psd_apply(Gs, E)
        sc = float(n / np.sqrt(max(float(np.real(np.vdot(E, SE))), 1e-300)))
        E, SE = E * sc, SE * sc
        max_rho = psd_residual(E, SE, w, n)
        converged = max_rho < tol
        obj_hist.append(-psd_loglik(E, SE, w, n))
        rho_hist.append(max_rho)
        SE_end = SE
    elif converged:
        SE_end = SE
    return E, SE, n_sweeps, converged, max_rho, rho_hist, obj_hist, safeguard, aa_acc, SE_end


def lbfgs_psd(Gs, w, e0, tol, max_iter, mem):
    """solver='lbfgs' of the rank-K / complex counts fit (blml_trivial.py:2590-2651, called at blml_gpu.py:605-609)
    on dense Grams: L-BFGS ascent of L = sum w log F - n log Q from e0 scaled to Q = n^2, Armijo c1 = 1e-4 with at
    most 40 halvings from the unit step (the first direction 1e-3 |e| / |g| times g), pairs kept with
    s.y > 1e-12 |s| |y|, one memory reset on a failed line search, the 1e-11 ll stop over 10 iterations, then the
    canonical form e = (Q / n) (w / F) S e kept unless it loses likelihood.  Returns
    (E, SE, iterations, converged, max_rho, rho_hist, obj_hist)."""
    w = np.asarray(w, float)
    n = float(w.sum())
    SE_of = lambda e: psd_apply(Gs, e)
    dot = lambda a, b: float(np.real(np.vdot(a, b)))
    loglik = lambda e, Se: psd_loglik(e, Se, w, n)
    residual = lambda e, Se: psd_residual(e, Se, w, n)
    e = np.array(e0)
    Se = SE_of(e)
    sc = n / np.sqrt(max(dot(e, Se), 1e-300))
    e, Se = e * sc, Se * sc

    def grad(e, Se):
        F = (np.abs(Se) ** 2).sum(1)
        Q = dot(e, Se)
        return 2 * (SE_of((w / np.maximum(F, 1e-300))[:, None] * Se) - (n / Q) * Se)

    def nrm(e, Se):
        s = n / np.sqrt(max(dot(e, Se), 1e-300))
        return e * s, Se * s

    L = loglik(e, Se) * n
    g = grad(e, Se)
    S_, Y_, R_ = [], [], []
    it = 0
    cap = np.inf if max_iter is None else max_iter
    max_rho = residual(*nrm(e, Se))
    rho_hist, obj_hist = [max_rho], [-L / n]
    reset = llstop = False
    while max_rho >= tol and it < cap:
        if len(obj_hist) > 10 and obj_hist[-11] - obj_hist[-1] < 1e-11:
            llstop = True
            break
        it += 1
        q = g.copy()
        al = []
        for s_, y_, r_ in zip(reversed(S_), reversed(Y_), reversed(R_)):
            a = r_ * dot(s_, q)
            al.append(a)
            q = q - a * y_
        q = q * (dot(S_[-1], Y_[-1]) / dot(Y_[-1], Y_[-1]) if S_ else 1e-3 * np.sqrt(dot(e, e) / max(dot(g, g), 1e-300)))
        for (s_, y_, r_), a in zip(zip(S_, Y_, R_), reversed(al)):
            q = q + (a - r_ * dot(y_, q)) * s_
        slope = dot(g, q)
        if not slope > 0:
            S_, Y_, R_ = [], [], []
            q = g * (1e-3 * np.sqrt(dot(e, e) / max(dot(g, g), 1e-300)))
            slope = dot(g, q)
        t = 1.0
        ok = False
        for _ in range(40):
            en = e + t * q
            Sen = SE_of(en)
            Ln = loglik(en, Sen) * n
            if Ln >= L + 1e-4 * t * slope:
                ok = True
                break
            t *= 0.5
        if not ok:
            if reset or not S_:
                break
            S_, Y_, R_ = [], [], []
            reset = True
            continue
        reset = False
        gn = grad(en, Sen)
        s_, y_ = en - e, g - gn
        sy = dot(s_, y_)
        if sy > 1e-12 * np.sqrt(dot(s_, s_) * dot(y_, y_)):
            S_.append(s_)
            Y_.append(y_)
            R_.append(1.0 / sy)
            if len(S_) > mem:
                S_.pop(0)
                Y_.pop(0)
                R_.pop(0)
        e, Se, L, g = en, Sen, Ln, gn
        max_rho = residual(*nrm(e, Se))
        rho_hist.append(max_rho)
        obj_hist.append(-L / n)
    F = (np.abs(Se) ** 2).sum(1)
    Q = dot(e, Se)
    ec = (Q / n) * (w / np.maximum(F, 1e-300))[:, None] * Se
    Sec = SE_of(ec)
    if loglik(ec, Sec) * n >= L - 1e-9 * n:
        e, Se = ec, Sec
    e, Se = nrm(e, Se)
    max_rho = residual(e, Se)
    converged = max_rho < tol or llstop
    rho_hist.append(max_rho)
    obj_hist.append(-loglik(e, Se))
    return e, Se, it, converged, max_rho, rho_hist, obj_hist


def psd_fit(Gs, w, tol=1e-4, block_size=4096, seed=0, max_sweeps=None, solver='cd', lbfgs_m=20, anderson=5,
            phase_noise=0.1, K=1, complex_g=False, ws=None):
    """The whole rank-K / complex counts fit (blml_gpu.py:332-441 then 461-486 with n_starts = (0, 0)): the real
    rank-1 fit under mode 0's Gram (quick_fit, whose sweeps draw from the same default_rng(seed) first), then
    psd_start and psd_cd or lbfgs_psd.  ws: a warm start (B, K') replacing the real start's column 0."""
    G0 = np.asarray(Gs[0], float)
    w = np.asarray(w, float)
    B = len(w)
    rng = np.random.default_rng(seed)
    Bs = max(1, min(int(block_size), -(-B // 8)))
    e_real = quick_fit(G0, w, tol, block_size, max_sweeps=max_sweeps, rng=rng)[0]
    e0 = psd_start(e_real, K, complex_g, rng, phase_noise, ws)
    if solver == 'lbfgs':
        return lbfgs_psd(Gs, w, e0, tol, max_sweeps, lbfgs_m)
    return psd_cd(Gs, w, e0, Bs, rng, tol, max_sweeps, anderson)



def quick_fit_dep(G, Gm, w, tol=1e-4, block_size=4096, seed=0, max_sweeps=None):
    """The conditional counts fit of the reference (blml_gpu.py:226-441, nfit = 2): fit 0 the joint (Gram G), fit 1
    the marginal on the same bins (Gram Gm).  Returns [(e, Se, n_sweeps, converged, max_rho), ...] per fit."""
    Gs = [np.asarray(G, float), np.asarray(Gm, float)]
    w = np.asarray(w, float)
    B = len(w)
    n = float(w.sum())
    rng = np.random.default_rng(seed)
    Bs = max(1, min(int(block_size), -(-B // 8)))
    es = [np.ones(B), np.ones(B)]
    Ses = [Gs[f] @ es[f] for f in range(2)]
    sbbs = [float(Gs[f][0, 0]) for f in range(2)]

    def resid(e, Se):
        rho = np.abs(Se * e / (n * w) - 1.0)
        return float(np.max(np.where((e > 0) & (w > 0), rho, 0.0)))

    def obj(e, Se):
        return float(0.5 * e @ Se - n * np.sum(w * np.log(np.maximum(e, 1e-300)))) / n

    def run(fits, safe):
        perm = rng.permutation(B)
        for f in fits:
            sweep(Gs[f], w, n, sbbs[f], es[f], Ses[f], perm, Bs, safe[f])

    safe, done, conv = [False, False], [False, False], [False, False]
    sweeps, rhos, hists = [0, 0], [np.inf, np.inf], [[], []]
    prev = [obj(es[f], Ses[f]) for f in range(2)]
    total = 0
    while not all(done):
        if max_sweeps is not None and total >= max_sweeps:
            break
        total += 1
        active = [f for f in range(2) if not done[f]]
        saves = {f: (es[f].copy(), Ses[f].copy()) for f in active}
        run(active, safe)
        now = {f: obj(es[f], Ses[f]) for f in active}
        redo = [f for f in active if not safe[f] and now[f] > prev[f] + 1e-10 * abs(prev[f])]
        for f in redo:
            es[f][:], Ses[f][:] = saves[f]
            safe[f] = True
        if redo:
            run(redo, safe)
            for f in redo:
                now[f] = obj(es[f], Ses[f])
        for f in active:
            sweeps[f] += 1
            prev[f] = now[f]
            if not np.all(np.isfinite(es[f])):
                rhos[f] = np.inf
                hists[f].append(np.inf)
                done[f] = True
                continue
            rho = resid(es[f], Ses[f])
            if rho < tol:
                Ses[f][:] = Gs[f] @ es[f]
                rho = resid(es[f], Ses[f])
            rhos[f] = rho
            hists[f].append(rho)
            if rho < tol:
                conv[f] = done[f] = True
            elif len(hists[f]) > 50 and hists[f][-1] > 0.99 * min(hists[f][:-50]):
                done[f] = True
    return [(es[f], Ses[f], sweeps[f], conv[f], rhos[f]) for f in range(2)]
