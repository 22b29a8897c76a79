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

Pinning (tests/test_oracle_quick.py): root by its quadratic's residual and the 1-D objective's derivative; the
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


def quick_fit(G, w, tol=1e-4, block_size=4096, seed=0, max_sweeps=None, log_steps=None, e0=None):
    """(e, Se, n_sweeps, converged, max_rho, rho_hist, obj_hist, safeguard) of the real rank-1 counts fit with Gram
    matrix G (B x B, the signs already in it: D S D) and counts w (B,).  log_steps: a list that receives
    (e_before, dlt, t, I) for every safeguarded block step.  e0: the start (warm_start), else ones."""
    G = np.asarray(G, float)
    w = np.asarray(w, float)
    B = len(w)
    n = float(w.sum())
    sbb = float(G[0, 0])
    rng = np.random.default_rng(seed)
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
