"""The engine's fit path compiled for the CPU under AddressSanitizer (tests/cpu_build: the device headers with a CPU
stand-in for HIP, every buffer, LDS object and workspace region bounded).  A replay stops at the first access
outside an allocation, so these runs check every index of the paths they take; their results are checked against
the oracle (real rank 1: the unique optimum, loglik 1e-9 relative), against the one-workgroup run (a spread fit on
a grid of 4 workgroups, separable tile passes on emulated matrix cores: loglik 1e-9) and against dense eigvalsh
(certificate, 1e-6).  No GPU."""
import os
import shutil
import sys

import numpy as np
import pytest

from oracle import blml_oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, 'cpu_build'))

CXX = os.environ.get('BLML_REPLAY_CXX', '/opt/rocm/llvm/bin/clang++')
pytestmark = pytest.mark.skipif(not shutil.which(CXX) and not os.path.exists(CXX),
                                reason='no clang++ for the CPU build')

PARAMS = dict(tol=1e-8, max_sweeps=None, anderson=5, stall=1000, phase_noise=0.1, ll_tol=None, ll_window=100,
              solver='cd', lbfgs_m=20)


@pytest.fixture(scope='module')
def replay():
    import replay as R
    R.build()
    return R


def _batch(counts, modes, cplx, dtype, spread, params, h=1.0):
    from blml_amd.binned import _FitSpec, FitBatch
    fs = _FitSpec(counts, modes, h, 0.0, cplx, None, None, 0)
    return FitBatch([fs], params, dtype, spread=spread)


def _field_counts(rng, shape, n):
    g = np.meshgrid(*[np.arange(m, dtype=float) for m in shape], indexing='ij')
    p = np.exp(-0.5 * sum(((x - m / 2.2) / (m / 4.5)) ** 2 for x, m in zip(g, shape))) + 0.05
    return rng.multinomial(n, (p / p.sum()).ravel()).reshape(shape).astype(float)


def test_real_rank1_fit_matches_oracle(replay):
    import blml_amd
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(-1.5, 0.5, 3000), rng.normal(1.8, 0.8, 2000)])
    counts = np.histogram(x, bins=512, range=(-6, 6))[0].astype(float)
    out = replay.run(_batch(counts, blml_amd.make_modes([0.05], ['sinc']), False, np.float64, -1, PARAMS))
    r = out['results'][0]
    ref = O.BLMLBinnedOracle(tol=1e-8).fit(counts, O.make_modes([0.05], ['sinc']))
    assert abs(r['loglik'] - ref.loglik_) < 1e-9 * abs(ref.loglik_)
    assert int(r['stop_reason']) == 0 and ref.stop_reason_ == 'tol'
    occ = counts > 0
    assert np.max(np.abs(out['density'][:counts.size] - ref.density_)[occ] / ref.density_[occ]) < 1e-6


def test_complex_slits_lbfgs_one_workgroup_in_bounds(replay):
    """The FP64 one-workgroup L-BFGS fit of a 40 x 36 complex slits lattice with a dirichlet background (padded
    power-of-two FFT lines of 128 values, band-restricted DFTs, two modes), capped at 30 iterations: every access
    of its start solve, L-BFGS iterations and outputs stays inside its allocation."""
    import blml_amd
    rng = np.random.default_rng(12)
    counts = _field_counts(rng, (40, 36), 20000)
    modes = (blml_amd.make_modes([0.35, 0.3], ['slits:0.5', 'slits:0.4'])
             + blml_amd.make_modes([3 / 40, 3 / 36], ['dirichlet', 'dirichlet'], periods=[40.0, 36.0]))
    params = dict(PARAMS, solver='lbfgs', max_sweeps=30, ll_tol=1e-10)
    out = replay.run(_batch(counts, modes, True, np.float64, -1, params))
    r = out['results'][0]
    assert int(r['n_sweeps']) == 30 and np.isfinite(r['loglik'])
    assert np.all(np.isfinite(out['density'][:counts.size]))


@pytest.mark.parametrize('dtype', [np.float64, np.float32], ids=['f64', 'f32'])
def test_fft_stage_order_gives_the_same_fit(replay, dtype):
    """The autotuner's FFT knob (the remainder radix stage first instead of last) on padded FFT lines of 128 and 64
    values, band DFTs and a group factor: every access in bounds and the same fit to the transforms' rounding."""
    import blml_amd
    rng = np.random.default_rng(5)
    counts = _field_counts(rng, (40, 30), 20000)
    modes = (blml_amd.make_modes([0.35, 0.3], ['slits:0.5', 'slits:0.4'])
             + blml_amd.make_modes([0.3, 0.3], ['pball', 'pball'], periods=[40.0, 30.0]))
    params = dict(PARAMS, solver='lbfgs', max_sweeps=12, ll_tol=1e-10)
    a = replay.run(_batch(counts, modes, True, dtype, -1, params))
    b = replay.run(_batch(counts, modes, True, dtype, -1, params), fft_sched=1)
    tol = 1e-12 if dtype == np.float64 else 1e-5
    la, lb = a['results']['loglik'][0], b['results']['loglik'][0]
    assert abs(la - lb) < tol * abs(la)
    assert np.max(np.abs(a['density'] - b['density'])) < tol * np.max(a['density']) * 10


def test_spread_separable_fit_equals_one_workgroup(replay):
    """A 4-D sinc fit spread over a grid of 4 workgroups (FP64 separable tile passes on emulated matrix cores, grid
    barriers, rank-ordered team sums) against the same fit on one workgroup (transform passes)."""
    import blml_amd
    rng = np.random.default_rng(3)
    counts = _field_counts(rng, (12, 10, 9, 8), 20000)
    modes = blml_amd.make_modes([0.35] * 4, ['sinc'] * 4)
    a = replay.run(_batch(counts, modes, False, np.float64, 1, PARAMS), ranks=4)
    b = replay.run(_batch(counts, modes, False, np.float64, -1, PARAMS))
    assert 'separable tile passes' in a['log']
    la, lb = a['results'][0]['loglik'], b['results'][0]['loglik']
    assert abs(la - lb) < 1e-9 * abs(lb)


def test_pipelined_matvec_centrosymmetric_split(replay):
    """The stand-alone mat-vec of an FP32 pipelined plan (matvec_pipe_kernel's work on emulated matrix cores, one
    32^3 tile: line strides 1, 32 and 1024): the centrosymmetric split of the line operators (two 16 x 16 products
    per line on v_mfma_f32_16x16x4f32) and the dense 32 x 32 products both equal the oracle's FFT mat-vec to FP32
    rounding (2e-6 of max |S e| times log2 B), and each other to the same bound."""
    import blml_amd
    shape, fc = (32, 32, 32), [0.3, 0.2, 0.35]
    counts = np.ones(shape)
    E, Ud, Uc = replay.run_pipe(_batch(counts, blml_amd.make_modes(fc, ['sinc'] * 3), False, np.float32, 1, PARAMS),
                                ranks=2)
    om = O.BLMLBinnedOracle()
    om._setup(counts, O.make_modes(fc, ['sinc'] * 3), 1.0, 0.0)
    want = om._S(0, E.astype(np.float64).reshape(shape)).ravel()
    bound = 2e-6 * np.abs(want).max() * np.log2(want.size)
    assert np.abs(Ud - want).max() < bound
    assert np.abs(Uc - want).max() < bound
    assert np.abs(Uc - Ud).max() < bound


def test_fp32_plan_and_certificate(replay):
    """An FP32 plan (FP32 transforms, the FP64 finishing pass) and the certificate of an FP64 fit (power iteration
    with the eigen-residual stop) against dense eigvalsh."""
    import blml_amd
    rng = np.random.default_rng(2)
    N = _field_counts(rng, (16, 16), 4000)
    modes = blml_amd.make_modes([0.4, 0.4], ['ball', 'ball'])
    out32 = replay.run(_batch(N, modes, False, np.float32, -1, dict(PARAMS, tol=1e-5)))
    out = replay.run(_batch(N, modes, False, np.float64, -1, PARAMS), cert_iters=20000, cert_tol=1e-9)
    r32, r = out32['results'][0], out['results'][0]
    assert abs(r32['loglik'] - r['loglik']) < 1e-6 * abs(r['loglik'])
    om = O.BLMLBinnedOracle()
    om._setup(N, O.make_modes([0.4, 0.4], ['ball', 'ball']), 1.0, 0.0)
    occ = N.ravel() > 0
    S = O.dense_gram(om.specs_[0], N.shape, om.h_)[np.ix_(occ, occ)]
    E = out['e'][:N.size][occ]
    D = E ** 2 / N.ravel()[occ]
    want = np.linalg.eigvalsh(np.sqrt(D)[:, None] * S * np.sqrt(D)[None])[-1] / N.sum()
    assert out['iters'][0] > 0 and abs(out['lam'][0] - want) < 1e-6 * want


# ------------------------------------------------------------------------------------------------ BLMLEmpirical
def _emp_replay(replay, est, counts, modes, **fk):
    from blml_amd.binned import FitBatch
    fs, params, ep, aux = est._prepare(counts, modes, 1.0, 0.0, fk.get('init'), fk.get('complex_g', False))
    b = FitBatch([fs], params, est.dtype, hist_cap=int(est.max_outer) + 2, spread=-1)
    return replay.run(b, emp=(ep, aux[None]))


def _emp_counts(bins=64):
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(-1.5, 0.5, 3000), rng.normal(1.8, 0.8, 2000)])
    return np.histogram(x, bins=bins, range=(-6, 6))[0].astype(float)


@pytest.mark.parametrize('reg,lam,max_iter', [('ridge', 1e-2, 20000), ('l2', 1e-3, 20000), ('l1', 1e-2, 400)])
def test_empirical_linear_fits_match_oracle(replay, reg, lam, max_iter):
    """blml_empirical's device code (CG / FISTA, targets, normalisation) on the CPU against the oracle: the same
    iteration counts and the same fit to rounding (FP64 transforms), every access in bounds."""
    import blml_amd
    from blml_amd.empirical import BLMLEmpirical
    from oracle import blml_empirical as OE
    counts = _emp_counts()
    kw = dict(reg=reg, lam=lam, tol=1e-10, max_iter=max_iter)
    out = _emp_replay(replay, BLMLEmpirical(**kw), counts, blml_amd.make_modes([0.2], ['sinc']))
    ref = OE.BLMLEmpiricalOracle(**kw).fit(counts, O.make_modes([0.2], ['sinc']))
    r = out['results'][0]
    assert int(r['n_sweeps']) == ref.n_iter_
    assert abs(r['loglik'] - ref.loglik_) < 1e-12 * abs(ref.loglik_)
    assert np.max(np.abs(out['density'][:counts.size] - ref.density_)) < 1e-10 * ref.density_.max()
    Zw, Zp, _, sc = out['emp'][0]
    assert Zw == pytest.approx(ref.window_norm_, rel=1e-12) and Zp == pytest.approx(ref.plane_norm_, rel=1e-12)


def test_empirical_retrieve_matches_reference_golden(replay):
    """The hidden-phase alternation with Anderson mixing (complex, ridge) on the CPU against the reference's own
    fit (tests/golden/empirical_fits.npz): the same fixed point (loglik 1e-8); the trajectory's length differs with
    the summation order of the dots (Anderson's accept test compares nearly equal J values)."""
    import blml_amd
    from blml_amd.empirical import BLMLEmpirical
    from cases_empirical import empirical_cases
    z = np.load(os.path.join(HERE, 'golden', 'empirical_fits.npz'))
    name, counts, mm, est_kw, fit_kw = empirical_cases()[0]
    modes = [md for kw in mm for md in blml_amd.make_modes(**kw)]
    out = _emp_replay(replay, BLMLEmpirical(**est_kw), counts, modes, **fit_kw)
    r = out['results'][0]
    ref = float(z['c0_loglik'])
    assert abs(r['loglik'] - ref) < 1e-8 * abs(ref)
    assert np.max(np.abs(out['density'][:counts.size] - z['c0_density'])) < 1e-5 * z['c0_density'].max()
    J = out['merit_history'][0]
    assert len(J) == int(r['start_sweeps']) + 1 and np.all(np.diff(J) <= 1e-12 * np.abs(J[:-1]))


def test_start_search_runs_match_oracle(replay):
    """The start search's device code (hio_run: FFT line passes, the data and band constraints, the score) on the
    CPU under AddressSanitizer, on the 2-D 24 x 20 case's 64 x 64 grid: three random runs and the eight symmetry-
    breaking runs of 3 HIO + 2 ER steps (few enough that rounding has not grown: the iteration is chaotic over
    hundreds) equal the oracle's runs -- spectra to 1e-9 of their scale, scores to 1e-10 relative."""
    sys.path.insert(0, os.path.join(HERE, 'golden'))
    from cases_starts import starts_cases
    from oracle import blml_starts as OS
    name, Z, w, h, fc, kinds = starts_cases()[1]
    G, band, runs = OS.start_runs(Z, w, h, fc, kinds, hio_pool=3, n_hio_iter=3, n_er_iter=2, seed=0)
    M = int(np.prod(G['shape']))
    rng = np.random.default_rng(0)
    phase = np.stack([rng.random(G['shape']).ravel() for _ in range(3)])
    U, _ = OS.band_grid(G['shape'], G['hc'], fc, kinds)
    kick = np.stack([(sg * 1.0 * p).ravel() for p in OS.even_kicks(U) for sg in (1, -1)])
    occ = np.ravel_multi_index(G['idx'], G['shape'])
    P, g, score = replay.run_hio(G['shape'], G['L'], 3, 2, 0.9, G['Ab2'], band.ravel(), G['A'].ravel(), occ,
                                 G['wg'], phase, kick)
    assert len(runs) == len(P) == 11
    for r, (fam, _, Po, so) in enumerate(runs):
        assert np.abs(P[r] - Po.ravel()).max() < 1e-9 * np.abs(Po).max(), (r, fam)
        assert score[r] == pytest.approx(so, rel=1e-10)
        assert np.abs(g[r] - np.fft.ifftn(Po)[G['idx']]).max() < 1e-9 * np.abs(np.fft.ifftn(Po)).max()


@pytest.mark.parametrize('cplx', [True, False], ids=['complex', 'real'])
def test_direct_line_sums_match_the_oracle_matvec(replay, cplx):
    """The batched mat-vec of an FP32 plan whose padded line factors are summed directly (OP_TOEP: lines of 37 and
    9 values, strided and contiguous, real lines paired into complex ones, steps past the line's end; the 70-value
    axis left to the FFT by the length bound) and of the same plan by FFT only: both equal the oracle's FP64
    mat-vec of every mode to FP32 rounding (1e-5 of max |S e|), with every access in bounds."""
    import blml_amd
    shape = (37, 9, 70)
    counts = _field_counts(np.random.default_rng(8), shape, 30000)
    fc1, kinds1 = [0.35, 0.3, 0.3], ['slits:0.5', 'sinc', 'sinc']
    fc2, per = [3 / 37, 3 / 9, 3 / 70], [37.0, 9.0, 70.0]
    modes = blml_amd.make_modes(fc1, kinds1) + blml_amd.make_modes(fc2, ['dirichlet'] * 3, periods=per)
    batch = _batch(counts, modes, cplx, np.float32, -1, PARAMS)
    E, Uf = replay.run_matvec(batch, toep_max=0)
    _, Ud = replay.run_matvec(batch, toep_max=64)
    om = O.BLMLBinnedOracle()
    om._setup(counts, O.make_modes(fc1, kinds1) + O.make_modes(fc2, ['dirichlet'] * 3, periods=per), 1.0, 0.0)
    B, nc = counts.size, 2 if cplx else 1
    for k in range(2):
        e = E[k * B * nc:(k + 1) * B * nc].astype(np.float64).reshape(B, nc)
        want = np.stack([om._S(k, e[:, c].reshape(shape)).ravel() for c in range(nc)], 1).ravel()
        scale = np.max(np.abs(want))
        for U in (Uf, Ud):
            got = U[k * B * nc:(k + 1) * B * nc].astype(np.float64)
            assert np.max(np.abs(got - want)) < 1e-5 * scale


@pytest.mark.parametrize('cplx', [True, False], ids=['complex', 'real'])
def test_direct_line_sums_in_bounds_through_a_fit(replay, cplx):
    """A whole FP32 fit (start solve, 12 iterations, outputs, the FP64 finishing pass) with the padded lines of 37
    and 9 values summed directly: every access in bounds, finite results, and the loglik of the FFT run's fit to
    the FP32 iteration's own spread (the two differ in rounding only; 12 sweeps of FP32 CD amplify a change of one
    unit in the last place to about 1e-4, as FP32 against FP64 does; the mat-vecs themselves are pinned above)."""
    import blml_amd
    rng = np.random.default_rng(8)
    counts = _field_counts(rng, (37, 9, 70), 30000)
    modes = (blml_amd.make_modes([0.35, 0.3, 0.3], ['slits:0.5', 'sinc', 'sinc'])
             + blml_amd.make_modes([3 / 37, 3 / 9, 3 / 70], ['dirichlet'] * 3, periods=[37.0, 9.0, 70.0]))
    params = dict(PARAMS, solver='lbfgs', max_sweeps=12, ll_tol=1e-10)
    a = replay.run(_batch(counts, modes, cplx, np.float32, -1, params))
    b = replay.run(_batch(counts, modes, cplx, np.float32, -1, params), toep_max=64)
    la, lb = a['results']['loglik'][0], b['results']['loglik'][0]
    assert np.isfinite(lb) and np.all(np.isfinite(b['density']))
    assert abs(la - lb) < 1e-3 * abs(la)


@pytest.mark.parametrize('case', ['complex_1mode_odd_bins', 'real_3modes_odd_bins'])
def test_lbfgs_history_passes_in_bounds_with_ragged_lengths(replay, case):
    """The compact L-BFGS form's history passes read W consecutive FP32 values per load (16-byte loads when the
    slot length K B nc is even, else 8-byte ones) with a scalar tail: a complex one-mode fit of 37 x 9 bins
    (K B nc = 666, two values past the last whole group of 4) and a real three-mode fit of 37 x 9 bins (K B nc =
    999, odd: 8-byte loads and a one-value tail), FP32 and 15 L-BFGS iterations each: every access in bounds and
    finite results."""
    import blml_amd
    rng = np.random.default_rng(21)
    counts = _field_counts(rng, (37, 9), 20000)
    if case.startswith('complex'):
        modes, cplx = blml_amd.make_modes([0.35, 0.3], ['sinc', 'sinc']), True
    else:
        modes = (blml_amd.make_modes([0.35, 0.3], ['sinc', 'sinc'])
                 + blml_amd.make_modes([3 / 37, 3 / 9], ['dirichlet'] * 2, periods=[37.0, 9.0])
                 + blml_amd.make_modes([0.2, 0.2], ['slits:0.5', 'slits:0.5']))
        cplx = False
    params = dict(PARAMS, solver='lbfgs', max_sweeps=15, ll_tol=1e-12, tol=1e-12)
    out = replay.run(_batch(counts, modes, cplx, np.float32, -1, params))
    r = out['results'][0]
    assert np.isfinite(r['loglik']) and np.all(np.isfinite(out['density'][:counts.size]))
    assert int(r['n_sweeps']) >= 5
