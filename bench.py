"""Benchmark of the binned BLML fit on one MI355X.

Workload (the headline configuration): 2048 independent 2-D binned fits of uneven lattices, one persistent launch
per step.  Sides round(exp(U[ln 16, ln 128])) per axis; half the fits are a periodic-ball signal mode
(fc ~ U[0.15, 0.45], period = the lattice) plus a periodic-ball background (fc = 3 / sqrt(m1 m2)), the other half
a two-slit signal per column (fc ~ U[0.2, 0.5], fill ~ U[0.3, 0.7]) plus a dirichlet x dirichlet background with
the lattice as period (fc = 3 / m per column: one harmonic each way); half the fits complex_g.  Fields: CN(0, 1)
coefficients on the signal band's harmonics, density |g|^2 normalised with a 5% uniform background, Poisson
counts with totals log-uniform in [1e3, 1e6].  numpy default_rng(3).  L-BFGS (lbfgs_m 20), FP32 transforms with
FP64 reductions; the plain FP64 method is run on the same GPU as the reference point of the metric.

A step = the learned warm starts of all 2048 fits (host features, the network on the device, resampling, upload),
one persistent FP32 launch that fits every one to the solver's tolerance, the judge -- the plain FP64 method
started at every result, one launch (the reference's polish test: a fit it moves by >= 1e-6 relative is no local
maximum of the plain method) -- and the cold refit of each rejected fit.  value = fits / s of that step.  Beside
it: the cold FP32 launch and the plain FP64 method (cold) on the same GPU, the fraction of kept fits whose loglik is
within 1e-6 of the FP64 run's and the fraction that is a local maximum to 1e-6 by the polish test, the plans' build
times; the FP64 oracle on the host's cores (cpu_baseline); the large single fit (32^5 bins, product sinc, 1e7 draws of a 4-component Gaussian mixture,
spread over every CU) in both precisions; the pair path's rows over the occupied bins of a 128^3 lattice (pairs/s).
roofline: the Gram mat-vec of the large fit (the memory-bound case); roofline_batched: the batched mat-vec of the
2048 fits.  Prints one JSON line."""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, 'band-limited-ml_amd')):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK = 8.0e12
MFMA_FP32_PEAK = 157.3e12      # dense FP32 matrix-core peak of the MI355X
N_FITS = 2048
SOLVER = dict(solver='lbfgs', lbfgs_m=20, tol=1e-5, ll_tol=1e-9, max_sweeps=3000)


# ------------------------------------------------------------------------------------------------ workload
def _band_field(rng, shape, harmonics):
    """g(x) = sum_k a_k exp(2 pi i k.x / m), a_k ~ CN(0, 1), on the lattice."""
    ks = np.asarray(harmonics, float)
    a = (rng.standard_normal(len(ks)) + 1j * rng.standard_normal(len(ks))) / np.sqrt(2)
    g = np.zeros(shape, complex)
    e = [np.exp(2j * np.pi * np.outer(ks[:, j], np.arange(m)) / m) for j, m in enumerate(shape)]
    for i in range(len(ks)):
        g += a[i] * np.multiply.outer(e[0][i], e[1][i])
    return g


def _pball_harmonics(fc, P):
    R = fc / 2.0
    N = [int(np.floor(R * p + 1e-9)) for p in P]
    K1, K2 = np.meshgrid(np.arange(-N[0], N[0] + 1), np.arange(-N[1], N[1] + 1), indexing='ij')
    keep = (K1 / P[0]) ** 2 + (K2 / P[1]) ** 2 <= R * R * (1 + 1e-9) + 1e-12
    return np.stack([K1[keep], K2[keep]], 1)


def _slits_band(fc, fill, m):
    """Integer harmonics k of an m-periodic column inside the two-slit band (1 - fill) fc / 2 <= |k| / m <= fc / 2
    (the nearest pair +-k when no integer falls inside)."""
    k = np.arange(-(m // 2), m // 2 + 1)
    nu = np.abs(k) / m
    band = k[(nu <= fc / 2) & (nu >= (1 - fill) * fc / 2)]
    if len(band) == 0:                      # a lattice too coarse for the band: its nearest harmonic pair
        kc = max(1, int(round(m * fc * (2 - fill) / 4)))
        band = np.array([-kc, kc])
    return band


def uneven_fits(n_fits=N_FITS, seed=3):
    """[(counts, make_modes groups, fit kwargs, description)] of the batched uneven-lattice workload."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_fits):
        shape = tuple(int(round(math.exp(rng.uniform(math.log(16), math.log(128))))) for _ in range(2))
        pball = rng.uniform() < 0.5
        cplx = rng.uniform() < 0.5
        total = math.exp(rng.uniform(math.log(1e3), math.log(1e6)))
        P = [float(shape[0]), float(shape[1])]
        if pball:
            fc = float(rng.uniform(0.15, 0.45))
            H = _pball_harmonics(fc, P)
            fbg = 3.0 / math.sqrt(shape[0] * shape[1])
            groups = [dict(fc=[[fc, fc], [fbg, fbg]], kinds=['pball', 'pball'], periods=P)]
        else:
            fc = [float(rng.uniform(0.2, 0.5)) for _ in range(2)]
            fill = [float(rng.uniform(0.3, 0.7)) for _ in range(2)]
            bands = [_slits_band(fc[j], fill[j], shape[j]) for j in range(2)]
            H = np.stack(np.meshgrid(*bands, indexing='ij'), -1).reshape(-1, 2)
            groups = [dict(fc=fc, kinds=[f'slits:{fill[0]!r}', f'slits:{fill[1]!r}']),
                      dict(fc=[3.0 / shape[0], 3.0 / shape[1]], kinds=['dirichlet', 'dirichlet'], periods=P)]
        g = _band_field(rng, shape, H)
        p = np.abs(g) ** 2
        p = p / p.sum()
        p = 0.95 * p + 0.05 / p.size
        counts = rng.poisson(total * p).astype(float)
        if counts.sum() == 0:
            counts.flat[0] = 1.0
        out.append((counts, groups, dict(complex_g=bool(cplx)), dict(kind='pball' if pball else 'slits', shape=shape)))
    return out


LARGE_SHAPE = (32,) * 5
LARGE_H = 0.25                 # [-4, 4]^5 in 32 bins per axis
LARGE_FC = 0.12 / LARGE_H      # 0.12 cycles per bin, in 1 / coordinate units (the fit runs with h = 0.25)


def large_fit_counts(n=10_000_000, seed=2):
    """Counts of the large single fit: n draws of a 4-component 5-D Gaussian mixture (means U[-2, 2]^5, diagonal
    sds U[0.3, 0.8], equal weights; numpy default_rng(seed)), binned on 32^5 bins over [-4, 4]^5 (draws outside are
    dropped).  Returns (counts, make_modes groups, fit kwargs)."""
    rng = np.random.default_rng(seed)
    d, nb = len(LARGE_SHAPE), LARGE_SHAPE[0]
    means = rng.uniform(-2.0, 2.0, (4, d))
    sds = rng.uniform(0.3, 0.8, (4, d))
    comp = rng.integers(0, 4, n)
    x = means[comp] + sds[comp] * rng.standard_normal((n, d))
    idx = np.floor((x + 4.0) / LARGE_H).astype(np.int64)
    keep = np.all((idx >= 0) & (idx < nb), axis=1)
    flat = np.ravel_multi_index(tuple(idx[keep].T), LARGE_SHAPE)
    counts = np.bincount(flat, minlength=nb ** d).astype(float).reshape(LARGE_SHAPE)
    groups = [dict(fc=[LARGE_FC] * d, kinds=['sinc'] * d)]
    return counts, groups, dict(h=LARGE_H, origin=-4.0)


PAIR_BINS = 128
QUICK_SWEEPS = 5               # sweeps of the companion fit


def pair_companion(n=2_000_000, seed=2):
    """The pair path's lattice: n draws of the large fit's mixture in 3-D (means U[-2, 2]^3, sds U[0.3, 0.8], 4 equal
    components, numpy default_rng(seed)) binned on 128^3 bins over [-4, 4]^3; returns the occupied bins' integer
    coordinates (B, 3), their counts and the step."""
    rng = np.random.default_rng(seed)
    d, h = 3, 8.0 / PAIR_BINS
    means = rng.uniform(-2.0, 2.0, (4, d))
    sds = rng.uniform(0.3, 0.8, (4, d))
    comp = rng.integers(0, 4, n)
    x = means[comp] + sds[comp] * rng.standard_normal((n, d))
    idx = np.floor((x + 4.0) / h).astype(np.int64)
    idx = idx[np.all((idx >= 0) & (idx < PAIR_BINS), axis=1)]
    Z, cnt = np.unique(idx, axis=0, return_counts=True)
    return Z, cnt.astype(float), h


def fit_params():
    """The solver keywords of the batched workload (FitBatch form)."""
    return dict(tol=SOLVER['tol'], max_sweeps=SOLVER['max_sweeps'], anderson=5, stall=1000, phase_noise=0.1,
                ll_tol=SOLVER['ll_tol'], ll_window=100, solver=SOLVER['solver'], lbfgs_m=SOLVER['lbfgs_m'])


def modes_of(groups, lib):
    return [md for kw in groups for md in lib.make_modes(**kw)]


# ------------------------------------------------------------------------------------------------ bench
def cpu_baseline(fits, seconds=20.0):
    """The FP64 oracle (numpy + scipy.fft, workers=-1) on a random sample of the workload's fits (a fixed
    permutation, so the sample's sizes and kinds follow the workload's) until ~seconds of work."""
    from oracle import blml_oracle as O
    cores = len(os.sched_getaffinity(0))
    done, t0 = 0, time.perf_counter()
    bins = 0
    for i in np.random.default_rng(0).permutation(len(fits)):
        counts, groups, kw, _ = fits[i]
        O.BLMLBinnedOracle(**SOLVER).fit(counts, modes_of(groups, O), **kw)
        done += 1
        bins += counts.size
        if time.perf_counter() - t0 > seconds:
            break
    dt = time.perf_counter() - t0
    mean_all = sum(c.size for c, *_ in fits) / len(fits)
    return dict(value=done / dt, unit='fits/s', cores=cores, kind='port',
                sample=f'{done} of the {len(fits)} fits drawn at random ({bins / done:.0f} bins per fit; the '
                       f'workload averages {mean_all:.0f}), FP64 numpy/scipy.fft workers=-1, {dt:.1f} s')


PROFILES = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'profiles')


def measured_traffic(kernel, workload='32x5_sinc'):
    """HBM bytes per launch of a mat-vec kernel from its rocprofv3 counter summary
    (profiles/<kernel>_<workload>_traffic.txt: FETCH_SIZE + WRITE_SIZE, each scaled by the same counter's reading for
    a copy of known size), or None when no summary of that kernel exists."""
    name = kernel.split('<')[0]
    try:
        with open(os.path.join(PROFILES, f'{name}_{workload}_traffic.txt')) as f:
            for line in f:
                if line.startswith('traffic_bytes_per_launch:'):
                    return float(line.split(':', 1)[1])
    except OSError:
        return None
    return None


def measured_counter(stem, key):
    """A per-launch counter figure from a rocprofv3 summary in profiles/<stem>.txt (a line 'key: value'), or None."""
    try:
        with open(os.path.join(PROFILES, f'{stem}.txt')) as f:
            for line in f:
                if line.startswith(key + ':'):
                    return float(line.split(':', 1)[1])
    except OSError:
        return None
    return None


# VALU issue peak of the MI355X: 256 CUs x 4 SIMDs, one wave64 instruction per SIMD every 4 cycles (16 lanes per
# cycle), 2.4 GHz
VALU_PEAK_WAVE_INST = 256 * 4 / 4 * 2.4e9


def large_matvec(torch, blml_amd, _FitSpec, FitBatch, params, event_time, reps=20):
    """The 32^5 single fit's plan and its mat-vec timed with HIP events on the launch stream.  Returns
    (spec, fit params, description, batch, roofline dict, seconds per mat-vec)."""
    lc, lg, lkw = large_fit_counts()
    lfs = _FitSpec(lc, modes_of(lg, blml_amd), lkw['h'], lkw['origin'], False, None, None, 0)
    lparams = dict(params, solver='lbfgs', lbfgs_m=10, ll_tol=1e-10, max_sweeps=3000)
    large = dict(bins=int(lfs.B), shape=list(LARGE_SHAPE), modes='sinc fc=0.12 cycles/bin per axis, h=0.25',
                 counts=int(lc.sum()), solver='cd (real rank 1; the reference ignores lbfgs there)')
    bl = FitBatch([lfs], lparams, np.float32, hist_cap=0).build()
    El = torch.randn(lfs.B, device='cuda', dtype=torch.float32)
    bl.matvec(El)
    t_l = event_time(lambda: bl.matvec(El), reps)
    alg_l = 2 * lfs.B * 4
    kname = bl.info()['matvec_kernel'] + '<float>'
    # the dense line operators' arithmetic: one multiply-add per input value of the line, per value and axis; the
    # FP32 pipelined kernel applies each (centrosymmetric) operator as two half-size products, half those FLOPs
    flops = 2 * lfs.B * sum(LARGE_SHAPE)
    done = flops // 2 if kname == 'matvec_pipe_kernel<float>' else flops
    roof = dict(bound='hbm', achieved=alg_l / t_l / 1e9, peak=HBM_PEAK / 1e9, unit='GB/s', frac=alg_l / t_l / HBM_PEAK,
                traffic=measured_traffic(kname), kernel=kname, bytes_per_launch=alg_l,
                us_per_launch=t_l * 1e6, workload='32^5 product-sinc mat-vec (separable tile passes, FP32)',
                flops_per_launch=flops, mfma_flops_per_launch=done, mfma_achieved=done / t_l / 1e12,
                mfma_peak=MFMA_FP32_PEAK / 1e12, mfma_unit='TFLOP/s', mfma_frac=done / t_l / MFMA_FP32_PEAK)
    return lfs, lparams, large, bl, roof, t_l


def batched_matvec(torch, batch, specs, event_time, reps):
    """The batched Gram mat-vec of every fit and mode of a plan (one workgroup per fit) timed with HIP events on the
    launch stream; its roofline dict."""
    g = torch.Generator(device='cuda').manual_seed(3)
    E = torch.randn(batch.sizes['out'], device='cuda', dtype=torch.float32, generator=g)
    batch.matvec(E)
    t_mv = event_time(lambda: batch.matvec(E), reps)
    alg = sum(2 * fs.K * fs.nc * fs.B * 4 for fs in specs)
    kname = batch.info()['matvec_kernel'] + '<float>'
    return dict(bound='hbm', achieved=alg / t_mv / 1e9, peak=HBM_PEAK / 1e9, unit='GB/s', frac=alg / t_mv / HBM_PEAK,
                traffic=measured_traffic(kname, 'c3_2048_fits'), kernel=kname, bytes_per_launch=alg,
                us_per_launch=t_mv * 1e6)


def copy_known_size(torch):
    """A plain device copy of 256 MiB (its FETCH_SIZE / WRITE_SIZE readings scale the counters)."""
    cp_src = torch.empty(64 << 20, dtype=torch.float32, device='cuda')
    cp_dst = torch.empty_like(cp_src)
    cp_dst.copy_(cp_src)
    torch.cuda.synchronize()
    del cp_src, cp_dst


def pair_leg(torch, st, event_time):
    """The pair path on the 128^3 companion: the rows of 65,536 occupied bins against every one (pairs/s) and the
    companion BLMLQuick fit through the same kernels."""
    # the pair path: rows s_r = sum_j S(z_r, z_j) w_j over the companion's occupied bins (slits per column),
    # FP32 tables and products, FP64 sums; a row subset against every source
    import ctypes
    from blml_amd import _lib, local
    Zp, wp, hp = pair_companion()
    tabp = local.lattice_tables(['slits:0.5'] * 3, [0.2 / hp] * 3, None, [hp] * 3, Zp)
    n_rows = min(len(Zp), 1 << 16)
    pt, dA, dB, keep = local._prepare(Zp[:n_rows], Zp, tabp, True, False)
    dw = torch.from_numpy(np.ascontiguousarray(wp)).to('cuda')
    ds = torch.zeros(n_rows, dtype=torch.float64, device='cuda')
    lib = _lib.load()
    call = lambda: _lib.check(lib.blml_pair_rows(ctypes.byref(pt), ctypes.c_void_p(dA.data_ptr()),
                                                 ctypes.c_void_p(dB.data_ptr()), ctypes.c_void_p(dw.data_ptr()),
                                                 n_rows, len(Zp), ctypes.c_void_p(ds.data_ptr()),
                                                 ctypes.c_void_p(st.cuda_stream)))
    call()
    t_p = event_time(call, 3)
    pairs = dict(bins=int(len(Zp)), rows=int(n_rows), sources=int(len(Zp)), draws=2_000_000, lattice=[PAIR_BINS] * 3,
                 kernel='slits fill 0.5, fc 0.2 cycles/bin per column (offset rows)', precision='f32 tables, f64 sums',
                 seconds=t_p, pairs_per_s=n_rows * len(Zp) / t_p)
    # the rows kernel's roofline: VALU issue (the kernel is bound by its per-pair index arithmetic, LDS lookups and
    # products; HBM carries only the sources).  VALU wave instructions per launch from rocprofv3 --pmc SQ_INSTS_VALU
    # (profiles/pair_rows_off_kernel_128cubed_slits_counters.txt), over this run's launch time
    vi = measured_counter('pair_rows_off_kernel_128cubed_slits_counters', 'valu_insts_per_launch')
    li = measured_counter('pair_rows_off_kernel_128cubed_slits_counters', 'lds_insts_per_launch')
    if vi:
        pairs['roofline'] = dict(bound='valu', kernel='pair_rows_off_kernel<float, 3, 4>',
                                 valu_insts_per_launch=vi, lds_insts_per_launch=li,
                                 valu_insts_per_pair=vi * 64 / (n_rows * len(Zp)),
                                 achieved=vi / t_p / 1e9, peak=VALU_PEAK_WAVE_INST / 1e9,
                                 unit='G wave64 VALU instructions/s', frac=vi / t_p / VALU_PEAK_WAVE_INST)
    del keep
    print(f'bench: pair rows {n_rows} x {len(Zp)} in {t_p * 1e3:.1f} ms', file=sys.stderr, flush=True)
    # derivative rows on the same bins (grad=True): dS/dlog fc of each column, three FP64 accumulators per row
    tabg = local.lattice_tables(['slits:0.5'] * 3, [0.2 / hp] * 3, None, [hp] * 3, Zp,
                                want=[(k, 'fc') for k in range(3)])
    n_g = 1 << 14
    launch_g, result_g = local.fused_rows_launcher(Zp[:n_g], Zp, wp[:, None], tabg, None, True, True)
    launch_g()
    t_g = event_time(launch_g, 3)
    result_g()
    pairs['grad_rows'] = dict(rows=n_g, sources=int(len(Zp)), entries=3, seconds=t_g,
                              pairs_per_s=n_g * len(Zp) / t_g)
    print(f'bench: derivative rows {n_g} x {len(Zp)} (3 entries) in {t_g * 1e3:.1f} ms', file=sys.stderr,
          flush=True)
    # the companion fit through the pair path: BLMLQuick on the lattice's counts (real rank 1, FP32 tables, block
    # 4096), a capped number of sweeps -- a band-pass slits kernel (fill 0.5) leaves the real rank-1 objective
    # without an interior optimum, so the fit runs to the cap and is scaled to e'Se = n^2
    from blml_amd.quick import BLMLQuick
    modes_p = [{j: dict(kind='slits', fc=0.2 / hp, fill=0.5, h=hp) for j in range(3)}]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    qf = BLMLQuick(tol=1e-4, max_sweeps=QUICK_SWEEPS, seed=0).fit(Zp * hp, modes_p, sample_weight=wp)
    t_q = time.perf_counter() - t0
    passes = qf.n_sweeps_ + 2                     # the sweeps, the first exact S e and the capped one
    pairs['quick_fit'] = dict(estimator='BLMLQuick(tol=1e-4, block_size=4096, dtype=float32)', bins=int(qf.n_bins_),
                              sweeps=int(qf.n_sweeps_), seconds=t_q, full_passes=passes,
                              pairs_per_s=passes * float(qf.n_bins_) ** 2 / t_q, converged=bool(qf.converged_),
                              residual_history=[float(r) for r in qf.rho_history_])
    print(f'bench: companion fit {qf.n_sweeps_} sweeps in {t_q:.2f} s', file=sys.stderr, flush=True)
    return pairs, Zp, wp, tabp


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=3)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--fits', type=int, default=N_FITS)
    ap.add_argument('--no-fp64', action='store_true', help='skip the plain FP64 reference run')
    ap.add_argument('--cpu-seconds', type=float, default=20.0)
    ap.add_argument('--only', choices=['all', 'matvec', 'batched', 'pairs'], default='all',
                    help='matvec: only the 32^5 mat-vec (timed) and the 256 MiB copy, for counter runs; batched: '
                         'only the batched mat-vec of the 2048 fits (timed) and the copy; pairs: only the pair '
                         'rows and the companion fit')
    args = ap.parse_args()
    if args.gpus != 1:
        print(f'bench.py: this build measures one GPU; --gpus {args.gpus} is not supported (independent fits '
              'shard across GPUs with no exchange; see the design notes)', file=sys.stderr)
        sys.exit(2)
    import torch
    import blml_amd
    from blml_amd.binned import _FitSpec, FitBatch
    if args.only != 'all':
        st0 = torch.cuda.current_stream()

        def ev_time(fn, reps):
            ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            ev0.record(st0)
            for _ in range(reps):
                fn()
            ev1.record(st0)
            torch.cuda.synchronize()
            return ev0.elapsed_time(ev1) / 1e3 / reps

        if args.only == 'pairs':
            pairs, *_ = pair_leg(torch, st0, ev_time)
            print(json.dumps(dict(pairs=pairs)), flush=True)
            return
        if args.only == 'matvec':
            *_, bl, roof, _ = large_matvec(torch, blml_amd, _FitSpec, FitBatch, fit_params(), ev_time, args.steps)
            bl.close()
        else:
            fits = uneven_fits(args.fits)
            specs = [_FitSpec(c, modes_of(g, blml_amd), 1.0, 0.0, kw['complex_g'], None, None, 0)
                     for c, g, kw, _ in fits]
            bb = FitBatch(specs, fit_params(), np.float32, hist_cap=0).build().autotune()
            roof = batched_matvec(torch, bb, specs, ev_time, args.steps)
            roof['autotune'] = bb.tuned
            bb.close()
        copy_known_size(torch)
        print(json.dumps(dict(roofline=roof)), flush=True)
        return
    fits = uneven_fits(args.fits)
    print(f'bench: {len(fits)} fits, {sum(c.size for c, *_ in fits)} bins', file=sys.stderr, flush=True)
    specs = [_FitSpec(c, modes_of(g, blml_amd), 1.0, 0.0, kw['complex_g'], None, None, 0) for c, g, kw, _ in fits]
    params = fit_params()
    from blml_amd import warm as W
    net = W.WarmNet.load()

    # the warm-start plan and its judge: counts, tables and draws resident on the device before the clock starts
    # (their build time is reported); a step is the network's warm starts (host features, the network on the device,
    # resampling), their upload, one batched FP32 launch to the solver's tolerance, the judge -- the plain FP64
    # method started at every result (Polish, one launch; the reference's polish test) -- and the cold refit of
    # every fit the judge rejects (moved by >= 1e-6, or not converged), its plan built inside the step
    from blml_amd.binned import Polish
    W.apply(net, specs)
    torch.cuda.synchronize()
    tb0 = time.perf_counter()
    bw = FitBatch(specs, params, np.float32, hist_cap=0).build().autotune()
    torch.cuda.synchronize()
    t_build = time.perf_counter() - tb0
    tuned = dict(warm_fp32=bw.tuned)
    pol = Polish(bw, params)
    torch.cuda.synchronize()
    t_build_judge = time.perf_counter() - tb0 - t_build
    judged = {}

    def warm_step():
        W.apply(net, specs)
        bw.refresh_init().run()
        sw = bw.scalars()
        rel, _ = pol.run().rel(sw)
        redo = [i for i in range(len(specs)) if abs(rel[i]) >= 1e-6 or
                _lib.STOP_REASONS[int(sw['stop_reason'][i])] in ('stall', 'max_sweeps')]
        ll = sw['loglik'].copy()
        if redo:
            cold = [_FitSpec(fits[i][0], modes_of(fits[i][1], blml_amd), 1.0, 0.0, fits[i][2]['complex_g'], None,
                             None, 0) for i in redo]
            bc = FitBatch(cold, params, np.float32, hist_cap=0).build().run()
            ll[redo] = bc.scalars()['loglik']
            bc.close()
        judged.update(loglik=ll, fallbacks=len(redo), polish_rel=rel, warm=sw)

    from blml_amd import _lib
    for _ in range(args.warmup):
        warm_step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        warm_step()
    torch.cuda.synchronize()
    dtw = time.perf_counter() - t0
    rw = bw.results()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    bw.run()
    torch.cuda.synchronize()
    t_launch = time.perf_counter() - t1
    t2 = time.perf_counter()
    pol.run().rel()
    t_judge = time.perf_counter() - t2
    pol.close()
    bw.close()
    print(f'bench: warm fp32 {dtw / args.steps * 1e3:.1f} ms/step ({t_launch * 1e3:.1f} ms of it the launch, '
          f'{t_judge * 1e3:.1f} ms the FP64 judge, {judged["fallbacks"]} cold fallbacks)', file=sys.stderr, flush=True)
    for fs in specs:                         # back to the cold start (the draws are the cold start's already)
        fs.init, fs.init_K, fs.warm = None, 0, False

    def timed(dtype):
        b = FitBatch(specs, params, dtype, hist_cap=0).build().autotune()
        tuned['cold_' + np.dtype(dtype).name] = b.tuned
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b.run()
        torch.cuda.synchronize()
        return b, time.perf_counter() - t0

    b32, dt32 = timed(np.float32)
    r32 = b32.results()
    print(f'bench: cold fp32 {dt32 * 1e3:.1f} ms', file=sys.stderr, flush=True)
    ll32 = np.array([r['loglik'] for r in r32])
    llw = np.asarray(judged['loglik'])         # the kept results: warm where the judge passed them, else cold
    relw = (llw - ll32) / np.abs(ll32)
    extra = dict(cold_fp32_fits_per_s=len(fits) / dt32, warm_speedup_vs_cold_fp32=dt32 / (dtw / args.steps),
                 warm_launch_seconds=t_launch, judge_seconds=t_judge,
                 warm_start_seconds=dtw / args.steps - t_launch - t_judge,
                 cold_fallbacks=int(judged['fallbacks']),
                 judge_pass_fraction=float(np.mean(np.abs(judged['polish_rel']) < 1e-6)),
                 plan_build_seconds=dict(warm_fp32=t_build, judge_fp64=t_build_judge), autotune=tuned,
                 warm_iterations=float(np.mean([r['n_sweeps'] + r['start_sweeps'] for r in rw])),
                 cold_iterations=float(np.mean([r['n_sweeps'] + r['start_sweeps'] for r in r32])),
                 warm_vs_cold_better=int((relw > 1e-6).sum()), warm_vs_cold_worse=int((relw < -1e-6).sum()),
                 warm_vs_cold_median_rel_ll=float(np.median(relw)),
                 warm_unconverged=int(sum(r['stop_reason'] in ('stall', 'max_sweeps') for r in rw)))

    def polished(res, rel):
        """Fraction of fits within 1e-6 of the FP64 run's loglik or, where not, local maxima of their own: the FP64
        method warm-started at the result moves loglik by < 1e-6 (the reference's polish criterion,
        results/pball_parity.py:92-113; a non-convex fit can end at another local maximum)."""
        from types import SimpleNamespace
        far = [i for i in range(len(fits)) if rel[i] > 1e-6]
        ok = rel <= 1e-6
        if far:
            pol = []
            for i in far:
                c, g, kw, _ = fits[i]
                e0 = SimpleNamespace(e_hat_=np.asarray(res[i]['e_hat']).astype(
                    np.complex128 if kw['complex_g'] else np.float64), shape_=c.shape)
                pol.append(_FitSpec(c, modes_of(g, blml_amd), 1.0, 0.0, kw['complex_g'], None, e0, 0))
            bp = FitBatch(pol, params, np.float64, hist_cap=0).build().run()
            rp = bp.results()
            bp.close()
            ok[far] = [abs(r['loglik'] - res[i]['loglik']) / abs(res[i]['loglik']) < 1e-6 for r, i in zip(rp, far)]
        return float(np.mean(ok))

    if not args.no_fp64:
        b64, dt64 = timed(np.float64)
        r64 = b64.results()
        b64.close()
        ll64 = np.array([r['loglik'] for r in r64])
        rel = np.abs(ll32 - ll64) / np.abs(ll64)
        rw64 = np.abs(llw - ll64) / np.abs(ll64)
        extra.update(fp64_fits_per_s=len(fits) / dt64, speedup_vs_fp64=dt64 / (dtw / args.steps),
                     cold_fp32_speedup_vs_fp64=dt64 / dt32, frac_within_1e6=float(np.mean(rw64 <= 1e-6)),
                     cold_frac_within_1e6=float(np.mean(rel <= 1e-6)), median_rel_ll=float(np.median(rw64)))
        print(f'bench: fp64 {dt64 * 1e3:.1f} ms/step; cold fp32 within 1e-6: {extra["cold_frac_within_1e6"]:.3f}',
              file=sys.stderr, flush=True)
        extra['frac_local_max_1e6'] = polished(rw, rw64)
        extra['cold_frac_local_max_1e6'] = polished(r32, rel)
        print(f'bench: local maxima within 1e-6 (polish): warm {extra["frac_local_max_1e6"]:.3f}, cold '
              f'{extra["cold_frac_local_max_1e6"]:.3f}', file=sys.stderr, flush=True)
    st = torch.cuda.current_stream()

    def event_time(fn, reps):
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record(st)
        for _ in range(reps):
            fn()
        ev1.record(st)
        torch.cuda.synchronize()
        return ev0.elapsed_time(ev1) / 1e3 / reps

    # the batched Gram mat-vec of the workload (all fits, all modes), timed with HIP events on the launch stream
    roof_batched = batched_matvec(torch, b32, specs, event_time, 20)
    b32.close()
    # the large single fit: 32^5 bins spread over every CU; its Gram mat-vec is the memory-bound roofline
    lfs, lparams, large, bl, roof, t_l = large_matvec(torch, blml_amd, _FitSpec, FitBatch, params, event_time)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rl = bl.run().results()[0]
    large.update(fp32_seconds=time.perf_counter() - t0, loglik_fp32=rl['loglik'], sweeps_fp32=rl['n_sweeps'],
                 stop_fp32=rl['stop_reason'], residual_fp32=rl['residual'])
    # the certificate lambda_max(D^1/2 S D^1/2) / n of the result (blml_binned.py:501-518; <= 1 + tol only at a
    # stationary point whose rank-1 lift is optimal)
    lam, its = bl.certificate(max_iter=2000, tol=1e-9)
    large.update(certificate_fp32=float(lam[0]), certificate_iters_fp32=int(its[0]))
    bl.close()
    if not args.no_fp64:
        bl64 = FitBatch([lfs], lparams, np.float64, hist_cap=0).build()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rl64 = bl64.run().results()[0]
        t64 = time.perf_counter() - t0
        lam64, _ = bl64.certificate(max_iter=2000, tol=1e-9)
        d64 = np.asarray(rl64['density'])
        large.update(fp64_seconds=t64, loglik_fp64=rl64['loglik'], sweeps_fp64=rl64['n_sweeps'],
                     rel_loglik=abs(rl['loglik'] - rl64['loglik']) / abs(rl64['loglik']),
                     density_max_rel_err=float(np.abs(np.asarray(rl['density']) - d64).max() / np.abs(d64).max()),
                     certificate_fp64=float(lam64[0]))
        bl64.close()
    # a plain device copy of known size (256 MiB), the scale of rocprof's FETCH_SIZE / WRITE_SIZE counters
    copy_known_size(torch)
    print(f'bench: large fit fp32 {large["fp32_seconds"]:.2f} s; mat-vec {t_l * 1e6:.1f} us', file=sys.stderr,
          flush=True)
    pairs, Zp, wp, tabp = pair_leg(torch, st, event_time)
    cpu = cpu_baseline(fits, args.cpu_seconds)
    # the pair path's CPU baseline: the C restatement of the rows (oracle/pairs_cpu.c, OpenMP over every host core;
    # FP32 table products and FP64 sums as the device's) on the same 65,536 rows x 303,017 sources
    from oracle import blml_pairs as OP
    n_cpu = min(len(Zp), 1 << 16)
    cores = len(os.sched_getaffinity(0))
    t0 = time.perf_counter()
    OP.pair_rows_c(Zp[:n_cpu], Zp, wp, tabp, f32=True, threads=cores)
    pairs['cpu_pairs_per_s'] = n_cpu * len(Zp) / (time.perf_counter() - t0)
    pairs['cpu_sample'] = (f'{n_cpu} rows x {len(Zp)} sources (the timed launch), the C/OpenMP restatement '
                           f'(oracle/pairs_cpu.c), {cores} threads')
    ms = dtw / args.steps * 1e3
    line = dict(metric='fits/sec at loglik tol 1e-6 (warm-start FP32 vs plain FP64) + mat-vec HBM GB/s',
                value=len(fits) / (dtw / args.steps), unit='fits/s', n_gpus=1, steps=args.steps,
                warmup=args.warmup, ms_per_step=ms, higher_is_better=True, scaling='weak', vs_baseline=None,
                dtype='f32 transforms, f64 reductions', data='synthetic',
                config=dict(workload=f'{len(fits)} uneven 2-D binned fits (pball / slits + background, '
                                     'L-BFGS, learned warm start; a step includes the network and the host features)',
                            fits=len(fits), solver=SOLVER),
                roofline=roof, cpu_baseline=cpu, roofline_batched=roof_batched, large_fit=large, pairs=pairs, **extra)
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
