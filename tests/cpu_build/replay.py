"""The CPU replay of the engine's fit path (replay.cpp) from Python: build() compiles it with clang++ and
AddressSanitizer (the engine's device headers copied with their LDS declarations turned into per-workgroup objects of
the CPU stand-in hip/hip_runtime.h, which comes first on the include path); run(batch) serializes a
blml_amd.binned.FitBatch (its blml_fit_t / blml_factor_t rows, tables, knobs and blml_fit's inputs), runs the
replay in a child process and returns its results.  Test infrastructure: nothing here runs on a GPU."""
import ctypes
import os
import re
import shutil
import struct
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, 'band-limited-ml_amd', 'csrc')
BUILD = os.path.join(HERE, '_build')
EXE = os.path.join(BUILD, 'replay')
CXX = os.environ.get('BLML_REPLAY_CXX', '/opt/rocm/llvm/bin/clang++')
HEADERS = ('common.hpp', 'separable.hpp', 'transform.hpp', 'solver.hpp', 'fit.hpp', 'plan_layout.hpp',
           'empirical.hpp', 'starts.hpp')
# every LDS declaration of the device headers -> a per-workgroup object of the CPU stand-in (one heap block each)
LDS_REWRITES = {
    'common.hpp': [('extern __shared__ __align__(16) unsigned char g_dyn_lds[];',
                    '#define g_dyn_lds (cpuhip::dyn_lds_base())')],
    'solver.hpp': [('__shared__ SolverLds g_sl;', '#define g_sl (*cpuhip::lds_obj<SolverLds>(0))')],
    'transform.hpp': [('__shared__ int lofs_in[BLML_NT], lofs_out[BLML_NT], lofs_in2[BLML_NT], lofs_out2[BLML_NT];',
                       'int *lofs_in = cpuhip::lds_obj<int>(1, BLML_NT), *lofs_out = cpuhip::lds_obj<int>(2, BLML_NT), '
                       '*lofs_in2 = cpuhip::lds_obj<int>(3, BLML_NT), *lofs_out2 = cpuhip::lds_obj<int>(4, BLML_NT);')],
    'fit.hpp': [('__shared__ double red[BLML_NWAVE * BLML_NRED];',
                 'double *red = cpuhip::lds_obj<double>(5, BLML_NWAVE * BLML_NRED);')],
}


def build(force=False, asan=True):
    """Compiles replay.cpp against copies of the engine's device headers; returns the executable's path."""
    srcs = [os.path.join(CSRC, h) for h in HEADERS] + [os.path.join(HERE, 'replay.cpp'),
                                                       os.path.join(HERE, 'hip', 'hip_runtime.h')]
    if not force and os.path.exists(EXE) and all(os.path.getmtime(s) <= os.path.getmtime(EXE) for s in srcs):
        return EXE
    src = os.path.join(BUILD, 'src')
    os.makedirs(src, exist_ok=True)
    for h in HEADERS:
        text = open(os.path.join(CSRC, h)).read()
        for old, new in LDS_REWRITES.get(h, ()):
            if old not in text:
                raise RuntimeError(f"{h}: the LDS declaration {old!r} the CPU build rewrites is gone")
            text = text.replace(old, new)
        code = re.sub(r'//.*', '', text)
        if re.search(r'\b__shared__\b', code):
            raise RuntimeError(f"{h}: an LDS declaration the CPU build does not know")
        with open(os.path.join(src, h), 'w') as f:
            f.write(text)
    flags = ['-std=c++17', '-O1', '-g', '-fno-omit-frame-pointer', '-Wno-unused-result', '-Wno-psabi']
    if asan:
        flags += ['-fsanitize=address', '-fsanitize-address-use-after-scope']
    cmd = [CXX] + flags + ['-I', HERE, '-I', src, '-I', os.path.join(ROOT, 'include'),
                           os.path.join(HERE, 'replay.cpp'), '-o', EXE + '.tmp']
    subprocess.run(cmd, check=True)
    os.replace(EXE + '.tmp', EXE)
    return EXE


def _bytes(x):
    return bytes(memoryview(x)) if x is not None else b''


def write_plan(batch, path, cert_iters=0, cert_tol=1e-6, ranks=4, emp=None):
    """The plan of a FitBatch (host tables only: no device needed) in replay.cpp's input format; ranks: the
    workgroups of a spread plan's cooperative grid; emp: (blml_emp_params_t, aux) to replay blml_empirical's fit
    in place of blml_fit's."""
    from blml_amd import _lib
    prm = batch.params_struct()
    fits = batch._Fits
    facs = batch._Facs
    nfacs = batch._nfacs
    fdata = np.ascontiguousarray(batch._fdata, np.float64)
    freqs = np.ascontiguousarray(batch._freqs, np.int32)
    counts = np.ascontiguousarray(batch.counts, np.float64)
    draws = np.ascontiguousarray(batch.draws, np.float64)
    init = None if batch.inits is None else np.ascontiguousarray(batch.inits, np.float64)
    signs = None if batch.signs is None else np.ascontiguousarray(batch.signs, np.float64)
    with open(path, 'wb') as f:
        f.write(b'BLMLRPL1')
        flags = (1 if cert_iters else 0) | (2 if emp is not None else 0)
        f.write(struct.pack('<iiiiiid', batch.code, len(batch.fits), nfacs, flags, int(cert_iters),
                            int(ranks), float(cert_tol)))
        f.write(struct.pack('<qqqqqq', len(fdata), len(freqs), len(counts), len(draws),
                            0 if init is None else len(init), 0 if signs is None else len(signs)))
        f.write(bytes(prm))
        f.write(bytes(fits))
        f.write(bytes(facs)[:nfacs * ctypes.sizeof(_lib.Factor)])
        for a in (fdata, freqs, counts, draws, init, signs):
            f.write(_bytes(a))
        if emp is not None:
            f.write(bytes(emp[0]))
            f.write(_bytes(np.ascontiguousarray(emp[1], np.float64).ravel()))


def run(batch, cert_iters=0, cert_tol=1e-6, timeout=1800, exe=None, ranks=4, emp=None, fft_sched=0, toep_max=0):
    """Replays the batch's fits on the CPU; returns dict(results (structured array), e, density, lam, iters, log).
    Raises RuntimeError with the replay's output when it fails (an AddressSanitizer report included)."""
    from blml_amd import _lib
    exe = exe or build()
    tmp = tempfile.mkdtemp(prefix='blml_replay_')
    try:
        pin, pout = os.path.join(tmp, 'plan.bin'), os.path.join(tmp, 'out.bin')
        write_plan(batch, pin, cert_iters, cert_tol, ranks, emp)
        env = dict(os.environ, ASAN_OPTIONS='detect_leaks=0:abort_on_error=0:detect_stack_use_after_return=0',
                   BLML_REPLAY_FFT_SCHED=str(int(fft_sched)), BLML_REPLAY_TOEP=str(int(toep_max)))
        p = subprocess.run([exe, pin, pout], capture_output=True, text=True, timeout=timeout, env=env)
        log = p.stdout + p.stderr
        if p.returncode != 0:
            raise RuntimeError(f"replay exited with {p.returncode}:\n{log[-6000:]}")
        raw = open(pout, 'rb').read()
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    nf = len(batch.fits)
    es = 4 if batch.code == _lib.F32 else 8
    out_elems = max(1, max(int(F.out_off) + fs.K * fs.B * fs.nc for F, fs in zip(batch._Fits, batch.fits)))
    dens_elems = max(1, max(int(F.dens_off) + fs.B for F, fs in zip(batch._Fits, batch.fits)))
    o = 0
    res = np.frombuffer(raw, _lib.RESULT_DTYPE, nf, o); o += nf * _lib.RESULT_DTYPE.itemsize
    e = np.frombuffer(raw, np.float32 if es == 4 else np.float64, out_elems, o); o += out_elems * es
    dens = np.frombuffer(raw, np.float64, dens_elems, o); o += dens_elems * 8
    lam = np.frombuffer(raw, np.float64, nf, o); o += nf * 8
    its = np.frombuffer(raw, np.int32, nf, o); o += nf * 4
    hc = batch.hist_cap
    mh = np.frombuffer(raw, np.float64, nf * hc, o).reshape(nf, hc); o += nf * hc * 8
    rh = np.frombuffer(raw, np.float64, nf * hc, o).reshape(nf, hc); o += nf * hc * 8
    nh = [min(int(r['n_hist']), hc) for r in res]
    out = dict(results=res, e=e, density=dens, lam=lam, iters=its, log=log,
               merit_history=[mh[i, :nh[i]] for i in range(nf)], rho_history=[rh[i, :nh[i]] for i in range(nf)])
    if emp is not None:
        out['emp'] = np.frombuffer(raw, np.float64, nf * _lib.EMP_OUT_LEN, o).reshape(nf, _lib.EMP_OUT_LEN)
    return out


def pipe_input(n):
    """replay.cpp's pipe_value: the input E of the pipelined mat-vec replay."""
    x = np.arange(n, dtype=np.float64) * 0.6180339887498949
    return (2.0 * (x - np.floor(x)) - 1.0).astype(np.float32)


def run_pipe(batch, ranks=4, timeout=1800, exe=None):
    """Replays matvec_pipe_kernel's work for an FP32 spread plan of one fit on pipe_input(n): returns (E, U with
    the dense line operators, U with their centrosymmetric split)."""
    exe = exe or build()
    fs = batch.fits[0]
    n = fs.K * fs.B * fs.nc
    tmp = tempfile.mkdtemp(prefix='blml_replay_')
    try:
        pin, pout = os.path.join(tmp, 'plan.bin'), os.path.join(tmp, 'out.bin')
        write_plan(batch, pin, 0, 1e-6, ranks, None)
        env = dict(os.environ, ASAN_OPTIONS='detect_leaks=0:abort_on_error=0:detect_stack_use_after_return=0',
                   BLML_REPLAY_PIPE='1')
        p = subprocess.run([exe, pin, pout], capture_output=True, text=True, timeout=timeout, env=env)
        log = p.stdout + p.stderr
        if p.returncode != 0:
            raise RuntimeError(f"replay exited with {p.returncode}:\n{log[-6000:]}")
        raw = np.fromfile(pout, np.float32)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    return pipe_input(n), raw[:n], raw[n:2 * n]


def run_matvec(batch, toep_max=0, timeout=1800, exe=None):
    """Replays matvec_kernel's work (every fit on one workgroup) on pipe_input(n) over the plan's outputs: returns
    (E, U).  toep_max: the plan's bound for the directly summed padded lines (FP32 plans)."""
    exe = exe or build()
    n = max(int(F.out_off) + fs.K * fs.B * fs.nc for F, fs in zip(batch._Fits, batch.fits))
    es = np.float32 if batch.code == 0 else np.float64
    tmp = tempfile.mkdtemp(prefix='blml_replay_')
    try:
        pin, pout = os.path.join(tmp, 'plan.bin'), os.path.join(tmp, 'out.bin')
        write_plan(batch, pin, 0, 1e-6, 1, None)
        env = dict(os.environ, ASAN_OPTIONS='detect_leaks=0:abort_on_error=0:detect_stack_use_after_return=0',
                   BLML_REPLAY_MATVEC='1', BLML_REPLAY_TOEP=str(int(toep_max)))
        p = subprocess.run([exe, pin, pout], capture_output=True, text=True, timeout=timeout, env=env)
        log = p.stdout + p.stderr
        if p.returncode != 0:
            raise RuntimeError(f"replay exited with {p.returncode}:\n{log[-6000:]}")
        raw = np.fromfile(pout, es)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    return pipe_input(n).astype(es), raw[:n]


def run_hio(shape, box, n_hio, n_er, beta, ab2, band, amp, occ, wocc, phase, kick, timeout=1800, exe=None):
    """Replays the start search's runs (hio_run) on the CPU: returns (P (runs, ncell) complex, g (runs, nocc)
    complex, score (runs,)).  phase (n_rand, ncell), kick (n_kick, ncell)."""
    exe = exe or build()
    d = len(shape)
    M = int(np.prod(shape))
    phase = np.asarray(phase, np.float64).reshape(-1, M)
    kick = np.asarray(kick, np.float64).reshape(-1, M)
    runs = len(phase) + len(kick)
    hdr = np.zeros(15, np.int32)
    hdr[0] = d
    hdr[1:1 + d] = shape
    hdr[6:6 + d] = box
    hdr[11:15] = n_hio, n_er, len(phase), len(kick)
    tmp = tempfile.mkdtemp(prefix='blml_replay_')
    try:
        pin, pout = os.path.join(tmp, 'hio.bin'), os.path.join(tmp, 'out.bin')
        with open(pin, 'wb') as f:
            f.write(hdr.tobytes())
            f.write(np.array([beta, ab2], np.float64).tobytes())
            f.write(np.array([len(occ)], np.int64).tobytes())
            f.write(np.ascontiguousarray(band, np.uint8).tobytes())
            for a, dt in ((amp, np.float64), (occ, np.int64), (wocc, np.float64), (phase, np.float64),
                          (kick, np.float64)):
                f.write(np.ascontiguousarray(a, dt).tobytes())
        env = dict(os.environ, ASAN_OPTIONS='detect_leaks=0:abort_on_error=0:detect_stack_use_after_return=0')
        p = subprocess.run([exe, '--hio', pin, pout], capture_output=True, text=True, timeout=timeout, env=env)
        if p.returncode != 0:
            raise RuntimeError(f"replay exited with {p.returncode}:\n{(p.stdout + p.stderr)[-6000:]}")
        raw = np.fromfile(pout, np.float64)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    nocc = len(occ)
    P = raw[:runs * M * 2].reshape(runs, M, 2)
    g = raw[runs * M * 2:runs * (M + nocc) * 2].reshape(runs, nocc, 2)
    score = raw[runs * (M + nocc) * 2:]
    return P[..., 0] + 1j * P[..., 1], g[..., 0] + 1j * g[..., 1], score
