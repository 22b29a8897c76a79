"""ctypes binding of libblml_hip.so (include/blml_hip.h).  Loading fails loudly: there is no CPU fallback."""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'libblml_hip.so')

MAXD, MAXK = 5, 8
F32, F64 = 0, 1
FAC_LINE, FAC_GROUP = 0, 1
STOP_REASONS = ('tol', 'stall', 'll_tol', 'max_sweeps')


class Factor(ctypes.Structure):
    _fields_ = [('kind', ctypes.c_int32), ('naxes', ctypes.c_int32), ('axes', ctypes.c_int32 * MAXD),
                ('circ', ctypes.c_int32 * MAXD), ('tlen', ctypes.c_int32 * MAXD),
                ('freq_off', ctypes.c_int64 * MAXD), ('data_off', ctypes.c_int64), ('spec_off', ctypes.c_int64)]


class Fit(ctypes.Structure):
    _fields_ = [('d', ctypes.c_int32), ('K', ctypes.c_int32), ('complex_g', ctypes.c_int32),
                ('shared', ctypes.c_int32), ('equal_s', ctypes.c_int32), ('B', ctypes.c_int32),
                ('shape', ctypes.c_int32 * MAXD), ('fac_begin', ctypes.c_int32 * MAXK),
                ('fac_count', ctypes.c_int32 * MAXK), ('init_K', ctypes.c_int32), ('has_signs', ctypes.c_int32),
                ('n', ctypes.c_double), ('s', ctypes.c_double * MAXK), ('counts_off', ctypes.c_int64),
                ('draws_off', ctypes.c_int64), ('init_off', ctypes.c_int64), ('signs_off', ctypes.c_int64),
                ('out_off', ctypes.c_int64), ('dens_off', ctypes.c_int64), ('hist_off', ctypes.c_int64),
                ('ws_off', ctypes.c_int64), ('ws_len', ctypes.c_int64), ('scratch_len', ctypes.c_int64)]


class Params(ctypes.Structure):
    _fields_ = [('tol', ctypes.c_double), ('max_sweeps', ctypes.c_int64), ('anderson', ctypes.c_int32),
                ('stall', ctypes.c_int32), ('phase_noise', ctypes.c_double), ('ll_tol', ctypes.c_double),
                ('ll_window', ctypes.c_int32), ('solver', ctypes.c_int32), ('lbfgs_m', ctypes.c_int32),
                ('hist_cap', ctypes.c_int32), ('spread', ctypes.c_int32)]


class Result(ctypes.Structure):
    _fields_ = [('loglik', ctypes.c_double), ('residual', ctypes.c_double), ('n_sweeps', ctypes.c_int32),
                ('stop_reason', ctypes.c_int32), ('converged', ctypes.c_int32), ('n_anderson', ctypes.c_int32),
                ('n_damped', ctypes.c_int32), ('n_gradient', ctypes.c_int32), ('start_sweeps', ctypes.c_int32),
                ('n_hist', ctypes.c_int32), ('n_matvec', ctypes.c_int32), ('reserved', ctypes.c_int32),
                ('seconds', ctypes.c_double * 5)]


class EmpParams(ctypes.Structure):
    _fields_ = [('reg', ctypes.c_int32), ('normalise', ctypes.c_int32), ('anscombe', ctypes.c_int32),
                ('retrieve', ctypes.c_int32), ('lam', ctypes.c_double), ('tol', ctypes.c_double),
                ('outer_tol', ctypes.c_double), ('max_iter', ctypes.c_int32), ('max_outer', ctypes.c_int32),
                ('anderson', ctypes.c_int32), ('reserved', ctypes.c_int32)]


class HioParams(ctypes.Structure):
    _fields_ = [('d', ctypes.c_int32), ('shape', ctypes.c_int32 * MAXD), ('box', ctypes.c_int32 * MAXD),
                ('n_hio', ctypes.c_int32), ('n_er', ctypes.c_int32), ('n_rand', ctypes.c_int32),
                ('n_kick', ctypes.c_int32), ('beta', ctypes.c_double), ('ab2', ctypes.c_double),
                ('nocc', ctypes.c_int64)]


EMP_AUX_LEN, EMP_OUT_LEN = MAXK + 1, 4
TUNE_MAX = 4

RESULT_DTYPE = np.dtype([(name, (np.float64, 5) if name == 'seconds' else
                          (np.float64 if t is ctypes.c_double else np.int32)) for name, t in Result._fields_])
assert RESULT_DTYPE.itemsize == ctypes.sizeof(Result)

EXPORTS = ('blml_plan_create', 'blml_plan_destroy', 'blml_matvec', 'blml_fit', 'blml_certificate',
           'blml_plan_info', 'blml_pair_rows', 'blml_pair_rows_grad', 'blml_pair_rows_multi', 'blml_pair_cols', 'blml_net_conv', 'blml_net_conv_grad_w',
           'blml_net_flip', 'blml_net_gelu_grad', 'blml_net_mse', 'blml_net_adam', 'blml_net_to_bf16',
           'blml_quick_sweep', 'blml_quick_exact', 'blml_quick_stats', 'blml_plan_matvec_kernel',
           'blml_pair_rows_noalloc', 'blml_point_rows', 'blml_grad_dots', 'blml_quick_psd_sweep',
           'blml_quick_psd_exact', 'blml_quick_psd_stats', 'blml_quick_psd_grad', 'blml_quick_psd_wscale',
           'blml_vec_gram', 'blml_vec_lincomb', 'blml_lbfgs_twoloop', 'blml_ift_elem', 'blml_empirical',
           'blml_plan_autotune', 'blml_complex_starts', 'blml_band_eval', 'blml_last_error')

_lib = None


def load(path=None):
    """The library handle; raises (ImportError / OSError) when it is not built -- never falls back to the CPU."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise ImportError(f"libblml_hip.so not built ({p}); run __graft_entry__.build() "
                          "or python band-limited-ml_amd/build.py")
    lib = ctypes.CDLL(p)
    vp, i32, i64, dp = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_void_p
    lib.blml_plan_create.argtypes = [ctypes.POINTER(Fit), i32, ctypes.POINTER(Factor), i32, ctypes.c_void_p, i64,
                                     ctypes.c_void_p, i64, i32, ctypes.POINTER(Params), ctypes.POINTER(vp)]
    lib.blml_plan_destroy.argtypes = [vp]
    lib.blml_matvec.argtypes = [vp, dp, dp, vp]
    lib.blml_fit.argtypes = [vp, dp, dp, dp, dp, dp, dp, dp, dp, dp, dp, vp]
    lib.blml_certificate.argtypes = [vp, dp, dp, i32, ctypes.c_double, dp, dp, vp]
    lib.blml_plan_autotune.argtypes = [vp, i32, dp, ctypes.POINTER(i32), vp]
    lib.blml_empirical.argtypes = [vp, ctypes.POINTER(EmpParams), dp, dp, dp, dp, dp, dp, dp, dp, dp, dp, vp]
    lib.blml_plan_info.argtypes = [vp, ctypes.POINTER(i64), ctypes.POINTER(i64), ctypes.POINTER(i32),
                                   ctypes.POINTER(i32)]
    lib.blml_plan_matvec_kernel.argtypes = [vp, ctypes.POINTER(i32)]
    lib.blml_pair_rows.argtypes = [vp, vp, vp, vp, i64, i64, vp, vp]
    lib.blml_pair_rows_noalloc.argtypes = [vp, vp, vp, vp, i64, i64, vp, vp]
    lib.blml_point_rows.argtypes = [vp, i32, i32, vp, vp, vp, vp, i32, i64, i64, vp, vp]
    lib.blml_grad_dots.argtypes = [i32, i32, i32, i64, vp, vp, vp, i32, vp, vp, ctypes.c_double, vp, vp]
    lib.blml_pair_cols.argtypes = [vp, vp, vp, i64, i64, i64, vp, vp, vp]
    lib.blml_pair_rows_grad.argtypes = [vp, vp, vp, vp, vp, i64, i64, vp, vp, vp]
    lib.blml_pair_rows_multi.argtypes = [vp, vp, vp, vp, vp, vp, i32, i64, i64, vp, vp, vp]
    f32 = ctypes.c_float
    lib.blml_net_conv.argtypes = [vp, vp, vp, vp, vp, i32, i32, i32, i32, i32, i32, vp]
    lib.blml_net_conv_grad_w.argtypes = [vp, vp, vp, vp, i32, i32, i32, i32, i32, vp]
    lib.blml_net_flip.argtypes = [vp, vp, i32, i32, vp]
    lib.blml_net_gelu_grad.argtypes = [vp, vp, vp, i64, vp]
    lib.blml_net_mse.argtypes = [vp, vp, vp, vp, vp, i64, f32, vp]
    lib.blml_net_adam.argtypes = [vp, vp, vp, vp, vp, i64, f32, f32, f32, f32, i32, vp]
    lib.blml_net_to_bf16.argtypes = [vp, vp, i64, vp]
    f64 = ctypes.c_double
    lib.blml_quick_sweep.argtypes = [vp, vp, i64, vp, vp, vp, i64, f64, f64, i32, vp, vp, vp]
    lib.blml_quick_exact.argtypes = [vp, vp, i64, vp, vp, vp, vp]
    lib.blml_quick_stats.argtypes = [vp, vp, vp, i64, f64, vp, vp]
    lib.blml_quick_psd_sweep.argtypes = [vp, i32, vp, i64, i32, i32, vp, vp, i64, f64, vp, i32, vp, vp, vp]
    lib.blml_quick_psd_exact.argtypes = [vp, i32, vp, i64, i32, i32, vp, vp, vp]
    lib.blml_quick_psd_stats.argtypes = [vp, vp, vp, i64, i32, i32, f64, f64, vp, vp]
    lib.blml_quick_psd_grad.argtypes = [vp, i32, vp, i64, i32, i32, vp, f64, vp, vp, vp, vp]
    lib.blml_quick_psd_wscale.argtypes = [vp, vp, i64, i32, i32, f64, vp, vp]
    lib.blml_vec_gram.argtypes = [i32, vp, i32, vp, i64, vp, vp]
    lib.blml_vec_lincomb.argtypes = [i32, vp, vp, i32, i64, vp, vp]
    lib.blml_lbfgs_twoloop.argtypes = [i32, vp, vp, vp, f64, i64, vp, vp, vp]
    lib.blml_ift_elem.argtypes = [i32, i32, i32, i64, i32, f64, f64, f64, f64, vp, vp, vp, vp, vp, vp, vp, vp]
    lib.blml_complex_starts.argtypes = [ctypes.POINTER(HioParams), vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]
    lib.blml_band_eval.argtypes = [i32, vp, i64, i64, vp, vp, vp, vp, vp]
    lib.blml_last_error.restype = ctypes.c_char_p
    for name in EXPORTS[:-1]:
        getattr(lib, name).restype = ctypes.c_int
    if path is None:
        _lib = lib
    return lib


def check(rc):
    """Maps a status code to the reference's exception types (ValueError / RuntimeError / MemoryError)."""
    if rc == 0:
        return
    msg = load().blml_last_error().decode(errors='replace')
    if rc == 1:
        raise ValueError(msg)
    if rc == 3:
        raise MemoryError(msg)
    raise RuntimeError(msg)
