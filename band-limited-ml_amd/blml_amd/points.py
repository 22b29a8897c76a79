"""The continuous kernel at arbitrary points on the MI355X (blml_point_rows): the units of a mode spec as the C-ABI's
blml_pointunit_t rows, and s_q = sum_j K(x_q - z_j) w_j for query points against lattice bins -- the sums behind
BLMLBinned.sqrt_density / predict / predict_grid (blml_binned.py:452-499)."""
import ctypes

import numpy as np

from . import _lib
from .modes import split_kind, units_of
from .units import dirichlet_Mi, dirichlet_N, pball_harmonics

__all__ = ['PointUnit', 'point_units', 'point_rows']

MAXU = 8
KIND = {'sinc': 0, 'slits': 1, 'dirichlet': 2, 'ball': 3, 'pball': 4}


class PointUnit(ctypes.Structure):
    _fields_ = [('kind', ctypes.c_int32), ('ncol', ctypes.c_int32), ('col', ctypes.c_int32 * _lib.MAXD),
                ('M', ctypes.c_int32), ('Mi', ctypes.c_int32), ('nharm', ctypes.c_int32),
                ('reserved', ctypes.c_int32), ('hoff', ctypes.c_int64), ('fc', ctypes.c_double * _lib.MAXD),
                ('fi', ctypes.c_double), ('period', ctypes.c_double * _lib.MAXD), ('scale', ctypes.c_double)]


def point_units(spec):
    """(units array, harmonics int32 array) of one mode spec (parse_modes: per column 'kinds', 'fc', 'periods'; fc in
    cycles per coordinate unit), in the order of the reference's kernel_units."""
    us = units_of(spec['kinds'])
    if len(us) > MAXU:
        raise ValueError(f"at most {MAXU} kernel units per mode on this engine")
    arr = (PointUnit * len(us))()
    harm = []
    at = 0
    fc = np.asarray(spec['fc'], float)
    per = spec['periods']
    for i, (kind, dims) in enumerate(us):
        name, tag = split_kind(kind)
        u = arr[i]
        u.kind = KIND[name]
        u.ncol = len(dims)
        for a, j in enumerate(dims):
            u.col[a] = j
            u.fc[a] = fc[j]
            u.period[a] = 0.0 if per[j] is None else float(per[j])
        if name == 'slits':
            u.fi = (1.0 - float(tag)) * fc[dims[0]]
        elif name == 'dirichlet':
            P = float(per[dims[0]])
            u.M = 2 * dirichlet_N(fc[dims[0]], P) + 1
            u.Mi = dirichlet_Mi(fc[dims[0]], P, float(tag) if tag else None)
            u.scale = 1.0 / P
        elif name == 'ball':
            if len(dims) > 5:
                raise ValueError("ball: at most 5 columns in a group on this engine")
        elif name == 'pball':
            H, sc = pball_harmonics(fc[dims[0]], [float(per[j]) for j in dims])
            H = np.asarray(H, np.int32).reshape(-1, len(dims))
            u.hoff, u.nharm, u.scale = at, len(H), sc
            harm.append(H.ravel())
            at += H.size
    return arr, (np.concatenate(harm) if harm else np.zeros(1, np.int32)).astype(np.int32)


def point_rows(units, harm, X, Z, W):
    """S (Nq, nw) = sum_j K(X_q - Z_j) W_j (host arrays in and out; W (M,) real or (M, 2) as real / imaginary)."""
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("the point kernel needs a ROCm GPU (torch.cuda.is_available() is False)")
    X = np.ascontiguousarray(np.atleast_2d(np.asarray(X, np.float64)))
    Z = np.ascontiguousarray(np.atleast_2d(np.asarray(Z, np.float64)))
    W = np.asarray(W, np.float64)
    nw = 1 if W.ndim == 1 else W.shape[1]
    W = np.ascontiguousarray(W.reshape(len(Z), nw))
    d = X.shape[1]
    if Z.shape[1] != d:
        raise ValueError("point rows: X and Z need the same number of columns")
    dX, dZ, dW = (torch.from_numpy(a).to('cuda') for a in (X, Z, W))
    dH = torch.from_numpy(np.ascontiguousarray(harm, np.int32)).to('cuda')
    S = torch.empty((max(len(X), 1), nw), dtype=torch.float64, device='cuda')
    lib = _lib.load()
    _lib.check(lib.blml_point_rows(units, len(units), d, ctypes.c_void_p(dH.data_ptr()), ctypes.c_void_p(dX.data_ptr()),
                                   ctypes.c_void_p(dZ.data_ptr()), ctypes.c_void_p(dW.data_ptr()), nw, len(X), len(Z),
                                   ctypes.c_void_p(S.data_ptr()),
                                   ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
    return S[:len(X)].cpu().numpy()
