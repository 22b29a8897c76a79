"""blml_pair_rows_noalloc (the pair rows without any device allocation: partial sums in the library's fixed array on
each GPU) against blml_pair_rows through the C-ABI: the same kernels over the same slices, so the same sums to the
bit, on two calls in a row (the second smaller, reusing the array the first left filled); and a call beyond the
array's capacity is refused with ValueError before anything runs."""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rows(lib, fn, t, dA, dB, w, R, M):
    import torch
    from blml_amd import _lib
    s = torch.full((max(R, 1),), float('nan'), dtype=torch.float64, device='cuda')
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.check(getattr(lib, fn)(ctypes.byref(t), ctypes.c_void_p(dA.data_ptr()), ctypes.c_void_p(dB.data_ptr()),
                                ctypes.c_void_p(w.data_ptr()), R, M, ctypes.c_void_p(s.data_ptr()), stream))
    torch.cuda.synchronize()
    return s[:R].cpu().numpy()


def test_noalloc_rows_equal_allocating_rows():
    import torch
    from blml_amd import _lib, local
    lib = _lib.load()
    rng = np.random.default_rng(31)
    shape = (40, 30, 16)
    Z = np.stack([rng.integers(0, m, 6000) for m in shape], 1)
    Z = np.unique(Z, axis=0)
    tab = local.lattice_tables(['sinc', 'slits:0.5', 'dirichlet'], [0.3, 0.25, 0.4], [None, None, 16.0],
                               [1.0, 1.0, 1.0], Z)
    for R, M in ((len(Z), len(Z)), (700, 333)):          # the second call smaller, with fewer rows and sources
        A, B = Z[:R], Z[-M:]
        t, dA, dB, keep = local._prepare(A, B, tab, False, False)
        w = torch.from_numpy(rng.standard_normal(M)).to('cuda')
        _note(f'noalloc rows: R {R} M {M}')
        want = _rows(lib, 'blml_pair_rows', t, dA, dB, w, R, M)
        got = _rows(lib, 'blml_pair_rows_noalloc', t, dA, dB, w, R, M)
        assert np.all(np.isfinite(want)) and np.array_equal(got, want), float(np.abs(got - want).max())
    # M = 0: zeros (a fill kernel); R beyond the fixed partials: ValueError, nothing launched
    t, dA, dB, keep = local._prepare(Z[:5], Z[:5], tab, False, False)
    w = torch.zeros(5, dtype=torch.float64, device='cuda')
    assert np.array_equal(_rows(lib, 'blml_pair_rows_noalloc', t, dA, dB, w, 5, 0), np.zeros(5))
    Rbig = (1 << 21) + 1
    Zbig = np.zeros((Rbig, 3), np.int64)
    t, dA, dB, keep = local._prepare(Zbig, Z[:5], tab, False, True)
    _note(f'noalloc rows: R {Rbig} refused')
    with pytest.raises(ValueError, match='partial sums'):
        _rows(lib, 'blml_pair_rows_noalloc', t, dA, dB, w, Rbig, 5)
