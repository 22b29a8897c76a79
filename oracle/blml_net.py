"""CPU FP64 restatement of the warm-start network's kernels -- TEST INFRASTRUCTURE ONLY (the checker of the HIP
convolution, gradient, GELU, loss and Adam kernels; no counterpart in the reference, which has no network).

  conv3x3(X, W, b)          Z[n, co, y, x] = sum_{ci, dy, dx} W[co, ci, dy, dx] X[n, ci, y + dy - 1, x + dx - 1] + b[co]
                            (zero padding)
  conv3x3_grad_w(X, dZ)     dW = sum over n, y, x of dZ[n, co, y, x] X[n, ci, y + dy - 1, x + dx - 1]; db = sum dZ
  conv3x3_grad_x(dZ, W)     the data gradient (transposed convolution)
  net_loss_and_grads(...)   4 layers, GELU between, masked squared loss sum M (Z - Y)^2 / sum M
  bf16(x)                   rounding to bfloat16 (nearest even), as the device stores operands
Pinned by brute-force loops and central finite differences on small inputs (tests/test_oracle_net.py)."""
import math

import numpy as np
from scipy.special import erf


def bf16(x):
    """Round float values to the nearest bfloat16 (ties to even); returns float32."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def _pad(X):
    return np.pad(X, ((0, 0), (0, 0), (1, 1), (1, 1)))


def conv3x3(X, W, b=None):
    X = np.asarray(X, float)
    W = np.asarray(W, float)
    N, C, H, Wd = X.shape
    Xp = _pad(X)
    Z = np.zeros((N, W.shape[0], H, Wd))
    for dy in range(3):
        for dx in range(3):
            Z += np.einsum('oc,nchw->nohw', W[:, :, dy, dx], Xp[:, :, dy:dy + H, dx:dx + Wd])
    if b is not None:
        Z += np.asarray(b, float)[None, :, None, None]
    return Z


def conv3x3_grad_w(X, dZ):
    X = np.asarray(X, float)
    dZ = np.asarray(dZ, float)
    N, C, H, Wd = X.shape
    Xp = _pad(X)
    dW = np.zeros((dZ.shape[1], C, 3, 3))
    for dy in range(3):
        for dx in range(3):
            dW[:, :, dy, dx] = np.einsum('nohw,nchw->oc', dZ, Xp[:, :, dy:dy + H, dx:dx + Wd])
    return dW, dZ.sum((0, 2, 3))


def conv3x3_grad_x(dZ, W):
    """dX[n, ci, q] = sum_{co, dy, dx} W[co, ci, dy, dx] dZ[n, co, q - (dy - 1, dx - 1)]."""
    Wf = np.asarray(W, float)[:, :, ::-1, ::-1].transpose(1, 0, 2, 3)
    return conv3x3(dZ, Wf)


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / math.sqrt(2.0)))


def gelu_d(x):
    return 0.5 * (1.0 + erf(x / math.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / math.sqrt(2.0 * math.pi)


def net_forward(X, Ws, bs):
    A, Z = [np.asarray(X, float)], [None]
    for l in range(4):
        z = conv3x3(A[l], Ws[l], bs[l])
        Z.append(z)
        A.append(gelu(z) if l < 3 else None)
    return Z, A


def net_loss_and_grads(X, Y, M, Ws, bs):
    """(loss, dWs, dbs) of the masked squared loss sum M (Z4 - Y)^2 / sum M (Y, M: (N, H, W))."""
    Z, A = net_forward(X, Ws, bs)
    msum = float(np.sum(M))
    r = Z[4][:, 0] - Y
    loss = float(np.sum(M * r * r)) / msum
    dZ = (2.0 * M * r / msum)[:, None]
    dWs, dbs = [None] * 4, [None] * 4
    for l in range(3, -1, -1):
        dWs[l], dbs[l] = conv3x3_grad_w(A[l], dZ)
        if l > 0:
            dZ = conv3x3_grad_x(dZ, Ws[l]) * gelu_d(Z[l])
    return loss, dWs, dbs


def adam(p, g, m, v, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    p = p - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return p, m, v
