"""The warm-start network's HIP kernels against the CPU restatement (oracle/blml_net.py) on the same BF16 operands:
convolution, data gradient and weight gradient to 1e-5 of their scale (exact BF16 products, FP32 sums); GELU to one
BF16 step; the whole network's parameter gradients to 5e-2 (relative L2: its activations and gradients are stored
in BF16 between layers); Adam to 1e-6; and a short training run lowers the loss."""
import ctypes
import sys

import numpy as np
import pytest

from oracle import blml_net as ON

pytestmark = pytest.mark.gpu


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _dev():
    import torch
    from blml_amd import _lib, warm
    return torch, _lib, warm


def _bf(torch, warm, a):
    return warm.to_bf16(torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda())


def _st(torch):
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


@pytest.mark.parametrize('cin,cout', [(6, 32), (32, 32), (32, 1), (1, 32)])
def test_conv_and_gradients_match_oracle(cin, cout):
    torch, _lib, warm = _dev()
    lib = _lib.load()
    rng = np.random.default_rng(cin * 100 + cout)
    N, H, W = 3, 64, 64
    X = ON.bf16(rng.standard_normal((N, cin, H, W)))
    Wt = ON.bf16(rng.standard_normal((cout, cin, 3, 3)) * 0.3)
    b = rng.standard_normal(cout).astype(np.float32)
    _note(f'net conv {cin}->{cout}')
    Xd, Wd = _bf(torch, warm, X), _bf(torch, warm, Wt)
    bd = torch.from_numpy(b).cuda()
    Z = torch.empty((N, cout, H, W), dtype=torch.float32, device='cuda')
    A = torch.empty((N, cout, H, W), dtype=torch.int16, device='cuda')
    P = lambda t: ctypes.c_void_p(t.data_ptr())
    _lib.check(lib.blml_net_conv(P(Xd), P(Wd), P(bd), P(Z), P(A), N, cin, cout, H, W, 1, _st(torch)))
    Zg = Z.cpu().numpy()
    Zo = ON.conv3x3(X, Wt, b)
    assert np.abs(Zg - Zo).max() < 1e-5 * np.abs(Zo).max()
    Ag = (A.cpu().numpy().astype(np.int32) & 0xFFFF).astype(np.uint32) << 16
    Ag = Ag.view(np.float32)
    Ao = ON.bf16(ON.gelu(Zg))
    assert np.abs(Ag - Ao).max() <= 2 ** -7 * np.abs(Ao).max() + 1e-30
    # weight gradient and data gradient for a BF16 dZ
    dZ = ON.bf16(rng.standard_normal((N, cout, H, W)) * 0.1)
    dZd = _bf(torch, warm, dZ)
    dW = torch.empty((cout, cin, 3, 3), dtype=torch.float32, device='cuda')
    db = torch.empty(cout, dtype=torch.float32, device='cuda')
    _lib.check(lib.blml_net_conv_grad_w(P(Xd), P(dZd), P(dW), P(db), N, cin, cout, H, W, _st(torch)))
    dWo, dbo = ON.conv3x3_grad_w(X, dZ)
    assert np.abs(dW.cpu().numpy() - dWo).max() < 1e-5 * np.abs(dWo).max()
    assert np.abs(db.cpu().numpy() - dbo).max() < 1e-5 * np.abs(dbo).max()
    Wf = torch.empty((cin, cout, 3, 3), dtype=torch.int16, device='cuda')
    _lib.check(lib.blml_net_flip(P(Wd), P(Wf), cin, cout, _st(torch)))
    dX = torch.empty((N, cin, H, W), dtype=torch.float32, device='cuda')
    _lib.check(lib.blml_net_conv(P(dZd), P(Wf), None, P(dX), None, N, cout, cin, H, W, 0, _st(torch)))
    dXo = ON.conv3x3_grad_x(dZ, Wt)
    assert np.abs(dX.cpu().numpy() - dXo).max() < 1e-5 * np.abs(dXo).max()


def test_network_gradients_and_adam_match_oracle():
    torch, _lib, warm = _dev()
    rng = np.random.default_rng(5)
    N = 4
    X = ON.bf16(rng.standard_normal((N, 6, 64, 64)) * 0.5)
    Y = rng.standard_normal((N, 64, 64)).astype(np.float32)
    M = (rng.uniform(size=(N, 64, 64)) > 0.4).astype(np.float32)
    net = warm.WarmNet(seed=3)
    W0 = [ON.bf16(w.cpu().numpy()) for w in net.W]
    Wm = [w.cpu().numpy().copy() for w in net.W]
    b0 = [b.cpu().numpy().copy() for b in net.b]
    _note('net full step')
    loss = net.step(_bf(torch, warm, X), torch.from_numpy(Y).cuda(), torch.from_numpy(M).cuda(), float(M.sum()))
    lo, dWo, dbo = ON.net_loss_and_grads(X, Y, M, W0, b0)
    assert abs(loss - lo) < 1e-2 * lo, (loss, lo)
    for l in range(4):
        g = net.gW[l].cpu().numpy()
        rel = np.linalg.norm(g - dWo[l]) / np.linalg.norm(dWo[l])
        relb = np.linalg.norm(net.gb[l].cpu().numpy() - dbo[l]) / max(np.linalg.norm(dbo[l]), 1e-30)
        _note(f'   layer {l}: dW rel {rel:.2e}, db rel {relb:.2e}')
        assert rel < 5e-2 and relb < 5e-2, (l, rel, relb)
        p1, _, _ = ON.adam(Wm[l], g, 0.0, 0.0, 1)
        assert np.abs(net.W[l].cpu().numpy() - p1).max() < 1e-6 * max(1.0, np.abs(p1).max())


def test_training_lowers_the_loss():
    torch, _lib, warm = _dev()
    rng = np.random.default_rng(6)
    n = 512
    X = rng.standard_normal((n, 6, 64, 64)).astype(np.float32) * 0.5
    Y = (0.7 * X[:, 0] - 0.3 * X[:, 1]).astype(np.float32)    # a local map the network can represent
    M = np.ones((n, 64, 64), np.float32)
    _note('net training run')
    net, hist = warm.train(X, Y, M, epochs=6, batch=128, lr=3e-3)
    _note(f'   losses {["%.4f" % h for h in hist]}')
    assert hist[-1] < 0.3 * hist[0]
