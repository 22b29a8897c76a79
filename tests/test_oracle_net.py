"""The warm-start network's CPU restatement (oracle/blml_net.py), pinned by brute-force loops and central finite
differences (float64)."""
import numpy as np

from oracle import blml_net as ON


def test_conv_matches_loops():
    rng = np.random.default_rng(0)
    X = rng.standard_normal((2, 3, 5, 6))
    W = rng.standard_normal((4, 3, 3, 3))
    b = rng.standard_normal(4)
    Z = ON.conv3x3(X, W, b)
    want = np.zeros_like(Z)
    for n in range(2):
        for co in range(4):
            for y in range(5):
                for x in range(6):
                    acc = b[co]
                    for ci in range(3):
                        for dy in range(3):
                            for dx in range(3):
                                yy, xx = y + dy - 1, x + dx - 1
                                if 0 <= yy < 5 and 0 <= xx < 6:
                                    acc += W[co, ci, dy, dx] * X[n, ci, yy, xx]
                    want[n, co, y, x] = acc
    assert np.abs(Z - want).max() < 1e-12


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(1)
    chans = (2, 3, 3, 3, 1)
    X = rng.standard_normal((2, 2, 6, 5))
    Y = rng.standard_normal((2, 6, 5))
    M = (rng.uniform(size=(2, 6, 5)) > 0.3).astype(float)
    Ws = [rng.standard_normal((chans[l + 1], chans[l], 3, 3)) * 0.5 for l in range(4)]
    bs = [rng.standard_normal(chans[l + 1]) * 0.1 for l in range(4)]
    loss, dWs, dbs = ON.net_loss_and_grads(X, Y, M, Ws, bs)
    h = 1e-6
    for l in range(4):
        for idx in [(0, 0, 1, 1), (chans[l + 1] - 1, chans[l] - 1, 2, 0)]:
            Wp = [w.copy() for w in Ws]
            Wm = [w.copy() for w in Ws]
            Wp[l][idx] += h
            Wm[l][idx] -= h
            fd = (ON.net_loss_and_grads(X, Y, M, Wp, bs)[0] - ON.net_loss_and_grads(X, Y, M, Wm, bs)[0]) / (2 * h)
            assert abs(fd - dWs[l][idx]) < 1e-6 * max(1.0, abs(fd)), (l, idx, fd, dWs[l][idx])
        bp = [b.copy() for b in bs]
        bm = [b.copy() for b in bs]
        bp[l][0] += h
        bm[l][0] -= h
        fd = (ON.net_loss_and_grads(X, Y, M, Ws, bp)[0] - ON.net_loss_and_grads(X, Y, M, Ws, bm)[0]) / (2 * h)
        assert abs(fd - dbs[l][0]) < 1e-6 * max(1.0, abs(fd))


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.14159265], np.float32)
    r = ON.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0                 # halfway: ties to even
    assert r[2] == np.float32(1.0 + 2 ** -7)           # above halfway: up
    assert r[3] == -2.5 and abs(r[4] - 3.140625) < 1e-7
