"""Pins the numpy CNN oracle: its analytic gradients agree with central finite
differences of its own loss, its convolution agrees with a direct loop, and
the bf16 rounding is round-to-nearest-even."""
import numpy as np

import cnn_oracle as cnn


def small_net(rng, ch=(2, 4, 4, 1), dils=(1, 2, 1)):
    Ws = [rng.normal(0, 0.3, size=(27, ch[i], ch[i + 1])) for i in range(len(ch) - 1)]
    return Ws, list(dils)


def test_conv_matches_direct_loop():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(3, 4, 5, 2))
    W = rng.normal(size=(27, 2, 3))
    out = cnn.conv(x, W, 2)
    Z, Y, X = x.shape[:3]
    ref = np.zeros((Z, Y, X, 3))
    for z in range(Z):
        for y in range(Y):
            for xx in range(X):
                for t, (dx, dy, dz) in enumerate(cnn.TAPS):
                    zz, yy, x2 = z + 2 * dz, y + 2 * dy, xx + 2 * dx
                    if 0 <= zz < Z and 0 <= yy < Y and 0 <= x2 < X:
                        ref[z, y, xx] += x[zz, yy, x2] @ W[t]
    assert np.abs(out - ref).max() < 1e-12


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(1)
    Ws, dils = small_net(rng)
    x = rng.normal(size=(4, 3, 5, 2))
    t = rng.normal(size=(4, 3, 5, 1))
    loss, grads, _ = cnn.loss_and_grads(x, t, Ws, dils)
    h = 1e-6
    for l in range(len(Ws)):
        for _ in range(12):
            idx = tuple(rng.integers(0, s) for s in Ws[l].shape)
            Wp = [w.copy() for w in Ws]
            Wm = [w.copy() for w in Ws]
            Wp[l][idx] += h
            Wm[l][idx] -= h
            fd = (cnn.loss_and_grads(x, t, Wp, dils)[0] - cnn.loss_and_grads(x, t, Wm, dils)[0]) / (2 * h)
            assert abs(fd - grads[l][idx]) <= 1e-6 * max(1.0, abs(fd)), (l, idx, fd, grads[l][idx])


def test_bf16_round_to_nearest_even():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.1415926], dtype=np.float32)
    r = cnn.bf16(vals)
    assert r[0] == 1.0
    assert r[1] == 1.0                   # tie -> even
    assert r[2] == 1.0 + 4 * 2 ** -8     # tie -> even (up)
    assert r[3] == -2.5
    assert abs(r[4] - 3.140625) < 1e-12
