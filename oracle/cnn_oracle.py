"""CPU oracle of the coarse-correction CNN (numpy, float64).

TEST INFRASTRUCTURE ONLY: the checker for the GPU forward/backward kernels.
There is no reference counterpart (the learned preconditioner is new); the
oracle itself is pinned by central finite differences of the loss
(tests/test_cnn_oracle.py).

Network: 3x3x3 convolutions with "same" zero padding and per-layer dilation,
no biases, ReLU after every layer but the last.  Activations are channel-last
[z][y][x][c]; weights are [27][Cin][Cout] with tap t = (dx+1) + 3(dy+1) +
9(dz+1).  ``bf16=True`` rounds weights, the network input and every hidden
activation to bfloat16 (round to nearest even) as the GPU path does, with
float64 accumulation here.
"""
import numpy as np

TAPS = [(dx, dy, dz) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


def bf16(a):
    a = np.asarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def shift(x, dx, dy, dz, dil):
    """s[z,y,x] = x[z + dz*dil, y + dy*dil, x + dx*dil] with zero outside"""
    Z, Y, X = x.shape[:3]
    out = np.zeros_like(x)
    oz, oy, ox = dz * dil, dy * dil, dx * dil
    zs, ys, xs = slice(max(0, -oz), min(Z, Z - oz)), slice(max(0, -oy), min(Y, Y - oy)), slice(max(0, -ox), min(X, X - ox))
    zt, yt, xt = slice(max(0, oz), min(Z, Z + oz)), slice(max(0, oy), min(Y, Y + oy)), slice(max(0, ox), min(X, X + ox))
    out[zs, ys, xs] = x[zt, yt, xt]
    return out


def conv(x, W, dil):
    out = 0.0
    for t, (dx, dy, dz) in enumerate(TAPS):
        out = out + shift(x, dx, dy, dz, dil) @ W[t]
    return out


def conv_bwd_data(dz_, W, dil):
    dx_ = 0.0
    for t, (dx, dy, dz) in enumerate(TAPS):
        dx_ = dx_ + shift(dz_ @ W[t].T, -dx, -dy, -dz, dil)
    return dx_


def conv_bwd_weights(x, dz_, dil):
    Ci, Co = x.shape[3], dz_.shape[3]
    dW = np.zeros((27, Ci, Co))
    for t, (dx, dy, dz) in enumerate(TAPS):
        dW[t] = shift(x, dx, dy, dz, dil).reshape(-1, Ci).T @ dz_.reshape(-1, Co)
    return dW


def forward(x, Ws, dils, bf=False):
    """returns the list of activations [a0=input, a1, ..., aL=output]"""
    acts = [bf16(x) if bf else np.asarray(x, dtype=np.float64)]
    L = len(Ws)
    for l in range(L):
        W = bf16(Ws[l]) if bf else Ws[l]
        z = conv(acts[-1], W, dils[l])
        if l < L - 1:
            z = np.maximum(z, 0.0)
            if bf:
                z = bf16(z)
        acts.append(z)
    return acts


def loss_and_grads(x, target, Ws, dils, bf=False):
    """MSE loss mean((y - t)^2) over the grid and dL/dW for every layer"""
    acts = forward(x, Ws, dils, bf)
    y = acts[-1]
    n = y.size
    loss = float(np.sum((y - target) ** 2) / n)
    g = 2.0 * (y - target) / n
    grads = [None] * len(Ws)
    for l in range(len(Ws) - 1, -1, -1):
        if l < len(Ws) - 1:
            g = g * (acts[l + 1] > 0)
        grads[l] = conv_bwd_weights(acts[l], g, dils[l])
        if l > 0:
            W = bf16(Ws[l]) if bf else Ws[l]
            g = conv_bwd_data(g, W, dils[l])
    return loss, grads, y
