"""CPU reference of the closure network (TEST INFRASTRUCTURE ONLY: the
checker of the HIP CNN kernels, never shipped or measured).

Volume v[y, z, x, c]: periodic in z and x, zero beyond the y range.  Layer
l: z = sum_t shift_t(a_{l-1}) @ W[t] + b with shift_t(v)[y, z, x] =
v[y + dy, z + dz, x + dx], t = 9 (dy+1) + 3 (dz+1) + (dx+1); a = gelu(z)
(exact, Phi by erf) for the first three layers, the fourth linear.  Loss:
mean over points and the 3 output channels of (y - target)^2.  The backward
pass is written out by hand; its pin is a central finite difference in
float64.  round_bf16 emulates the device's BF16 storage of inputs, weights,
activations and gradients (round to nearest even on the float32 bits).
"""
import math

import numpy as np

try:
    from scipy.special import erf as _erf
except ImportError:  # pragma: no cover
    _erf = np.vectorize(math.erf)


def round_bf16(a):
    a = np.asarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def taps():
    return [(t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1) for t in range(27)]


def shift(v, dy, dz, dx):
    """s[y, z, x] = v[y + dy, z + dz, x + dx]; periodic z, x; zero in y"""
    s = np.roll(v, shift=(-dz, -dx), axis=(1, 2))
    out = np.zeros_like(s)
    ny = v.shape[0]
    if dy >= 0:
        out[:ny - dy] = s[dy:]
    else:
        out[-dy:] = s[:ny + dy]
    return out


def conv(v, W, b):
    out = np.zeros(v.shape[:3] + (W.shape[2],)) + b
    for t, (dy, dz, dx) in enumerate(taps()):
        out += shift(v, dy, dz, dx) @ W[t]
    return out


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / math.sqrt(2.0)))


def gelu_d(z):
    return 0.5 * (1.0 + _erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def forward(layers, x, bf16=False):
    """layers [(W, b)]; returns output y and the cache [(input, z)]"""
    rnd = round_bf16 if bf16 else (lambda a: a)
    cache = []
    a = rnd(x)
    for l, (W, b) in enumerate(layers):
        z = conv(a, rnd(W), b)
        cache.append((a, z))
        if l < len(layers) - 1:
            # the device applies GELU to the FP32 sum and stores z and a in BF16
            a = rnd(gelu(z))
            cache[-1] = (cache[-1][0], rnd(z))
        else:
            a = z
    return a, cache


def loss(y, target):
    return float(np.mean((y - target) ** 2))


def backward(layers, cache, y, target, bf16=False):
    """gradients [(dW, db)] of the mean squared error"""
    rnd = round_bf16 if bf16 else (lambda a: a)
    n = y.size
    d = rnd(2.0 * (y - target) / n)
    grads = [None] * len(layers)
    for l in range(len(layers) - 1, -1, -1):
        W, b = layers[l]
        a_in, _ = cache[l]
        dW = np.zeros_like(W, dtype=np.float64)
        flat_d = d.reshape(-1, d.shape[-1])
        for t, (dy, dz, dx) in enumerate(taps()):
            dW[t] = shift(a_in, dy, dz, dx).reshape(-1, a_in.shape[-1]).T @ flat_d
        grads[l] = (dW, flat_d.sum(axis=0))
        if l == 0:
            break
        da = np.zeros(a_in.shape)
        Wr = rnd(W)
        for t, (dy, dz, dx) in enumerate(taps()):
            da += shift(d @ Wr[t].T, -dy, -dz, -dx)
        zprev = cache[l - 1][1]
        d = rnd(da * gelu_d(zprev))
    return grads
