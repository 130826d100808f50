"""Pins of the CPU reference of the closure network (oracle/cnn_ref.py):
the hand-written backward pass against central finite differences in
float64 on a tiny volume, the shift against a brute-force index loop, and
the BF16 rounding against the bit pattern rule."""
import numpy as np

from oracle import cnn_ref


def _tiny(seed=0, ny=3, nz=4, nx=5):
    rng = np.random.default_rng(seed)
    cin, cout = (4, 5, 5, 5), (5, 5, 5, 3)
    layers = [(rng.standard_normal((27, ci, co)) * 0.3, rng.standard_normal(co) * 0.1) for ci, co in zip(cin, cout)]
    x = rng.standard_normal((ny, nz, nx, 4))
    t = rng.standard_normal((ny, nz, nx, 3))
    return layers, x, t


def test_shift_matches_index_loop():
    rng = np.random.default_rng(1)
    v = rng.standard_normal((3, 4, 5, 2))
    for dy, dz, dx in cnn_ref.taps():
        s = cnn_ref.shift(v, dy, dz, dx)
        for y in range(3):
            for z in range(4):
                for x in range(5):
                    yy = y + dy
                    e = v[yy, (z + dz) % 4, (x + dx) % 5] if 0 <= yy < 3 else np.zeros(2)
                    assert np.array_equal(s[y, z, x], e)


def test_backward_matches_finite_differences():
    layers, x, t = _tiny()
    y, cache = cnn_ref.forward(layers, x)
    grads = cnn_ref.backward(layers, cache, y, t)
    rng = np.random.default_rng(2)
    h = 1e-6
    for l, (W, b) in enumerate(layers):
        for _ in range(6):
            idx = tuple(rng.integers(0, s) for s in W.shape)
            W0 = W[idx]
            W[idx] = W0 + h
            lp = cnn_ref.loss(cnn_ref.forward(layers, x)[0], t)
            W[idx] = W0 - h
            lm = cnn_ref.loss(cnn_ref.forward(layers, x)[0], t)
            W[idx] = W0
            fd = (lp - lm) / (2 * h)
            assert abs(grads[l][0][idx] - fd) <= 1e-6 * max(1.0, abs(fd)), (l, idx, grads[l][0][idx], fd)
        k = int(rng.integers(0, b.shape[0]))
        b0 = b[k]
        b[k] = b0 + h
        lp = cnn_ref.loss(cnn_ref.forward(layers, x)[0], t)
        b[k] = b0 - h
        lm = cnn_ref.loss(cnn_ref.forward(layers, x)[0], t)
        b[k] = b0
        assert abs(grads[l][1][k] - (lp - lm) / (2 * h)) <= 1e-6


def test_gelu_derivative_and_bf16_rounding():
    z = np.linspace(-6, 6, 101)
    fd = (cnn_ref.gelu(z + 1e-6) - cnn_ref.gelu(z - 1e-6)) / 2e-6
    assert np.allclose(cnn_ref.gelu_d(z), fd, atol=1e-8)
    # bf16: 8 bits of mantissa, ties to even
    assert cnn_ref.round_bf16(np.float32(1.0 + 2 ** -8)) == 1.0
    assert cnn_ref.round_bf16(np.float32(1.0 + 3 * 2 ** -8)) == 1.0 + 2 ** -6
    assert cnn_ref.round_bf16(np.float32(1.0 + 2 ** -7)) == 1.0 + 2 ** -7
