"""The closure network's BF16 MFMA kernels (kernels_cnn.hip) against the CPU
reference oracle/cnn_ref.py, which emulates the device's BF16 storage of
inputs, weights, activations and gradients (accumulation FP32 on the device,
float64 on the host):

* forward output to 1e-3 relative L2;
* gradients of every layer's weights and biases to 5e-3 relative L2 and
  cosine >= 0.9999;
* training: Adam on a linear-teacher target lowers the loss 5x in 150 steps.
"""
import numpy as np
import pytest

from cfd_amd import cnn
from oracle import cnn_ref

pytestmark = pytest.mark.gpu

SHAPE = (3, 8, 32)


def _data(seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(SHAPE + (4,)).astype(np.float32)
    x[..., 3] = (rng.random(SHAPE) < 0.2).astype(np.float32)
    t = rng.standard_normal(SHAPE + (3,)).astype(np.float32)
    return x, t


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_forward_and_gradients_match_reference():
    print("case cnn forward/backward", SHAPE, flush=True)
    w = cnn.init_params(seed=3, last_scale=1.0)
    x, t = _data(4)
    net = cnn.CNN(*SHAPE)
    try:
        net.set_params(w)
        net.set_input(x)
        net.set_target(t)
        y = net.forward()
        loss = net.backward()
        g = net.grads()
    finally:
        net.close()
    layers = [(W.astype(np.float64), b.astype(np.float64)) for W, b in cnn.split(w.astype(np.float64))]
    yr, cache = cnn_ref.forward(layers, x.astype(np.float64), bf16=True)
    tr = cnn_ref.round_bf16(t)
    assert _rel(y, yr) <= 1e-3, _rel(y, yr)
    assert abs(loss - cnn_ref.loss(yr, tr)) <= 1e-3 * cnn_ref.loss(yr, tr)
    gr = cnn_ref.backward(layers, cache, yr, tr, bf16=True)
    for l, ((dW, db), (gW, gb)) in enumerate(zip(gr, cnn.split(g.astype(np.float64)))):
        e = _rel(gW, dW)
        cos = float(np.dot(gW.ravel(), dW.ravel()) / (np.linalg.norm(gW) * np.linalg.norm(dW)))
        assert e <= 5e-3 and cos >= 0.9999, (l, e, cos)
        assert _rel(gb, db) <= 5e-3, (l, _rel(gb, db))


def test_adam_fits_a_linear_teacher():
    print("case cnn training", SHAPE, flush=True)
    rng = np.random.default_rng(5)
    x, _ = _data(6)
    Wt = rng.standard_normal((27, 4, 3)) * 0.2
    t = cnn_ref.conv(x.astype(np.float64), Wt, np.zeros(3)).astype(np.float32)
    net = cnn.CNN(*SHAPE)
    try:
        net.set_params(cnn.init_params(seed=7))
        net.set_input(x)
        net.set_target(t)
        losses = []
        for _ in range(150):
            net.forward()
            losses.append(net.backward())
            net.adam(3e-3)
    finally:
        net.close()
    assert losses[-1] < 0.2 * losses[0], (losses[0], losses[-1])
