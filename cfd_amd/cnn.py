"""The closure network (cfd_cnn_* of the C ABI): a 3D CNN over a channel
volume, BF16 MFMA kernels on the device.  Host arrays are float32,
point-major: input (ny, nz, nx, 4) = u, v, w, mask; target and output
(ny, nz, nx, 3)."""
import ctypes

import numpy as np

from . import _lib

_F = ctypes.POINTER(ctypes.c_float)
CIN = (4, 32, 32, 32)
COUT = (32, 32, 32, 3)


def nparams():
    return _lib.load().cfd_cnn_nparams()


def split(flat):
    """flat parameters -> [(W (27, cin, cout), b (cout,))] per layer"""
    out, o = [], 0
    for ci, co in zip(CIN, COUT):
        W = flat[o:o + 27 * ci * co].reshape(27, ci, co)
        o += 27 * ci * co
        b = flat[o:o + co]
        o += co
        out.append((W, b))
    assert o == len(flat)
    return out


def join(layers):
    return np.concatenate([np.concatenate([W.ravel(), b.ravel()]) for W, b in layers]).astype(np.float32)


def init_params(seed=0, last_scale=0.1):
    """He-normal weights for the GELU layers, a small last layer, zero biases."""
    rng = np.random.default_rng(seed)
    layers = []
    for l, (ci, co) in enumerate(zip(CIN, COUT)):
        std = np.sqrt(2.0 / (27 * ci)) * (last_scale if l == len(CIN) - 1 else 1.0)
        layers.append((rng.standard_normal((27, ci, co)) * std, np.zeros(co)))
    return join(layers)


class CNN:
    def __init__(self, ny, nz, nx, device=0):
        self._lib = _lib.load()
        h = ctypes.c_void_p()
        rc = self._lib.cfd_cnn_create(device, ny, nz, nx, ctypes.byref(h))
        if rc:
            raise _lib.CfdError("cfd_cnn_create: %s" % self._lib.cfd_cnn_last_error(None).decode())
        self._h = h
        self.shape = (ny, nz, nx)
        self.npts = ny * nz * nx

    def _call(self, name, *args):
        rc = getattr(self._lib, name)(self._h, *args)
        if rc:
            raise _lib.CfdError("%s: %s" % (name, self._lib.cfd_cnn_last_error(self._h).decode()))

    def close(self):
        if getattr(self, "_h", None) is not None:
            self._lib.cfd_cnn_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def set_params(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        assert w.size == nparams()
        self._call("cfd_cnn_set_params", w.ctypes.data_as(_F))

    def params(self):
        w = np.zeros(nparams(), np.float32)
        self._call("cfd_cnn_get_params", w.ctypes.data_as(_F))
        return w

    def grads(self):
        g = np.zeros(nparams(), np.float32)
        self._call("cfd_cnn_get_grads", g.ctypes.data_as(_F))
        return g

    def set_input(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        assert x.shape == self.shape + (4,)
        self._call("cfd_cnn_set_input", x.ctypes.data_as(_F))

    def set_target(self, t):
        t = np.ascontiguousarray(t, dtype=np.float32)
        assert t.shape == self.shape + (3,)
        self._call("cfd_cnn_set_target", t.ctypes.data_as(_F))

    def forward(self):
        self._call("cfd_cnn_forward")
        y = np.zeros(self.shape + (3,), np.float32)
        self._call("cfd_cnn_get_output", y.ctypes.data_as(_F))
        return y

    def backward(self):
        loss = ctypes.c_double()
        self._call("cfd_cnn_backward", ctypes.byref(loss))
        return loss.value

    def adam(self, lr, b1=0.9, b2=0.999, eps=1e-8):
        self._call("cfd_cnn_adam", ctypes.c_float(lr), ctypes.c_float(b1), ctypes.c_float(b2), ctypes.c_float(eps))

    def dataset(self, nslots):
        self._call("cfd_cnn_dataset", int(nslots))

    def train_slot(self, slot, lr, b1=0.9, b2=0.999, eps=1e-8):
        loss = ctypes.c_double()
        self._call("cfd_cnn_train_slot", int(slot), ctypes.c_float(lr), ctypes.c_float(b1), ctypes.c_float(b2),
                   ctypes.c_float(eps), ctypes.byref(loss))
        return loss.value

    def eval_slot(self, slot):
        loss = ctypes.c_double()
        self._call("cfd_cnn_eval_slot", int(slot), ctypes.byref(loss))
        return loss.value

    def load_slot(self, slot):
        self._call("cfd_cnn_load_slot", int(slot))

    def set_scales(self, s6):
        s = np.ascontiguousarray(s6, dtype=np.float32)
        self._call("cfd_cnn_set_scales", s.ctypes.data_as(_F))

    def scales(self):
        s = np.zeros(6, np.float32)
        self._call("cfd_cnn_get_scales", s.ctypes.data_as(_F))
        return s

    def device_bytes(self):
        v = ctypes.c_double()
        self._call("cfd_cnn_device_bytes", ctypes.byref(v))
        return v.value


def flops_per_forward(npts):
    """Multiply-adds x 2 of one forward pass as the kernels run it: 27 taps x
    (padded input channels) x (padded output channels) per point and layer
    (layer 0 pads its 4 inputs to 16, layer 3 its 3 outputs to 32)."""
    cin_pad, cout_pad = (16, 32, 32, 32), (32, 32, 32, 32)
    return 2.0 * 27 * sum(a * b for a, b in zip(cin_pad, cout_pad)) * npts


def bench(ny=209, nz=96, nx=192, reps=10, seed=0):
    """Times the network on one volume of the closure's coarse grid (random
    input and target): a forward pass (ms, wall clock over ``reps`` passes,
    one output read at the end, less one output read alone) and a training
    step (forward, backward with the loss read, Adam)."""
    import time
    net = CNN(ny, nz, nx)
    try:
        rng = np.random.default_rng(seed)
        net.set_params(init_params(seed))
        net.set_input(rng.standard_normal((ny, nz, nx, 4)).astype(np.float32))
        net.set_target(0.1 * rng.standard_normal((ny, nz, nx, 3)).astype(np.float32))
        net.forward()
        net.backward()
        net.adam(1e-4)
        t0 = time.perf_counter()
        net.forward()
        t_read = time.perf_counter() - t0  # one pass + the read
        t0 = time.perf_counter()
        for _ in range(reps):
            net._call("cfd_cnn_forward")
        net.forward()
        t_fwd = (time.perf_counter() - t0 - t_read) / reps
        t0 = time.perf_counter()
        for _ in range(reps):
            net._call("cfd_cnn_forward")
            net.backward()
            net.adam(1e-4)
        t_train = (time.perf_counter() - t0) / reps
        fl = flops_per_forward(net.npts)
        return {"volume": [ny, nz, nx], "points": net.npts, "params": nparams(), "reps": reps,
                "forward_ms": 1e3 * t_fwd, "train_step_ms": 1e3 * t_train,
                "forward_tflop": fl / 1e12, "forward_tflops_per_s": fl / t_fwd / 1e12,
                "device_bytes": net.device_bytes()}
    finally:
        net.close()


def main(argv=None):
    import argparse
    import json
    ap = argparse.ArgumentParser(description="time the closure network on one coarse-grid volume")
    ap.add_argument("--ny", type=int, default=209)
    ap.add_argument("--nz", type=int, default=96)
    ap.add_argument("--nx", type=int, default=192)
    ap.add_argument("--reps", type=int, default=10)
    a = ap.parse_args(argv)
    print(json.dumps(bench(a.ny, a.nz, a.nx, a.reps)))


if __name__ == "__main__":
    main()
