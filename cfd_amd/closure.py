"""Learned closure of the nonlinear term on a 2x coarser x/z grid.

The network (cfd_amd.cnn, BF16 MFMA kernels) sees the coarse velocity and
the IB mask and predicts filt(N(u)) - N_coarse(filt(u)), the part of the
box-filtered fine nonlinear term that the coarse grid misses.  Samples come
from a fine FP64 run (cfd_closure_snapshot); the trained network is attached
to the coarse solver, which then adds the correction in every nonlinear term.

    python -m cfd_amd.closure [--snapshots 2000] [--steps 5000] [--out DIR]

runs the whole experiment at Re_tau 395: fine DNS 384 x 192 x 192 as the
data source and the judge, coarse 192 x 192 x 96 with and without the
closure, and writes closure_weights.npz and closure_report.json.
"""
import argparse
import ctypes
import json
import os
import time

import numpy as np

from . import _lib, cnn, ic, workloads
from .channel import Channel


def snapshot(fine, coarse, net, slot):
    """One sample into dataset slot `slot`; returns the six RMS values."""
    rms = np.zeros(6)
    rc = _lib.load().cfd_closure_snapshot(fine._h, coarse._h, net.handle, int(slot),
                                          rms.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    _lib.check(rc, coarse._h)
    return rms


def start_fields(ch, params, seed=0, amp=0.1):
    """Seeded laminar + perturbation field, projected, loaded into ch."""
    g = ch.ygrid()
    f = ic.project(ch, ic.initial_fields(params, g, seed=seed, amp=amp))
    for k in ("u1", "u2", "u3", "p"):
        ch.set_field(k, f[k])
    ch.set_scalars(t=0.0, dt=0.0, mpgx=0.0, QxT=ic.flowrate(f["u1"], g, params.ny))
    return f


def filtered_start(coarse_params, fine_fields, fine_params):
    """The fine start field box-filtered onto the coarse modes (host side,
    the same filter as the device's)."""
    N1f, N2f = 2 * (fine_params.ngalx // 3), 2 * (fine_params.ngalz // 3)
    N1c, N2c = 2 * (coarse_params.ngalx // 3), 2 * (coarse_params.ngalz // 3)
    Mxf, Mxc = N1f // 2 + 1, N1c // 2 + 1
    kz = np.arange(N2c)
    kz = np.where(kz < N2c // 2, kz, kz - N2c)
    kf = np.where(kz >= 0, kz, N2f + kz)
    i = np.arange(Mxc)
    h = np.sinc(i / coarse_params.ngalx)[None, :] * np.sinc(np.abs(kz) / coarse_params.ngalz)[:, None]
    out = {}
    for name, f in fine_fields.items():
        ny = f.shape[1]
        a = f.reshape(N2f, Mxf, ny)[kf][:, :Mxc, :] * h[:, :, None]
        out[name] = a.reshape(N2c * Mxc, ny)
    return out


def drag(ch):
    """Mean pressure gradient (the drag per unit volume at constant flow rate)."""
    return -ch.scalars()["mpgx"]


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--snapshots", type=int, default=2000)
    ap.add_argument("--every", type=int, default=2, help="fine steps between samples")
    ap.add_argument("--slots", type=int, default=256, help="device ring buffer of samples")
    ap.add_argument("--train-per-sample", type=int, default=1)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--steps", type=int, default=5000, help="steps of the a-posteriori runs")
    ap.add_argument("--out", default=".")
    ap.add_argument("--seed", type=int, default=0)
    a = ap.parse_args(argv)
    pf, _ = workloads.get("rt395_384x192x192")
    pc, _ = workloads.get("rt395_coarse_192x192x96")
    fine, coarse = Channel(pf), Channel(pc)
    net = cnn.CNN(coarse.nyu, pc.ngalz, pc.ngalx)
    net.set_params(cnn.init_params(seed=a.seed))
    net.dataset(a.slots)
    f0 = start_fields(fine, pf, seed=0)
    rng = np.random.default_rng(a.seed)
    report = {"snapshots": a.snapshots, "every": a.every, "slots": a.slots}
    t0 = time.perf_counter()
    losses, held = [], []
    filled = 0
    for k in range(a.snapshots):
        fine.step(a.every)
        slot = k % a.slots
        rms = snapshot(fine, coarse, net, slot)
        filled = min(filled + 1, a.slots)
        if k == 0:
            report["rms_first"] = rms.tolist()
            report["scales"] = net.scales().tolist()
        # held-out check before training on the new sample
        if k >= 10 and k % 50 == 0:
            held.append((k, net.eval_slot(slot)))
            print("sample %5d  t %.4f  held-out loss %.4f  train loss %.4f" % (
                k, fine.scalars()["t"], held[-1][1], np.mean(losses[-50:]) if losses else float("nan")), flush=True)
        for _ in range(a.train_per_sample):
            s = slot if rng.random() < 0.5 else int(rng.integers(0, filled))
            losses.append(net.train_slot(s, a.lr))
    report["train_seconds"] = time.perf_counter() - t0
    report["held_out"] = held
    report["train_loss_last100"] = float(np.mean(losses[-100:]))
    w = net.params()
    np.savez(os.path.join(a.out, "closure_weights.npz"), params=w, scales=net.scales())
    fine.close()
    coarse.close()
    print(json.dumps(report), flush=True)
    net.close()
    return report


if __name__ == "__main__":
    main()
