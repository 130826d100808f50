"""Learned closure of the nonlinear term on a 2x coarser x/z grid.

The network (cfd_amd.cnn, BF16 MFMA kernels) sees the coarse velocity and
the IB mask and predicts filt(N(u)) - N_coarse(filt(u)), the part of the
box-filtered fine nonlinear term that the coarse grid misses.  Samples come
from a fine FP64 run (cfd_closure_snapshot); the trained network is attached
to the coarse solver, which then adds the correction in its nonlinear terms.

    python -m cfd_amd.closure [--samples 2000] [--every 5] [--out DIR]

runs the experiment at Re_tau 395: the fine DNS on 384 x 192 x 192 produces
the samples (one every `every` steps, trained on online from a device ring
buffer) and is the reference; the coarse solver on 192 x 192 x 96 then runs from
the box-filtered start field over the same simulated time without and with
the closure (as predicted, and backscatter-clipped: components that would feed
energy into the resolved field set to zero), and once more with the closure
and the full DNS as fallback (run_with_fallback).  Writes closure_weights.npz and
closure_report.json: a-priori
loss on held-out samples against predicting no correction, drag (the mean
pressure gradient at constant flow rate) and mean-profile errors against the
fine run, and simulated time per wall-clock second of each run.
"""
import argparse
import dataclasses
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
    q0 = ic.flowrate(f["u1"], g, params.ny)
    for k in ("u1", "u2", "u3", "p"):
        ch.set_field(k, f[k])
    ch.set_scalars(t=0.0, dt=0.0, mpgx=0.0, QxT=q0)
    return f, q0


def filtered_start(coarse_params, fine_fields, fine_params):
    """Fine spectral fields box-filtered onto the coarse modes (host side,
    the same filter as the device's: H(k) = sinc(k dx_c / 2))."""
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


def prolong(fine_params, coarse_fields, coarse_params):
    """Coarse spectral fields zero-padded onto the fine modes: the column
    (i, kz) of the coarse grid goes to the column (i, kz) of the fine grid
    (the mode selection of filtered_start inverted; the box filter is not
    undone), every other fine mode is zero."""
    N1f, N2f = 2 * (fine_params.ngalx // 3), 2 * (fine_params.ngalz // 3)
    N1c, N2c = 2 * (coarse_params.ngalx // 3), 2 * (coarse_params.ngalz // 3)
    Mxf, Mxc = N1f // 2 + 1, N1c // 2 + 1
    kz = np.arange(N2c)
    kz = np.where(kz < N2c // 2, kz, kz - N2c)
    kf = np.where(kz >= 0, kz, N2f + kz)
    out = {}
    for name, f in coarse_fields.items():
        ny = f.shape[1]
        a = np.zeros((N2f, Mxf, ny), dtype=complex)
        a[kf, :Mxc, :] = f.reshape(N2c, Mxc, ny)
        out[name] = a.reshape(N2f * Mxf, ny)
    return out


def run_with_fallback(coarse, fine, t_end, max_steps, drag_cap, checkpoint_every=50, drag_lo=None):
    """The coarse run with the full DNS as its fallback.  The coarse solver
    (closure attached or not) steps until t >= t_end; every checkpoint_every
    steps its state is copied to the host.  When its drag leaves
    [-drag_cap, drag_cap] (and [drag_lo, inf) when drag_lo is given) or
    stops being finite, the fine solver takes the
    last checkpoint (zero-padded onto the fine modes, with its t, dt and mean
    pressure gradient) and runs on to t_end (or to a drag that is not
    finite, the last entry then).  Returns the (t, drag, on_fine)
    history, the t of the switch (None without one) and the wall time."""
    names = ("u1", "u2", "u3", "p")

    def snap(ch):
        return {n: ch.get_field(n) for n in names}, ch.scalars()

    hist = []
    w0 = time.perf_counter()
    ck = snap(coarse)
    switch = None
    for k in range(max_steps):
        coarse.step(1)
        s = coarse.scalars()
        d = -s["mpgx"]
        if not (np.isfinite(d) and abs(d) <= drag_cap and (drag_lo is None or d >= drag_lo)):
            switch = ck[1]["t"]
            break
        hist.append((s["t"], d, 0))
        if s["t"] >= t_end:
            break
        if (k + 1) % checkpoint_every == 0:
            ck = snap(coarse)
    if switch is not None:
        hist = [h for h in hist if h[0] <= switch]
        f = prolong(fine.params, ck[0], coarse.params)
        for n in names:
            fine.set_field(n, f[n])
        sc = ck[1]
        fine.set_scalars(t=sc["t"], dt=sc["dt"], dti=sc["dti"], mpgx=sc["mpgx"], QxT=sc["QxT"])
        for _ in range(max_steps):
            fine.step(1)
            s = fine.scalars()
            hist.append((s["t"], -s["mpgx"], 1))
            if s["t"] >= t_end or not np.isfinite(s["mpgx"]):
                break
    fine.sync()
    coarse.sync()
    return np.array(hist).reshape(-1, 3), switch, time.perf_counter() - w0


def run_until(ch, t_end, max_steps, drag_cap):
    """Step ch until t >= t_end, or until the drag leaves [-drag_cap,
    drag_cap] or stops being finite (a blown-up run); returns (t, drag) per
    step and the wall time."""
    hist = []
    ch.sync()
    w0 = time.perf_counter()
    for _ in range(max_steps):
        ch.step(1)
        s = ch.scalars()
        d = -s["mpgx"]
        if not (np.isfinite(d) and abs(d) <= drag_cap):
            break
        hist.append((s["t"], d))
        if s["t"] >= t_end:
            break
    ch.sync()
    return np.array(hist).reshape(-1, 2), time.perf_counter() - w0


def compare(hist, ref, t0, t1):
    """Drag of hist against ref (both (t, drag) arrays) over t0 <= t <= t1:
    relative error of the mean and RMS of the pointwise difference."""
    m = (hist[:, 0] >= t0) & (hist[:, 0] <= t1)
    mr = (ref[:, 0] >= t0) & (ref[:, 0] <= t1)
    if not m.any():
        return {"window": [t0, t1], "mean_drag": None, "drag_error": None, "drag_rms_error": None}
    a, b = hist[m, 1].mean(), ref[mr, 1].mean()
    rb = np.interp(hist[m, 0], ref[:, 0], ref[:, 1])
    return {"window": [float(t0), float(t1)], "mean_drag": float(a), "mean_drag_ref": float(b),
            "drag_error": float(abs(a - b) / abs(b)),
            "drag_rms_error": float(np.sqrt(np.mean((hist[m, 1] - rb) ** 2)) / abs(b))}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--samples", type=int, default=2000)
    ap.add_argument("--every", type=int, default=5, help="fine steps between samples")
    ap.add_argument("--slots", type=int, default=256, help="device ring buffer of samples")
    ap.add_argument("--train-per-sample", type=int, default=2)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--held-out", type=int, default=8)
    ap.add_argument("--closure-every", default="1,2,4",
                    help="coarse steps between network evaluations, one coarse run per value (comma list)")
    ap.add_argument("--max-coarse-steps", type=int, default=20000)
    ap.add_argument("--coarse-start", type=float, default=0.0,
                    help="fraction of the fine run's time at which the coarse runs start from its filtered state")
    ap.add_argument("--closure-gain", type=float, default=0.35,
                    help="the network's output scales times this gain in the coarse runs")
    ap.add_argument("--gain-scan", default="0,0.2,0.35,0.5,0.7",
                    help="gains of the scan (closure every step) from the start and the validation start")
    ap.add_argument("--validation-start", type=float, default=0.75,
                    help="fraction of the fine run's time of the scan's second start")
    ap.add_argument("--filter", default="36,32,0",
                    help="modal filter of the coarse runs, alpha,order,eta_c (alpha 0: none)")
    ap.add_argument("--weights", default=None,
                    help="evaluate this trained network (closure_weights.npz) instead of training one")
    ap.add_argument("--out", default=".")
    ap.add_argument("--seed", type=int, default=0, help="network initialisation and sample draws")
    ap.add_argument("--coarse-precision", type=int, default=0,
                    help="precision of the evaluated coarse runs (0 FP64, 1 mixed); samples and the fine run are FP64")
    ap.add_argument("--fine-seed", type=int, default=0,
                    help="seed of the fine run's initial field (another trajectory for choosing the gain)")
    a = ap.parse_args(argv)
    fa, fo, fe = a.filter.split(",")
    filt = (float(fa), int(fo), float(fe))
    every_list = [int(e) for e in a.closure_every.split(",")]
    gains = [float(g) for g in a.gain_scan.split(",")] if a.gain_scan else []
    pf, _ = workloads.get("rt395_384x192x192")
    pc, _ = workloads.get("rt395_coarse_192x192x96")
    fine, scratch = Channel(pf), Channel(pc)
    pr = dataclasses.replace(pc, precision=a.coarse_precision)  # the evaluated coarse runs
    net = cnn.CNN(scratch.nyu, pc.ngalz, pc.ngalx)
    net.set_params(cnn.init_params(seed=a.seed))
    net.dataset(a.slots + a.held_out)
    rng = np.random.default_rng(a.seed)
    report = {"coarse_precision": a.coarse_precision, "fine_seed": a.fine_seed, "seed": a.seed, "samples": a.samples, "every": a.every, "slots": a.slots, "train_per_sample": a.train_per_sample,
              "closure_every": every_list, "closure_gain": a.closure_gain, "coarse_filter": {"alpha": filt[0], "order": filt[1], "eta_c": filt[2]}}
    train = a.weights is None
    if train:
        # pass 1: the fine run alone, probing the sample magnitudes along the
        # trajectory; the normalisation takes the largest RMS of each quantity
        f0, q0 = start_fields(fine, pf, seed=a.fine_seed)
        probe = cnn.CNN(scratch.nyu, pc.ngalz, pc.ngalx)
        probe.dataset(1)
        nsteps = a.samples * a.every
        rmax = np.zeros(6)
        for k in range(nsteps):
            fine.step(1)
            if (k + 1) % max(nsteps // 24, 1) == 0:
                rmax = np.maximum(rmax, snapshot(fine, scratch, probe, 0))
        probe.close()
        report["scales_max_rms"] = rmax.tolist()
        net.set_scales(rmax.astype(np.float32))
        print("scales", rmax, flush=True)
    else:
        wts = np.load(a.weights)
        net.set_params(wts["params"])
        net.set_scales(np.asarray(wts["scales"], np.float32))
        report["weights"] = a.weights
    # pass 2: the same fine run from the same start: samples, online
    # training, and the reference drag history
    f0, q0 = start_fields(fine, pf, seed=a.fine_seed)
    k_start = int(a.coarse_start * a.samples)
    k_val = int(a.validation_start * a.samples)
    fs, ts = (f0, 0.0), None
    fv = None
    hist_f = []
    losses = []
    filled = 0
    w0 = time.perf_counter()
    wfine = 0.0
    for k in range(a.samples):
        t1 = time.perf_counter()
        for _ in range(a.every):
            fine.step(1)
            s = fine.scalars()
            hist_f.append((s["t"], -s["mpgx"]))
        wfine += time.perf_counter() - t1
        if k + 1 == k_start:
            fs = ({n: fine.get_field(n) for n in ("u1", "u2", "u3", "p")}, fine.scalars()["t"])
        if k + 1 == k_val:
            fv = ({n: fine.get_field(n) for n in ("u1", "u2", "u3", "p")}, fine.scalars()["t"])
        if not train:
            continue
        slot = k % a.slots
        rms = snapshot(fine, scratch, net, slot)
        filled = min(filled + 1, a.slots)
        if k == 0:
            report["rms_first"] = rms.tolist()
        for _ in range(a.train_per_sample):
            s = slot if rng.random() < 0.5 else int(rng.integers(0, filled))
            losses.append(net.train_slot(s, a.lr))
        if k % 100 == 0:
            print("sample %5d  t %.3f  drag %.5f  train loss %.4f  (%.0f s)" % (
                k, hist_f[-1][0], hist_f[-1][1], np.mean(losses[-100:]), time.perf_counter() - w0), flush=True)
    if train:
        report["train_seconds"] = time.perf_counter() - w0 - wfine
        report["train_loss_last200"] = float(np.mean(losses[-200:]))
    # held-out samples further along the fine run
    zero = cnn.CNN(scratch.nyu, pc.ngalz, pc.ngalx)
    zero.set_params(np.zeros(cnn.nparams(), np.float32))
    zero.set_scales(net.scales())
    zero.dataset(a.held_out)
    held = []
    for k in range(a.held_out):
        for _ in range(a.every):
            fine.step(1)
            s = fine.scalars()
            hist_f.append((s["t"], -s["mpgx"]))
        snapshot(fine, scratch, net, a.slots + k)
        snapshot(fine, scratch, zero, k)
        held.append((net.eval_slot(a.slots + k), zero.eval_slot(k)))
    held = np.array(held)
    report["held_out_loss"] = float(held[:, 0].mean())
    report["held_out_loss_no_correction"] = float(held[:, 1].mean())
    report["held_out_gain"] = float(held[:, 1].mean() / held[:, 0].mean())
    print("held-out loss %.4f vs %.4f without correction" % (held[:, 0].mean(), held[:, 1].mean()), flush=True)
    hist_f = np.array(hist_f)
    t_end = float(hist_f[-1, 0])
    # simulated time per second of the fine steps of the sampling loop alone
    report["fine"] = {"steps": len(hist_f), "t_end": t_end, "wall_s": wfine,
                      "sim_time_per_s": float(hist_f[a.samples * a.every - 1, 0] / max(wfine, 1e-9))}
    w = net.params()
    scales = net.scales()
    np.savez(os.path.join(a.out, "closure_weights.npz"), params=w, scales=scales, gain=a.closure_gain)

    def set_gain(g):
        sc = np.array(scales, np.float32).copy()
        sc[3:] *= g
        net.set_scales(sc)
    zero.close()
    scratch.close()
    # the coarse runs from the filtered start field, without and with the
    # closure, over the fine run's simulated time (or until they blow up)
    fc = filtered_start(pc, fs[0], pf)
    t_start = fs[1]
    report["coarse_start_t"] = t_start
    cap = 50.0 * float(np.abs(hist_f[:, 1]).max())
    runs, prof = {}, {}
    set_gain(a.closure_gain)
    cases = [("coarse", None, 0)] + [("coarse_closure_every%d" % e, "plain", e) for e in every_list]
    cases.append(("coarse_closure_clip_every%d" % every_list[0], "clip", every_list[0]))
    for label, attach, every in cases:
        ch = Channel(pr)
        g = ch.ygrid()
        for kname in ("u1", "u2", "u3", "p"):
            ch.set_field(kname, fc[kname])
        ch.set_scalars(t=t_start, dt=0.0, mpgx=-float(np.interp(t_start, hist_f[:, 0], hist_f[:, 1])) if t_start else 0.0,
                       QxT=ic.flowrate(fc["u1"], g, pc.ny))
        ch.set_mode_filter(*filt)
        if attach:
            ch.closure_attach(net, every=every, clip=attach == "clip")
        hist, wall = run_until(ch, t_end, a.max_coarse_steps, cap)
        t_last = float(hist[-1, 0]) if len(hist) else 0.0
        runs[label] = {"hist": hist, "steps": len(hist), "wall_s": wall, "t_reached": t_last,
                       "completed": bool(t_last >= t_end),
                       "sim_time_per_s": float((t_last - t_start) / wall) if wall > 0 else 0.0}
        runs[label]["speedup_vs_fine"] = runs[label]["sim_time_per_s"] / report["fine"]["sim_time_per_s"]
        prof[label] = ch.get_field("u1")[0].real if runs[label]["completed"] else None
        ch.closure_attach(None)
        ch.close()
    t1 = min(r["t_reached"] for r in runs.values())
    for label, r in runs.items():
        hist = r.pop("hist")
        r["drag"] = compare(hist, hist_f, t_start + 0.25 * (t1 - t_start), t1)
        r["drag_own_horizon"] = compare(hist, hist_f, t_start + 0.25 * (r["t_reached"] - t_start), r["t_reached"])
        if prof[label] is not None:
            uf = fine.get_field("u1")[0].real
            r["mean_profile_error"] = float(np.linalg.norm(prof[label] - uf) / np.linalg.norm(uf))
        print(label, json.dumps(r), flush=True)
    report.update(runs)
    # the gain scan: closure every step, from the start and from the
    # validation start, against the fine run over the last three quarters
    scan = {}
    for name, st in (("start", fs), ("validation_start", fv)):
        if st is None or not gains:
            continue
        fcs = filtered_start(pc, st[0], pf)
        res = []
        for gval in gains:
            ch = Channel(pr)
            g = ch.ygrid()
            for kname in ("u1", "u2", "u3", "p"):
                ch.set_field(kname, fcs[kname])
            ch.set_scalars(t=st[1], dt=0.0, mpgx=-float(np.interp(st[1], hist_f[:, 0], hist_f[:, 1])),
                           QxT=ic.flowrate(fcs["u1"], g, pc.ny))
            ch.set_mode_filter(*filt)
            if gval:
                set_gain(gval)
                ch.closure_attach(net, every=1)
            hist, _ = run_until(ch, t_end, a.max_coarse_steps, cap)
            r = compare(hist, hist_f, st[1] + 0.25 * (t_end - st[1]), t_end)
            r["gain"] = gval
            r["t_reached"] = float(hist[-1, 0]) if len(hist) else st[1]
            if r["t_reached"] >= t_end:
                uf = fine.get_field("u1")[0].real
                r["mean_profile_error"] = float(np.linalg.norm(ch.get_field("u1")[0].real - uf) / np.linalg.norm(uf))
            res.append(r)
            print("gain scan", name, json.dumps(r), flush=True)
            ch.closure_attach(None)
            ch.close()
        scan[name] = {"t": st[1], "runs": res}
    report["gain_scan"] = scan
    set_gain(a.closure_gain)
    # the learned path with the full DNS as its fallback: the coarse run with
    # the closure until its drag leaves the bound, then the fine solver from
    # its last checkpoint, zero-padded onto the fine modes, to the same t_end
    # (the network every max(closure_every) steps)
    ch, fb = Channel(pr), Channel(pf)
    g = ch.ygrid()
    for kname in ("u1", "u2", "u3", "p"):
        ch.set_field(kname, fc[kname])
    ch.set_scalars(t=t_start, dt=0.0, mpgx=-float(np.interp(t_start, hist_f[:, 0], hist_f[:, 1])) if t_start else 0.0,
                   QxT=ic.flowrate(fc["u1"], g, pc.ny))
    ch.set_mode_filter(*filt)
    ch.closure_attach(net, every=max(every_list))
    # the trust band of the coarse drag: at least 0 (a mean pressure gradient
    # that opposes the constant flow rate drives the flow backwards) and at
    # most 4 times the largest drag of the fine run before the switch-on
    pre = hist_f[:, 0] <= max(t_start, float(hist_f[0, 0]))
    hi = 4.0 * float(np.abs(hist_f[pre, 1]).max())
    hist, switch, wall = run_with_fallback(ch, fb, t_end, a.max_coarse_steps, hi, drag_lo=0.0)
    ch.closure_attach(None)
    r = {"closure_every": max(every_list), "drag_band": [0.0, hi], "switch_t": switch, "coarse_steps": int((hist[:, 2] == 0).sum()), "fine_steps": int((hist[:, 2] == 1).sum()),
         "wall_s": wall, "t_reached": float(hist[-1, 0]) if len(hist) else t_start}
    r["completed"] = bool(r["t_reached"] >= t_end and np.isfinite(hist[:, 1]).all())
    bad = ~np.isfinite(hist[:, 1])
    r["first_nonfinite_drag_t"] = float(hist[bad, 0][0]) if bad.any() else None
    r["sim_time_per_s"] = float((r["t_reached"] - t_start) / wall) if wall > 0 else 0.0
    r["speedup_vs_fine"] = r["sim_time_per_s"] / report["fine"]["sim_time_per_s"]
    r["drag"] = compare(hist[:, :2], hist_f, t_start + 0.25 * (t_end - t_start), t_end)
    print("coarse_closure_fallback", json.dumps(r), flush=True)
    report["coarse_closure_fallback"] = r
    ch.close()
    fb.close()
    fine.close()
    net.close()
    with open(os.path.join(a.out, "closure_report.json"), "w") as fh:
        json.dump(report, fh, indent=1)
    print(json.dumps(report), flush=True)
    return report


if __name__ == "__main__":
    main()
