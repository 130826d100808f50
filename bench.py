"""Throughput of the RK3/CN substep on one MI355X (metric: RK3 substeps/s).

A step is one RK3 substep of the rough-channel DNS (RHS, nonlinear term with
its FFTs, IB forcing, the three implicit y-solves, flow-rate correction,
pressure solve and projection) over the whole device-resident field.  The
workload is Re_tau = 395 on 384 x 192 x 192 (x, y, z) with 16 x 8 posts per
wall, FP64.  Fields are generated on the host (seeded), projected and uploaded
before timing; the timed region runs on device-resident data only.

Prints one JSON line.  --gpus N > 1 exits: this build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0


def xpass_bytes(ch):
    """Algorithmic bytes of one x-pass launch: each Z plane row read or
    written once (3 u-plane + 3 v-plane inputs; 4 interior u-plane + 2
    v-plane product outputs), double2 elements."""
    zrow = ch.params.ngalz * ch.Mx * 16
    return zrow * (3 * ch.nyu + 3 * ch.nyv + 4 * (ch.nyu - 2) + 2 * ch.nyv)


def class_bytes(ch):
    """Algorithmic bytes per launch of each kernel class."""
    S = lambda ny: ny * ch.Ncol * 16  # noqa: E731
    Z = lambda ny: ny * ch.params.ngalz * ch.Mx * 16  # noqa: E731
    nyu, nyv, nyp = ch.nyu, ch.nyv, ch.nyp
    return {
        # 6 jobs families: read spectral planes once (u1,u3,v + interpolation
        # reuse), write 6 Z stacks
        "zinv": 2 * S(nyu) + S(nyv) + 3 * Z(nyu) + 3 * Z(nyv),
        "xpass": xpass_bytes(ch),
        # read 4 interior u-plane + 2 v-plane product stacks, write 6 spectral
        "zfwd": 4 * Z(nyu - 2) + 2 * Z(nyv) + 3 * S(nyu - 2) + 2 * S(nyv) + S(nyv - 2),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30, help="timed RK3 substeps")
    ap.add_argument("--warmup", type=int, default=6, help="untimed RK3 substeps")
    ap.add_argument("--workload", default="rt395_384x192x192")
    ap.add_argument("--cpu-substeps", type=int, default=3, help="oracle substeps for the CPU baseline (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs the substep on one GPU; --gpus %d is not supported" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    from cfd_amd import Channel, ic, workloads

    params, _ = workloads.get(args.workload)
    ch = Channel(params)
    g = ch.ygrid()
    f = ic.initial_fields(params, g, seed=0, amp=0.1)
    f = ic.project(ch, f)
    q0 = ic.flowrate(f["u1"], g, params.ny)
    ch.set_scalars(t=0.0, dt=0.0, mpgx=0.0, QxT=q0)
    k = [0]

    def run(n):
        for _ in range(n):
            ch.substep(k[0] % 3 + 1)
            k[0] += 1

    run(args.warmup)
    ch.sync()
    t0 = time.perf_counter()
    run(args.steps)
    ch.sync()
    t1 = time.perf_counter()
    dt_wall = t1 - t0
    value = args.steps / dt_wall

    # per-kernel-class durations with HIP events on the solver's stream
    nprof = 6
    ch.profile(True)
    run(nprof)
    prof = ch.profile_read()
    ch.profile(False)
    cb = class_bytes(ch)
    per_launch = {c: (ms / cnt if cnt else 0.0) for c, (ms, cnt) in prof.items()}
    dom = max(("zinv", "xpass", "zfwd"), key=lambda c: prof[c][0])
    avg_ms = per_launch[dom]
    achieved = cb[dom] / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic_%s_%s.json" % (dom, args.workload))
    if os.path.exists(tf):
        with open(tf) as fh:
            traffic = json.load(fh).get("hbm_bytes_per_launch")
    sc = ch.scalars()
    u1 = ch.get_field("u1")
    finite = bool(np.isfinite(u1).all())

    # CPU baseline: the FP64 oracle on the host cores, a bounded sample
    cpu = None
    if args.cpu_substeps > 0:
        from oracle.oracle import Oracle, set_threads
        ncores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
        set_threads(ncores)
        orc = Oracle(params)
        for key in ("u1", "u2", "u3", "p"):
            orc.set_field(key, f[key])
        orc.set_scalars(t=0.0, dt=0.0, mpgx=0.0, QxT=q0)
        c0 = time.perf_counter()
        for s in range(args.cpu_substeps):
            orc.substep(s % 3 + 1)
        c1 = time.perf_counter()
        cpu = {"value": args.cpu_substeps / (c1 - c0), "unit": "substeps/s", "cores": ncores, "kind": "port",
               "sample": "%d RK3 substeps of the same %s workload from the same initial field (oracle/, FP64, "
                         "OpenMP)" % (args.cpu_substeps, args.workload)}
        orc.close()

    out = {
        "metric": "RK3 substeps/s at 384x192x192 FP64",
        "value": value,
        "unit": "substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * dt_wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": args.workload, "grid": [params.ngalx, params.ny, params.ngalz],
                   "dsty": params.dsty, "posts_per_wall": [params.ntilex, params.ntilez], "Re_bulk": params.Re,
                   "precision": "fp64", "parallelism": "single-gpu"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "traffic": traffic, "avg_launch_ms": avg_ms,
                     "algorithmic_bytes": cb[dom]},
        "kernel_ms_per_substep": {c: ms / nprof for c, (ms, cnt) in prof.items()},
        "cpu_baseline": cpu,
        "sim": {"t": sc["t"], "dt": sc["dt"], "mpgx": sc["mpgx"], "finite": finite},
    }
    print(json.dumps(out))
    ch.close()


if __name__ == "__main__":
    main()
