"""Throughput of the RK3/CN substep on one MI355X (metric: RK3 substeps/s).

A step is one RK3 substep of the rough-channel DNS (RHS, nonlinear term with
its FFTs, IB forcing, the three implicit y-solves, flow-rate correction,
pressure solve and projection) over the whole device-resident field.  The
workload is Re_tau = 395 on 384 x 192 x 192 (x, y, z) with 16 x 8 posts per
wall.  Fields are generated on the host (seeded), projected and uploaded
before timing; the timed region runs on device-resident data only.

`value` is the FP64 path.  The same run times the mixed path (precision 1:
FP32 transforms, products and nonlinear scratch; FP64 fields, solves,
projection and flow-rate correction) over the same substeps from the same
field and reports its throughput and its error against the FP64 run; the
learned path (the 192 x 192 x 96 coarse grid with the trained BF16 closure
network, simulated time per second against the FP64 run's); and the
ensemble of 16 uneven channels stepped by captured HIP graphs.

Prints one JSON line.  --gpus N > 1 exits: this build runs on one GPU.
"""
import argparse
import dataclasses
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0


def class_bytes(ch, zb):
    """Algorithmic bytes per launch of the three FFT kernel classes: every
    element of a kernel's inputs and outputs counted once.  S = one spectral
    plane stack (Ncol complex per plane), Z = one half-transformed stack (Ngz x
    Mx complex per plane); spectral velocities are FP64 (16 B), Z and
    nonlinear scratch elements are zb bytes (16 FP64, 8 mixed)."""
    S = lambda ny, b=16: ny * ch.Ncol * b  # noqa: E731
    Z = lambda ny: ny * ch.params.ngalz * ch.Mx * zb  # noqa: E731
    nyu, nyv = ch.nyu, ch.nyv
    return {
        # read u1, u3 (u planes) and v (v planes) once, write 6 Z stacks
        # (u1, u2 interpolated, u3 on u planes; u1, u3 interpolated, v on v planes)
        "zinv": 2 * S(nyu) + S(nyv) + 3 * Z(nyu) + 3 * Z(nyv),
        # read the 6 input stacks, write 4 interior u-plane + 2 v-plane product stacks
        "xpass": 3 * Z(nyu) + 3 * Z(nyv) + 4 * Z(nyu - 2) + 2 * Z(nyv),
        # read 4 interior u-plane + 2 v-plane product stacks, write nxz1, nxz3,
        # vv (interior u planes), uv, vw (v planes) and nxz2 (interior v planes)
        "zfwd": 4 * Z(nyu - 2) + 2 * Z(nyv) + 3 * S(nyu - 2, zb) + 2 * S(nyv, zb) + S(nyv - 2, zb),
        # k_rhs_seg, three components: read u, the old Nu and the x/z part of
        # the nonlinear term of each, the three y fluxes (uv, vw on v planes,
        # vv on u planes) and p once; write u and Nu of each
        "ysolve": 4 * (2 * S(nyu) + S(nyv)) + S(ch.nyp) + zb / 16 * (2 * S(nyu) + S(nyv) + S(nyu) + 2 * S(nyv)),
        # k_press_seg: read u1, u3, v, p; write psi, p, u1, u3, v and div
        "press": 2 * (2 * S(nyu) + S(nyv)) + 4 * S(ch.nyp),
    }


def tuned(params, f, q0, args):
    """Launch shapes from the autotuner (cfd_amd.ensemble.tune_params) on a
    scratch solver started from the same field; (tuning, summary)."""
    if args.no_tune:
        return None, None
    from cfd_amd import ensemble

    def load(ch):
        for k in ("u1", "u2", "u3", "p"):
            ch.set_field(k, f[k])
        ch.set_scalars(t=0.0, dt=0.0, mpgx=0.0, QxT=q0)

    best, table = ensemble.tune_params(params, load)
    dflt = [ms for t, ms in table if t == dict(fft=0, seg_rows=13, solve_threads=256)][0]
    return best, {"tuning": best, "kernel_ms_per_substep_tuned": min(ms for _, ms in table),
                  "kernel_ms_per_substep_default_shapes": dflt, "candidates": len(table)}


def timed(params, f, q0, args, nprof, tuning=None):
    """Create a Channel, load the field, run warmup + timed + profiled
    substeps; returns (substeps/s, ms per substep, per-class profile, final state)."""
    from cfd_amd import Channel

    ch = Channel(params, tuning=tuning)
    for k in ("u1", "u2", "u3", "p"):
        ch.set_field(k, f[k])
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
    # per-kernel-class durations with HIP events on the solver's stream
    ch.profile(True)
    run(nprof)
    prof = ch.profile_read()
    ch.profile(False)
    sc = ch.scalars()
    state = {key: ch.get_field(key) for key in ("u1", "u2", "u3")}
    state["mpgx"] = sc["mpgx"]
    state["scalars"] = sc
    state["device_bytes"] = ch.device_bytes()
    return ch, args.steps / dt_wall, 1e3 * dt_wall / args.steps, prof, state


def roofline_classes(ch, prof, zb, nprof):
    """HBM roofline fraction of every kernel class with an algorithmic byte
    count: bytes per substep / measured class time per substep."""
    out = {}
    for c, b in class_bytes(ch, zb).items():
        ms = prof[c][0] / nprof
        gbs = b / (ms * 1e-3) / 1e9
        out[c] = {"algorithmic_bytes": b, "ms": ms, "achieved_GBps": gbs, "frac": gbs / PEAK_HBM_GBS}
    return out


def roofline(ch, prof, zb, workload, label):
    cb = class_bytes(ch, zb)
    per_launch = {c: (ms / cnt if cnt else 0.0) for c, (ms, cnt) in prof.items()}
    dom = max(("zinv", "xpass", "zfwd"), key=lambda c: prof[c][0])
    avg_ms = per_launch[dom]
    achieved = cb[dom] / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic_%s_%s%s.json" % (dom, workload, label))
    if os.path.exists(tf):
        with open(tf) as fh:
            traffic = json.load(fh).get("hbm_bytes_per_launch")
    return {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBS, "traffic": traffic, "avg_launch_ms": avg_ms,
            "algorithmic_bytes": cb[dom]}


def copy_calibration():
    """1 GiB device copy, three times: a kernel of known HBM traffic (1 GiB read
    + 1 GiB written per launch) and the measured copy bandwidth."""
    import torch

    n = 1 << 27  # doubles
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    y.copy_(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    print(json.dumps({"copy_bytes_per_launch": 2 * n * 8, "launches": 4,
                      "copy_GBps": 3 * 2 * n * 8 / (t1 - t0) / 1e9}))


def realised_dt(report):
    """Mean time step of the evaluated runs of the closure report (simulated
    time span / steps): the fine FP64 run over its whole span from t = 0 and
    each coarse run from its start to its end.  At the seeded near-laminar
    start the coarse dt is twice the fine dt, but in the developed flow both
    grids share the y grid and its wall-normal limit, and their mean dt is
    about equal; throughput ratios use these realised values."""
    out = {"fine": report["fine"]["t_end"] / report["fine"]["steps"]}
    t0 = report["coarse_start_t"]
    for lab in report:
        if lab == "coarse" or lab.startswith("coarse_closure_every"):
            r = report[lab]
            out[lab] = (r["t_reached"] - t0) / r["steps"]
    return out


def learned_leg(args, fine_value, fine_dt):
    """The learned path: the 192 x 192 x 96 coarse grid (modal filter alpha
    36, order 32, as in the closure experiment) with the trained closure
    network (cfd_amd/data/closure_rt395.npz) evaluated every ``every`` steps,
    timed like the FP64 leg from a seeded start, on the FP64 path and on the
    mixed path (keys ending in ``_mixed``).  Simulated time per second against
    the fine FP64 run's = (substeps/s x realised mean dt of the evaluated
    coarse run) / (fine substeps/s x realised mean dt of the evaluated fine
    run), from the closure report of that precision (the mixed one evaluates
    its coarse runs on the mixed path against the same FP64 fine run); the
    report's own wall-clock ratio of the evaluated runs (speedup_vs_fine) is
    quoted beside it.  The dt of the timed substeps (a near-laminar start,
    where the coarse dt is twice the fine dt) is recorded but not used for
    the ratio."""
    from cfd_amd import Channel, closure, cnn, workloads

    pc, _ = workloads.get("rt395_coarse_192x192x96")
    w = np.load(os.path.join(ROOT, "cfd_amd", "data", "closure_rt395.npz"))
    out = {"workload": "rt395_coarse_192x192x96", "grid": [pc.ngalx, pc.ny, pc.ngalz],
           "weights": "cfd_amd/data/closure_rt395.npz",
           "dtype": "bf16 network (fp32 accumulation); f64 solver, or the mixed path (keys ending in _mixed)"}
    net = None
    for prec, sfx in ((0, ""), (1, "_mixed")):
        rname = "profiles/closure_rt395_384x192x192_to_192x192x96%s.json" % sfx
        report = None
        if os.path.exists(os.path.join(ROOT, rname)):
            with open(os.path.join(ROOT, rname)) as fh:
                report = json.load(fh)
        rdt = realised_dt(report) if report else None
        out["realised_mean_dt" + sfx] = rdt
        for every in (0, 1, 2, 4):
            ch = Channel(dataclasses.replace(pc, precision=prec))
            closure.start_fields(ch, pc, seed=0)
            ch.set_mode_filter(36.0, 32, 0.0)  # the coarse runs' modal filter
            if every:
                if net is None:
                    net = cnn.CNN(ch.nyu, pc.ngalz, pc.ngalx)
                    net.set_params(w["params"])
                    sc = np.array(w["scales"], np.float32).copy()
                    if "gain" in w.files:  # the damped correction of the closure experiment
                        sc[3:] *= float(w["gain"])
                    net.set_scales(sc)
                ch.closure_attach(net, every=every)
            for s in range(args.warmup):
                ch.substep(s % 3 + 1)
            ch.sync()
            t0 = time.perf_counter()
            for s in range(args.steps):
                ch.substep((args.warmup + s) % 3 + 1)
            ch.sync()
            t1 = time.perf_counter()
            v = args.steps / (t1 - t0)
            dt = ch.scalars()["dt"]
            key = ("closure_every_%d" % every if every else "no_closure") + sfx
            lab = "coarse_closure_every%d" % every if every else "coarse"
            out[key] = {"substeps_per_s": v, "substeps_per_s_vs_fine_fp64": v / fine_value,
                        "dt_of_timed_substeps": dt, "fine_dt_of_timed_substeps": fine_dt}
            if rdt and lab in rdt:
                out[key]["sim_time_per_s_vs_fine_fp64"] = (v * rdt[lab]) / (fine_value * rdt["fine"])
                out[key]["report_speedup_vs_fine"] = report[lab].get("speedup_vs_fine")
                out[key]["report_drag_error"] = report[lab]["drag"]["drag_error"]
            if every:
                ch.closure_attach(None)
            ch.close()
        if report:
            r = report
            out["accuracy_from_report" + sfx] = {
                "report": rname,
                "held_out_gain": r.get("held_out_gain"),
                **{lab: {k: r[lab].get(k) for k in ("completed", "t_reached", "drag")}
                   for lab in r if lab == "coarse" or lab.startswith("coarse_closure_every")},
                "coarse_closure_fallback": {k: r.get("coarse_closure_fallback", {}).get(k)
                                            for k in ("closure_every", "switch_t", "completed", "t_reached", "drag")},
                "closure_every": r.get("closure_every"), "closure_gain": r.get("closure_gain"),
                "coarse_filter": r.get("coarse_filter"), "gain_scan": r.get("gain_scan")}
    if net is not None:
        net.close()
    return out


def fortran_leg(params, steps, fine_value):
    """The drop-in Fortran driver (cfd_amd/fortran/littleharsh_gpu) on the same
    workload: input.in written in the reference's layout with the driver's
    optional records (tanh grid, no 24pt extras, FP64, fused stepping), field
    files of the seeded projected field (flag_init = 1), `steps` time steps;
    the driver times its own loop steps (those without a field record or a
    statistics sample) and prints their rate."""
    import shutil
    import subprocess
    import tempfile
    from cfd_amd import Channel, ic, inputfile

    exe = os.path.join(ROOT, "cfd_amd", "fortran", "littleharsh_gpu")
    if not os.path.exists(exe):
        return {"skipped": "driver not built"}
    p = inputfile.representable(params)
    d = tempfile.mkdtemp(prefix="cfd_fortran_leg_")
    try:
        for sub in ("in", "out"):
            os.makedirs(os.path.join(d, sub))
        ch = Channel(p)
        g = ch.ygrid()
        f = ic.project(ch, ic.initial_fields(p, g, seed=0, amp=0.1))
        for k in ("u1", "u2", "u3", "p"):
            ch.set_field(k, f[k])
        ch.set_scalars(t=0.0, dt=0.0, mpgx=0.0, QxT=ic.flowrate(f["u1"], g, p.ny))
        fn = os.path.basename(ch.record_out(os.path.join(d, "in") + "/", 0))
        ch.close()
        inputfile.write(d, p, flag_init=1, fnameimb=fn, dirin="./in/", dirout="./out/", nwrite=100000)
        out = subprocess.run([exe, str(steps)], cwd=d, capture_output=True, text=True, timeout=600)
        if out.returncode != 0:
            return {"error": (out.stdout + out.stderr)[-2000:]}
        line = [ln for ln in out.stdout.splitlines() if ln.strip().startswith("loop:")][0].split()
        v = float(line[5].replace("D", "E"))
        return {"substeps_per_s": v, "timed_steps": int(line[1]), "seconds": float(line[3].replace("D", "E")),
                "vs_python_driven": v / fine_value,
                "run": "littleharsh_gpu %d from input.in (flag_init 1, fused cfd_step, FP64)" % steps}
    finally:
        shutil.rmtree(d, ignore_errors=True)


def multiband_leg(args, f, q0, fine_value, nprof, tuning):
    """The 384 x 192 x 192 case with a 192 x 96 central band (multi-band
    resolution), FP64, the launch shapes of the one-resolution run, from the
    same projected field (the short columns' rows between their bands
    dropped)."""
    from cfd_amd import fieldfile, workloads

    pmb, _ = workloads.get("rt395_mb_384x192x192_c192x96")
    ch, v, ms, prof, st = timed(pmb, f, q0, args, nprof, tuning)
    ch.close()
    N = fieldfile.band_table(pmb)
    return {"workload": "rt395_mb_384x192x192_c192x96", "value": v, "unit": "substeps/s", "ms_per_step": ms,
            "speedup_vs_one_resolution": v / fine_value, "central_u_planes": N[11] - N[7], "u_planes": pmb.nyu,
            "central_grid": [pmb.ngalx2, pmb.ngalz2],
            "kernel_ms_per_substep": {c: t / nprof for c, (t, cnt) in prof.items()},
            "finite": bool(np.isfinite(st["u1"]).all()), "dt": st["scalars"]["dt"]}


def ensemble_leg(fine_value, fine_points):
    """16 uneven channels, tuned, 50 RK3 steps (the first eager, 49 by the
    captured graphs, one launch per member kernel); the same members stepped
    by the graphs with batched launches over the member table (the tunings'
    column-sweep shapes, four-step transforms), for comparison."""
    from cfd_amd import ensemble

    rep = ensemble.run(16, 1, 50, autotune=True, log=None)
    bat = ensemble.run(16, 1, 50, log=None, batched=True,
                       tunings=[dict(m["tuning"], fft=0) for m in rep["members"]])
    return {"members": 16, "seed": 1, "steps": 50,
            "member_substeps_per_s": rep["member_substeps_per_s"],
            "point_substeps_per_s_batched": bat["point_substeps_per_s"],
            "batched_over_per_member": bat["point_substeps_per_s"] / rep["point_substeps_per_s"],
            "ensemble_substeps_per_s": rep["ensemble_substeps_per_s"],
            "point_substeps_per_s": rep["point_substeps_per_s"],
            "point_substeps_per_s_single_channel": fine_value * fine_points,
            "points": rep["points"], "tuning_seconds": rep["tuning_seconds"],
            "all_finite": all(m["finite"] for m in rep["members"]),
            "all_healthy": all(m["healthy"] for m in rep["members"]),
            "min_dt_over_first_step_dt": min(m["dt"] / m["dt_after_first_step"] for m in rep["members"]),
            "max_div_over_max_u": max(m["max_div_over_max_u"] for m in rep["members"]),
            "tunings": [m["tuning"] for m in rep["members"]]}


def host_cores():
    """(threads this process may run on, physical cores of the host): the
    affinity mask and the distinct (socket, core) pairs of /proc/cpuinfo."""
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    phys = set()
    try:
        sock = core = None
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("physical id"):
                    sock = line.split(":")[1].strip()
                elif line.startswith("core id"):
                    core = line.split(":")[1].strip()
                elif not line.strip():
                    if core is not None:
                        phys.add((sock, core))
                    sock = core = None
    except OSError:
        pass
    return aff, (len(phys) or None)


def cpu_baseline(args, params, f, q0):
    """The FP64 oracle (oracle/, OpenMP over columns and planes) on the host
    cores: cpu_warmup untimed substeps, then cpu_substeps timed substeps of
    the same workload from the same field, driven operator by operator (the
    reference's call sequence, littleharsh.f90:165-187) so each operator
    class is timed on its own."""
    from oracle.oracle import Oracle, set_threads
    ncores, phys = host_cores()
    set_threads(ncores)
    orc = Oracle(params)
    for key in ("u1", "u2", "u3", "p"):
        orc.set_field(key, f[key])
    orc.set_scalars(t=0.0, dt=0.0, mpgx=0.0, QxT=q0)
    classes = ("RHS0", "nonlinear", "solveU_V", "meanflow_ctU", "solveP", "v_corr")
    acc = dict.fromkeys(classes, 0.0)

    def sub(kRK, clk):
        seq = (("RHS0", lambda: orc.RHS0(kRK)), ("nonlinear", lambda: orc.nonlinear(kRK)),
               ("solveU_V", lambda: (orc.solveU(1, kRK), orc.solveV(kRK), orc.solveU(3, kRK))),
               ("meanflow_ctU", lambda: orc.meanflow_ctU(kRK)), ("solveP", lambda: orc.solveP(kRK)),
               ("v_corr", lambda: orc.v_corr(kRK)))
        for name, fn in seq:
            if name == "meanflow_ctU" and params.flag_ctpress:
                continue
            a = time.perf_counter()
            fn()
            if clk:
                acc[name] += time.perf_counter() - a

    for s in range(args.cpu_warmup):
        sub(s % 3 + 1, False)
    c0 = time.perf_counter()
    for s in range(args.cpu_substeps):
        sub((args.cpu_warmup + s) % 3 + 1, True)
    c1 = time.perf_counter()
    orc.close()
    return {"value": args.cpu_substeps / (c1 - c0), "unit": "substeps/s", "cores": ncores,
            "host_physical_cores": phys, "kind": "port",
            "ms_per_substep_by_class": {k: 1e3 * v / args.cpu_substeps for k, v in acc.items()},
            "sample": "%d warm-up then %d timed RK3 substeps of the same %s workload from the same initial "
                      "field (oracle/, FP64, OpenMP on %d threads, operator by operator)"
                      % (args.cpu_warmup, args.cpu_substeps, args.workload, ncores)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30, help="timed RK3 substeps")
    ap.add_argument("--warmup", type=int, default=6, help="untimed RK3 substeps")
    ap.add_argument("--workload", default="rt395_384x192x192")
    ap.add_argument("--cpu-substeps", type=int, default=30,
                    help="timed oracle substeps for the CPU baseline (0: skip)")
    ap.add_argument("--cpu-warmup", type=int, default=5, help="untimed oracle substeps before them")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision leg")
    ap.add_argument("--no-tune", action="store_true", help="default launch shapes instead of the autotuner's")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-closure leg")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the 16-channel ensemble leg")
    ap.add_argument("--no-fortran", action="store_true", help="skip the Fortran driver leg")
    ap.add_argument("--no-multiband", action="store_true", help="skip the coarser-central-band leg")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device-to-device copy 3 times and print its bandwidth (the reference "
                         "kernel for scaling HBM byte counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs the substep on one GPU; --gpus %d is not supported" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    from cfd_amd import Channel, diagnostics, ic, workloads

    params, _ = workloads.get(args.workload)
    probe = Channel(params)
    g = probe.ygrid()
    f = ic.initial_fields(params, g, seed=0, amp=0.1)
    f = ic.project(probe, f)
    probe.close()
    q0 = ic.flowrate(f["u1"], g, params.ny)
    nprof = 6

    tun64, tsum64 = tuned(params, f, q0, args)
    ch, value, ms_step, prof, st64 = timed(params, f, q0, args, nprof, tun64)
    rl = roofline(ch, prof, 16, args.workload, "")
    sc = st64["scalars"]
    finite = bool(np.isfinite(st64["u1"]).all())
    kms = {c: ms / nprof for c, (ms, cnt) in prof.items()}
    ch.close()

    mixed = None
    if not args.no_mixed:
        pm = dataclasses.replace(params, precision=1)
        tunm, tsumm = tuned(pm, f, q0, args)
        chm, vm, msm, profm, stm = timed(pm, f, q0, args, nprof, tunm)
        mixed = {
            "value": vm, "unit": "substeps/s", "ms_per_step": msm, "speedup_vs_fp64": vm / value,
            "dtype": "f32 transforms/products/nonlinear scratch; f64 fields, solves, projection, flow rate",
            "roofline": roofline(chm, profm, 8, args.workload, "_mixed"),
            "kernel_ms_per_substep": {c: ms / nprof for c, (ms, cnt) in profm.items()},
            "roofline_per_class": roofline_classes(chm, profm, 8, nprof),
            "autotune": tsumm,
            "device_bytes": stm["device_bytes"],
            "err_vs_fp64": dict(diagnostics.compare(stm, st64, g, params.Re),
                                substeps=args.warmup + args.steps + nprof),
        }
        chm.close()

    learned = None if args.no_learned else learned_leg(args, value, sc["dt"])
    ens = None if args.no_ensemble else ensemble_leg(value, params.ngalx * params.ny * params.ngalz)
    fortran = None if args.no_fortran else fortran_leg(params, 20, value)
    multiband = None
    if not args.no_multiband and args.workload == "rt395_384x192x192":
        multiband = multiband_leg(args, f, q0, value, nprof, tun64)

    cpu = cpu_baseline(args, params, f, q0) if args.cpu_substeps > 0 else None

    out = {
        "metric": "RK3 substeps/s at %dx%dx%d FP64" % (params.ngalx, params.ny, params.ngalz),
        "value": value,
        "unit": "substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": args.workload, "grid": [params.ngalx, params.ny, params.ngalz],
                   "dsty": params.dsty, "posts_per_wall": [params.ntilex, params.ntilez], "Re_bulk": params.Re,
                   "precision": "fp64", "parallelism": "single-gpu"},
        "roofline": rl,
        "kernel_ms_per_substep": kms,
        "roofline_per_class": roofline_classes(ch, prof, 16, nprof),
        "autotune": tsum64,
        "device_bytes": st64["device_bytes"],
        "mixed": mixed,
        "learned": learned,
        "ensemble": ens,
        "fortran_driver": fortran,
        "multiband": multiband,
        "cpu_baseline": cpu,
        "sim": {"t": sc["t"], "dt": sc["dt"], "mpgx": sc["mpgx"], "finite": finite},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
