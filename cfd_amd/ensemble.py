"""Ensembles of independent rough channels of uneven grids on one GPU.

The reference runs one channel per MPI job.  Here a set of channels, each a
full solver (its own grid, Reynolds number, posts and fields), is stepped
together: ``Ensemble`` captures one HIP graph per RK substep that runs the
substep of every member on the member's own stream (cfd_ensemble_create),
so one graph launch advances all of them and the members' kernels overlap
on the device.

``members`` draws the uneven set: grids from 96 x 96 x 64 to 256 x 160 x 128
(x, y, z), Re_tau uniform in [180, 395], post diameter D+ and height k+
uniform in [8, 20], 4 x 2 to 16 x 8 posts per wall, a seeded perturbation
per member.  The box is Lx = 2 pi, Lz = 2 pi Ngalz / Ngalx (square cells in
x-z, so that the posts, whose radius is in grid units, are circular).  A post
narrower than 1.2 cells is widened to 1.2 cells, the smallest post that the
5-point forcing cross resolves, and a radius whose forcing stencils
extrapolate (lin_interp, geom_circ_rough_mixed_forcing_24pt.f90:634-738:
weights above 1 in magnitude, e.g. 4.3 at r = 1.507 cells and 450 just above
sqrt(2) for a post centred between grid points, where the forced point's two
fluid neighbours line up with its boundary point) is moved to the nearest
radius whose stencils interpolate: with extrapolating weights the forcing
amplifies the velocity every substep and the run blows up within a few
steps, on the CPU oracle as on the GPU.  ``members`` records the D+ drawn
and the D+ simulated.

``tune`` is the autotuner: per member it times every launch-shape choice
(FFT engine, rows per segment and threads per block of the segmented y
solves) on a scratch copy of the member and keeps the fastest.  Every
choice gives the same solution to rounding.
"""
import ctypes
import dataclasses
import math

import numpy as np

from . import _lib, ic
from .channel import Channel, build_iblist, build_ygrid
from .params import Params

X_SIZES = (96, 128, 144, 192, 256)
Z_SIZES = (64, 72, 96, 128)
Y_RANGE = (96, 160)
ARRAYS = ((4, 2), (8, 4), (16, 8))
R_MIN = 1.2
FFT_CHOICES = (0, 1)
SEG_ROWS = (5, 9, 13, 21)
SOLVE_THREADS = (64, 128, 256)


def _rb(re_tau):
    """Bulk Reynolds number 2h Ub / nu for Re_tau (Dean's correlation)."""
    return 0.5 * (re_tau / 0.09) ** (1.0 / 0.88)


def y_folds(g):
    """True when the y grid folds back: yu(j) < yv(j) < yu(j+1) fails."""
    yu, yv = g["yu"], g["yv"]
    return not (np.all(yu[:-1] < yv) and np.all(yv < yu[1:]))


def _pick_dsty(p):
    """Canopy points below the tips: about a third of posth / (tip spacing),
    the first of the nearby counts whose grid does not fold."""
    g0 = build_ygrid(p.replace(dsty=0))
    tip = g0["yv"][1] - g0["yv"][0]
    d0 = int(max(2, min(24, round(p.posth / tip / 3.0))))
    for d in [d0] + [d0 + s * k for k in range(1, 12) for s in (1, -1)]:
        if d < 1:
            continue
        q = p.replace(dsty=d)
        if not y_folds(build_ygrid(q)):
            return d
    raise ValueError("no canopy point count gives a monotonic y grid for posth = %g" % p.posth)


def _fits(p):
    try:
        build_iblist(p, 0, 0)
        return True
    except _lib.CfdError:
        return False


def ib_max_weight(p):
    """Largest |weight| of the forcing stencils of both grids (0 without
    posts); above 1 the stencil extrapolates."""
    w = 0.0
    for grid in (0, 1):
        _, w3 = build_iblist(p, grid, 0)
        if len(w3):
            w = max(w, float(np.abs(w3[:, 1:]).max()))
    return w


def _stable_radius(p, step=0.005, span=1.0):
    """p with the radius nearest to p.r (not below R_MIN) whose post fits
    its tile and whose forcing stencils interpolate; None when no radius
    within span does."""
    for n in range(int(span / step) + 1):
        for d in ((0.0,) if n == 0 else (n * step, -n * step)):
            r = p.r + d
            if r < R_MIN:
                continue
            q = p.replace(r=r)
            if _fits(q) and ib_max_weight(q) <= 1.0:
                return q
    return None


@dataclasses.dataclass
class Member:
    params: Params
    re_tau: float
    dplus: float       # D+ drawn
    dplus_sim: float   # D+ of the simulated posts (after the 1.2-cell floor)
    kplus: float
    seed: int


def member_params(ngalx, ny, ngalz, re_tau, dplus, kplus, arrays):
    """Params of one member, with the largest post array in ``arrays`` (in
    order of preference) whose tiles hold the post."""
    alp, bet = 1.0, ngalx / ngalz
    dxp = 2.0 * math.pi / alp * re_tau / ngalx  # dx+ (= dz+)
    r = max(0.5 * dplus / dxp, R_MIN)
    base = Params(ngalx=ngalx, ngalz=ngalz, ny=ny, dsty=4, ygrid=1, ytanh_beta=2.0, geometry_type=3,
                  Re=float(round(_rb(re_tau))), alp=alp, bet=bet, posth=kplus / re_tau, r=r, stag=0.5,
                  extras24=0, CFL=0.5, flag_ctpress=0, precision=0)
    base = base.replace(dsty=_pick_dsty(base))
    for ntx, ntz in arrays:
        if ngalx % ntx or ngalz % ntz:
            continue
        dnx, dnz = ngalx // ntx, ngalz // ntz
        p = _stable_radius(base.replace(ntilex=ntx, ntilez=ntz, c_i=(dnx + 1) / 2.0, c_k=(dnz + 1) / 2.0))
        if p is not None:
            return p, 2.0 * p.r * dxp
    raise ValueError("no post array of %s fits a post of r = %.3g cells on %dx%d" % (arrays, r, ngalx, ngalz))


def members(n=16, seed=1):
    """The uneven ensemble: member 0 is 96 x 96 x 64 and member n-1
    256 x 160 x 128 (x, y, z), the others drawn; all draws from one
    generator seeded with ``seed``."""
    rng = np.random.default_rng(seed)
    out = []
    for m in range(n):
        if m == 0:
            nx, ny, nz = X_SIZES[0], Y_RANGE[0], Z_SIZES[0]
        elif m == n - 1:
            nx, ny, nz = X_SIZES[-1], Y_RANGE[1], Z_SIZES[-1]
        else:
            nx = int(rng.choice(X_SIZES))
            nz = int(rng.choice([z for z in Z_SIZES if z <= nx]))
            ny = int(rng.integers(Y_RANGE[0] // 2, Y_RANGE[1] // 2 + 1)) * 2
        re_tau = float(rng.uniform(180.0, 395.0))
        dplus = float(rng.uniform(8.0, 20.0))
        kplus = float(rng.uniform(8.0, 20.0))
        a = int(rng.integers(len(ARRAYS)))
        s = int(rng.integers(1 << 30))
        p, dsim = member_params(nx, ny, nz, re_tau, dplus, kplus, ARRAYS[a::-1])
        out.append(Member(p, re_tau, dplus, dsim, kplus, s))
    return out


def start(ch, member, amp=0.1):
    """Seeded laminar + perturbation field of the member, projected, loaded."""
    p = member.params
    g = ch.ygrid()
    f = ic.project(ch, ic.initial_fields(p, g, seed=member.seed, amp=amp))
    for k in ("u1", "u2", "u3", "p"):
        ch.set_field(k, f[k])
    ch.set_scalars(t=0.0, dt=0.0, mpgx=0.0, QxT=ic.flowrate(f["u1"], g, p.ny))
    return f


def _gpu_ms_per_substep(ch, nsteps):
    ch.sync()
    ch.profile(True)
    ch.step(nsteps)
    ch.sync()
    prof = ch.profile_read()
    ch.profile(False)
    return sum(m for m, _ in prof.values()) / (3.0 * nsteps)


def tune(member, nsteps=2, repeats=2, ffts=None):
    """Time every launch-shape choice for ``member`` on scratch solvers (one
    per FFT engine) started from the member's initial field; kernel time per
    substep from the solvers' event timers.  Returns (best tuning, table of
    (tuning, ms per substep))."""
    return tune_params(member.params, lambda ch: start(ch, member), nsteps, repeats, ffts)


def tune_params(params, load, nsteps=2, repeats=2, ffts=None):
    """The autotuner for any solver: ``load(ch)`` puts the starting state
    into a scratch solver.  The mixed path (precision 1) runs the four-step
    engine only, so only its column-solve shapes are tuned; ``ffts`` limits
    the FFT engines tried (the batched ensemble kernels are four-step)."""
    table = []
    for fft in (ffts if ffts is not None else FFT_CHOICES if params.precision == 0 else (0,)):
        ch = Channel(params, tuning=dict(fft=fft))
        try:
            load(ch)
            ch.step(1)  # dt from the CFL condition, first-use set-up
            for sr in SEG_ROWS:
                for th in SOLVE_THREADS:
                    ch.set_tuning(seg_rows=sr, solve_threads=th)
                    ms = min(_gpu_ms_per_substep(ch, nsteps) for _ in range(repeats))
                    table.append((dict(fft=fft, seg_rows=sr, solve_threads=th), ms))
        finally:
            ch.close()
    best = min(table, key=lambda e: e[1])[0]
    return best, table


class Ensemble:
    """One HIP graph per RK substep over the members' substeps.

    Every channel must have run a substep (``step(1)``) before, with
    profiling off, no closure attached and no statistics request pending.
    The channels stay usable: call ``sync`` before reading or driving one
    directly.  A member retuned after the capture keeps its captured shapes
    here.

    ``batched``: the nonlinear transforms and the two column sweeps of all
    members run as one launch per kernel variant over a member table
    (cfd_ensemble_create_batched) instead of one launch per member; needs
    single-resolution members on the four-step transforms."""

    def __init__(self, channels, batched=False):
        self._lib = _lib.load()
        self.channels = list(channels)
        self.batched = bool(batched)
        hs = (ctypes.c_void_p * len(self.channels))(*[c._h.value for c in self.channels])
        e = ctypes.c_void_p()
        create = self._lib.cfd_ensemble_create_batched if self.batched else self._lib.cfd_ensemble_create
        rc = create(hs, len(self.channels), ctypes.byref(e))
        if rc:
            raise _lib.CfdError("cfd_amd error %d: %s" % (rc, self._lib.cfd_ensemble_last_error(None).decode()))
        self._e = e

    def _call(self, name, *args):
        if self._e is None:
            raise _lib.CfdError("ensemble is closed")
        rc = getattr(self._lib, name)(self._e, *args)
        if rc:
            raise _lib.CfdError("cfd_amd error %d: %s" % (rc, self._lib.cfd_ensemble_last_error(self._e).decode()))

    def substep(self, kRK):
        self._call("cfd_ensemble_substep", int(kRK))

    def step(self, nsteps=1):
        self._call("cfd_ensemble_step", int(nsteps))

    def sync(self):
        self._call("cfd_ensemble_sync")

    def close(self):
        if getattr(self, "_e", None) is not None:
            self._lib.cfd_ensemble_destroy(self._e)
            self._e = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def run(n=16, seed=1, steps=50, autotune=True, log=print, batched=False, tunings=None):
    """The ensemble run: members drawn, each tuned (or left at the default
    shapes; ``tunings``: given ones), started, its first step eager (the
    capture needs one substep run on every member), the other ``steps - 1``
    steps by the captured graphs, batched (one launch per kernel variant over
    the members) or one launch per member.  Returns a report: members,
    tunings, throughput, final state per member."""
    import time
    from . import diagnostics
    ms = members(n, seed)
    rows = []
    t0 = time.perf_counter()
    for i, m in enumerate(ms):
        if tunings is not None:
            rows.append(dict(tuning=tunings[i], ms_default=None, ms_tuned=None))
            continue
        if autotune:
            best, table = tune(m, ffts=(0,) if batched else None)
            dflt = [x for t, x in table if t == dict(fft=0, seg_rows=13, solve_threads=256)][0]
            bms = min(x for _, x in table)
        else:
            best, dflt, bms = None, None, None
        rows.append(dict(tuning=best, ms_default=dflt, ms_tuned=bms))
        if log and autotune:
            log("member %2d %dx%dx%d tuned %s: %.4f ms/substep (default shapes %.4f)"
                % (i, m.params.ngalx, m.params.ny, m.params.ngalz, best, bms, dflt))
    t_tune = time.perf_counter() - t0
    chs = []
    try:
        for m, r in zip(ms, rows):
            ch = Channel(m.params, tuning=r["tuning"])
            start(ch, m)
            chs.append(ch)
        for ch in chs:
            ch.step(1)
        for ch in chs:
            ch.sync()
        dt1 = [ch.scalars()["dt"] for ch in chs]
        with Ensemble(chs, batched=batched) as ens:
            t1 = time.perf_counter()
            ens.step(steps - 1)
            ens.sync()
            T = time.perf_counter() - t1
        out = []
        for m, ch, r, d1 in zip(ms, chs, rows, dt1):
            p = m.params
            u1 = ch.get_field("u1")
            sc = ch.scalars()
            div = np.abs(ch.get_field("div")).max()
            bw, tw = diagnostics.wall_shear(u1, ch.ygrid(), p.Re)
            out.append(dict(grid=[p.ngalx, p.ny, p.ngalz], dsty=p.dsty, re_tau=m.re_tau, Re_bulk=p.Re,
                            posts_per_wall=[p.ntilex, p.ntilez], dplus=m.dplus, dplus_sim=m.dplus_sim,
                            kplus=m.kplus, r_cells=p.r, Lz=2 * math.pi / p.bet, seed=m.seed, **r,
                            t=sc["t"], dt=sc["dt"], mpgx=sc["mpgx"], wall_shear=[bw, tw],
                            max_div_over_max_u=float(div / np.abs(u1).max()),
                            device_bytes=ch.device_bytes(),
                            finite=bool(np.isfinite(u1).all()), dt_after_first_step=d1,
                            # a run that blows up drives dt towards zero
                            healthy=bool(np.isfinite(u1).all() and 0.2 * d1 <= sc["dt"] <= 5.0 * d1)))
    finally:
        for ch in chs:
            ch.close()
    pts = sum(m.params.ngalx * m.params.ny * m.params.ngalz for m in ms)
    sub = n * 3 * (steps - 1)
    return dict(members=out, steps=steps, graph_steps=steps - 1, batched=batched, seconds=T, tuning_seconds=t_tune,
                member_substeps_per_s=sub / T, ensemble_substeps_per_s=3 * (steps - 1) / T,
                point_substeps_per_s=pts * 3 * (steps - 1) / T, points=pts, autotune=autotune,
                device_bytes=sum(m["device_bytes"] for m in out))


def main(argv=None):
    import argparse
    import json
    ap = argparse.ArgumentParser(description="16 uneven rough channels stepped by captured HIP graphs")
    ap.add_argument("--members", type=int, default=16)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--batched", action="store_true",
                    help="batched launches over the member table instead of one launch per member kernel")
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    rep = run(a.members, a.seed, a.steps, autotune=not a.no_tune, batched=a.batched)
    s = json.dumps(rep, indent=1)
    if a.out:
        with open(a.out, "w") as fh:
            fh.write(s + "\n")
    print(json.dumps({k: v for k, v in rep.items() if k != "members"}))


if __name__ == "__main__":
    main()
