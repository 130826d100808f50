"""Multi-band resolution on the GPU (start.f90:89-111, 1420-1557) against the
CPU FP64 oracle.

The wall bands are transformed at Ngalx x Ngalz, the central band at the
coarser Ngalx2 x Ngalz2 over its long columns; the short columns (the wall
bands' modes outside the central band's N1c x N2c) are closed systems per
wall band with boundary rows at the band interface.  The oracle's multi-band
restatement is pinned on the CPU by per-band dense solves, the projection
invariant, a numpy nonlinear term and closed-form modes; here the HIP path
(fused column sweeps with the short columns' band rows, the central solver's
transforms, the per-band time step) is compared with it through the C ABI:
three substeps from one random state to rel-L2 <= 1e-11 per field, ten RK3
steps from a projected state to <= 1e-10, the time step to 1e-13.

Cases: a rough channel with posts (24 x 48 wall bands, 12 x 24 central band,
tanh grid, ny 33), the same grid smooth, and the grid of the reference's
shipped input.in (Ngalx 24/12, Ngalz 12/6, ny 25, alp 128, bet 256, constant
pressure gradient) with a post that fits its 6 x 6-point tiles (the shipped
post, r = 5.4 grid points centred at (12, 12), stops the reference's own
geometry check, "Circ too close to edge").
"""
import ctypes
import os

import numpy as np
import pytest

from cfd_amd import Channel, Params, _lib, ic
from cfd_amd._lib import CfdError
from oracle.oracle import Oracle

from .helpers import max_rel, random_field, rel_l2

pytestmark = pytest.mark.gpu

MB = Params(ngalx=24, ngalz=48, ny=33, dsty=3, ygrid=1, ytanh_beta=2.0, geometry_type=3, ntilex=4, ntilez=8, r=1.2,
            c_i=3.5, c_k=3.5, stag=0.5, extras24=0, Re=2000.0, alp=1.0, bet=4.0, posth=0.06, CFL=0.5, ngalx2=12,
            ngalz2=24)
MB_SMOOTH = MB.replace(geometry_type=0, dsty=0, posth=0.0)
SHIPPED_GRID = Params(ngalx=24, ngalz=12, ngalx2=12, ngalz2=6, ny=25, dsty=3, dsty1=0, ppp=5, ygrid=0,
                      geometry_type=3, ntilex=4, ntilez=2, extras24=0, flag_ctpress=1, Re=11180.0, alp=128.0,
                      bet=256.0, dyq=10.0, posth=0.089, r=1.0, c_i=3.0, c_k=3.0, stag=0.5, CFL=0.7,
                      mpgx=-2.4201e-3)
CASES = {"mb_rough": MB, "mb_smooth": MB_SMOOTH, "shipped_grid": SHIPPED_GRID}
STATE = ("u1", "u2", "u3", "p", "Nu1", "Nu2", "Nu3", "du1", "du2", "du3")


def bands(p):
    """N(3, 0:3), N(4, 0:3) of the band table (host-only C ABI call)."""
    N = (ctypes.c_int * 20)()
    assert _lib.load().cfd_band_table(ctypes.byref(p.to_c()), N) == 0
    N = list(N)
    return [N[4 * b + 2] for b in range(4)], [N[4 * b + 3] for b in range(4)]


def _grid(name):
    return {"u2": "v", "Nu2": "v", "du2": "v", "p": "p"}.get(name, "u")


def mask_bands(p, name, a):
    """Zero a short column's rows between its wall bands (and, for v, its
    interface faces, where the band systems hold v = 0)."""
    N3, N4 = bands(p)
    g = _grid(name)
    jlo = {"u": -p.dsty, "v": -p.dsty, "p": -p.dsty + 1}[g]
    t1 = {"u": N4[1] + 1, "v": N3[1] + 1, "p": N4[1]}[g]
    b3 = {"u": N4[2], "v": N3[2], "p": N4[2] + 1}[g]
    a = a.copy()
    sh = p.short_columns()
    lo, hi = (t1, b3) if g == "v" else (t1 + 1, b3 - 1)
    a[np.ix_(sh, np.arange(lo - jlo, hi - jlo + 1))] = 0.0
    return a


def _pair(p, seed):
    print("multi-band case", p.ngalx, p.ngalz, p.ngalx2, p.ngalz2, p.ny, "seed", seed, flush=True)
    ch = Channel(p)
    orc = Oracle(p)
    rng = np.random.default_rng(seed)
    for name in STATE:
        f = random_field(p, name, rng, amp=1.0 if name in ("u1", "u2", "u3") else 0.3)
        if name == "u1":
            f[0, :] += 1.0
        if name == "u2":
            f[:, 0] = 0.0
            f[:, -1] = 0.0
        f = mask_bands(p, name, f)
        ch.set_field(name, f)
        orc.set_field(name, f)
    sc = dict(t=0.0, dt=2e-4, dti=5e3, mpgx=p.mpgx if p.flag_ctpress else -1e-3, QxT=orc.flowrate() + 0.01)
    ch.set_scalars(**sc)
    orc.set_scalars(**sc)
    return ch, orc


def plane_sizes(path):
    """(j, nx, nz) of every plane record of a field file (Fortran sequential
    unformatted, 4-byte markers; three header records first)."""
    out = []
    with open(path, "rb") as fh:
        raw = fh.read()
    pos, rec = 0, 0
    while pos < len(raw):
        n = int(np.frombuffer(raw, np.int32, 1, pos)[0])
        if rec >= 3:
            j, _, nx, nz = np.frombuffer(raw, np.int32, 4, pos + 4)
            assert n == 16 + 8 + 8 * nx * nz
            out.append((int(j), int(nx), int(nz)))
        pos += n + 8
        rec += 1
    return out


def _cmp(ch, orc, names, tol):
    for n in names:
        a, b = ch.get_field(n), orc.get_field(n)
        e = rel_l2(a, b)
        assert e <= tol, "%s rel-L2 %.3e > %.1e (max rel %.3e)" % (n, e, tol, max_rel(a, b))


@pytest.mark.parametrize("case", list(CASES))
def test_multiband_substeps(case):
    """Three fused substeps (cfd_substep) from one random state: velocity,
    pressure and divergence to 1e-11, dt and t to 1e-13."""
    p = CASES[case]
    ch, orc = _pair(p, seed=70)
    # the same state on the wall bands' grid in every band: the comparison
    # tells the two resolutions apart by far more than its tolerance
    uni = Oracle(p.replace(ngalx2=0, ngalz2=0))
    for name in STATE:
        uni.set_field(name, orc.get_field(name))
    uni.set_scalars(**orc.scalars())
    for kRK in (1, 2, 3):
        ch.substep(kRK)
        orc.substep(kRK)
        uni.substep(kRK)
    _cmp(ch, orc, ("u1", "u2", "u3", "p"), 1e-11)
    assert rel_l2(ch.get_field("u2"), uni.get_field("u2")) > 1e-6
    d1, d0 = ch.get_field("div"), orc.get_field("div")
    assert np.abs(d1 - d0).max() <= 1e-11 * np.abs(orc.get_field("u1")).max()
    a, b = ch.scalars(), orc.scalars()
    for k in ("t", "dt", "dtc"):
        assert abs(a[k] - b[k]) <= 1e-13 * abs(b[k]), (k, a[k], b[k])


@pytest.mark.parametrize("case", list(CASES))
def test_multiband_ten_steps(case):
    """Ten RK3 steps (cfd_step) from a projected state: rel-L2 <= 1e-10 per
    velocity component, divergence at round-off in every band."""
    p = CASES[case]
    print("multi-band ten steps", case, flush=True)
    ch = Channel(p)
    orc = Oracle(p)
    f = ic.initial_fields(p, ch.ygrid(), seed=3, amp=0.1)
    f = {k: mask_bands(p, k, v) for k, v in f.items()}
    f = ic.project(orc, f)
    for s in (ch, orc):
        for k in ("u1", "u2", "u3", "p"):
            s.set_field(k, f[k])
        s.set_scalars(t=0.0, dt=0.0, mpgx=p.mpgx if p.flag_ctpress else 0.0,
                      QxT=ic.flowrate(f["u1"], ch.ygrid(), p.ny))
    ch.step(10)
    orc.step(10)
    _cmp(ch, orc, ("u1", "u2", "u3"), 1e-10)
    a, b = ch.scalars(), orc.scalars()
    assert abs(a["t"] - b["t"]) <= 1e-12 * b["t"]
    d = ch.get_field("div")
    assert np.abs(d).max() <= 1e-10 * np.abs(ch.get_field("u1")).max()


def test_multiband_per_operator_entry_points_refused():
    """The operator-by-operator entry points exist for one resolution; with a
    coarser central band they return CFD_EINVAL with a message."""
    ch = Channel(MB_SMOOTH)
    for call in (lambda: ch.nonlinear(1), lambda: ch.solveP(1), lambda: ch.RHS0(1), lambda: ch.four_to_phys("u1")):
        with pytest.raises(CfdError, match="central band"):
            call()


# four-step lengths in both resolutions, as the mixed path needs
MB96 = Params(ngalx=96, ngalz=96, ngalx2=48, ngalz2=48, ny=33, dsty=3, ygrid=1, ytanh_beta=2.0, geometry_type=3,
              ntilex=8, ntilez=8, r=2.5, c_i=6.0, c_k=6.0, stag=0.5, extras24=0, Re=2000.0, alp=1.0, bet=4.0,
              posth=0.06, CFL=0.5)


def test_multiband_mixed_path():
    """The mixed path (FP32 transforms, products and nonlinear scratch in
    both resolutions) with a coarser central band: ten RK3 steps from a
    projected state within 1e-5 rel-L2 of the FP64 oracle and of the FP64
    GPU path, per velocity component, flow rate exact to 1e-12 relative,
    divergence at FP64 round-off; FP64 itself to 1e-10."""
    p = MB96
    print("multi-band mixed path", flush=True)
    chs = {prec: Channel(p.replace(precision=prec)) for prec in (0, 1)}
    orc = Oracle(p)
    g = chs[0].ygrid()
    f = ic.initial_fields(p, g, seed=4, amp=0.1)
    f = ic.project(orc, {k: mask_bands(p, k, v) for k, v in f.items()})
    q0 = ic.flowrate(f["u1"], g, p.ny)
    for s in (chs[0], chs[1], orc):
        for k in ("u1", "u2", "u3", "p"):
            s.set_field(k, f[k])
        s.set_scalars(t=0.0, dt=0.0, mpgx=0.0, QxT=q0)
    for s in (chs[0], chs[1], orc):
        s.step(10)
    _cmp(chs[0], orc, ("u1", "u2", "u3"), 1e-10)
    for n in ("u1", "u2", "u3"):
        a, b, c = chs[1].get_field(n), orc.get_field(n), chs[0].get_field(n)
        assert rel_l2(a, b) <= 1e-5 and rel_l2(a, c) <= 1e-5, (n, rel_l2(a, b), rel_l2(a, c))
    assert abs(ic.flowrate(chs[1].get_field("u1"), g, p.ny) - q0) <= 1e-12 * abs(q0)
    d = np.abs(chs[1].get_field("div")).max() / np.abs(chs[1].get_field("u1")).max()
    assert d <= 1e-12, d


def test_multiband_record_out_read_in(tmp_path):
    """record_out writes each plane at its band's size (N(1,b)+2 x N(2,b),
    FOU3D.f90:2172-2345): the central planes at N1c+2 x N2c; read_in puts
    them back, so a field of long columns on the central planes returns
    bit for bit."""
    p = MB
    ch, _ = _pair(p, seed=72)
    ch.substep(1)
    u = {n: ch.get_field(n) for n in ("u1", "u2", "u3", "p")}
    ch.record_out(str(tmp_path), iter=7)
    N3, N4 = bands(p)
    files = sorted(os.listdir(tmp_path))
    assert len(files) == 4, files
    name = [f for f in files if f.startswith("u1_")][0]
    for j, nx, nz in plane_sizes(os.path.join(tmp_path, name)):
        central = N4[1] + 1 <= j <= N4[2]
        assert (nx, nz) == ((p.N1c + 2, p.N2c) if central else (p.N1 + 2, p.N2)), (j, nx, nz)
    ch2 = Channel(p)
    ch2.read_in(str(tmp_path), "u1" + name[2:], flag_init=2)
    sh = p.short_columns()
    for n in ("u1", "u2", "u3", "p"):
        a, b = ch2.get_field(n), u[n]
        g = _grid(n)
        jlo = {"u": -p.dsty, "v": -p.dsty, "p": -p.dsty + 1}[g]
        c0, c1 = {"u": (N4[1] + 1, N4[2]), "v": (N3[1] + 1, N3[2]), "p": (N4[1], N4[2] + 1)}[g]
        rows = np.arange(c0 - jlo, c1 - jlo + 1)
        # the short columns have no modes on the central planes of the files
        b = b.copy()
        b[np.ix_(sh, rows)] = 0.0
        assert np.array_equal(a, b), n


@pytest.mark.parametrize("tuning", [dict(fft=1), dict(seg_rows=5, solve_threads=64),
                                    dict(seg_rows=21, solve_threads=128)])
def test_multiband_launch_shapes(tuning):
    """The Stockham engine and other segment lengths / block sizes of the
    column solves (the short columns' band rows fall in different segments)
    give the oracle's three substeps to 1e-11."""
    p = MB
    print("multi-band launch shapes", tuning, flush=True)
    ch0, orc = _pair(p, seed=74)
    ch = Channel(p, tuning=tuning)
    for name in STATE:
        ch.set_field(name, ch0.get_field(name))
    ch.set_scalars(**ch0.scalars())
    for kRK in (1, 2, 3):
        ch.substep(kRK)
        orc.substep(kRK)
    _cmp(ch, orc, ("u1", "u2", "u3", "p"), 1e-11)


def test_multiband_workload_384x192x192_two_steps():
    """The 384 x 192 x 192 workload with a 192 x 96 central band (17 central
    u planes of 209) at full size: two RK3 steps from a projected start
    against the oracle, rel-L2 <= 1e-10 per velocity component, divergence at
    round-off, the same time and mean pressure gradient."""
    from cfd_amd import workloads
    from oracle.oracle import set_threads
    p, _ = workloads.get("rt395_mb_384x192x192_c192x96")
    print("case rt395_mb_384x192x192_c192x96 two steps vs oracle", flush=True)
    set_threads(min(16, len(os.sched_getaffinity(0))))
    ch = Channel(p)
    orc = Oracle(p)
    g = ch.ygrid()
    f = ic.initial_fields(p, g, seed=0, amp=0.1)
    f = ic.project(orc, {k: mask_bands(p, k, v) for k, v in f.items()})
    q0 = ic.flowrate(f["u1"], g, p.ny)
    for s in (ch, orc):
        for k in ("u1", "u2", "u3", "p"):
            s.set_field(k, f[k])
        s.set_scalars(t=0.0, dt=0.0, mpgx=0.0, QxT=q0)
    ch.step(2)
    orc.step(2)
    _cmp(ch, orc, ("u1", "u2", "u3"), 1e-10)
    sa, sb = ch.scalars(), orc.scalars()
    assert abs(sa["t"] - sb["t"]) <= 1e-12 * sb["t"]
    assert abs(sa["mpgx"] - sb["mpgx"]) <= 1e-9 * abs(sb["mpgx"])
    d = np.abs(ch.get_field("div")).max() / np.abs(ch.get_field("u1")).max()
    assert d <= 1e-12, d


def test_multiband_member_in_an_ensemble():
    """A multi-band channel stepped by the ensemble's captured graphs (its
    central solver's kernels are captured on its stream) beside a
    one-resolution channel: every field and scalar bit-identical to a copy
    stepped eagerly, after a substep of each kRK and a whole step, and both
    against the oracle to 1e-11."""
    from cfd_amd import ensemble
    print("multi-band member in an ensemble", flush=True)
    a, orc = _pair(MB, seed=78)
    b = Channel(MB)
    for name in STATE:
        b.set_field(name, a.get_field(name))
    b.set_scalars(**a.scalars())
    c, _ = _pair(MB_SMOOTH.replace(ngalx2=0, ngalz2=0), seed=79)
    for ch in (a, b, c):
        ch.substep(1)
    orc.substep(1)
    with ensemble.Ensemble([a, c]) as ens:
        for kRK in (2, 3):
            ens.substep(kRK)
            ens.sync()
            b.substep(kRK)
            orc.substep(kRK)
            for n in ("u1", "u2", "u3", "p"):
                assert np.array_equal(a.get_field(n), b.get_field(n)), (kRK, n)
        ens.step(1)
        ens.sync()
    b.step(1)
    orc.step(1)
    for n in ("u1", "u2", "u3", "p"):
        assert np.array_equal(a.get_field(n), b.get_field(n)), n
    assert a.scalars() == b.scalars()
    _cmp(a, orc, ("u1", "u2", "u3", "p"), 1e-11)


def _fold_long(raw, p):
    """recvwrspec (spectra.f90:129-170) for the midband's columns: (Ncol, ny)
    per-column sums of the long columns -> buffSP (ny, N2c/2+1, N1c/2+1);
    kz >= 0 at row kz, kz < 0 at row -kz (N(2,1)+2-k with the wall bands'
    N2), kx > 0 doubled."""
    ny = raw.shape[1]
    r = raw.T.reshape(ny, p.N2, p.Mx)
    mx, h = p.N1c // 2 + 1, p.N2c // 2
    out = np.zeros((ny, h + 1, mx))
    out[:, :h] = r[:, :h, :mx]
    for k in range(p.N2 - h, p.N2):
        out[:, p.N2 - k] += r[:, k, :mx]
    out[:, :, 1:] *= 2.0
    return out


def _check_stats(p, d, tag, dims, ref, istat, it, tol_sp, tol):
    from cfd_amd import fieldfile, statsfile
    from .test_gpu_stats import _scale
    N = fieldfile.band_table(p)
    h, prof, cond = statsfile.read_stats(d + "stats_" + tag, dims["dnx"], dims["dnz"])
    assert h["istat"] == istat and h["iter"] == it
    ngal = list(N)
    for b in (1, 2, 3):
        ngal[4 * b:4 * b + 2] = [p.ngalx2, p.ngalz2] if b == 2 else [p.ngalx, p.ngalz]
    assert h["table"] == ngal
    for s_ in ("spU", "spV", "spW", "spUV", "spP"):
        hs, sp = statsfile.read_spectrum(d + s_ + "_" + tag)
        assert hs["table"] == N and hs["istat"] == istat
        e = _fold_long(ref["spec"][s_], p)
        assert sp.shape == e.shape, s_
        t = tol_sp if tol_sp is not None else (1e-14 if s_ == "spUV" else 0.0)
        assert np.abs(sp - e).max() <= t * np.abs(e).max(), (s_, np.abs(sp - e).max() / np.abs(e).max())
    R = ref["prof"]
    npts = p.ngalx * p.ngalz
    for k, v in prof.items():
        sc = _scale(R, k, npts)
        assert np.abs(v - R[k]).max() <= tol * sc, (k, np.abs(v - R[k]).max() / sc)
    C = {k[:-1]: cb + ct for k, (cb, ct) in ref["cond"].items()}
    for k, v in cond.items():
        e = C[k[:-1]]
        assert v.shape == e.shape, k
        sc = _scale(C, k[:-1], 2 * p.ntilex * p.ntilez)
        assert np.abs(v - e).max() <= tol * sc, (k, np.abs(v - e).max() / sc)


@pytest.mark.parametrize("case", ["mb_rough", "shipped_grid"])
def test_multiband_statistics_against_oracle(tmp_path, case):
    """stats.f90 / spectra.f90 with a coarser central band, written and read
    back.  One sample from identical state: profiles on every plane (the
    central planes from the central band's grid) and the wall bands'
    conditional sums within 1e-12 of their natural scale, spectra of the long
    columns in the central band's N1c/2+1 x N2c/2+1 layout bit for bit (spUV
    to 1e-14); the stats file's Ngal table carries Ngalx2, Ngalz2 in band 2
    and the spectrum files' N table N1c, N2c.  Then, the write having zeroed
    both resolutions' sums, one RK3 step and a second sample (the two sides'
    fields then differ at the 1e-12 level): within 1e-10."""
    from cfd_amd import fieldfile

    p = CASES[case]
    ch, orc = _pair(p, 400)
    try:
        N = fieldfile.band_table(p)
        assert N[8:10] == [p.N1c, p.N2c]
        orc.stats_init(N)
        ch.stats()
        orc.stats()
        t = ch.scalars()["t"]
        d = str(tmp_path) + "/"
        ch.write_stats(d, iter=5)
        tag = "%04dx%04dx%04d_%05d.dat" % (ch.N1, ch.N2, p.ny, int(10.0 * t))
        dims = ch.stats_dims()
        _check_stats(p, d, tag, dims, orc.stats_get(), 1, 5, None, 1e-12)
        ch.step(1)
        orc.step(1)
        orc.stats_reset()
        ch.stats()
        orc.stats()
        t = ch.scalars()["t"]
        d2 = str(tmp_path / "second") + "/"
        os.makedirs(d2)
        ch.write_stats(d2, iter=6)
        tag = "%04dx%04dx%04d_%05d.dat" % (ch.N1, ch.N2, p.ny, int(10.0 * t))
        _check_stats(p, d2, tag, dims, orc.stats_get(), 1, 6, 1e-10, 1e-10)
    finally:
        ch.close()
        orc.close()
