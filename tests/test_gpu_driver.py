"""The Fortran driver (cfd_amd/fortran/littleharsh_gpu) reads input.in in
the reference layout and calls the hot path through the ISO_C_BINDING
interface; its result must equal the same run driven from Python."""
import dataclasses
import os
import shutil
import subprocess

import numpy as np
import pytest

from cfd_amd import Channel, Params

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FDIR = os.path.join(os.path.dirname(HERE), "cfd_amd", "fortran")
EXE = os.path.join(FDIR, "littleharsh_gpu")


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran driver not built")
@pytest.mark.parametrize("fused", [1, 0])
def test_fortran_driver_matches_python(tmp_path, fused):
    """The example input.in (the reference's layout, parabolic start, constant
    pressure gradient) for 3 steps, stepped by cfd_step (fused = 1, the
    default when the optional record is absent) or by the reference's call
    sequence of entry points (fused = 0), against the Python run."""
    for f in ("input.in", "geom_circ_rough.txt"):
        shutil.copy(os.path.join(FDIR, f), tmp_path / f)
    if not fused:
        with open(tmp_path / "input.in", "a") as fh:
            fh.write("fused =         0 ! one entry point per subroutine\n")
    print("fortran driver, 3 steps, fused", fused, flush=True)
    out = subprocess.run([EXE, "3"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    line = [l for l in out.stdout.splitlines() if "centreline" in l][0]
    fort = float(line.split()[-1].replace("D", "E"))

    p = Params(ngalx=48, ngalz=48, ny=25, dsty=4, dsty1=0, ppp=5, dyq=10.0, ygrid=0, geometry_type=3, ntilex=2,
               ntilez=2, extras24=1, flag_ctpress=1, Re=2800.0, alp=1.0, bet=2.0, posth=0.089, posth1=0.0,
               CFL=0.5, mpgx=-1.0714e-3, r=5.43723, c_i=12.0, c_k=12.0, stag=0.5)
    ch = Channel(p)
    g = ch.ygrid()
    u1 = np.zeros(ch.shape("u1"), complex)
    u2 = np.zeros(ch.shape("u2"), complex)
    u3 = np.zeros(ch.shape("u3"), complex)
    pp = np.zeros(ch.shape("p"), complex)
    u1[0] = 0.5 * p.Re * p.mpgx * (g["yu"] ** 2 - 1.0)
    u3[0] = 1e-5
    u2[0] = 1e-6
    pp[0] = 1.0
    for k, v in (("u1", u1), ("u2", u2), ("u3", u3), ("p", pp)):
        ch.set_field(k, v)
    yv = g["yv"]
    q = 0.0
    for j in range(1, p.ny):
        q = q + (yv[j + p.dsty] - yv[j - 1 + p.dsty]) * u1[0, j + p.dsty].real
    ch.set_scalars(mpgx=p.mpgx, QxT=q, Qx=q)
    ch.step(3)
    py = ch.get_field("u1")[0, ch.nyu // 2].real
    # the fused sweeps solve the same systems by substructuring: the two
    # call sequences agree to round-off; the fused run is the Python run's,
    # up to the 16 digits the driver prints
    assert abs(py - fort) <= (1e-15 if fused else 1e-13) * abs(py), (py, fort)
    assert any(l.strip().startswith("loop:") for l in out.stdout.splitlines()), out.stdout


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran driver not built")
def test_fortran_driver_restart(tmp_path):
    """3 steps, then a second run from the files it wrote (flag_init = 2) for 2
    more: the centreline value equals an uninterrupted 5-step run (prescribed
    pressure gradient in this input.in), and the resumed run starts at iter 3."""
    full, part = tmp_path / "full", tmp_path / "part"
    for d in (full, part):
        d.mkdir()
        for f in ("input.in", "geom_circ_rough.txt"):
            shutil.copy(os.path.join(FDIR, f), d / f)
    print("fortran driver, 5 steps", flush=True)
    out = subprocess.run([EXE, "5"], cwd=full, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    ref = [l for l in out.stdout.splitlines() if "centreline" in l][0].split()[-1]
    print("fortran driver, 3 steps", flush=True)
    out = subprocess.run([EXE, "3"], cwd=part, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    files = sorted(f for f in os.listdir(part) if f.startswith("u1_") and f.endswith(".dat"))
    assert files, "no field file written"
    lines = open(part / "input.in").read().splitlines()
    for i, l in enumerate(lines):
        if l.startswith("flgin"):
            lines[i] = "flgin =         2 ! continue from field files"
        if l.startswith("fnmeb"):
            lines[i] = "fnmeb =" + files[-1]
    open(part / "input.in", "w").write("\n".join(lines) + "\n")
    print("fortran driver, restart for 2 steps", flush=True)
    out = subprocess.run([EXE, "2"], cwd=part, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    got = [l for l in out.stdout.splitlines() if "centreline" in l][0].split()[-1]
    assert got == ref, (got, ref)
    assert any(" iter       5" in l for l in out.stdout.splitlines()), out.stdout


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran driver not built")
def test_fortran_driver_statistics(tmp_path):
    """Smooth walls (a band table with wall bands), nwrite = 4, 9 steps: the
    driver samples every nstat = min(20, nwrite) = 4 steps and writes the
    statistics with each record (iter 0, 4, 8), each write zeroing the sums.
    Files of equal int(10 t) overwrite each other, as the reference's do, so
    only the surviving ones are checked: iter 0 holds no sample, iter 4 and
    8 one each, and a statistics file holds the sample of the state that the
    record of the same iteration holds: a Channel loaded from those field
    files and sampled once writes the same sums bit for bit."""
    from cfd_amd import fieldfile, statsfile
    for f in ("input.in", "geom_circ_rough.txt"):
        shutil.copy(os.path.join(FDIR, f), tmp_path / f)
    lines = open(tmp_path / "input.in").read().splitlines()
    for i, l in enumerate(lines):
        if l.startswith("type "):
            lines[i] = "type  =         0 ! smooth"
        if l.startswith("nwrite"):
            lines[i] = "nwrite=         4 ! time steps between outputs"
    open(tmp_path / "input.in", "w").write("\n".join(lines) + "\n")
    print("fortran driver, smooth, 9 steps with statistics", flush=True)
    out = subprocess.run([EXE, "9"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "statistics off" not in out.stdout
    p = Params(ngalx=48, ngalz=48, ny=25, dsty=4, dsty1=0, ppp=5, dyq=10.0, ygrid=0, geometry_type=0, ntilex=2,
               ntilez=2, extras24=1, flag_ctpress=1, Re=2800.0, alp=1.0, bet=2.0, posth=0.089, posth1=0.0,
               CFL=0.5, mpgx=-1.0714e-3, r=5.43723, c_i=12.0, c_k=12.0, stag=0.5)
    stats = sorted(f for f in os.listdir(tmp_path) if f.startswith("stats_"))
    assert stats
    ch = Channel(p)
    try:
        dims = ch.stats_dims()
        checked = 0
        for f in stats:
            h, prof, cond = statsfile.read_stats(str(tmp_path / f), dims["dnx"], dims["dnz"])
            assert h["iter"] in (0, 4, 8) and h["istat"] == (0 if h["iter"] == 0 else 1), (f, h["iter"], h["istat"])
            for s in ("spU", "spV", "spW", "spUV", "spP"):
                assert os.path.exists(tmp_path / (s + "_" + f[len("stats_"):]))
            tag = f[len("stats_"):]
            u1 = "u1_" + tag[:15] + "t" + tag[15:]
            hh = fieldfile.read(str(tmp_path / u1), ch.nyu, ch.N2, ch.Mx, p.ngalz)[0]
            if h["istat"] != 1 or hh["iter"] != h["iter"]:
                continue  # the record of this name was overwritten by a later one
            for name, ny_g in (("u1", ch.nyu), ("u2", ch.nyv), ("u3", ch.nyu), ("p", ch.nyp)):
                hh, planes, mm = fieldfile.read(str(tmp_path / (name + u1[2:])), ny_g, ch.N2, ch.Mx, p.ngalz)
                assert mm == 0 and hh["iter"] == h["iter"]
                ch.set_field(name, planes.reshape(ny_g, ch.Ncol).T)
            ch.set_scalars(t=hh["t"], mpgx=hh["mpgx"])
            ch.stats()
            d = str(tmp_path / ("py%d" % h["iter"])) + "/"
            os.makedirs(d)
            ch.write_stats(d, iter=h["iter"])
            hp, prof_p, cond_p = statsfile.read_stats(d + "stats_" + tag, dims["dnx"], dims["dnz"])
            for k in prof:
                assert (prof[k] == prof_p[k]).all(), k
            for k in cond:
                assert (cond[k] == cond_p[k]).all(), k
            # the driver's header carries t after the step's time-step update
            assert h["t"] > hh["t"] and h["mpgx"] == hh["mpgx"]
            checked += 1
        assert checked >= 1, stats
    finally:
        ch.close()


def _shipped_params():
    return Params(ngalx=48, ngalz=48, ny=25, dsty=4, dsty1=0, ppp=5, dyq=10.0, ygrid=0, geometry_type=3, ntilex=2,
                  ntilez=2, extras24=1, flag_ctpress=1, Re=2800.0, alp=1.0, bet=2.0, posth=0.089, posth1=0.0,
                  CFL=0.5, mpgx=-1.0714e-3, r=5.43723, c_i=12.0, c_k=12.0, stag=0.5)


def test_hist_record_against_numpy():
    """cfd_hist: err = max |div| over the p planes, Qx = the tip-to-tip flow
    rate of the (0,0) u1 column (and stored as the Qx scalar), Umax its largest
    real part between the tips, Uslip its value interpolated to yv(0)."""
    from cfd_amd import ic
    p = _shipped_params()
    ch = Channel(p)
    g = ch.ygrid()
    f = ic.project(ch, ic.initial_fields(p, g, seed=5, amp=0.1))
    for k in ("u1", "u2", "u3", "p"):
        ch.set_field(k, f[k])
    ch.set_scalars(t=0.0, dt=0.0, mpgx=p.mpgx, QxT=ic.flowrate(f["u1"], g, p.ny))
    ch.step(2)
    h = ch.hist()
    sc = ch.scalars()
    u = ch.get_field("u1")[0].real
    yu, yv, d = g["yu"], g["yv"], p.dsty
    q = 0.0
    for j in range(1, p.ny):
        q = q + (yv[j + d] - yv[j - 1 + d]) * u[j + d]
    assert h["Qx"] == q and ch.scalars()["Qx"] == q
    assert h["Umax"] == max(0.0, u[d:d + p.ny + 1].max())
    uslip = (u[1 + d] * (yv[d] - yu[d]) + u[d] * (yu[1 + d] - yv[d])) / (yu[1 + d] - yu[d])
    assert abs(h["Uslip"] - uslip) <= 1e-15 * abs(uslip)
    # |div| of a complex mode on the device and in numpy: the two moduli may
    # differ in the last bit, so two units in the last place
    err = np.abs(ch.get_field("div")).max()
    assert abs(h["err"] - err) <= 4.5e-16 * err and err > 0
    for k in ("t", "dt", "dtc", "mpgx", "dgx", "QxT"):
        assert h[k] == sc[k], k
    ch.close()


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran driver not built")
def test_fortran_driver_history_file(tmp_path):
    """nwrite = 2, 5 steps of the shipped input (prescribed pressure gradient):
    hist_<N1>x<N2>x<ny>.dat holds the header (t = 0, Re, alp, bet, mpgx, nband)
    and the band table, then one record per field record (iter 0, 2, 4 and the
    final 5): flag_ctpress, iter, t, dtv, dtc, dt, err, Qx, mpgx, Umax, Uslip,
    equal to the same run driven from Python; a second run appends."""
    import struct
    from cfd_amd import fieldfile, statsfile
    for f in ("input.in", "geom_circ_rough.txt"):
        shutil.copy(os.path.join(FDIR, f), tmp_path / f)
    lines = open(tmp_path / "input.in").read().splitlines()
    for i, l in enumerate(lines):
        if l.startswith("nwrite"):
            lines[i] = "nwrite=         2 ! time steps between outputs"
    open(tmp_path / "input.in", "w").write("\n".join(lines) + "\n")
    print("fortran driver, 5 steps with history", flush=True)
    out = subprocess.run([EXE, "5"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    p = _shipped_params()
    ch = Channel(p)
    name = "hist_%04dx%04dx%04d.dat" % (ch.N1, ch.N2, p.ny)
    recs = statsfile.records(str(tmp_path / name))
    assert len(recs) == 2 + 4
    t0, Re, alp, bet, mpgx = struct.unpack_from("<5d", recs[0], 0)
    (nband,) = struct.unpack_from("<i", recs[0], 40)
    assert len(recs[0]) == 44 and (t0, Re, alp, bet, nband) == (0.0, p.Re, p.alp, p.bet, 3)
    assert abs(mpgx - p.mpgx) <= 1e-15 * abs(p.mpgx)
    assert list(struct.unpack("<20i", recs[1])) == list(fieldfile.band_table(p))
    rows = []
    for r in recs[2:]:
        assert len(r) == 8 + 9 * 8
        fl, it = struct.unpack_from("<2i", r, 0)
        rows.append((fl, it) + struct.unpack_from("<9d", r, 8))
    assert [r[1] for r in rows] == [0, 2, 4, 5] and all(r[0] == 1 for r in rows)
    # the same run from Python: the initial condition of flag_init = 3
    g = ch.ygrid()
    u1 = np.zeros(ch.shape("u1"), complex)
    u1[0] = 0.5 * p.Re * p.mpgx * (g["yu"] ** 2 - 1.0)
    u2 = np.zeros(ch.shape("u2"), complex)
    u2[0] = 1e-6
    u3 = np.zeros(ch.shape("u3"), complex)
    u3[0] = 1e-5
    pp = np.zeros(ch.shape("p"), complex)
    pp[0] = 1.0
    for k, v in (("u1", u1), ("u2", u2), ("u3", u3), ("p", pp)):
        ch.set_field(k, v)
    q = 0.0
    for j in range(1, p.ny):
        q = q + (g["yv"][j + p.dsty] - g["yv"][j - 1 + p.dsty]) * u1[0, j + p.dsty].real
    ch.set_scalars(mpgx=p.mpgx, QxT=q, Qx=q)
    it = 0
    for row in rows:
        ch.step(row[1] - it)
        it = row[1]
        h = ch.hist()
        want = (h["t"], h["dtv"], h["dtc"], h["dt"], h["err"], h["Qx"], h["mpgx"], h["Umax"], h["Uslip"])
        for k, (a, b) in enumerate(zip(row[2:], want)):
            if k == 4:  # err: a divergence at round-off, path-dependent in its digits
                assert 0 <= a <= 1e-10 * abs(h["Umax"]) and 0 <= b <= 1e-10 * abs(h["Umax"]), (row[1], a, b)
            else:
                assert abs(a - b) <= 1e-12 * abs(b), (row[1], k, a, b)
    # a second run appends to the same file
    out = subprocess.run([EXE, "1"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert len(statsfile.records(str(tmp_path / name))) == 6 + 2
    ch.close()


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran driver not built")
def test_fortran_driver_lastplane(tmp_path):
    """CFD_LASTPLANE_NP=3 makes the driver run the as-shipped nonlinear term of
    a 3-rank run (cfd_set_compat_lastplane).  From a perturbed field (smooth
    walls, whose band table has a partition; flag_init = 2 from field files
    written by record_out), 3 steps: the centreline value equals the same run
    driven from Python with the replay, and differs from the per-plane run."""
    from cfd_amd import ic
    p = dataclasses.replace(_shipped_params(), geometry_type=0)
    w = Channel(p)
    g = w.ygrid()
    f = ic.project(w, ic.initial_fields(p, g, seed=0, amp=0.1))
    for k in ("u1", "u2", "u3", "p"):
        w.set_field(k, f[k])
    w.set_scalars(t=0.0, dt=0.0, mpgx=p.mpgx, QxT=ic.flowrate(f["u1"], g, p.ny))
    u1 = os.path.basename(w.record_out(str(tmp_path) + "/", iter=0))
    w.close()
    for fn in ("input.in", "geom_circ_rough.txt"):
        shutil.copy(os.path.join(FDIR, fn), tmp_path / fn)
    lines = open(tmp_path / "input.in").read().splitlines()
    for i, l in enumerate(lines):
        if l.startswith("type "):
            lines[i] = "type  =         0 ! smooth"
        if l.startswith("flgin"):
            lines[i] = "flgin =         2 ! continue from field files"
        if l.startswith("fnmeb"):
            lines[i] = "fnmeb =" + u1
    open(tmp_path / "input.in", "w").write("\n".join(lines) + "\n")
    vals = {}
    for tag, env in (("replay", "3"), ("per_plane", None)):
        run_dir = tmp_path / tag
        run_dir.mkdir()
        for fn in os.listdir(tmp_path):
            if os.path.isfile(tmp_path / fn):
                shutil.copy(tmp_path / fn, run_dir / fn)
        e = dict(os.environ)
        e.pop("CFD_LASTPLANE_NP", None)
        if env:
            e["CFD_LASTPLANE_NP"] = env
        print("fortran driver, 3 steps,", tag, flush=True)
        out = subprocess.run([EXE, "3"], cwd=run_dir, capture_output=True, text=True, timeout=120, env=e)
        assert out.returncode == 0, out.stdout + out.stderr
        assert ("last-plane replay" in out.stdout) == (env is not None)
        line = [l for l in out.stdout.splitlines() if "centreline" in l][0]
        vals[tag] = float(line.split()[-1].replace("D", "E"))
    ch = Channel(p)
    ch.set_scalars(t=0.0, dt=0.0, dti=0.0, mpgx=p.mpgx, dgx=0.0, QxT=0.0, Qx=0.0)
    ch.read_in(str(tmp_path) + "/", u1, flag_init=2)
    ch.set_compat_lastplane(3)
    ch.step(3)
    py = ch.get_field("u1")[0, ch.nyu // 2].real
    ch.close()
    assert abs(py - vals["replay"]) <= 1e-13 * abs(py), (py, vals)
    assert vals["replay"] != vals["per_plane"], vals


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran driver not built")
@pytest.mark.parametrize("precision", [0, 1])
def test_fortran_driver_rt395_384x192x192(tmp_path, precision):
    """The 384 x 192 x 192 workload (tanh grid, 16 x 8 posts without the 24pt
    extras, FP64 or mixed) launched from input.in through the driver's
    optional records, started from field files of the seeded projected field
    (flag_init = 1): after 3 time steps the driver's final field files equal,
    byte for byte, those of the same run driven from Python (read_in of the
    same files, cfd_step, record_out)."""
    from cfd_amd import ic, inputfile, workloads
    p0, _ = workloads.get("rt395_384x192x192")
    p = inputfile.representable(dataclasses.replace(p0, precision=precision))
    ind, outd, pyd = tmp_path / "in", tmp_path / "out", tmp_path / "py"
    for d in (ind, outd, pyd):
        d.mkdir()
    print("case rt395_384x192x192 driver precision", precision, flush=True)
    ch = Channel(p)
    g = ch.ygrid()
    f = ic.project(ch, ic.initial_fields(p, g, seed=0, amp=0.1))
    for k in ("u1", "u2", "u3", "p"):
        ch.set_field(k, f[k])
    ch.set_scalars(t=0.0, dt=0.0, mpgx=0.0, QxT=ic.flowrate(f["u1"], g, p.ny))
    fn = os.path.basename(ch.record_out(str(ind) + "/", 0))
    ch.close()
    inputfile.write(str(tmp_path), p, flag_init=1, fnameimb=fn, dirin="./in/", dirout="./out/", nwrite=1000)
    out = subprocess.run([EXE, "3"], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert any(l.strip().startswith("loop:") for l in out.stdout.splitlines()), out.stdout
    py = Channel(p)
    py.read_in(str(ind) + "/", fn, 1)
    py.step(3)
    last = os.path.basename(py.record_out(str(pyd) + "/", 3))
    py.close()
    for name in ("u1", "u2", "u3", "p"):
        a = (outd / (name + last[2:])).read_bytes()
        b = (pyd / (name + last[2:])).read_bytes()
        assert a == b, name


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran driver not built")
def test_fortran_driver_multiband_shipped_grid(tmp_path):
    """The grid of the reference's shipped input.in with its coarser central
    band (Ngalx 24/12, Ngalz 12/6, ny 25, alp 128, bet 256, constant pressure
    gradient; a post that fits its tiles), started from field files of a
    seeded projected field (flag_init = 1; the central planes stored at the
    central band's size): after 3 time steps the driver's field files equal,
    byte for byte, those of the same run driven from Python."""
    from cfd_amd import ic, inputfile
    from oracle.oracle import Oracle
    from .test_gpu_multiband import SHIPPED_GRID, mask_bands
    p = inputfile.representable(SHIPPED_GRID)
    ind, outd, pyd = tmp_path / "in", tmp_path / "out", tmp_path / "py"
    for d in (ind, outd, pyd):
        d.mkdir()
    print("case multi-band shipped grid driver", flush=True)
    ch = Channel(p)
    g = ch.ygrid()
    f = ic.initial_fields(p, g, seed=1, amp=0.1)
    f = ic.project(Oracle(p), {k: mask_bands(p, k, v) for k, v in f.items()})
    for k in ("u1", "u2", "u3", "p"):
        ch.set_field(k, f[k])
    ch.set_scalars(t=0.0, dt=0.0, mpgx=p.mpgx, QxT=ic.flowrate(f["u1"], g, p.ny))
    fn = os.path.basename(ch.record_out(str(ind) + "/", 0))
    ch.close()
    inputfile.write(str(tmp_path), p, flag_init=1, fnameimb=fn, dirin="./in/", dirout="./out/", nwrite=1000)
    out = subprocess.run([EXE, "3"], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    py = Channel(p)
    py.read_in(str(ind) + "/", fn, 1)
    py.step(3)
    last = os.path.basename(py.record_out(str(pyd) + "/", 3))
    py.close()
    for name in ("u1", "u2", "u3", "p"):
        a = (outd / (name + last[2:])).read_bytes()
        b = (pyd / (name + last[2:])).read_bytes()
        assert a == b, name


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran driver not built")
def test_fortran_driver_multiband_statistics(tmp_path):
    """The shipped grid's coarser central band with statistics: nwrite = 2,
    3 steps (a write of empty sums at iter 0, a sample and a write at iter
    2).  Every name collides (int(10 t) = 0), so the surviving files are those
    of iter 2; they equal, byte for byte, the statistics and spectrum files of
    the same requests and steps driven from Python, hold one sample and carry
    Ngalx2, Ngalz2 in band 2 of the Ngal table."""
    from cfd_amd import ic, inputfile, statsfile
    from oracle.oracle import Oracle
    from .test_gpu_multiband import SHIPPED_GRID, mask_bands
    p = inputfile.representable(SHIPPED_GRID)
    ind, outd, pyd = tmp_path / "in", tmp_path / "out", tmp_path / "py"
    for d in (ind, outd, pyd):
        d.mkdir()
    print("case multi-band shipped grid driver statistics", flush=True)
    ch = Channel(p)
    g = ch.ygrid()
    f = ic.initial_fields(p, g, seed=2, amp=0.1)
    f = ic.project(Oracle(p), {k: mask_bands(p, k, v) for k, v in f.items()})
    for k in ("u1", "u2", "u3", "p"):
        ch.set_field(k, f[k])
    ch.set_scalars(t=0.0, dt=0.0, mpgx=p.mpgx, QxT=ic.flowrate(f["u1"], g, p.ny))
    fn = os.path.basename(ch.record_out(str(ind) + "/", 0))
    ch.close()
    inputfile.write(str(tmp_path), p, flag_init=1, fnameimb=fn, dirin="./in/", dirout="./out/", nwrite=2)
    out = subprocess.run([EXE, "3"], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "statistics off" not in out.stdout
    py = Channel(p)
    try:
        py.read_in(str(ind) + "/", fn, 1)
        for it in range(3):
            if it == 0:
                py.stats_request(sample=False, write_dir=str(pyd) + "/", iter=0)
            if it == 2:
                py.stats_request(sample=True, write_dir=str(pyd) + "/", iter=2)
            py.step(1)
        dims = py.stats_dims()
    finally:
        py.close()
    names = sorted(x for x in os.listdir(outd) if x.startswith(("stats_", "sp")))
    assert len(names) == 6, names
    assert names == sorted(os.listdir(pyd)), (names, sorted(os.listdir(pyd)))
    for x in names:
        assert (outd / x).read_bytes() == (pyd / x).read_bytes(), x
    st = [x for x in names if x.startswith("stats_")][0]
    h, prof, _ = statsfile.read_stats(str(outd / st), dims["dnx"], dims["dnz"])
    assert h["istat"] == 1 and h["iter"] == 2
    assert h["table"][8:10] == [p.ngalx2, p.ngalz2] and h["table"][4:6] == [p.ngalx, p.ngalz]
    assert all(np.isfinite(v).all() for v in prof.values()) and np.abs(prof["U2m"]).max() > 0


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran driver not built")
def test_fortran_driver_multiband_refuses_operator_sequence(tmp_path):
    """fused = 0 (one entry point per subroutine) has no multi-band path: the
    driver stops with a message before creating a solver."""
    from cfd_amd import inputfile
    from .test_gpu_multiband import SHIPPED_GRID
    inputfile.write(str(tmp_path), SHIPPED_GRID, flag_init=3, fused=0)
    out = subprocess.run([EXE, "1"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert out.returncode != 0
    assert "fused = 1 only" in out.stdout + out.stderr
