"""The Fortran driver (cfd_amd/fortran/littleharsh_gpu) reads input.in in
the reference layout and calls the hot path through the ISO_C_BINDING
interface; its result must equal the same run driven from Python."""
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
def test_fortran_driver_matches_python(tmp_path):
    for f in ("input.in", "geom_circ_rough.txt"):
        shutil.copy(os.path.join(FDIR, f), tmp_path / f)
    print("fortran driver, 3 steps", flush=True)
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
    assert abs(py - fort) <= 1e-13 * abs(py), (py, fort)


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
