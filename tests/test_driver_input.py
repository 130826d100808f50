"""The Fortran driver's input.in reader (CPU): the record sequence and edit
descriptors of start.f90:62-227 and the implied-decimal rule of (7X,D10.1)
(a real written without a decimal point is scaled by 1/10)."""
import os
import shutil
import subprocess

import pytest

FDIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "cfd_amd", "fortran")
EXE = os.path.join(FDIR, "littleharsh_gpu")


def parse(tmp_path, edit=None):
    for f in ("input.in", "geom_circ_rough.txt"):
        shutil.copy(os.path.join(FDIR, f), tmp_path / f)
    if edit:
        txt = (tmp_path / "input.in").read_text().split("\n")
        for key, val in edit.items():
            i = [k for k, l in enumerate(txt) if l.startswith(key)][0]
            txt[i] = txt[i][:7] + val + txt[i][17:]
        (tmp_path / "input.in").write_text("\n".join(txt))
    out = subprocess.run([EXE, "parse"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    vals = {}
    for line in out.stdout.splitlines():
        parts = line.split()
        if parts:
            vals[parts[0]] = [float(x.replace("D", "E")) for x in parts[1:]]
    return vals


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran driver not built")
def test_reads_example(tmp_path):
    v = parse(tmp_path)
    assert v["ints"] == [48, 48, 25, 4, 0, 5, 3, 2, 2]
    assert v["flag_ctpress"] == [1]
    assert v["reals"] == [2800.0, 1.0, 2.0, 10.0, 0.089, 0.5, -1.0714e-3]
    assert v["post"] == [5.43723, 12.0, 12.0, 0.5]


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran driver not built")
def test_implied_decimal(tmp_path):
    # 'alp = 128d0' reads as 12.8: format D10.1 on a real without a decimal
    # point implies one decimal place (start.f90:434, read(40,10))
    v = parse(tmp_path, {"alp": "     128d0", "bet": "     256d0"})
    assert abs(v["reals"][1] - 12.8) < 1e-12 and abs(v["reals"][2] - 25.6) < 1e-12


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran driver not built")
@pytest.mark.parametrize("workload", ["rt180_64x96x64", "rt395_384x192x192", "rt550_768x384x384"])
def test_inputfile_round_trip(tmp_path, workload):
    """cfd_amd.inputfile writes a workload's input.in and geom_circ_rough.txt
    in the reference layout plus the optional records; the driver reads back
    exactly the values inputfile.representable reports (every real within
    the 10-character field, the tanh grid, no 24pt extras)."""
    import dataclasses

    from cfd_amd import inputfile, workloads
    p, _ = workloads.get(workload)
    p = dataclasses.replace(p, precision=1)
    q = inputfile.write(str(tmp_path), p)
    out = subprocess.run([EXE, "parse"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    v = {}
    for line in out.stdout.splitlines():
        parts = line.split()
        if parts:
            v[parts[0]] = [float(x.replace("D", "E")) for x in parts[1:]]
    assert v["ints"] == [q.ngalx, q.ngalz, q.ny, q.dsty, q.dsty1, q.ppp, q.geometry_type, q.ntilex, q.ntilez]
    assert v["reals"] == [q.Re, q.alp, q.bet, q.dyq, q.posth, q.CFL, q.mpgx]
    assert v["post"] == [q.r, q.c_i, q.c_k, q.stag]
    assert v["ext"] == [1, 0, 1, 1, 2.0]
    # the 10-character fields keep 8 significant digits of these values
    for k in ("posth", "r"):
        assert abs(getattr(q, k) - getattr(p, k)) <= 1e-7 * abs(getattr(p, k))


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran driver not built")
def test_optional_records_default(tmp_path):
    """Without the optional records the driver keeps the reference's choices
    (polynomial grid, 24pt extras, FP64) and fused stepping; a record that is
    present sets its value, in any order."""
    v = parse(tmp_path)
    assert v["ext"] == [0, 1, 0, 1, 2.0]
    with open(tmp_path / "input.in", "a") as fh:
        fh.write("fused =         0 ! per entry point\nybeta =     3.5d0 ! tanh\nygrid =         1 ! tanh\n")
    out = subprocess.run([EXE, "parse"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    ext = [float(x.replace("D", "E")) for x in [l for l in out.stdout.splitlines() if l.split()[:1] == ["ext"]][0].split()[1:]]
    assert ext == [1, 1, 0, 0, 3.5]
