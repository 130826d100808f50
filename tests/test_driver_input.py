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
