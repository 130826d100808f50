"""Field files in the reference layout (record_out / read_in), host side.

Pins:
* the band table N(4,0:4) against the compiled reference y_grid_canopy
  (golden tests/golden/bands.npz; rows 1-2 from start.f90:89-111);
* the byte layout, by an independent parse of the Fortran sequential records;
* the Fortran runtime itself, both ways: a test program using the read
  statements of read_in reads our file, and one using the write statements of
  record_out writes a file our reader reads (amdflang, skipped without it);
* the buff_to_u copy rule for planes of another size (FOU3D.f90:2631-2654),
  restated in numpy.
"""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

from cfd_amd import Params, fieldfile

HERE = os.path.dirname(os.path.abspath(__file__))
FC = shutil.which("amdflang") or ("/opt/rocm/bin/amdflang" if os.path.exists("/opt/rocm/bin/amdflang") else None)

YGRIDS = {"shipped": (25, 4, 0, 10.0, 5, 0.089, 0.0), "twolayer": (33, 6, 3, 8.0, 5, 0.12, 0.05),
          "smooth": (25, 0, 0, 10.0, 5, 0.0, 0.0), "p7": (41, 5, 0, 14.0, 7, 0.06, 0.0)}
BANDS = {"shipped_h25": ("shipped", 0, 2), "smooth_h25": ("smooth", 0, 2), "p7_ntz8": ("p7", 3, 8),
         "twolayer_ntz12": ("twolayer", 3, 12)}


@pytest.mark.parametrize("case", list(BANDS))
def test_band_table_matches_reference(case):
    gname, geom, ntz = BANDS[case]
    ny, dsty, dsty1, dyq, ppp, posth, posth1 = YGRIDS[gname]
    p = Params(ngalx=48, ngalz=72, ny=ny, dsty=dsty, dsty1=dsty1, dyq=dyq, ppp=ppp, posth=posth, posth1=posth1,
               ygrid=0, geometry_type=geom, ntilex=2, ntilez=ntz, bet=2.0, alp=1.0)
    N = np.array(fieldfile.band_table(p)).reshape(5, 4)  # N[band, row]
    ref = np.load(os.path.join(HERE, "golden", "bands.npz"))["%s_N" % case].reshape(5, 4)
    assert (N[:, 2:] == ref[:, 2:]).all(), (N, ref)
    # x/z mode counts: 0 in band 0, dealiased in bands 1-3, (-2, 0) in band 4
    assert N[0, :2].tolist() == [0, 0] and N[4, :2].tolist() == [-2, 0]
    assert (N[1:4, 0] == 32).all() and (N[1:4, 1] == 48).all()


def _records(path):
    """Independent parse: list of record payloads (4-byte little-endian
    markers before and after each)."""
    b = open(path, "rb").read()
    out, o = [], 0
    while o < len(b):
        (n,) = struct.unpack_from("<i", b, o)
        payload = b[o + 4:o + 4 + n]
        (m,) = struct.unpack_from("<i", b, o + 4 + n)
        assert m == n
        out.append(payload)
        o += 8 + n
    return out


def _sample(seed=0, ny_g=5, n2=6, mx=4):
    rng = np.random.default_rng(seed)
    planes = rng.standard_normal((ny_g, n2, mx)) + 1j * rng.standard_normal((ny_g, n2, mx))
    y = np.linspace(-1.2, 1.3, ny_g + 2)
    dthdy = 1.0 / (1.0 + y ** 2)
    return planes, y, dthdy


def test_written_bytes_parse_independently(tmp_path):
    planes, y, dthdy = _sample()
    N = list(range(101, 121))
    fn = str(tmp_path / "u1_test.dat")
    # p-like file: grid array of 7 centres from j = -3, planes j = -2 .. 2
    fieldfile.write(fn, planes, y, dthdy, field_id=4, j0=-2, jy0=-3, t=1.25, Re=2800.0, alp=1.0, bet=2.0,
                    mpgx=-3e-4, iter=77, N=N)
    recs = _records(fn)
    assert len(recs) == 3 + 5
    assert len(recs[0]) == 5 * 8 + 2 * 4 + 88 * 4
    t, Re, alp, bet, mpgx = struct.unpack_from("<5d", recs[0], 0)
    nband, it = struct.unpack_from("<2i", recs[0], 40)
    assert (t, Re, alp, bet, mpgx, nband, it) == (1.25, 2800.0, 1.0, 2.0, -3e-4, 3, 77)
    assert struct.unpack_from("<88i", recs[0], 48) == (0,) * 88
    assert list(struct.unpack("<20i", recs[1])) == N
    r3 = np.frombuffer(recs[2], dtype="<f8")
    assert len(r3) == 2 * 7 + 1
    assert (r3[:7] == y).all() and r3[7] == 0.0 and (r3[8:] == dthdy).all()
    for q in range(5):
        j, fid, nx, nz = struct.unpack_from("<4i", recs[3 + q], 0)
        (yj,) = struct.unpack_from("<d", recs[3 + q], 16)
        buf = np.frombuffer(recs[3 + q], dtype="<f8", offset=24).reshape(nz, nx)  # buff(nx, nz) column-major
        assert (j, fid, nx, nz) == (-2 + q, 4, 8, 6)
        assert yj == y[q + 1]
        assert (buf[:, 0::2] == planes[q].real).all() and (buf[:, 1::2] == planes[q].imag).all()


def test_read_back_is_bit_exact(tmp_path):
    planes, y, dthdy = _sample(seed=3)
    fn = str(tmp_path / "u2_x.dat")
    fieldfile.write(fn, planes, y[:5], dthdy[:5], field_id=2, j0=0, jy0=0, t=0.5, iter=9,
                    N=list(range(20)))
    h, back, mm = fieldfile.read(fn, 5, 6, 4, ngz=9)
    assert mm == 0 and h["iter"] == 9 and h["t"] == 0.5 and h["nband"] == 3 and h["N"] == list(range(20))
    assert (back == planes).all()


def _buff_to_u_restated(buf, nzt, mxt, ngz):
    """FOU3D.f90:2631-2654 into a plane of ngz z-slots, then the slots of the
    nzt dealiased target modes (bottom nzt/2 slots, top nzt - nzt/2 slots)."""
    nzf, mxf = buf.shape
    slots = np.zeros((ngz, 2 * mxt))
    real = np.zeros((nzf, 2 * mxf))
    real[:, 0::2], real[:, 1::2] = buf.real, buf.imag
    kh, ix = min(nzf // 2, ngz // 2), min(2 * mxf, 2 * mxt)
    for kf in range(1, kh + 1):
        slots[kf - 1, :ix] = real[kf - 1, :ix]
    for kf in range(nzf - kh + 1, nzf + 1):
        slots[kf + ngz - nzf - 1, :ix] = real[kf - 1, :ix]
    tgt = np.concatenate([slots[:nzt // 2], slots[ngz - (nzt - nzt // 2):]])
    return tgt[:, 0::2] + 1j * tgt[:, 1::2]


@pytest.mark.parametrize("shape", [((8, 5), (4, 3), 6), ((4, 3), (8, 5), 12), ((6, 4), (6, 4), 9)])
def test_plane_size_change_follows_buff_to_u(tmp_path, shape):
    (nzf, mxf), (nzt, mxt), ngz = shape
    rng = np.random.default_rng(5)
    planes = rng.standard_normal((2, nzf, mxf)) + 1j * rng.standard_normal((2, nzf, mxf))
    fn = str(tmp_path / "u3_x.dat")
    fieldfile.write(fn, planes, np.zeros(2), np.ones(2), field_id=3, j0=0, jy0=0)
    h, back, mm = fieldfile.read(fn, 2, nzt, mxt, ngz=ngz)
    assert mm == (0 if (nzf, mxf) == (nzt, mxt) else 2)
    for q in range(2):
        assert (back[q] == _buff_to_u_restated(planes[q], nzt, mxt, ngz)).all()


def _fortran(tmp_path, name):
    exe = str(tmp_path / name)
    subprocess.run([FC, "-O1", "-o", exe, os.path.join(HERE, "fortran", name + ".f90")], check=True,
                   cwd=str(tmp_path), capture_output=True)
    return exe


@pytest.mark.skipif(FC is None, reason="amdflang not available")
def test_fortran_reads_our_file(tmp_path):
    planes, y, dthdy = _sample(seed=11, ny_g=4, n2=6, mx=5)
    N = [3 * q - 7 for q in range(20)]
    fn = str(tmp_path / "u1_ours.dat")
    fieldfile.write(fn, planes, y[:4], dthdy[:4], field_id=1, j0=-1, jy0=-1, t=3.5, Re=6887.0, alp=1.0, bet=2.0,
                    mpgx=-2.5e-4, iter=42, N=N)
    exe = _fortran(tmp_path, "ff_read")
    out = subprocess.run([exe, fn, "4"], check=True, capture_output=True, text=True).stdout.split("\n")
    assert [float(v) for v in out[0].split()] == [3.5, 6887.0, 1.0, 2.0, -2.5e-4]
    assert [int(v) for v in out[1].split()] == [3, 42, 0]
    assert [int(v) for v in out[2].split()] == N
    i = np.arange(1, 11)[None, :]
    k = np.arange(1, 7)[:, None]
    for q in range(4):
        f = out[3 + q].split()
        assert [int(v) for v in f[:4]] == [q - 1, 1, 10, 6]
        assert float(f[4]) == y[q]
        buf = np.zeros((6, 10))
        buf[:, 0::2], buf[:, 1::2] = planes[q].real, planes[q].imag
        cks = float(np.sum(buf * (i + 7 * k)))
        assert abs(float(f[5]) - cks) <= 1e-12 * max(1.0, abs(cks))


@pytest.mark.skipif(FC is None, reason="amdflang not available")
def test_we_read_fortran_written_file(tmp_path):
    exe = _fortran(tmp_path, "ff_write")
    fn = str(tmp_path / "u1_fortran.dat")
    subprocess.run([exe, fn], check=True, capture_output=True)
    h, planes, mm = fieldfile.read(fn, 7, 8, 5, ngz=12)
    assert mm == 0
    assert (h["t"], h["Re"], h["alp"], h["bet"], h["mpgx"], h["nband"], h["iter"]) == \
        (12.5, 2800.0, 1.0, 2.0, -0.0625, 3, 1234)
    assert h["N"] == [10 * b + r for b in range(5) for r in range(1, 5)]
    for q in range(7):
        i = np.arange(1, 11)
        for k in range(1, 9):
            row = i + 1000.0 * k + 1e6 * q
            assert (planes[q, k - 1].real == row[0::2]).all() and (planes[q, k - 1].imag == row[1::2]).all()
    recs = _records(fn)  # the Fortran runtime's framing is the one parsed above
    assert len(recs) == 10 and len(recs[0]) == 400
