"""four_to_phys on the GPU (four_to_phys.f90:61-326): spectral field ->
physical values on the 3/2-padded grid.

* Channel.four_to_phys against numpy's inverse FFTs of the zero-padded
  spectrum, with the kz = -N2/2 mode (set non-zero here on purpose) left out
  as the reference's resize loop leaves it out, unscaled in both directions;
  FP64 in both precisions and with either FFT engine;
* the tool end to end: record_out files, input_FtP.txt, `python -m
  cfd_amd.four_to_phys`; record sequence, headers, ids, y values and the
  physical planes of every file.
"""
import dataclasses
import os
import struct
import subprocess
import sys

import numpy as np
import pytest

from cfd_amd import fieldfile
from cfd_amd.four_to_phys import plane_sequence

from .helpers import EDGE, SHIPPED, SMALLPOSTS, WIDE, random_field
from .test_fieldfile import _records
from .test_gpu_parity import _channel

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _spectral(p, name, seed):
    rng = np.random.default_rng(seed)
    f = random_field(p, name, rng)
    # a non-zero kz = -N2/2 line: four_to_phys must leave it out
    f.reshape(p.N2, p.Mx, -1)[p.N2 // 2] = 5.0 + 3.0j
    return f


def _numpy_phys(p, f):
    """Zero padding to (Ngalz, Ngalx/2 + 1) without kz = -N2/2, then
    sum_k u_k exp(+i k x) in z and in x (numpy's inverse FFTs times the
    lengths); c2r takes the real part of the kx = 0 and x-Nyquist terms."""
    N2, Mx, ngz, ngx = p.N2, p.Mx, p.ngalz, p.ngalx
    planes = f.T.reshape(-1, N2, Mx)
    C = np.zeros((planes.shape[0], ngz, ngx // 2 + 1), dtype=complex)
    C[:, :N2 // 2, :Mx] = planes[:, :N2 // 2]
    C[:, ngz - N2 // 2 + 1:, :Mx] = planes[:, N2 // 2 + 1:]
    return np.fft.irfft(np.fft.ifft(C, axis=1) * ngz, n=ngx, axis=2) * ngx


CASES = {"shipped": (SHIPPED, None), "smallposts": (SMALLPOSTS, None), "wide": (WIDE, None),
         "wide_stockham": (WIDE, "stockham"), "edge": (EDGE, None),
         "wide_mixed": (dataclasses.replace(WIDE, precision=1), None)}


@pytest.mark.parametrize("case", list(CASES))
def test_four_to_phys_matches_numpy(case):
    p, fft = CASES[case]
    print("case four_to_phys", case, flush=True)
    ch = _channel(p, fft)
    try:
        for s, name in enumerate(("u1", "u2", "u3", "p")):
            f = _spectral(p, name, 900 + s)
            ch.set_field(name, f)
            got = ch.four_to_phys(name)
            ref = _numpy_phys(p, f)
            assert got.shape == ref.shape
            err = np.abs(got - ref).max() / np.abs(ref).max()
            assert err <= 1e-13, (case, name, err)
    finally:
        ch.close()


def test_four_to_phys_tool_end_to_end(tmp_path):
    """record_out -> `python -m cfd_amd.four_to_phys input_FtP.txt`."""
    # tanh grid with dsty = 3 and a band table in order (smooth walls: h = 0.25)
    p = dataclasses.replace(SMALLPOSTS, geometry_type=0)
    print("case four_to_phys tool smallposts-smooth", flush=True)
    ch = _channel(p)
    try:
        src = {}
        for s, name in enumerate(("u1", "u2", "u3", "p")):
            src[name] = random_field(p, name, np.random.default_rng(950 + s))
            ch.set_field(name, src[name])
        ch.set_scalars(t=2.5, mpgx=-1e-3)
        os.makedirs(tmp_path / "fields")
        os.makedirs(tmp_path / "phys")
        ch.record_out(str(tmp_path / "fields"), iter=31)
        ref = {name: ch.four_to_phys(name) for name in src}
    finally:
        ch.close()
    # (7X,I10) x 3, (7X,D10.1), (8X,A30), (7X,A30)
    (tmp_path / "input_FtP.txt").write_text(
        "nx    =%10d\nnz    =%10d\nny    =%10d\nt     =%10s\ndirname=./fields\ndirout=./phys\n"
        % (p.N1, p.N2, p.ny, "2.5"))
    env = dict(os.environ, PYTHONPATH=ROOT + os.pathsep + os.environ.get("PYTHONPATH", ""))
    r = subprocess.run([sys.executable, "-m", "cfd_amd.four_to_phys", "input_FtP.txt"], cwd=str(tmp_path), env=env,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    tail = "%04dx%04dx%04d_t%05d.dat" % (p.N1, p.N2, p.ny, 25)
    N = fieldfile.band_table(p)
    for name, fid, grid in (("u1", 1, "u"), ("u2", 2, "v"), ("u3", 3, "u"), ("p", 4, "p")):
        spec = _records(str(tmp_path / "fields" / ("%s_%s" % (name, tail))))
        phys = _records(str(tmp_path / "phys" / ("%s_phys_%s" % (name, tail))))
        seq = plane_sequence(grid, N)
        assert len(phys) == 3 + len(seq), name
        assert phys[0] == spec[0] and phys[2] == spec[2]
        ngal = list(struct.unpack("<20i", phys[1]))
        assert ngal[4:6] == [p.ngalx, p.ngalz] and ngal[2:4] == N[2:4]
        numpy_ref = _numpy_phys(p, src[name])
        for (q, j, rid), rec in zip(seq, phys[3:]):
            jj, idd, nx2, nz2 = struct.unpack_from("<4i", rec, 0)
            (y,) = struct.unpack_from("<d", rec, 16)
            assert (jj, idd, nx2, nz2) == (j, fid if rid is None else rid, p.ngalx + 2, p.ngalz)
            assert y == struct.unpack_from("<d", spec[3 + q], 16)[0]
            buf = np.frombuffer(rec, dtype="<f8", offset=24).reshape(p.ngalz, p.ngalx + 2)
            assert (buf[:, p.ngalx:] == 0).all()
            # the tool's channel runs the same kernels on the same spectra
            assert (buf[:, :p.ngalx] == ref[name][q]).all(), (name, j)
        err = np.abs(ref[name] - numpy_ref).max() / np.abs(numpy_ref).max()
        assert err <= 1e-13, (name, err)


def _resize_ref(old, nx, nz):
    """resize of four_to_phys.f90:328-348 on one plane (old[k, i], x fastest):
    the x reals up to min(nx, nx_old), z-modes k < min(nz, nz_old)/2 from the
    bottom and the min(nz, nz_old)/2 - 1 highest-index modes from the top of
    the old plane."""
    nzo, nxo = old.shape
    out = np.zeros((nz, nx))
    m, ix = min(nz // 2, nzo // 2), min(nx, nxo)
    out[:m, :ix] = old[:m, :ix]
    if m > 1:
        out[nz - (m - 1):, :ix] = old[nzo - (m - 1):, :ix]
    return out


def test_four_to_phys_tool_multiband(tmp_path):
    """A multi-band run's files (central planes stored at the central band's
    N1c+2 x N2c): every output record of band b is the inverse transform on
    band b's grid (Ngal = 3/2 N(1:2, b)) of its plane resized to N(1:2, b),
    numpy's inverse FFTs as the reference; the interface repeats resize the
    previous band's last plane; the Ngal header carries each band's grid."""
    from cfd_amd.four_to_phys import band_of
    from .test_gpu_multiband import MB, _pair
    p = MB
    print("case four_to_phys tool multi-band", flush=True)
    ch, _ = _pair(p, seed=76)
    ch.substep(1)
    ch.set_scalars(t=1.5)
    os.makedirs(tmp_path / "fields")
    os.makedirs(tmp_path / "phys")
    ch.record_out(str(tmp_path / "fields"), iter=3)
    ch.close()
    (tmp_path / "input_FtP.txt").write_text(
        "nx    =%10d\nnz    =%10d\nny    =%10d\nt     =%10s\ndirname=./fields\ndirout=./phys\n"
        % (p.N1, p.N2, p.ny, "1.5"))
    env = dict(os.environ, PYTHONPATH=ROOT + os.pathsep + os.environ.get("PYTHONPATH", ""))
    r = subprocess.run([sys.executable, "-m", "cfd_amd.four_to_phys", "input_FtP.txt"], cwd=str(tmp_path), env=env,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    tail = "%04dx%04dx%04d_t%05d.dat" % (p.N1, p.N2, p.ny, 15)
    N = fieldfile.band_table(p)
    assert (N[8], N[9]) == (p.N1c, p.N2c) and (N[4], N[5]) == (p.N1, p.N2)
    seen_central = 0
    for name, fid, grid in (("u1", 1, "u"), ("u2", 2, "v"), ("u3", 3, "u"), ("p", 4, "p")):
        spec = _records(str(tmp_path / "fields" / ("%s_%s" % (name, tail))))
        phys = _records(str(tmp_path / "phys" / ("%s_phys_%s" % (name, tail))))
        seq, bands = plane_sequence(grid, N), band_of(grid, N)
        assert len(phys) == 3 + len(seq) and len(bands) == len(seq), name
        ngal = list(struct.unpack("<20i", phys[1]))
        assert ngal[4:6] == [p.ngalx, p.ngalz] and ngal[8:10] == [p.ngalx2, p.ngalz2]
        got, refs = [], []
        for (q, j, rid), b, rec in zip(seq, bands, phys[3:]):
            n1, n2 = N[4 * b], N[4 * b + 1]
            gx, gz = n1 * 3 // 2, n2 * 3 // 2
            jj, idd, nx2, nz2 = struct.unpack_from("<4i", rec, 0)
            assert (jj, nx2, nz2) == (j, gx + 2, gz), (name, j, b)
            sj, sid, snx, snz = struct.unpack_from("<4i", spec[3 + q], 0)
            old = np.frombuffer(spec[3 + q], dtype="<f8", offset=24).reshape(snz, snx)
            seen_central += snx == p.N1c + 2
            re = _resize_ref(old, n1 + 2, n2)
            c = (re[:, 0::2] + 1j * re[:, 1::2])  # (n2, n1/2+1)
            pp = dataclasses.replace(p, ngalx=gx, ngalz=gz)
            ref = _numpy_phys(pp, c.reshape(-1, 1))[0]
            buf = np.frombuffer(rec, dtype="<f8", offset=24).reshape(gz, gx + 2)
            assert (buf[:, gx:] == 0).all()
            got.append(buf[:, :gx])
            refs.append(ref)
        # the x-inverse transforms planes in pairs, so a plane's round-off
        # scales with its partner's: the bound is relative to the field
        scale = max(np.abs(r).max() for r in refs)
        for (q, j, rid), b, g, r in zip(seq, bands, got, refs):
            assert np.abs(g - r).max() <= 1e-13 * scale, (name, j, b, np.abs(g - r).max() / scale)
    assert seen_central > 0
