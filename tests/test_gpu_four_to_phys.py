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
