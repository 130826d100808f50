"""record_out / read_in through the C ABI on the GPU.

* record_out writes what the device holds: headers (t, mpgx, iter, the band
  table) and every plane, read back with the host reader;
* read_in restores the fields bit for bit, with the flag_init branches of
  start (t and iter0 from the header for flag_init = 2, zero for 1; QxT the
  flow rate of the read u1);
* kill and restart: with a prescribed pressure gradient (flag_ctpress = 1) the
  resumed run equals the uninterrupted one bit for bit.  With a constant flow
  rate the reference restarts with mpgx from input.in and QxT from the read
  field; mpgx is not a removable constant there (the flow-rate correction
  solves the tip-to-tip system of LUsol0, not the full column the forcing
  enters), so the resumed run equals the uninterrupted one bit for bit once
  the caller restores mpgx and QxT, and differs otherwise.
"""
import dataclasses

import numpy as np
import pytest

from cfd_amd import Channel, fieldfile, ic
from cfd_amd._lib import CfdError

from .helpers import SHIPPED, SMALLPOSTS
from .test_gpu_parity import _pair

pytestmark = pytest.mark.gpu


def _planes(ch, name):
    f = ch.get_field(name)  # (Ncol, ny) -> (ny, N2, Mx)
    return f.T.reshape(f.shape[1], ch.N2, ch.Mx)


def test_record_out_writes_device_state(tmp_path):
    p = SMALLPOSTS
    ch, _ = _pair(p, seed=80)
    ch.set_scalars(t=1.234, mpgx=-2e-4)
    u1path = ch.record_out(str(tmp_path), iter=17)
    tail = u1path.split("/u1")[-1]
    N = fieldfile.band_table(p)
    g = ch.ygrid()
    for name, ny_g, fid in (("u1", ch.nyu, 1), ("u2", ch.nyv, 2), ("u3", ch.nyu, 3), ("p", ch.nyp, 4)):
        h, planes, mm = fieldfile.read(str(tmp_path / (name + tail)), ny_g, ch.N2, ch.Mx, p.ngalz)
        assert mm == 0
        assert (h["t"], h["mpgx"], h["iter"], h["nband"]) == (1.234, -2e-4, 17, 3)
        assert (h["Re"], h["alp"], h["bet"]) == (p.Re, p.alp, p.bet)
        assert h["N"] == N
        assert (planes == _planes(ch, name)).all(), name
    assert tail == "_%04dx%04dx%04d_t%05d.dat" % (ch.N1, ch.N2, p.ny, 12)
    del g


@pytest.mark.parametrize("flag_init", [1, 2])
def test_read_in_restores_fields(tmp_path, flag_init):
    p = SMALLPOSTS
    a, _ = _pair(p, seed=81)
    a.set_scalars(t=0.75, mpgx=-3e-4)
    u1path = a.record_out(str(tmp_path), iter=40)
    b = Channel(p)
    iter0 = b.read_in(str(tmp_path), u1path.split("/")[-1], flag_init=flag_init)
    for name in ("u1", "u2", "u3", "p"):
        assert (b.get_field(name) == a.get_field(name)).all(), name
    sb = b.scalars()
    assert iter0 == (40 if flag_init == 2 else 0)
    assert sb["t"] == (0.75 if flag_init == 2 else 0.0)
    assert sb["mpgx"] == 0.0  # not restored, as in the reference
    q = ic.flowrate(a.get_field("u1"), a.ygrid(), p.ny)
    assert abs(sb["QxT"] - q) <= 1e-15 * abs(q)


def test_read_in_rejects_other_nband(tmp_path):
    p = SMALLPOSTS
    ch = Channel(p)
    z = np.zeros((ch.nyu, ch.N2, ch.Mx), complex)
    g = ch.ygrid()
    fieldfile.write(str(tmp_path / "u1_bad.dat"), z, g["yu"], g["dthdyu"], field_id=1, j0=ch.jlo_u,
                    jy0=ch.jlo_u, nband=2)
    with pytest.raises(CfdError):
        ch.read_in(str(tmp_path), "u1_bad.dat", flag_init=2)


def _start(p, seed):
    ch = Channel(p)
    f = ic.initial_fields(p, ch.ygrid(), seed=seed, amp=0.1)
    f = ic.project(ch, f)
    for k in ("u1", "u2", "u3", "p"):
        ch.set_field(k, f[k])
    q = ic.flowrate(f["u1"], ch.ygrid(), p.ny)
    ch.set_scalars(t=0.0, dt=0.0, mpgx=p.mpgx, QxT=q)
    return ch


@pytest.mark.parametrize("ctpress", [1, 0])
def test_restart_continues_the_run(tmp_path, ctpress):
    p = dataclasses.replace(SHIPPED, flag_ctpress=ctpress, mpgx=-1.07e-3 if ctpress else 0.0)
    a = _start(p, seed=5)
    a.step(4)
    b = _start(p, seed=5)
    b.step(2)
    sb = b.scalars()
    u1path = b.record_out(str(tmp_path), iter=2)
    c = Channel(p)
    c.set_scalars(mpgx=p.mpgx)  # the driver's mpgx from input.in
    assert c.read_in(str(tmp_path), u1path.split("/")[-1], flag_init=2) == 2
    if not ctpress:
        c.set_scalars(mpgx=sb["mpgx"], QxT=sb["QxT"])
    c.step(2)
    for name in ("u1", "u2", "u3", "p"):
        assert (c.get_field(name) == a.get_field(name)).all(), name
    assert c.scalars()["t"] == a.scalars()["t"]
