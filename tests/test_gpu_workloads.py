"""The benchmark workloads at their full size against the CPU FP64 oracle.

From the same seeded, projected initial field (laminar profile plus a 10 %
perturbation, seed 0) the GPU FP64 path and the oracle run 10 RK3 steps of
the 64 x 96 x 64 (Re_tau 180, 8 x 4 posts) and 384 x 192 x 192 (Re_tau 395,
16 x 8 posts) workloads: relative L2 <= 1e-10 per velocity component, the
divergence of the GPU field at round-off, IB lists bit-exact, the same time
and the same mean pressure gradient; the mixed path on the larger workload
within the errors it is reported with (wall shear <= 0.5 %).  The 768 x 384 x
384 workload (Re_tau 550, 32 x 16 posts of k+ 20, dsty 16) runs one RK3 step
against the oracle, FP64 and mixed.
"""
import ctypes
import os

import numpy as np
import pytest

from cfd_amd import Channel, diagnostics, ic, workloads
from oracle.oracle import Oracle, set_threads

from .helpers import max_rel, rel_l2

pytestmark = pytest.mark.gpu

WORKLOADS = ["rt180_64x96x64", "rt395_384x192x192"]


def _start(p):
    ch = Channel(p)
    orc = Oracle(p)
    g = ch.ygrid()
    f = ic.project(orc, ic.initial_fields(p, g, seed=0, amp=0.1))
    q0 = ic.flowrate(f["u1"], g, p.ny)
    for s in (ch, orc):
        for k in ("u1", "u2", "u3", "p"):
            s.set_field(k, f[k])
        s.set_scalars(t=0.0, dt=0.0, mpgx=0.0, QxT=q0)
    return ch, orc, f, q0


def _iblists_equal(ch, orc):
    from cfd_amd import _lib  # noqa: F401
    for g in (0, 1):
        for w in (0, 1):
            n = ctypes.c_int()
            ch._call("cfd_get_iblist", g, w, None, None, ctypes.byref(n))
            l9 = np.zeros((n.value, 9), dtype=np.int32)
            w3 = np.zeros((n.value, 3))
            ch._call("cfd_get_iblist", g, w, l9.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                     w3.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(n))
            ol, ow = orc.iblist(g, w)
            assert n.value > 0
            assert np.array_equal(l9, ol) and np.array_equal(w3, ow), (g, w)


@pytest.mark.parametrize("workload", WORKLOADS)
def test_ten_steps_against_oracle(workload):
    p, _ = workloads.get(workload)
    print("case", workload, "10 steps vs oracle", flush=True)
    set_threads(min(16, len(os.sched_getaffinity(0))))
    ch, orc, _, _ = _start(p)
    _iblists_equal(ch, orc)
    ch.step(10)
    orc.step(10)
    for n in ("u1", "u2", "u3"):
        a, b = ch.get_field(n), orc.get_field(n)
        e = rel_l2(a, b)
        print("  %s rel-L2 %.3e" % (n, e), flush=True)
        assert e <= 1e-10, "%s rel-L2 %.3e (max rel %.3e)" % (n, e, max_rel(a, b))
    sa, sb = ch.scalars(), orc.scalars()
    assert abs(sa["t"] - sb["t"]) <= 1e-12 * sb["t"]
    assert abs(sa["mpgx"] - sb["mpgx"]) <= 1e-9 * abs(sb["mpgx"])
    d = np.abs(ch.get_field("div")).max() / np.abs(ch.get_field("u1")).max()
    print("  max |div| / max |u1| %.3e" % d, flush=True)
    assert d <= 1e-12
    ch.close()
    orc.close()


def test_mixed_path_errors_at_384x192x192():
    """The mixed path against the FP64 path over 10 steps from the same field:
    velocity within 1e-5, wall shear within 0.5 % (the stated bound; it is
    ~1e-7 here), flow rate exact."""
    import dataclasses
    p, _ = workloads.get("rt395_384x192x192")
    print("case rt395_384x192x192 mixed vs FP64", flush=True)
    a = Channel(p)
    b = Channel(dataclasses.replace(p, precision=1))
    g = a.ygrid()
    f = ic.project(a, ic.initial_fields(p, g, seed=0, amp=0.1))
    q0 = ic.flowrate(f["u1"], g, p.ny)
    for s in (a, b):
        for k in ("u1", "u2", "u3", "p"):
            s.set_field(k, f[k])
        s.set_scalars(t=0.0, dt=0.0, mpgx=0.0, QxT=q0)
    a.step(10)
    b.step(10)
    st = {k: a.get_field(k) for k in ("u1", "u2", "u3")}
    st["mpgx"] = a.scalars()["mpgx"]
    sm = {k: b.get_field(k) for k in ("u1", "u2", "u3")}
    sm["mpgx"] = b.scalars()["mpgx"]
    err = diagnostics.compare(sm, st, g, p.Re)
    print("  ", err, flush=True)
    for k in ("rel_l2_u1", "rel_l2_u2", "rel_l2_u3"):
        assert err[k] <= 1e-5, (k, err[k])
    assert err["wall_shear_rel"] <= 5e-3
    assert abs(ic.flowrate(sm["u1"], g, p.ny) - q0) <= 1e-12 * abs(q0)
    a.close()
    b.close()


def test_rt550_768x384x384_one_step():
    """The full 768 x 384 x 384 workload (417 u planes, 16 canopy planes per
    wall, 32 x 16 posts): one RK3 step of the FP64 path against the oracle
    from the seeded projected field (rel-L2 <= 1e-10 per velocity component,
    IB lists bit-exact, divergence at round-off, same t and mean pressure
    gradient), and the mixed path's step against the FP64 path's (wall shear
    within 0.5 %, velocity within 1e-5).  Prints the device footprint of both
    handles."""
    import dataclasses
    p, _ = workloads.get("rt550_768x384x384")
    print("case rt550_768x384x384 one RK3 step vs oracle", flush=True)
    set_threads(min(16, len(os.sched_getaffinity(0))))
    ch, orc, f, q0 = _start(p)
    assert ch.nyu == 417
    _iblists_equal(ch, orc)
    b64 = ch.device_bytes()
    ch.step(1)
    orc.step(1)
    for n in ("u1", "u2", "u3"):
        a, b = ch.get_field(n), orc.get_field(n)
        e = rel_l2(a, b)
        print("  %s rel-L2 %.3e" % (n, e), flush=True)
        assert e <= 1e-10, "%s rel-L2 %.3e (max rel %.3e)" % (n, e, max_rel(a, b))
    sa, sb = ch.scalars(), orc.scalars()
    assert abs(sa["t"] - sb["t"]) <= 1e-12 * sb["t"]
    # after one step mpgx is the sum of three flow-rate corrections g = (QxT -
    # Q(u)) / Q(u11) (flowrate_corr, littleharsh.f90:452-484): a flow-rate
    # difference at round-off divided by a quantity proportional to dt, so its
    # error bound is absolute, 1e-13 QxT / dt
    assert abs(sa["mpgx"] - sb["mpgx"]) <= 1e-13 * abs(sb["QxT"]) / sb["dt"], (sa["mpgx"], sb["mpgx"])
    d = np.abs(ch.get_field("div")).max() / np.abs(ch.get_field("u1")).max()
    print("  max |div| / max |u1| %.3e" % d, flush=True)
    assert d <= 1e-12
    st = {k: ch.get_field(k) for k in ("u1", "u2", "u3")}
    st["mpgx"] = sa["mpgx"]
    g = ch.ygrid()
    orc.close()
    ch.close()
    m = Channel(dataclasses.replace(p, precision=1))
    bm = m.device_bytes()
    for k in ("u1", "u2", "u3", "p"):
        m.set_field(k, f[k])
    m.set_scalars(t=0.0, dt=0.0, mpgx=0.0, QxT=q0)
    m.step(1)
    sm = {k: m.get_field(k) for k in ("u1", "u2", "u3")}
    sm["mpgx"] = m.scalars()["mpgx"]
    err = diagnostics.compare(sm, st, g, p.Re)
    print("  device GB fp64 %.2f mixed %.2f; mixed vs fp64 %s" % (b64 / 1e9, bm / 1e9, err), flush=True)
    for k in ("rel_l2_u1", "rel_l2_u2", "rel_l2_u3"):
        assert err[k] <= 1e-5, (k, err[k])
    assert err["wall_shear_rel"] <= 5e-3
    assert abs(ic.flowrate(sm["u1"], g, p.ny) - q0) <= 1e-12 * abs(q0)
    assert b64 < 24e9 and bm < b64
    m.close()
