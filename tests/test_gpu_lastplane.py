"""The reference's as-shipped last-plane nonlinear term on np ranks
(cfd_set_compat_lastplane; ops_in_planes, FOU3D.f90:1553-1613) on the GPU.

* the closed forms of tests/lastplane_cases.py on the four-step engine
  (48 x 33 x 72, rough walls) and the Stockham engine (36 x 25 x 24, smooth):
  the x/z term of Nu1 appears on exactly the planes of the rank whose last
  plane holds u3 (the top ghost plane included), that of Nu2 likewise;
* random fields against the oracle with the same replay: the nonlinear term
  (relative L2 <= 1e-11 per field) and one RK3 step of the fused substep
  (<= 1e-10 per velocity component);
* the mixed path with the replay within 1e-5 of the FP64 path;
* np = 0 gives back the per-plane term bit for bit;
* one RK3 step of the 384 x 192 x 192 workload with the replay of 8 ranks
  against the oracle (<= 1e-10 per velocity component).
"""
import dataclasses

import numpy as np
import pytest

from cfd_amd import Channel, fieldfile
from oracle.oracle import Oracle

from . import lastplane_cases as lc
from .helpers import BANDED, SMOOTH, random_field, rel_l2

pytestmark = pytest.mark.gpu

CASES = {"banded": BANDED, "smooth": SMOOTH}
STATE = ("u1", "u2", "u3", "p", "Nu1", "Nu2", "Nu3", "du1", "du2", "du3")


def _load_random(solvers, p, seed):
    rng = np.random.default_rng(seed)
    for name in STATE:
        f = random_field(p, name, rng, amp=1.0 if name in ("u1", "u2", "u3") else 0.3)
        if name == "u1":
            f[0, :] += 1.0
        if name == "u2":
            f[:, 0] = 0.0
            f[:, -1] = 0.0
        for s in solvers:
            s.set_field(name, f)
    for s in solvers:
        s.set_scalars(t=0.0, dt=2e-3, dti=500.0, mpgx=-1e-3, QxT=2.0)


def _nplanes_u(p):
    N = fieldfile.band_table(p)
    return N[3 * 4 + 3] - N[3]


@pytest.mark.parametrize("case", list(CASES))
def test_closed_forms(case):
    p = CASES[case]
    print("case last-plane closed forms", case, flush=True)
    ch = Channel(p)
    jlo = {"u": ch.jlo_u, "v": ch.jlo_v}
    kx, kz = lc.wavenumbers(p)
    for np_ranks in (3, 5):
        ch.set_compat_lastplane(np_ranks)
        R = lc.ranks(p, np_ranks)
        for j0 in (R["u"][1][2], R["u"][-1][2], R["u"][0][0]):
            plane = lc.load_plane_field(ch, p, jlo, ch.nyu, ch.nyv, "u", j0, seed=j0 + 50)
            ch.nonlinear(2)
            nu1 = ch.get_field("Nu1")
            expect = 1j * kz * plane
            for a, z, src in R["u"]:
                for j in range(a, z + 1):
                    col = nu1[:, j - jlo["u"]]
                    if src == j0:
                        assert np.abs(col - expect).max() <= 1e-12 * np.abs(expect).max(), (np_ranks, j0, j)
                    else:
                        assert np.abs(col).max() <= 1e-14 * np.abs(expect).max(), (np_ranks, j0, j)
        ref = Channel(p)
        for j0 in (R["v"][0][2], R["v"][1][0]):
            plane = lc.load_plane_field(ch, p, jlo, ch.nyu, ch.nyv, "v", j0, seed=j0 + 7)
            lc.load_plane_field(ref, p, jlo, ref.nyu, ref.nyv, "v", j0, seed=j0 + 7)
            ch.nonlinear(2)
            ref.nonlinear(2)
            d = ch.get_field("Nu2") - ref.get_field("Nu2")
            xz = 1j * kx * plane
            for a, z, src in R["v"]:
                for j in range(a, z + 1):
                    e = xz * ((src == j0) - (j == j0))
                    assert np.abs(d[:, j - jlo["v"]] - e).max() <= 1e-12 * np.abs(xz).max(), (np_ranks, j0, j)
        ref.close()
    ch.close()


@pytest.mark.parametrize("case", list(CASES))
def test_nonlinear_and_step_against_oracle(case):
    p = CASES[case]
    N = fieldfile.band_table(p)
    for np_ranks in (3, 7, _nplanes_u(p)):
        print("case last-plane vs oracle", case, "np", np_ranks, flush=True)
        ch, orc = Channel(p), Oracle(p)
        ch.set_compat_lastplane(np_ranks)
        orc.set_compat_lastplane(N, np_ranks)
        _load_random((ch, orc), p, seed=np_ranks)
        ch.nonlinear(1)
        orc.nonlinear(1)
        for n in ("Nu1", "Nu2", "Nu3", "du1", "du2", "du3"):
            e = rel_l2(ch.get_field(n), orc.get_field(n))
            assert e <= 1e-11, (n, np_ranks, e)
        _load_random((ch, orc), p, seed=np_ranks + 100)
        ch.step(1)
        orc.step(1)
        for n in ("u1", "u2", "u3"):
            e = rel_l2(ch.get_field(n), orc.get_field(n))
            assert e <= 1e-10, (n, np_ranks, e)
        ch.close()
        orc.close()


def test_mixed_path_and_switch_off():
    p = BANDED
    print("case last-plane mixed and off", flush=True)
    a, b = Channel(p), Channel(dataclasses.replace(p, precision=1))
    for ch in (a, b):
        ch.set_compat_lastplane(5)
    _load_random((a, b), p, seed=11)
    a.step(1)
    b.step(1)
    for n in ("u1", "u2", "u3"):
        assert rel_l2(b.get_field(n), a.get_field(n)) <= 1e-5, n
    # np = 0: the per-plane term, bit for bit that of a solver never switched
    c = Channel(p)
    a.set_compat_lastplane(0)
    _load_random((a, c), p, seed=12)
    a.nonlinear(2)
    c.nonlinear(2)
    for n in ("Nu1", "Nu2", "Nu3"):
        assert np.array_equal(a.get_field(n), c.get_field(n)), n
    with pytest.raises(RuntimeError):
        a.set_compat_lastplane(2)
    for ch in (a, b, c):
        ch.close()


def test_workload_size_against_oracle():
    """One RK3 step of the 384 x 192 x 192 workload (four-step kernels at 384
    and 192 points) with the replay of an 8-rank run, against the oracle:
    relative L2 <= 1e-10 per velocity component."""
    import os

    from cfd_amd import ic, workloads
    from oracle.oracle import set_threads

    p, _ = workloads.get("rt395_384x192x192")
    print("case last-plane rt395_384x192x192 np 8", flush=True)
    set_threads(min(16, len(os.sched_getaffinity(0))))
    ch, orc = Channel(p), Oracle(p)
    ch.set_compat_lastplane(8)
    orc.set_compat_lastplane(fieldfile.band_table(p), 8)
    g = ch.ygrid()
    f = ic.project(orc, ic.initial_fields(p, g, seed=0, amp=0.1))
    q0 = ic.flowrate(f["u1"], g, p.ny)
    for s in (ch, orc):
        for k in ("u1", "u2", "u3", "p"):
            s.set_field(k, f[k])
        s.set_scalars(t=0.0, dt=0.0, mpgx=0.0, QxT=q0)
    ch.step(1)
    orc.step(1)
    for n in ("u1", "u2", "u3"):
        e = rel_l2(ch.get_field(n), orc.get_field(n))
        print("  %s rel-L2 %.3e" % (n, e), flush=True)
        assert e <= 1e-10, (n, e)
    ch.close()
    orc.close()
