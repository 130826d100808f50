"""Parity of the HIP path (through the C ABI) with the CPU FP64 oracle.

Each reference operator is run from identical seeded state on both sides and
every output compared.  Tolerances (FP64 arithmetic, different FFT
algorithms and FMA contraction on the GPU): relative L2 <= 1e-11 per field for
one operator, <= 1e-10 per velocity component after 10 RK3 steps (the bar the
project states for the GPU FP64 path).  IB lists compare bit-exactly.
"""
import os

import numpy as np
import pytest

from cfd_amd import Channel
from oracle.oracle import Oracle

from .helpers import COARSE, EDGE, SHIPPED, SMALLPOSTS, SMOOTH, WIDE, WIDE768, random_field, rel_l2, max_rel

pytestmark = pytest.mark.gpu

CASES = {"shipped": SHIPPED, "smallposts": SMALLPOSTS, "smooth": SMOOTH, "wide": WIDE, "edge": EDGE,
         "coarse": COARSE}
STATE = ("u1", "u2", "u3", "p", "Nu1", "Nu2", "Nu3", "du1", "du2", "du3")
TOL_OP = 1e-11


def _pair(p, seed=0):
    print("case", p.ngalx, p.ny, p.ngalz, "seed", seed, flush=True)
    ch = Channel(p)
    orc = Oracle(p)
    rng = np.random.default_rng(seed)
    for name in STATE:
        f = random_field(p, name, rng, amp=1.0 if name in ("u1", "u2", "u3") else 0.3)
        if name == "u1":
            f[0, :] += 1.0  # mean flow
        if name == "u2":
            f[:, 0] = 0.0  # no flow through the walls
            f[:, -1] = 0.0
        ch.set_field(name, f)
        orc.set_field(name, f)
    sc = dict(t=0.0, dt=2e-3, dti=500.0, mpgx=-1e-3, QxT=orc.flowrate() + 0.01)
    ch.set_scalars(**sc)
    orc.set_scalars(**sc)
    return ch, orc


def _cmp(ch, orc, names, tol=TOL_OP, interior=None):
    for n in names:
        a, b = ch.get_field(n), orc.get_field(n)
        if interior:
            a, b = a[:, 1:-1], b[:, 1:-1]
        e = rel_l2(a, b)
        assert e <= tol, "%s rel-L2 %.3e > %.1e (max rel %.3e)" % (n, e, tol, max_rel(a, b))


def test_folded_y_grid_is_refused():
    """posth = 0 with dsty = 3 folds the canopy extension back on itself
    (yu(j) < yv(j) < yu(j+1) fails); the solver refuses the grid."""
    import dataclasses
    from cfd_amd._lib import CfdError
    with pytest.raises(CfdError, match="folds"):
        Channel(dataclasses.replace(SMALLPOSTS, geometry_type=0, posth=0.0))


@pytest.mark.parametrize("case", list(CASES))
def test_iblists_match_oracle(case):
    p = CASES[case]
    ch = Channel(p)
    orc = Oracle(p)
    import ctypes
    from cfd_amd import _lib
    for g in (0, 1):
        for w in (0, 1):
            n = ctypes.c_int()
            ch._call("cfd_get_iblist", g, w, None, None, ctypes.byref(n))
            l9 = np.zeros((n.value, 9), dtype=np.int32)
            w3 = np.zeros((n.value, 3))
            ch._call("cfd_get_iblist", g, w, l9.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                     w3.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(n))
            ol, ow = orc.iblist(g, w)
            assert np.array_equal(l9, ol) and np.array_equal(w3, ow)


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("kRK", [1, 2, 3])
def test_rhs0(case, kRK):
    ch, orc = _pair(CASES[case], seed=kRK)
    ch.RHS0(kRK)
    orc.RHS0(kRK)
    _cmp(ch, orc, ("du1", "du2", "du3"), interior=True)


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("kRK", [1, 2])
def test_nonlinear(case, kRK):
    ch, orc = _pair(CASES[case], seed=10 + kRK)
    ch.nonlinear(kRK)
    orc.nonlinear(kRK)
    _cmp(ch, orc, ("Nu1", "Nu2", "Nu3", "du1", "du2", "du3"), interior=True)
    a, b = ch.scalars(), orc.scalars()
    for k in ("t", "dt", "dti", "dtc"):
        assert abs(a[k] - b[k]) <= 1e-13 * abs(b[k]), (k, a[k], b[k])


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("comp", [1, 2, 3])
def test_solve(case, comp):
    ch, orc = _pair(CASES[case], seed=20 + comp)
    if comp == 2:
        ch.solveV(2)
        orc.solveV(2)
        _cmp(ch, orc, ("u2", "du2"))
    else:
        ch.solveU(comp, 2)
        orc.solveU(comp, 2)
        _cmp(ch, orc, ("u%d" % comp, "du%d" % comp))


@pytest.mark.parametrize("case", list(CASES))
def test_meanflow(case):
    ch, orc = _pair(CASES[case], seed=30)
    ch.meanflow_ctU(3)
    orc.meanflow_ctU(3)
    _cmp(ch, orc, ("u1",))
    a, b = ch.scalars(), orc.scalars()
    for k in ("mpgx", "dgx"):
        assert abs(a[k] - b[k]) <= 1e-11 * max(abs(b[k]), 1e-12), (k, a[k], b[k])
    assert abs(orc.flowrate() - b["QxT"]) <= 1e-13 * abs(b["QxT"])


@pytest.mark.parametrize("case", list(CASES))
def test_solve_p_and_vcorr(case):
    ch, orc = _pair(CASES[case], seed=40)
    ch.solveP(1)
    orc.solveP(1)
    _cmp(ch, orc, ("div", "psi", "p"))
    ch.v_corr(1)
    orc.v_corr(1)
    _cmp(ch, orc, ("u1", "u2", "u3"), tol=1e-10)
    # the divergence left by the projection is round-off on both sides: compare
    # it against the velocity scale, not relative to itself
    u = ch.get_field("u1")
    assert np.abs(ch.get_field("div")).max() <= 1e-11 * np.abs(u).max()
    assert np.abs(orc.get_field("div")).max() <= 1e-11 * np.abs(u).max()


@pytest.mark.parametrize("case", list(CASES))
def test_substep(case):
    ch, orc = _pair(CASES[case], seed=50)
    for kRK in (1, 2, 3):
        ch.substep(kRK)
        orc.substep(kRK)
    _cmp(ch, orc, ("u1", "u2", "u3", "p"), tol=1e-11)


@pytest.mark.parametrize("case", ["shipped", "smallposts"])
def test_ten_steps(case):
    """10 RK3 steps from the same seeded projected field: rel-L2 <= 1e-10."""
    from cfd_amd import ic
    p = CASES[case]
    ch = Channel(p)
    orc = Oracle(p)
    f = ic.initial_fields(p, ch.ygrid(), seed=0, amp=0.1)
    f = ic.project(orc, f)
    for s in (ch, orc):
        for k in ("u1", "u2", "u3", "p"):
            s.set_field(k, f[k])
        s.set_scalars(t=0.0, dt=0.0, mpgx=0.0, QxT=ic.flowrate(f["u1"], ch.ygrid(), p.ny))
    ch.step(10)
    orc.step(10)
    _cmp(ch, orc, ("u1", "u2", "u3"), tol=1e-10)
    a, b = ch.scalars(), orc.scalars()
    assert abs(a["t"] - b["t"]) <= 1e-12 * b["t"]
    # divergence at round-off (v_corr leaves the divergence of the new field)
    d = ch.get_field("div")
    assert np.abs(d).max() <= 1e-10 * np.abs(ch.get_field("u1")).max()


@pytest.mark.parametrize("case", list(CASES))
def test_fused_substep_equals_operator_sequence(case):
    """cfd_substep (fused column sweeps) == the reference call sequence of the
    individual entry points, on the GPU.  The fused sweep solves the same
    Helmholtz systems by substructuring instead of one Thomas sweep, so the
    two differ by the round-off of the elimination order: fields to 1e-12
    relative L2; mpgx, which divides a flow-rate difference by a quantity
    proportional to dt (flowrate_corr, littleharsh.f90:452-484), to
    1e-13 * QxT / dt absolute."""
    p = CASES[case]
    a, _ = _pair(p, seed=60)
    b, _ = _pair(p, seed=60)
    for kRK in (1, 2, 3):
        a.substep(kRK)
        b.RHS0(kRK)
        b.nonlinear(kRK)
        b.solveU(1, kRK)
        b.solveV(kRK)
        b.solveU(3, kRK)
        b.meanflow_ctU(kRK)
        b.solveP(kRK)
        b.v_corr(kRK)
    for n in ("u1", "u2", "u3", "p", "Nu1", "Nu2", "Nu3"):
        x, y = a.get_field(n), b.get_field(n)
        if n.startswith("Nu"):
            x, y = x[:, 1:-1], y[:, 1:-1]
        assert rel_l2(x, y) <= 1e-12, (n, rel_l2(x, y))
    sa, sb = a.scalars(), b.scalars()
    for k in ("t", "dt"):
        assert abs(sa[k] - sb[k]) <= 1e-13 * abs(sb[k]), k
    assert abs(sa["mpgx"] - sb["mpgx"]) <= 1e-13 * abs(sb["QxT"]) / sb["dt"], (sa["mpgx"], sb["mpgx"])


CTP_CASES = ("shipped", "smallposts")


def _ctp(case):
    """The case with a constant mean pressure gradient (flag_ctpress = 1, the
    reference's shipped mode, input.in flpress = 1)."""
    import dataclasses
    return dataclasses.replace(CASES[case], flag_ctpress=1)


def _reference_sequence(s, kRK, ctpress):
    """The reference's RK loop body (littleharsh.f90:165-187) operator by
    operator: no meanflow_ctU with a constant pressure gradient."""
    s.RHS0(kRK)
    s.nonlinear(kRK)
    s.solveU(1, kRK)
    s.solveV(kRK)
    s.solveU(3, kRK)
    if not ctpress:
        s.meanflow_ctU(kRK)
    s.solveP(kRK)
    s.v_corr(kRK)


@pytest.mark.parametrize("case", CTP_CASES)
@pytest.mark.parametrize("fused", [True, False])
def test_substep_constant_pressure_gradient(case, fused):
    """flag_ctpress = 1: the fixed mean pressure gradient enters RHS0 of the
    (0,0) u1 column (littleharsh.f90:650-656), meanflow_ctU is skipped and
    meanflow_ctP records the flow rate after the third substep
    (littleharsh.f90:177-196, 504-520).  One RK3 step from identical state,
    GPU (fused substeps or the operator entry points) against the oracle's
    operator sequence: fields to 1e-11, mpgx unchanged, Qx to 1e-13."""
    ch, orc = _pair(_ctp(case), seed=55)
    mpgx0 = orc.scalars()["mpgx"]
    for kRK in (1, 2, 3):
        if fused:
            ch.substep(kRK)
        else:
            _reference_sequence(ch, kRK, True)
        _reference_sequence(orc, kRK, True)
    ch.meanflow_ctP()
    orc.meanflow_ctP()
    _cmp(ch, orc, ("u1", "u2", "u3", "p"), tol=1e-11)
    a, b = ch.scalars(), orc.scalars()
    assert a["mpgx"] == mpgx0 and b["mpgx"] == mpgx0
    assert abs(a["Qx"] - b["Qx"]) <= 1e-13 * abs(b["Qx"]), (a["Qx"], b["Qx"])
    assert abs(a["t"] - b["t"]) <= 1e-13 * abs(b["t"])


@pytest.mark.parametrize("case", CTP_CASES)
def test_ten_steps_constant_pressure_gradient(case):
    """10 RK3 steps (cfd_step: fused substeps and meanflow_ctP after each
    step) with a constant mean pressure gradient of -3/Re (the laminar value
    at unit bulk velocity) from the seeded projected field, against or_step:
    rel-L2 <= 1e-10 per velocity component, the same t and Qx."""
    from cfd_amd import ic
    p = _ctp(case)
    ch = Channel(p)
    orc = Oracle(p)
    f = ic.initial_fields(p, ch.ygrid(), seed=0, amp=0.1)
    f = ic.project(orc, f)
    mpgx = -3.0 / p.Re
    for s in (ch, orc):
        for k in ("u1", "u2", "u3", "p"):
            s.set_field(k, f[k])
        s.set_scalars(t=0.0, dt=0.0, mpgx=mpgx, QxT=0.0)
    ch.step(10)
    orc.step(10)
    _cmp(ch, orc, ("u1", "u2", "u3"), tol=1e-10)
    a, b = ch.scalars(), orc.scalars()
    assert abs(a["t"] - b["t"]) <= 1e-12 * b["t"]
    assert a["mpgx"] == mpgx and b["mpgx"] == mpgx
    assert abs(a["Qx"] - b["Qx"]) <= 1e-12 * abs(b["Qx"]), (a["Qx"], b["Qx"])
    # the flow rate moved: the run is not a trivial fixed point
    assert abs(b["Qx"] - ic.flowrate(f["u1"], ch.ygrid(), p.ny)) > 1e-9
    d = ch.get_field("div")
    assert np.abs(d).max() <= 1e-10 * np.abs(ch.get_field("u1")).max()


def _channel(p, fft=None):
    """A Channel whose FFT kernels are chosen by CFD_AMD_FFT at creation."""
    if fft:
        os.environ["CFD_AMD_FFT"] = fft
    try:
        return Channel(p)
    finally:
        os.environ.pop("CFD_AMD_FFT", None)


@pytest.mark.parametrize("workload", ["rt395_384x192x192", "rt395_coarse_192x192x96"])
def test_four_step_fft_equals_stockham_at_workload_size(workload):
    """The four-step register FFT kernels and the LDS Stockham kernels give
    the same three substeps on the full 384 x 192 x 192 workload and on the
    192 x 192 x 96 closure grid."""
    from cfd_amd import ic, workloads
    p, _ = workloads.get(workload)
    print("case %s four-step vs stockham" % workload, flush=True)
    a = _channel(p)
    b = _channel(p, "stockham")
    f = ic.initial_fields(p, a.ygrid(), seed=3, amp=0.1)
    f = ic.project(a, f)
    q0 = ic.flowrate(f["u1"], a.ygrid(), p.ny)
    for ch in (a, b):
        for k in ("u1", "u2", "u3", "p"):
            ch.set_field(k, f[k])
        ch.set_scalars(t=0.0, dt=0.0, mpgx=0.0, QxT=q0)
    a.nonlinear(1)
    b.nonlinear(1)
    for n in ("Nu1", "Nu2", "Nu3"):
        x, y = a.get_field(n)[:, 1:-1], b.get_field(n)[:, 1:-1]
        assert rel_l2(x, y) <= 1e-13, (n, rel_l2(x, y))
    for ch in (a, b):
        for k in ("u1", "u2", "u3", "p"):
            ch.set_field(k, f[k])
        ch.set_scalars(t=0.0, dt=0.0, mpgx=0.0, QxT=q0)
    for kRK in (1, 2, 3):
        a.substep(kRK)
        b.substep(kRK)
    for n in ("u1", "u2", "u3", "p"):
        x, y = a.get_field(n), b.get_field(n)
        assert rel_l2(x, y) <= 1e-13, (n, rel_l2(x, y))
    sa, sb = a.scalars(), b.scalars()
    for k in ("t", "dt", "mpgx"):
        assert abs(sa[k] - sb[k]) <= 1e-13 * abs(sb[k]), k


def test_substep_768x384_against_oracle():
    """One RK3 step on the 768 x 384 lateral grid (x-pass four-step at 768
    points, Stockham z-forward at 384) against the oracle."""
    ch, orc = _pair(WIDE768, seed=90)
    for kRK in (1, 2, 3):
        ch.substep(kRK)
        orc.substep(kRK)
    _cmp(ch, orc, ("u1", "u2", "u3", "p"), tol=1e-11)


def test_four_step_fft_equals_stockham_768x384():
    a, _ = _pair(WIDE768, seed=91)
    os.environ["CFD_AMD_FFT"] = "stockham"
    try:
        b, _ = _pair(WIDE768, seed=91)
    finally:
        os.environ.pop("CFD_AMD_FFT", None)
    for kRK in (1, 2, 3):
        a.substep(kRK)
        b.substep(kRK)
    # random full-spectrum fields: p comes from a divergence (cancellation), so
    # its round-off bound is one decade looser than the velocities'
    for n, tol in (("u1", 1e-13), ("u2", 1e-13), ("u3", 1e-13), ("p", 1e-12)):
        assert rel_l2(a.get_field(n), b.get_field(n)) <= tol, (n, rel_l2(a.get_field(n), b.get_field(n)))
