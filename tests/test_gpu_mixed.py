"""The mixed-precision path (precision = 1): FP32 z/x transforms, products and
nonlinear scratch; FP64 fields, RHS, y-solves, projection and flow-rate
correction.  Compared with the FP64 path on the GPU and with the FP64 oracle.

The nonlinear term is formed with ~1e-7 relative error and enters the
velocity through dt*gRK*Nu, so the tolerances below (1e-6 relative L2 after
one RK3 step, 1e-5 after ten) are FP32-level bounds; the projection stays FP64,
so the divergence is checked at the FP64 bound.
"""
import dataclasses

import numpy as np
import pytest

from cfd_amd import Channel, ic
from cfd_amd._lib import CfdError
from oracle.oracle import Oracle

from .helpers import COARSE, SHIPPED, SMALLPOSTS, SMOOTH, WIDE, WIDE768, rel_l2
from .test_gpu_parity import _pair

pytestmark = pytest.mark.gpu

CASES = {"shipped": SHIPPED, "smallposts": SMALLPOSTS, "wide": WIDE, "wide768": WIDE768, "coarse": COARSE}


def _mixed(p):
    return dataclasses.replace(p, precision=1)


@pytest.mark.parametrize("case", list(CASES))
def test_mixed_nonlinear_close_to_fp64(case):
    p = CASES[case]
    a, _ = _pair(p, seed=70)
    b, _ = _pair(_mixed(p), seed=70)
    a.nonlinear(2)
    b.nonlinear(2)
    for n in ("Nu1", "Nu2", "Nu3"):
        x, y = a.get_field(n)[:, 1:-1], b.get_field(n)[:, 1:-1]
        e = rel_l2(y, x)
        assert e <= 1e-5, (n, e)
        assert e > 0.0, "mixed path identical to FP64: FP32 kernels not in use"


@pytest.mark.parametrize("case", list(CASES))
def test_mixed_step_close_to_fp64(case):
    p = CASES[case]
    a, _ = _pair(p, seed=71)
    b, _ = _pair(_mixed(p), seed=71)
    for kRK in (1, 2, 3):
        a.substep(kRK)
        b.substep(kRK)
    for n in ("u1", "u2", "u3"):
        assert rel_l2(b.get_field(n), a.get_field(n)) <= 1e-6, n
    u = b.get_field("u1")
    assert np.abs(b.get_field("div")).max() <= 1e-11 * np.abs(u).max()


@pytest.mark.parametrize("case", ["shipped", "smallposts"])
def test_mixed_ten_steps_against_oracle(case):
    """10 RK3 steps from the same projected field: FP32-level agreement with
    the FP64 oracle, FP64-level divergence, same time and flow rate."""
    p = CASES[case]
    ch = Channel(_mixed(p))
    orc = Oracle(p)
    f = ic.initial_fields(p, ch.ygrid(), seed=0, amp=0.1)
    f = ic.project(orc, f)
    q0 = ic.flowrate(f["u1"], ch.ygrid(), p.ny)
    for s in (ch, orc):
        for k in ("u1", "u2", "u3", "p"):
            s.set_field(k, f[k])
        s.set_scalars(t=0.0, dt=0.0, mpgx=0.0, QxT=q0)
    ch.step(10)
    orc.step(10)
    for n in ("u1", "u2", "u3"):
        assert rel_l2(ch.get_field(n), orc.get_field(n)) <= 1e-5, n
    a, b = ch.scalars(), orc.scalars()
    assert abs(a["t"] - b["t"]) <= 1e-6 * b["t"]
    assert abs(ic.flowrate(ch.get_field("u1"), ch.ygrid(), p.ny) - q0) <= 1e-12 * abs(q0)
    assert np.abs(ch.get_field("div")).max() <= 1e-10 * np.abs(ch.get_field("u1")).max()


def test_mixed_rejects_lengths_without_four_step_kernels():
    with pytest.raises(CfdError):
        Channel(_mixed(SMOOTH))
