"""Ensembles of uneven channels in captured HIP graphs, launch-shape
tuning, and the four-step FFT lengths of the ensemble grids.

* graph replays against eager substeps: two copies of the 16-member
  ensemble from the same start, one stepped by the captured graphs (with
  per-member launches, and with batched launches over the member table),
  one member by member through cfd_step; every field and scalar
  bit-identical after each replay (host reads of every member between
  replays).
* tunings: the same member under every FFT engine / segment shape gives
  the same solution to FP64 rounding (relative L2 <= 1e-11 after 2 steps).
* the segmented y solves at every segment length against the oracle, on
  grids whose column lengths put a separator on the last row.
* the four-step lengths of the ensemble grids (x 96, 128, 144, 256; z 72,
  128) against the oracle, one substep.
"""
import numpy as np
import pytest

from cfd_amd import Channel, ensemble
from cfd_amd._lib import CfdError
from cfd_amd.params import Params

from .helpers import rel_l2, max_rel
from .test_gpu_parity import _cmp, _pair

pytestmark = pytest.mark.gpu

FIELDS = ("u1", "u2", "u3", "p")


def _started(members, tuning=None):
    out = []
    for m in members:
        ch = Channel(m.params, tuning=tuning)
        ensemble.start(ch, m)
        out.append(ch)
    return out


def _same(a, b, what):
    for n in FIELDS:
        x, y = a.get_field(n), b.get_field(n)
        assert np.array_equal(x, y), "%s: %s differs (max rel %.3e)" % (what, n, max_rel(x, y))
    sa, sb = a.scalars(), b.scalars()
    for k in sa:
        assert sa[k] == sb[k], "%s: scalar %s %r != %r" % (what, k, sa[k], sb[k])


@pytest.mark.parametrize("batched,count", [(False, 3), (True, 16)])
def test_graph_stepped_members_against_oracle(batched, count):
    """Members of the 16-member set (per-member launches: the three smallest,
    one of them the 96 x 96 x 64 end member; batched launches over the member
    table: all sixteen), one eager substep, then kRK = 2 and 3 and one full
    RK3 step from the captured graphs, against the oracle's substeps from the
    same state: relative L2 <= 1e-11 per field, the same t and mean pressure
    gradient."""
    from oracle.oracle import Oracle, set_threads
    import os
    set_threads(min(16, len(os.sched_getaffinity(0))))
    ms = sorted(ensemble.members(16, seed=1),
                key=lambda m: m.params.ngalx * m.params.ny * m.params.ngalz)[:count]
    print("%d members, graphs vs oracle, batched=%s:" % (count, batched),
          [(m.params.ngalx, m.params.ny, m.params.ngalz) for m in ms], flush=True)
    A = _started(ms)
    orcs = []
    for m, ch in zip(ms, A):
        o = Oracle(m.params)
        for n in FIELDS:
            o.set_field(n, ch.get_field(n))
        o.set_scalars(**{k: v for k, v in ch.scalars().items()})
        orcs.append(o)
    for ch, o in zip(A, orcs):
        ch.substep(1)  # the eager warm-up substep the graphs need
        o.substep(1)
        ch.sync()
    with ensemble.Ensemble(A, batched=batched) as ens:
        for kRK in (2, 3):
            ens.substep(kRK)
        ens.step(1)
        ens.sync()
    for o, m in zip(orcs, ms):
        for kRK in (2, 3, 1, 2, 3):
            o.substep(kRK)
        assert m.params.flag_ctpress == 0
    for i, (ch, o) in enumerate(zip(A, orcs)):
        for n in FIELDS:
            e = rel_l2(ch.get_field(n), o.get_field(n))
            assert e <= 1e-11, "member %d %s rel-L2 %.3e" % (i, n, e)
        sa, sb = ch.scalars(), o.scalars()
        assert abs(sa["t"] - sb["t"]) <= 1e-13 * sb["t"]
        assert abs(sa["mpgx"] - sb["mpgx"]) <= 1e-13 * abs(sb["QxT"]) / sb["dt"]
        ch.close()
        o.close()


@pytest.mark.parametrize("batched", [False, True])
def test_graph_replays_equal_eager_substeps(batched):
    """Both graph forms (per-member launches; batched launches over the member
    table) against the same members stepped eagerly, bit for bit."""
    ms = ensemble.members(16, seed=1)
    print("16 members, graph vs eager, batched=%s" % batched, flush=True)
    A, B = _started(ms), _started(ms)
    for ch in A + B:
        ch.step(1)  # warm-up substeps, eager on both sides
    for ch in A + B:
        ch.sync()
    dt1 = [ch.scalars()["dt"] for ch in A]
    with ensemble.Ensemble(A, batched=batched) as ens:
        for rep in range(3):
            ens.step(1)
            ens.sync()
            for ch in B:
                ch.step(1)
                ch.sync()
            for i, (a, b) in enumerate(zip(A, B)):
                _same(a, b, "replay %d member %d" % (rep, i))
        # substep by substep, with eager host work between the launches
        for kRK in (1, 2, 3):
            ens.substep(kRK)
            ens.sync()
            for ch in B:
                ch.substep(kRK)
                ch.sync()
            for i, (a, b) in enumerate(zip(A, B)):
                _same(a, b, "substep %d member %d" % (kRK, i))
    # no member blows up (its CFL time step would collapse within a few
    # steps, as it does for posts whose forcing stencils extrapolate)
    for i, (ch, d1) in enumerate(zip(A, dt1)):
        assert 0.2 * d1 <= ch.scalars()["dt"] <= 5.0 * d1, (i, d1, ch.scalars()["dt"])
    # the members run on by themselves after the ensemble is gone
    A[0].step(1)
    B[0].step(1)
    _same(A[0], B[0], "after the ensemble")
    for ch in A + B:
        ch.close()


@pytest.mark.parametrize("batched", [False, True])
def test_mixed_members_graph_replays_equal_eager_substeps(batched):
    """The sixteen members on the mixed path (FP32 transforms and products,
    the batched FP32 kernels): both graph forms against the same members
    stepped eagerly, bit for bit, and close to the FP64 members at the
    FP32 level (relative L2 <= 1e-5 after two RK3 steps)."""
    import dataclasses
    ms = ensemble.members(16, seed=1)
    mm = [dataclasses.replace(m, params=dataclasses.replace(m.params, precision=1)) for m in ms]
    print("16 mixed members, graph vs eager, batched=%s" % batched, flush=True)
    A, B, C = _started(mm), _started(mm), _started(ms)
    for ch in A + B + C:
        ch.step(1)
    for ch in A + B + C:
        ch.sync()
    with ensemble.Ensemble(A, batched=batched) as ens:
        ens.step(1)
        ens.sync()
    for ch in B + C:
        ch.step(1)
        ch.sync()
    for i, (a, b, c) in enumerate(zip(A, B, C)):
        _same(a, b, "mixed member %d" % i)
        for n in ("u1", "u2", "u3"):
            e = rel_l2(a.get_field(n), c.get_field(n))
            assert 0.0 < e <= 1e-5, "member %d %s rel-L2 vs FP64 %.3e" % (i, n, e)
    for ch in A + B + C:
        ch.close()


def test_ensemble_refusals():
    m = ensemble.members(2, seed=1)
    a, b = _started(m)
    with pytest.raises(CfdError, match="one substep"):
        ensemble.Ensemble([a, b])
    a.step(1)
    b.step(1)
    with pytest.raises(CfdError, match="twice"):
        ensemble.Ensemble([a, a])
    b.profile(True)
    with pytest.raises(CfdError, match="profiling"):
        ensemble.Ensemble([a, b])
    b.profile(False)
    with ensemble.Ensemble([a, b]) as e:
        e.step(1)
        e.sync()
        with pytest.raises(CfdError, match="kRK"):
            e.substep(4)
    # the batched kernels are four-step transforms: a Stockham member is
    # refused, by name
    c = Channel(m[1].params, tuning=dict(fft=1))
    ensemble.start(c, m[1])
    c.step(1)
    with pytest.raises(CfdError, match="member 1: .*four-step"):
        ensemble.Ensemble([a, c], batched=True)
    with ensemble.Ensemble([a, c]) as e:
        e.step(1)
        e.sync()
    a.close()
    b.close()
    c.close()


@pytest.mark.parametrize("which", [0, 15])
def test_every_tuning_gives_the_same_solution(which):
    m = ensemble.members(16, seed=1)[which]
    print("member", which, m.params.ngalx, m.params.ny, m.params.ngalz, flush=True)
    ref = _started([m])[0]
    assert ref.tuning() == dict(fft=0, seg_rows=13, solve_threads=256)
    ref.step(2)
    want = {n: ref.get_field(n) for n in FIELDS}
    for t in (dict(fft=1), dict(fft=0, seg_rows=5, solve_threads=64), dict(fft=1, seg_rows=9, solve_threads=128),
              dict(fft=0, seg_rows=21, solve_threads=256)):
        ch = _started([m], tuning=t)[0]
        assert all(ch.tuning()[k] == v for k, v in t.items())
        ch.step(2)
        for n in FIELDS:
            e = rel_l2(ch.get_field(n), want[n])
            assert e <= 1e-11, "tuning %s: %s rel-L2 %.3e" % (t, n, e)
        ch.close()
    ref.close()


def test_fft_engine_is_fixed_at_creation():
    m = ensemble.members(1, seed=1)[0]
    ch = Channel(m.params, tuning=dict(fft=1, seg_rows=9))
    assert ch.tuning() == dict(fft=1, seg_rows=9, solve_threads=256)
    ch.set_tuning(seg_rows=21, solve_threads=64)
    assert ch.tuning() == dict(fft=1, seg_rows=21, solve_threads=64)
    with pytest.raises(CfdError, match="fixed at creation"):
        ch.set_tuning(fft=0)
    with pytest.raises(CfdError, match="seg_rows"):
        ch.set_tuning(seg_rows=7)
    ch.close()


def test_tuner_times_every_choice():
    m = ensemble.members(1, seed=1)[0]
    best, table = ensemble.tune(m, nsteps=1, repeats=1)
    assert len(table) == len(ensemble.FFT_CHOICES) * len(ensemble.SEG_ROWS) * len(ensemble.SOLVE_THREADS)
    assert all(ms > 0 for _, ms in table)
    assert best == min(table, key=lambda e: e[1])[0]
    Channel(m.params, tuning=best).close()


def _seg_grids(R):
    """Smooth grids whose u columns (ny + 1 rows) and p columns (ny - 1
    rows) are whole multiples of R + 1 rows: a separator on the last row."""
    out = []
    for ny in (2 * (R + 1) - 1, 2 * (R + 1) + 1):
        out.append(Params(ngalx=24, ngalz=24, ny=ny, dsty=0, ygrid=0, geometry_type=0, ntilex=2, ntilez=2,
                          Re=1500.0, alp=1.0, bet=2.0, posth=0.0, CFL=0.5))
    return out


@pytest.mark.parametrize("R", [5, 9, 21])
@pytest.mark.parametrize("threads", [64, 128])
def test_segment_shapes_against_oracle(R, threads):
    for p in _seg_grids(R) + [Params(ngalx=72, ngalz=48, ny=33, dsty=3, ygrid=1, ytanh_beta=2.0, geometry_type=3,
                                     ntilex=6, ntilez=4, Re=2800.0, alp=1.0, bet=2.0, posth=0.06, r=1.6, c_i=6.0,
                                     c_k=6.0, stag=0.5, CFL=0.5)]:
        ch, orc = _pair(p, seed=60 + R)
        ch.set_tuning(seg_rows=R, solve_threads=threads)
        for kRK in (1, 2, 3):
            ch.substep(kRK)
            orc.substep(kRK)
        _cmp(ch, orc, ("u1", "u2", "u3", "p"))
        ch.close()
        orc.close()


def _fft_grid(nx, nz):
    ntx, ntz = 4, 2
    dnx, dnz = nx // ntx, nz // ntz
    return Params(ngalx=nx, ngalz=nz, ny=24, dsty=3, ygrid=1, ytanh_beta=2.0, geometry_type=3, ntilex=ntx,
                  ntilez=ntz, Re=5000.0, alp=1.0, bet=nx / nz, posth=0.12, r=1.2, c_i=(dnx + 1) / 2.0,
                  c_k=(dnz + 1) / 2.0, stag=0.5, CFL=0.5)


@pytest.mark.parametrize("nx,nz", [(96, 72), (128, 128), (144, 72), (256, 64)])
def test_ensemble_fft_lengths_against_oracle(nx, nz):
    p = _fft_grid(nx, nz)
    ch, orc = _pair(p, seed=70)
    assert ch.tuning()["fft"] == 0
    ch.nonlinear(1)
    orc.nonlinear(1)
    _cmp(ch, orc, ("Nu1", "Nu2", "Nu3"), interior=True)
    ch.close()
    orc.close()
    ch, orc = _pair(p, seed=71)
    for kRK in (1, 2, 3):
        ch.substep(kRK)
        orc.substep(kRK)
    _cmp(ch, orc, ("u1", "u2", "u3", "p"))
    ch.close()
    orc.close()
