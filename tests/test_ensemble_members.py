"""The uneven ensemble's members (host side, no GPU): ranges, determinism,
valid y grids and post lists; the tuning argument checks of the C ABI."""
import ctypes

import numpy as np
import pytest

from cfd_amd import _lib, ensemble
from cfd_amd.channel import _tuning_c, build_iblist, build_ygrid
from cfd_amd.params import Params

from .helpers import SMALLPOSTS


@pytest.fixture(scope="module")
def ms():
    return ensemble.members(16, seed=1)


def test_ranges(ms):
    assert len(ms) == 16
    p0, p1 = ms[0].params, ms[-1].params
    assert (p0.ngalx, p0.ny, p0.ngalz) == (96, 96, 64)
    assert (p1.ngalx, p1.ny, p1.ngalz) == (256, 160, 128)
    shapes = set()
    for m in ms:
        p = m.params
        assert 96 <= p.ngalx <= 256 and 96 <= p.ny <= 160 and 64 <= p.ngalz <= 128
        assert 180 <= m.re_tau <= 395 and 8 <= m.dplus <= 20 and 8 <= m.kplus <= 20
        assert (p.ntilex, p.ntilez) in ensemble.ARRAYS
        assert p.ngalx % p.ntilex == 0 and p.ngalz % p.ntilez == 0
        assert abs(p.posth * m.re_tau - m.kplus) < 1e-12 * m.kplus
        # square cells in x-z: Lx / Ngalx == Lz / Ngalz
        assert abs((2 * np.pi / p.alp) / p.ngalx - (2 * np.pi / p.bet) / p.ngalz) < 1e-12
        # the post diameter simulated is the one drawn unless the floor
        # applies or the drawn radius has extrapolating forcing stencils
        dxp = 2 * np.pi / p.alp * m.re_tau / p.ngalx
        assert abs(2 * p.r * dxp - m.dplus_sim) < 1e-9
        r0 = max(0.5 * m.dplus / dxp, ensemble.R_MIN)
        assert ensemble.ib_max_weight(p) <= 1.0
        if ensemble.ib_max_weight(p.replace(r=r0)) <= 1.0:
            assert p.r == pytest.approx(r0, rel=1e-14)
        else:
            assert p.r != r0 and abs(p.r - r0) < 0.2
        shapes.add((p.ngalx, p.ny, p.ngalz))
    assert len(shapes) >= 12  # uneven


def test_deterministic(ms):
    again = ensemble.members(16, seed=1)
    assert [m.params for m in again] == [m.params for m in ms]
    other = ensemble.members(16, seed=2)
    assert [m.params for m in other][1:-1] != [m.params for m in ms][1:-1]


def test_y_grids_monotonic_and_posts_listed(ms):
    for m in ms:
        p = m.params
        g = build_ygrid(p)
        assert not ensemble.y_folds(g)
        assert np.all(np.diff(g["yu"]) > 0) and np.all(np.diff(g["yv"]) > 0)
        for grid in (0, 1):
            for wall in (0, 1):
                l9, w3 = build_iblist(p, grid, wall)
                assert len(l9) > 0


def test_extrapolating_radius_is_moved():
    """A post centred between grid points with r just above sqrt(2) cells:
    the forced points at distance sqrt(2.5) take two fluid neighbours nearly
    in line with their boundary point, and lin_interp's weights extrapolate
    (450 at r = 1.415, 4.3 at r = 1.507); the member draw moves such a radius
    to the nearest one whose stencils interpolate."""
    p = ensemble.members(16, seed=1)[-1].params
    assert (p.ngalx // p.ntilex, p.ngalz // p.ntilez, p.c_i, p.c_k) == (16, 16, 8.5, 8.5)
    assert ensemble.ib_max_weight(p.replace(r=1.415)) > 100.0
    assert ensemble.ib_max_weight(p.replace(r=1.507)) > 4.0
    assert ensemble.ib_max_weight(p.replace(r=1.41)) < 0.1
    q = ensemble._stable_radius(p.replace(r=1.507))
    assert q.r == pytest.approx(1.582) and ensemble.ib_max_weight(q) <= 1.0
    # every member of a few other draws is stable by this measure
    for seed in (2, 3):
        for m in ensemble.members(16, seed=seed):
            assert ensemble.ib_max_weight(m.params) <= 1.0


def test_fold_check_catches_a_folded_grid():
    """posth = 0 with dsty = 3 extends the canopy grid back on itself."""
    g = build_ygrid(SMALLPOSTS.replace(geometry_type=0, posth=0.0))
    assert ensemble.y_folds(g)
    assert not ensemble.y_folds(build_ygrid(SMALLPOSTS))


def test_tuning_arguments_are_checked():
    lib = _lib.load()
    c = Params().to_c()
    h = ctypes.c_void_p()
    for t in (dict(fft=2), dict(seg_rows=7), dict(solve_threads=512)):
        rc = lib.cfd_create_tuned(ctypes.byref(c), _tuning_c(t), ctypes.byref(h))
        assert rc == 1 and not h.value
        assert b"tuning" in lib.cfd_last_error(None)
    with pytest.raises(ValueError):
        _tuning_c(dict(engine=1))
    assert list(_tuning_c({})) == [0, 13, 256]
