"""Host-side pieces of the learned closure (no GPU): the prolongation of
coarse fields onto the fine modes used by the full-DNS fallback."""
import numpy as np


def test_prolong_inverts_the_mode_selection():
    """filtered_start(prolong(c)) = H c: zero-padding the coarse modes onto
    the fine grid and filtering back returns them times the box filter, and
    every fine mode outside the coarse set is zero."""
    import dataclasses
    from cfd_amd import closure
    from cfd_amd.params import Params
    fine = Params(ngalx=96, ngalz=48, ny=20, dsty=0, geometry_type=0)
    coarse = dataclasses.replace(fine, ngalx=48, ngalz=24)
    N1c, N2c = 2 * (coarse.ngalx // 3), 2 * (coarse.ngalz // 3)
    Mxc = N1c // 2 + 1
    rng = np.random.default_rng(3)
    fc = {"u1": rng.standard_normal((N2c * Mxc, 21)) + 1j * rng.standard_normal((N2c * Mxc, 21))}
    ff = closure.prolong(fine, fc, coarse)
    N1f, N2f = 2 * (fine.ngalx // 3), 2 * (fine.ngalz // 3)
    Mxf = N1f // 2 + 1
    assert ff["u1"].shape == (N2f * Mxf, 21)
    back = closure.filtered_start(coarse, ff, fine)["u1"]
    kz = np.arange(N2c)
    kz = np.where(kz < N2c // 2, kz, kz - N2c)
    h = (np.sinc(np.arange(Mxc) / coarse.ngalx)[None, :] * np.sinc(np.abs(kz) / coarse.ngalz)[:, None]).reshape(-1)
    assert np.array_equal(back, fc["u1"] * h[:, None])
    # the fine modes outside the coarse set are zero: as many nonzero columns as coarse ones
    assert np.count_nonzero(np.abs(ff["u1"]).sum(axis=1)) == N2c * Mxc
