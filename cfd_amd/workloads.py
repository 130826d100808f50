"""Named workloads (grid, physics and post geometry) of this project.

Bulk Reynolds numbers follow Dean's correlation Re_tau ~ 0.09 Re_2h^0.88 for
the target friction Reynolds number.  Post radii are in grid units of a tile:
r = (D+/2) / dx+ with dx+ = Lx Re_tau / Ngalx.  The tanh map with beta = 2
spans the tips; the canopy grid below them is the reference's quintic.
"""
import math

from .params import Params


def _rb(re_tau):
    return 0.5 * (re_tau / 0.09) ** (1.0 / 0.88)


def _radius(Dplus, re_tau, ngalx, Lx=2 * math.pi):
    return 0.5 * Dplus / (Lx * re_tau / ngalx)


WORKLOADS = {
    # 64 x 96 x 64, Re_tau 180, 8 x 4 posts per wall.  The posts of D+ = 12
    # are 0.68 dx wide, below one grid cell; r = 1.2 cells is the smallest
    # post resolved by a 5-point cross and is used instead.
    "rt180_64x96x64": dict(
        params=Params(ngalx=64, ngalz=64, ny=96, dsty=2, ygrid=1, ytanh_beta=2.0, Re=round(_rb(180.0)),
                      alp=1.0, bet=2.0, posth=12.0 / 180.0, ntilex=8, ntilez=4, r=1.2, c_i=4.0, c_k=8.0,
                      stag=0.5, extras24=0, CFL=0.5, flag_ctpress=0),
        steps=100, re_tau=180.0),
    # 384 x 192 x 192, Re_tau 395, 16 x 8 posts, D+ = k+ = 16
    "rt395_384x192x192": dict(
        params=Params(ngalx=384, ngalz=192, ny=192, dsty=8, ygrid=1, ytanh_beta=2.0, Re=round(_rb(395.0)),
                      alp=1.0, bet=2.0, posth=16.0 / 395.0, ntilex=16, ntilez=8,
                      r=_radius(16.0, 395.0, 384), c_i=12.0, c_k=12.0, stag=0.5, extras24=0, CFL=0.5,
                      flag_ctpress=0),
        steps=200, re_tau=395.0),
    # the 384 x 192 x 192 case on a 2x coarser x/z grid (the learned
    # closure's grid): posts of the same diameter (r = 0.62 cells), centres a
    # quarter cell off the grid points so that no forcing point sits on a
    # centre
    "rt395_coarse_192x192x96": dict(
        params=Params(ngalx=192, ngalz=96, ny=192, dsty=8, ygrid=1, ytanh_beta=2.0, Re=round(_rb(395.0)),
                      alp=1.0, bet=2.0, posth=16.0 / 395.0, ntilex=16, ntilez=8,
                      r=_radius(16.0, 395.0, 192), c_i=6.25, c_k=6.25, stag=0.5, extras24=0, CFL=0.5,
                      flag_ctpress=0),
        steps=5000, re_tau=395.0),
    # the 384 x 192 x 192 case with a coarser central band (multi-band
    # resolution): 192 x 96 points in x/z between the wall bands, whose
    # height is two post spacings in z (17 of the 209 u planes central)
    "rt395_mb_384x192x192_c192x96": dict(
        params=Params(ngalx=384, ngalz=192, ny=192, dsty=8, ygrid=1, ytanh_beta=2.0, Re=round(_rb(395.0)),
                      alp=1.0, bet=2.0, posth=16.0 / 395.0, ntilex=16, ntilez=8,
                      r=_radius(16.0, 395.0, 384), c_i=12.0, c_k=12.0, stag=0.5, extras24=0, CFL=0.5,
                      flag_ctpress=0, ngalx2=192, ngalz2=96),
        steps=200, re_tau=395.0),
    # 768 x 384 x 384, Re_tau 550, 32 x 16 posts, D+ = 16, k+ = 20
    "rt550_768x384x384": dict(
        params=Params(ngalx=768, ngalz=384, ny=384, dsty=16, ygrid=1, ytanh_beta=2.0, Re=round(_rb(550.0)),
                      alp=1.0, bet=2.0, posth=20.0 / 550.0, ntilex=32, ntilez=16,
                      r=_radius(16.0, 550.0, 768), c_i=12.0, c_k=12.0, stag=0.5, extras24=0, CFL=0.5,
                      flag_ctpress=0),
        steps=100, re_tau=550.0),
}


def get(name):
    w = WORKLOADS[name]
    return w["params"], w["steps"]
