"""Run parameters of the rough-channel substep.

``Params`` mirrors the values ``start`` reads from ``input.in``
(start.f90:62-227) plus the post geometry of ``geom_circ_rough.txt``
(geom_circ_rough_mixed_forcing_24pt.f90:113-123).  ngalx/ngalz are the wall
bands' resolution (Ngalx1, Ngalz1); ngalx2/ngalz2 the central band's
(Ngalx2, Ngalz2), 0 for the same resolution in every band.
"""
import dataclasses
import math

from ._lib import CParams, INT_FIELDS, DBL_FIELDS


@dataclasses.dataclass
class Params:
    ngalx: int = 48
    ngalz: int = 48
    ny: int = 25
    dsty: int = 4
    dsty1: int = 0
    ppp: int = 5
    ygrid: int = 0            # 0 polynomial (y_grid_canopy.f90), 1 tanh
    geometry_type: int = 3    # 0 smooth, 3 canopies
    ntilex: int = 2
    ntilez: int = 2
    extras24: int = 0
    flag_ctpress: int = 0     # 0 constant flow rate, 1 constant mpgx
    precision: int = 0
    device: int = 0
    ngalx2: int = 0           # Ngalx2, Ngalz2 of the central y band (input.in:16, 18);
    ngalz2: int = 0           # 0 = as the wall bands (one resolution)
    Re: float = 2800.0
    alp: float = 1.0          # Lx = 2 pi / alp
    bet: float = 2.0          # Lz = 2 pi / bet
    dyq: float = 10.0
    posth: float = 0.089
    posth1: float = 0.0
    ytanh_beta: float = 2.0
    CFL: float = 0.5
    mpgx: float = 0.0
    r: float = 5.43723
    c_i: float = 12.0
    c_k: float = 12.0
    stag: float = 0.5

    def to_c(self):
        c = CParams()
        for n in INT_FIELDS:
            setattr(c, n, int(getattr(self, n)))
        for n in DBL_FIELDS:
            setattr(c, n, float(getattr(self, n)))
        return c

    def replace(self, **kw):
        return dataclasses.replace(self, **kw)

    # derived sizes (start.f90:108-111, 1420-1422; y_grid_canopy.f90:250-290)
    @property
    def N1(self):
        return 2 * (self.ngalx // 3)

    @property
    def N2(self):
        return 2 * (self.ngalz // 3)

    @property
    def Mx(self):
        return self.N1 // 2 + 1

    @property
    def Ncol(self):
        return self.Mx * self.N2

    @property
    def multiband(self):
        """True when the central band is transformed on a coarser grid."""
        return ((self.ngalx2 or self.ngalx) != self.ngalx) or ((self.ngalz2 or self.ngalz) != self.ngalz)

    @property
    def N1c(self):
        return 2 * ((self.ngalx2 or self.ngalx) // 3)

    @property
    def N2c(self):
        return 2 * ((self.ngalz2 or self.ngalz) // 3)

    def short_columns(self):
        """Boolean per column (column = k Mx + i, the wall bands' numbering):
        True for the columns outside the central band's N1c x N2c modes,
        which exist in the wall bands only (proc_lims_columns,
        start.f90:1424-1501)."""
        import numpy as np
        i = np.arange(self.Ncol) % self.Mx
        k = np.arange(self.Ncol) // self.Mx
        kin = (k < self.N2c // 2) | (k >= self.N2 - self.N2c // 2)
        return ~((i <= self.N1c // 2) & kin)

    @property
    def nyu(self):
        return self.ny + 2 * self.dsty + 1

    @property
    def nyv(self):
        return self.ny + 2 * self.dsty

    @property
    def nyp(self):
        return self.ny + 2 * self.dsty - 1

    @property
    def Lx(self):
        return 2 * math.pi / self.alp

    @property
    def Lz(self):
        return 2 * math.pi / self.bet
