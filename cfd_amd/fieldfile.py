"""Field files in the reference layout without a GPU (the handle-free
``cfd_field_file_write`` / ``cfd_field_file_read`` of the C ABI).

A field file (record_out, FOU3D.f90:2172-2345) holds one spectral field as
Fortran sequential unformatted records: a header (t, Re, alp, bet, mpgx,
nband, iter), the band table N(4, 0:nband+1), the y grid with its metric, and
one record per y-plane with the (Re, Im) pairs of the Mx x-modes for each of
the N2 z-modes in column order.  ``planes`` here is an array (ny_planes, N2,
Mx) complex128: plane-major, the device layout of a field.
"""
import ctypes
import os

import numpy as np

from . import _lib

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)


def write(path, planes, y, dthdy, *, field_id, j0, jy0, t=0.0, Re=1.0, alp=1.0, bet=1.0, mpgx=0.0, iter=0,
          N=None, nband=3, dthetavi=0.0):
    planes = np.ascontiguousarray(planes, dtype=np.complex128)
    ny_g, nz, mx = planes.shape
    y = np.ascontiguousarray(y, dtype=np.float64)
    dthdy = np.ascontiguousarray(dthdy, dtype=np.float64)
    hd = _lib.CFileHeader(t, Re, alp, bet, mpgx, nband, iter)
    for q, v in enumerate(N if N is not None else [0] * 20):
        hd.N[q] = int(v)
    rc = _lib.load().cfd_field_file_write(os.fsencode(path), ctypes.byref(hd), int(field_id), float(dthetavi),
                                          y.ctypes.data_as(_D), dthdy.ctypes.data_as(_D), len(y), int(jy0), ny_g,
                                          int(j0), 2 * mx, nz, planes.view(np.float64).ctypes.data_as(_D))
    if rc != 0:
        raise _lib.CfdError("cfd_field_file_write failed for %s" % path)


def read(path, ny_planes, n2, mx, ngz):
    """Returns (header dict, planes (ny_planes, n2, mx) complex128, number of
    planes whose stored size differed)."""
    hd = _lib.CFileHeader()
    planes = np.zeros((ny_planes, n2, mx), dtype=np.complex128)
    mm = ctypes.c_int(0)
    rc = _lib.load().cfd_field_file_read(os.fsencode(path), ctypes.byref(hd), int(ny_planes), 2 * mx, int(n2),
                                         int(ngz), planes.view(np.float64).ctypes.data_as(_D), ctypes.byref(mm))
    if rc != 0:
        raise _lib.CfdError("cfd_field_file_read failed for %s" % path)
    h = {k: getattr(hd, k) for k in ("t", "Re", "alp", "bet", "mpgx", "nband", "iter")}
    h["N"] = list(hd.N)
    return h, planes, mm.value


def band_table(params):
    """N(1:4, 0:4) (column-major, 20 ints) of a configuration, as record_out
    writes it."""
    c = params.to_c()
    out = (ctypes.c_int * 20)()
    rc = _lib.load().cfd_band_table(ctypes.byref(c), out)
    if rc != 0:
        raise _lib.CfdError("cfd_band_table failed")
    return list(out)


def plane_partition(params, np_ranks):
    """The reference's plane partition over np_ranks >= 3 ranks
    (start.f90:1652-1866): {"u": [(first, last), ...], "v": [...]} per rank,
    planes without walls in the reference's y indices."""
    c = params.to_c()
    out = (ctypes.c_int * (4 * np_ranks))()
    rc = _lib.load().cfd_plane_partition(ctypes.byref(c), int(np_ranks), out)
    if rc != 0:
        raise _lib.CfdError("cfd_plane_partition failed")
    lim = list(out)
    return {g: [(lim[(k * 2) * np_ranks + r], lim[(k * 2 + 1) * np_ranks + r]) for r in range(np_ranks)]
            for k, g in enumerate(("u", "v"))}
