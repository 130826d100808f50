"""Readers for the statistics and spectrum files (write_stats,
stats.f90:255-442; write_spect / recvwrspec, spectra.f90:62-182).

Both are Fortran sequential unformatted files (4-byte little-endian record
markers).  Records:

stats_<N1>x<N2>x<ny>_<int(10t)>.dat
    1  t, Re, alp, bet, mpgx (real*8), nband, iter (int*4), 88 int*4 zeros
    2  Ngal(4, 0:4) (int*4)
    3  yu, dthetai, dthdyu       4  yv, dthetai, dthdyv
    5  istat
    6-18   Um U2m Vm V2m Wm W2m Pm P2m UVm UWm VWm wxm wx2m, each x(j) over
           the planes of its grid
    19-31  the same quantities' tile-conditional sums xC(dnx, dnz, 0:nyc-1),
           bottom and top wall added, j1 = distance in planes from the wall

sp{U,V,W,UV,P}_<N1>x<N2>x<ny>_<int(10t)>.dat
    1-5 as above but record 2 is N(4, 0:4)
    6  buffSP(0:N1/2, 1:N2/2+1, planes): sum over samples of |u|^2 of the
       column (kx, +-kz), every kx > 0 doubled
"""
import struct

import numpy as np

STATS = ("Um", "U2m", "Vm", "V2m", "Wm", "W2m", "Pm", "P2m", "UVm", "UWm", "VWm", "wxm", "wx2m")
CONDS = ("UmC", "U2mC", "VmC", "V2mC", "WmC", "W2mC", "PmC", "P2mC", "UVmC", "UWmC", "VWmC", "wxmC", "wx2mC")


def records(path):
    """Payloads of every record of a Fortran sequential unformatted file."""
    out = []
    with open(path, "rb") as fh:
        while True:
            head = fh.read(4)
            if not head:
                break
            (n,) = struct.unpack("<i", head)
            payload = fh.read(n)
            (m,) = struct.unpack("<i", fh.read(4))
            if m != n or len(payload) != n:
                raise ValueError("broken record framing in %s" % path)
            out.append(payload)
    return out


def _header(recs, nyu):
    t, Re, alp, bet, mpgx = struct.unpack_from("<5d", recs[0], 0)
    nband, it = struct.unpack_from("<2i", recs[0], 40)
    tab = list(struct.unpack("<20i", recs[1]))
    g3 = np.frombuffer(recs[2], dtype="<f8")
    g4 = np.frombuffer(recs[3], dtype="<f8")
    nyv = (len(g4) - 1) // 2
    (istat,) = struct.unpack("<i", recs[4])
    return dict(t=t, Re=Re, alp=alp, bet=bet, mpgx=mpgx, nband=nband, iter=it, table=tab,
                yu=g3[:nyu].copy(), dthetai_u=float(g3[nyu]), dthdyu=g3[nyu + 1:].copy(),
                yv=g4[:nyv].copy(), dthetai_v=float(g4[nyv]), dthdyv=g4[nyv + 1:].copy(), istat=istat)


def read_stats(path, dnx, dnz):
    """Header dict (with Ngal as 'table'), profiles {name: (ny,)}, conditional
    sums {name: (nyc, dnz, dnx)}."""
    recs = records(path)
    if len(recs) != 5 + 26:
        raise ValueError("%s: %d records, a stats file has 31" % (path, len(recs)))
    nyu = (len(recs[2]) // 8 - 1) // 2
    h = _header(recs, nyu)
    prof = {n: np.frombuffer(recs[5 + q], dtype="<f8").copy() for q, n in enumerate(STATS)}
    cond = {}
    for q, n in enumerate(CONDS):
        a = np.frombuffer(recs[18 + q], dtype="<f8")
        if a.size % (dnx * dnz):
            raise ValueError("%s: conditional record %s does not hold whole %dx%d tiles" % (path, n, dnx, dnz))
        cond[n] = a.reshape(-1, dnz, dnx).copy()
    return h, prof, cond


def read_spectrum(path):
    """Header dict (with N as 'table') and buffSP as (planes, N2/2+1, N1/2+1)
    with N1, N2 of the central band (the midband, whose columns the spectra
    cover)."""
    recs = records(path)
    if len(recs) != 6:
        raise ValueError("%s: %d records, a spectrum file has 6" % (path, len(recs)))
    nyu = (len(recs[2]) // 8 - 1) // 2
    h = _header(recs, nyu)
    N1, N2 = h["table"][8], h["table"][9]  # the midband's modes (spectra.f90:80-82)
    sp = np.frombuffer(recs[5], dtype="<f8").reshape(-1, N2 // 2 + 1, N1 // 2 + 1).copy()
    return h, sp
