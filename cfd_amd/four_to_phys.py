"""Post-processing: spectral field files -> physical field files
(four_to_phys.f90:1-326), the transforms on the GPU (cfd_four_to_phys).

    python -m cfd_amd.four_to_phys [input_FtP.txt]

reads, as the reference program does, input_FtP.txt (nx, nz, ny as
(7X,I10); t as (7X,D10.1); the input directory as (8X,A30); the output
directory as (7X,A30)), opens <dir>/<f>_<nx>x<nz>x<ny>_t<int(10 t)>.dat for f in
u1, u2, u3, p and writes <dirout>/<f>_phys_<same>.dat: the header record
copied, the band table with Ngal = 3/2 N in bands 1-3, the grid record copied,
then one record per plane j, id, Ngalx+2, Ngalz, y(j), buff(Ngalx+2, Ngalz)
holding the physical values (x fastest, two trailing zeros per row).  The
plane sequence is the reference's, band by band: u/w and v files repeat the
plane at each of the two band interfaces (the u1 and u3 repeats carry id 1,
the u2 repeats the id of the plane), the p file has every plane once.  Each
band's planes are transformed on that band's grid, Ngal(1:2, b) = 3/2 N(1:2,
b) (multi-band files: the central band's planes are stored and transformed
at its own size); a plane is first brought to its band's mode counts by the
reference's resize (four_to_phys.f90:328-348), which keeps min(nz, nz_old)/2
z-modes from the bottom and one fewer from the top, and so also drops the
kz = -N2/2 mode; the interface repeat of a band is the previous band's last
plane, resized.
"""
import os
import struct
import sys

import numpy as np

from .channel import Channel
from .params import Params

FIELDS = (("u1", 1, "u"), ("u2", 2, "v"), ("u3", 3, "u"), ("p", 4, "p"))


def _d10_1(field):
    """A (D10.1) real: without a decimal point the digits carry one implied
    decimal, as Fortran formatted input reads them."""
    f = field.strip().lower().replace("d", "e")
    if "." in f:
        return float(f)
    mant, _, ex = f.partition("e")
    return int(mant) * 10.0 ** (int(ex) if ex else 0) / 10.0


def read_input(path):
    lines = open(path).read().splitlines()
    nx = int(lines[0][7:17])
    nz = int(lines[1][7:17])
    ny = int(lines[2][7:17])
    t = _d10_1(lines[3][7:17])
    dirname = lines[4][8:38].strip()
    dirout = lines[5][7:37].strip()
    return nx, nz, ny, t, dirname, dirout


def _raw_records(path, n):
    """The first n record payloads of a Fortran sequential unformatted file."""
    out = []
    with open(path, "rb") as fh:
        for _ in range(n):
            (m,) = struct.unpack("<i", fh.read(4))
            out.append(fh.read(m))
            fh.read(4)
    return out


def _plane_ys(path, nplanes):
    """y(j) of the first nplanes plane records (the records after the three
    header records), skipping the plane payloads."""
    ys = []
    with open(path, "rb") as fh:
        for r in range(3 + nplanes):
            (m,) = struct.unpack("<i", fh.read(4))
            if r < 3:
                fh.seek(m + 4, os.SEEK_CUR)
                continue
            head = fh.read(24)
            ys.append(struct.unpack_from("<d", head, 16)[0])
            fh.seek(m - 24 + 4, os.SEEK_CUR)
    return ys


def _rec(fh, payload):
    fh.write(struct.pack("<i", len(payload)))
    fh.write(payload)
    fh.write(struct.pack("<i", len(payload)))


def plane_sequence(grid, N):
    """(index into the field's planes, record j, record id or None = the
    plane's own id) in the order four_to_phys writes them; N is the 20-int
    band table N(1:4, 0:4).

    The reference reads the planes as a stream, band by band, and writes each
    interface record from the last plane it read.  That is the plane of the
    interface only when the band limits of the grid are in order; otherwise
    its loops read past the last plane of the file.  Such a table is refused
    here."""
    n = lambda r, b: N[4 * b + r - 1]  # noqa: E731  N(r, b)
    for r in (3, 4):
        lims = [n(r, b) for b in range(4)]
        if lims != sorted(lims):
            raise ValueError("band limits N(%d, 0:3) = %s are not in order: the plane stream of four_to_phys "
                             "does not fit this grid" % (r, lims))
    if grid == "p":
        j0 = n(4, 0) + 1
        return [(j - j0, j, None) for j in range(n(4, 0) + 1, n(4, 3) + 1)]
    r = 4 if grid == "u" else 3
    j0 = n(r, 0)
    seq = []
    for b in (1, 2, 3):
        if b == 1:
            seq.append((0, j0, None))
        else:
            ji = n(r, b - 1)
            seq.append((ji - j0, ji, 1 if grid == "u" else None))
        seq += [(j - j0, j, None) for j in range(n(r, b - 1) + 1, n(r, b) + 1)]
    seq.append((n(r, 3) + 1 - j0, n(r, 3) + 1, None))
    return seq


def band_of(grid, N):
    """The band whose grid transforms each record of plane_sequence(grid, N),
    in the same order (four_to_phys.f90:67-121 for u/w and v, 275-322 for p:
    a band's interface repeat belongs to the band it opens)."""
    n = lambda r, b: N[4 * b + r - 1]  # noqa: E731  N(r, b)
    if grid == "p":
        return [1 if j <= n(4, 1) - 1 else (2 if j <= n(4, 2) + 1 else 3) for j in range(n(4, 0) + 1, n(4, 3) + 1)]
    r = 4 if grid == "u" else 3
    out = []
    for b in (1, 2, 3):
        out += [b] * (1 + n(r, b) - n(r, b - 1))
    out.append(3)
    return out


def read_planes(path):
    """The plane records of a field file: (j, id, nx, nz, y, buff) with buff
    (nz, nx) reals, x fastest."""
    out = []
    with open(path, "rb") as fh:
        for r in range(3):
            (m,) = struct.unpack("<i", fh.read(4))
            fh.seek(m + 4, os.SEEK_CUR)
        while True:
            h = fh.read(4)
            if len(h) < 4:
                break
            (m,) = struct.unpack("<i", h)
            rec = fh.read(m)
            fh.read(4)
            j, fid, nx, nz = struct.unpack_from("<4i", rec, 0)
            (y,) = struct.unpack_from("<d", rec, 16)
            out.append((j, fid, nx, nz, y, np.frombuffer(rec, dtype="<f8", offset=24).reshape(nz, nx)))
    return out


def resize(old, nx, nz):
    """resize (four_to_phys.f90:328-348): a plane of nx_old x nz_old reals
    (old[k, i]) into nx x nz, x reals up to min(nx, nx_old), the first
    min(nz, nz_old)/2 z-modes from the bottom and one fewer from the top."""
    nzo, nxo = old.shape
    out = np.zeros((nz, nx))
    m, ix, dk = min(nz // 2, nzo // 2), min(nx, nxo), nzo - nz
    out[:m, :ix] = old[:m, :ix]
    for k in range(nz - m + 1, nz):
        out[k, :ix] = old[k + dk, :ix]
    return out


def write_phys(path, header_rec, N, grid_rec, grid, field_id, ys, phys, per_record=None):
    """Write one physical-field file.  header_rec, grid_rec: raw payloads of
    records 1 and 3 of the spectral file; ys: y(j) of each plane record of
    the spectral file; phys: (planes, Ngalz, Ngalx), or per_record: one
    (Ngalz_b, Ngalx_b) plane per record of plane_sequence (bands of their own
    size)."""
    ngal = list(N)
    for b in (1, 2, 3):
        ngal[4 * b + 0] = N[4 * b + 0] * 3 // 2
        ngal[4 * b + 1] = N[4 * b + 1] * 3 // 2
    with open(path, "wb") as fh:
        _rec(fh, header_rec)
        _rec(fh, struct.pack("<20i", *ngal))
        _rec(fh, grid_rec)
        for i, (q, j, rid) in enumerate(plane_sequence(grid, N)):
            pl = per_record[i] if per_record is not None else phys[q]
            ngz, ngx = pl.shape
            buf = np.zeros((ngz, ngx + 2))
            buf[:, :ngx] = pl
            _rec(fh, struct.pack("<4id", j, field_id if rid is None else rid, ngx + 2, ngz, ys[q]) +
                 buf.astype("<f8").tobytes())


def convert(nx, nz, ny, t, dirname, dirout):
    tail = "%04dx%04dx%04d_t%05d.dat" % (nx, nz, ny, int(10.0 * t))
    chans = {}
    for name, fid, grid in FIELDS:
        src = os.path.join(dirname, "%s_%s" % (name, tail))
        head = _raw_records(src, 3)
        N = list(struct.unpack("<20i", head[1][:80]))
        seq = plane_sequence(grid, N)  # refuses a band table the plane stream does not fit
        bands = band_of(grid, N)
        recs = read_planes(src)
        ys = [r[4] for r in recs]
        phys = [None] * len(seq)
        for b in sorted(set(bands)):
            N1b, N2b = N[4 * b], N[4 * b + 1]
            key = (N1b, N2b)
            if key not in chans:
                # the transforms need the plane counts and the x/z sizes only:
                # a grid without canopy points and ny + 2 dsty planes has the
                # same u, v and p plane counts, whatever the file's y grid was
                dsty = -N[2]
                p = Params(ngalx=N1b * 3 // 2, ngalz=N2b * 3 // 2, ny=ny + 2 * dsty, dsty=0, geometry_type=0,
                           posth=0.0, ygrid=0)
                chans[key] = Channel(p)
            ch = chans[key]
            ny_g = {"u": ch.nyu, "v": ch.nyv, "p": ch.nyp}[grid]
            # every plane at this band's size (the x-inverse transforms
            # planes in pairs, so each plane keeps the partner it has in a
            # one-resolution field)
            field = np.zeros((ny_g, ch.N2, ch.Mx), dtype=complex)
            for q in range(ny_g):
                re = resize(recs[q][5], N1b + 2, N2b)
                field[q] = re[:, 0::2] + 1j * re[:, 1::2]
            ch.set_field(name, field.reshape(ny_g, ch.Ncol).T)
            ph = ch.four_to_phys(name)
            for i, ((q, j, rid), bb) in enumerate(zip(seq, bands)):
                if bb == b:
                    phys[i] = ph[q]
        out = os.path.join(dirout, "%s_phys_%s" % (name, tail))
        write_phys(out, head[0], N, head[2], grid, fid, ys, None, per_record=phys)
        print("Writing", out)
    for ch in chans.values():
        ch.close()


def main(argv):
    path = argv[1] if len(argv) > 1 else "input_FtP.txt"
    convert(*read_input(path))


if __name__ == "__main__":
    main(sys.argv)
