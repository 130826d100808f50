"""Writers of input.in and geom_circ_rough.txt for the Fortran driver.

The records follow the sequence and edit descriptors that ``start`` reads
(start.f90:62-227: reals as (7X,D10.1), integers as (7X,I10), the pressure
flag as (12X,I5), strings as (7X,A90)) and the geometry file of
geom_circ_rough_mixed_forcing_24pt.f90:113-123; the five optional records the
driver reads after the reference's last one (ygrid, ybeta, extr24, precis,
fused) are appended.  A real field holds 10 characters, so a value is written
with as many digits as fit and always with a decimal point (without one the
D10.1 descriptor scales the value by 1/10); ``representable`` returns the
parameters as the driver reads them back, for runs that must match it bit
for bit.

    python -m cfd_amd.inputfile rt395_384x192x192 [dir]

writes both files of a named workload (flag_init 3, the parabolic profile)
into dir (default .).
"""
import dataclasses
import os
import sys

REAL_FIELDS = ("Re", "alp", "bet", "dyq", "posth", "posth1", "ytanh_beta", "CFL", "mpgx", "r", "c_i", "c_k",
               "stag")


def fmt_real(x):
    """The 10-character field of x with a decimal point and the most digits
    that fit, as the (7X,D10.1) descriptor reads it; returns (text, value)."""
    x = float(x)
    best = None
    for d in range(10, -1, -1):
        for style in ("f", "f0", "e"):
            t = ("%.*f" % (d, x)) if style != "e" else ("%.*e" % (d, x)).replace("e", "d")
            if style == "f0":  # |x| < 1 without the leading zero: one more digit
                if not (t.startswith("0.") or t.startswith("-0.")):
                    continue
                t = t.replace("0.", ".", 1)
            if style != "e" and "." not in t:
                t += "."
            if len(t) <= 10:
                v = float(t.replace("d", "e"))
                err = abs(v - x)
                if best is None or err < best[2]:
                    best = (t, v, err)
        if best is not None and best[2] == 0.0:
            break
    if best is None:
        raise ValueError("%r does not fit a 10-character real field" % x)
    return best[0].rjust(10), best[1]


def representable(p):
    """p with every real replaced by the value its input.in field carries."""
    return dataclasses.replace(p, **{k: fmt_real(getattr(p, k))[1] for k in REAL_FIELDS})


def _r(name, x, comment=""):
    return "%-6s=%s ! %s" % (name, fmt_real(x)[0], comment)


def _i(name, n, comment=""):
    return "%-6s=%10d ! %s" % (name, int(n), comment)


def input_text(p, flag_init=3, fnameimb="none", dirin="./", dirout="./", nwrite=1000, maxt=2000.0, fused=1):
    """The text of input.in for the parameters p (cfd_amd.Params)."""
    if p.ngalx <= 0 or p.ngalz <= 0:
        raise ValueError("grid sizes")
    blank = "                  !"
    L = ["! " + "-" * 82,
         "! input.in written by cfd_amd.inputfile: the record sequence of start.f90:62-227",
         "! and the optional records of the GPU driver after the last one.",
         "!", "!", "!", "!",
         "! " + "-" * 82,
         "! Fluid properties " + "-" * 64,
         _r("Re", p.Re, "bulk Reynolds number"),
         "",
         "! Box geometry and spacial discretisation " + "-" * 42,
         _r("alp", p.alp, "Lx = 2*pi/alp"),
         _r("bet", p.bet, "Lz = 2*pi/bet"),
         _i("Ngalx1", p.ngalx, "collocation points in x, band 1"),
         _i("Ngalx2", p.ngalx2 or p.ngalx, "collocation points in x, band 2"),
         _i("Ngalz1", p.ngalz, "collocation points in z, band 1"),
         _i("Ngalz2", p.ngalz2 or p.ngalz, "collocation points in z, band 2"),
         _i("ny", p.ny, "grid points in y, y(0)=-1, y(ny-1)=1"),
         _r("dyq", p.dyq, "dymax/dymin of the polynomial map"),
         _i("pp", p.ppp, "y polynomial exponent"),
         "",
         "! Initial conditions " + "-" * 62,
         _i("flgin", flag_init, "1 new from files, 2 continue, 3 parabolic profile"),
         blank, blank, blank,
         "fnmeb =" + fnameimb,
         blank,
         "dirin =" + dirin,
         "dirout=" + dirout,
         "",
         "! Parameters of the numerical method I " + "-" * 44,
         _r("CFL", p.CFL, "CFL"),
         _i("nwrite", nwrite, "time steps between outputs"),
         "flpress    =%5d ! 0 constant flow rate, 1 constant pressure gradient" % int(p.flag_ctpress),
         blank,
         _r("mpgx", p.mpgx, "prescribed mean pressure gradient (flpress = 1)"),
         blank, blank,
         "! Parameters of the numerical method II " + "-" * 43,
         _r("Kib", 1.0, "unused on the path"),
         _r("maxerr", 1e-8, "unused on the path"),
         _r("maxt", maxt, "maximum simulation time"),
         _r("maxA", 4.0, "unused on the path"),
         _i("FFTbot", 7, "unused on the path"),
         _i("FFTtop", 146, "unused on the path"),
         "",
         "! Surface Geometry " + "-" * 64,
         _i("type", p.geometry_type, "0 smooth wall, 3 canopies"),
         blank, blank, blank, blank,
         _i("ntilex", p.ntilex, "tiles in x (even)"),
         _i("ntilez", p.ntilez, "tiles in z"),
         _r("posth", p.posth, "canopy height in outer units"),
         _r("posth1", p.posth1, "turning point of the two-layer canopy grid"),
         _r("hny", 1.0, "unused on the path"),
         _i("npeakx", 13, "unused on the path"),
         _i("npeakz", 13, "unused on the path"),
         _r("Lfracx", 3.0, "unused on the path"),
         _r("Lfracz", 3.0, "unused on the path"),
         _i("dsty", p.dsty, "points in the canopy in y"),
         _i("dsty1", p.dsty1, "points in the upper canopy layer"),
         _i("dstx", 3, "unused on the path"),
         _i("dstz", 3, "unused on the path"),
         _i("ygrid", p.ygrid, "0 polynomial map, 1 tanh map"),
         _r("ybeta", p.ytanh_beta, "stretching of the tanh map"),
         _i("extr24", p.extras24, "1: the 4 fixed forcing points of the 24pt geometry"),
         _i("precis", p.precision, "0 FP64, 1 mixed"),
         _i("fused", fused, "1 cfd_step per time step, 0 one entry point per subroutine")]
    return "\n".join(L) + "\n"


def geom_text(p):
    """The text of geom_circ_rough.txt: radius, centre (tile grid units) and
    stagger."""
    return "\n".join([_r("r_p1", p.r, "radius of the post (tile grid units)"),
                      _r("c_i_p1", p.c_i, "centre x"),
                      _r("c_k_p1", p.c_k, "centre z"),
                      _r("stag", p.stag, "stagger of every other tile row")]) + "\n"


def write(directory, p, **kw):
    """Write input.in and geom_circ_rough.txt for p into directory; returns the
    parameters as the driver reads them (representable(p))."""
    with open(os.path.join(directory, "input.in"), "w") as fh:
        fh.write(input_text(p, **kw))
    with open(os.path.join(directory, "geom_circ_rough.txt"), "w") as fh:
        fh.write(geom_text(p))
    return representable(p)


def main(argv):
    from . import workloads
    if not argv:
        print(__doc__)
        return 2
    p, _ = workloads.get(argv[0])
    write(argv[1] if len(argv) > 1 else ".", p)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
