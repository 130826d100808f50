"""Writes the golden fixtures of tests/golden/ from the compiled reference
routines (oracle/_ref/libref.so, built from the reference sources by
oracle/Makefile.ref): FFTPACK wrappers cft/rft (cft_buff.f, rft_buff.f),
y_grid_canopy (y_grid_canopy.f90) and boundary_circ_rough
(geom_circ_rough_mixed_forcing_24pt.f90 + mergeSort.f90).

Run from the repository root:  python tests/golden/make_golden.py
"""
import ctypes
import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
_D = ctypes.POINTER(ctypes.c_double)

FFT_N = (6, 12, 24, 48, 64, 72, 96, 192, 384)

# y grids: (ny, dsty, dsty1, dyq, ppp, posth, posth1)
YGRIDS = {
    "shipped": (25, 4, 0, 10.0, 5, 0.089, 0.0),
    "twolayer": (33, 6, 3, 8.0, 5, 0.12, 0.05),
    "smooth": (25, 0, 0, 10.0, 5, 0.0, 0.0),
    "p7": (41, 5, 0, 14.0, 7, 0.06, 0.0),
}

# IB geometries: (ngx, ngz, ny, dsty, ntilex, ntilez, r, c_i, c_k, stag)
GEOMS = {
    "shipped": (48, 48, 25, 4, 2, 2, 5.43723, 12.0, 12.0, 0.5),
    "r4_offc": (96, 48, 25, 3, 4, 2, 4.1, 11.5, 12.25, 0.5),
    "r6_nostag": (48, 72, 25, 5, 2, 3, 6.3, 12.0, 12.0, 0.0),
    "r2_stag25": (96, 96, 29, 4, 4, 4, 2.2, 12.0, 11.0, 0.25),
}


def fft_kat():
    ref = O.load_ref()
    rng = np.random.default_rng(2024)
    out = {}
    for n in FFT_N:
        m = 3
        c = rng.standard_normal(2 * n * m)
        for iopt in (-1, 1):
            a = c.copy()
            ref.ref_cft(n, a.ctypes.data_as(_D), 2, 2 * n, m, iopt)
            out["cft_%d_%d_in" % (n, iopt)] = c
            out["cft_%d_%d_out" % (n, iopt)] = a
        isa = n + 2
        x = rng.standard_normal(isa * m)
        for iopt in (-1, 1):
            a = x.copy()
            ref.ref_rft(n, a.ctypes.data_as(_D), isa, m, iopt)
            out["rft_%d_%d_in" % (n, iopt)] = x
            out["rft_%d_%d_out" % (n, iopt)] = a
    np.savez_compressed(os.path.join(OUT, "fft_kat.npz"), **out)


def ygrids():
    ref = O.load_ref()
    out = {}
    td = tempfile.mkdtemp()
    for name, (ny, dsty, dsty1, dyq, ppp, posth, posth1) in YGRIDS.items():
        nyu, nyv = ny + 2 * dsty + 1, ny + 2 * dsty
        arr = [np.zeros(nyu), np.zeros(nyv), np.zeros(nyu), np.zeros(nyv), np.zeros(3 * nyu), np.zeros(3 * nyv)]
        ref.ref_ygrid(ny, dsty, dsty1, ctypes.c_double(dyq), ppp, ctypes.c_double(posth), ctypes.c_double(posth1),
                      (td + "/").encode(), *[a.ctypes.data_as(_D) for a in arr])
        for key, a in zip(("yu", "yv", "dthdyu", "dthdyv", "dyu2i", "dyv2i"), arr):
            out["%s_%s" % (name, key)] = a
        out["%s_params" % name] = np.array([ny, dsty, dsty1, dyq, ppp, posth, posth1])
    np.savez_compressed(os.path.join(OUT, "ygrid.npz"), **out)


# band tables: y grid case and the bottom-band height h_ny
BANDS = {"shipped_h25": ("shipped", 0.25), "smooth_h25": ("smooth", 0.25),
         "p7_ntz8": ("p7", 2.0 * (2.0 * np.pi / 2.0) / 8), "twolayer_ntz12": ("twolayer", 2.0 * (2.0 * np.pi / 2.0) / 12)}


def bands():
    ref = O.load_ref()
    out = {}
    td = tempfile.mkdtemp()
    for name, (gname, h) in BANDS.items():
        ny, dsty, dsty1, dyq, ppp, posth, posth1 = YGRIDS[gname]
        n = (ctypes.c_int * 20)()
        ref.ref_ygrid_bands(ny, dsty, dsty1, ctypes.c_double(dyq), ppp, ctypes.c_double(posth),
                            ctypes.c_double(posth1), ctypes.c_double(h), (td + "/").encode(), n)
        out["%s_N" % name] = np.array(list(n), dtype=np.int32)
        out["%s_params" % name] = np.array([ny, dsty, dsty1, dyq, ppp, posth, posth1, h])
    np.savez_compressed(os.path.join(OUT, "bands.npz"), **out)


def read_boundary(fn):
    b = open(fn, "rb").read()
    off = [0]

    def rd(dt, n):
        v = np.frombuffer(b, dtype=dt, count=n, offset=off[0])
        off[0] += v.nbytes
        return v

    rd("<f8", 3)
    rd("<i4", 20)
    h = rd("<i4", 10)
    nv, nu = int(h[0]), int(h[1])
    L = {k: rd("<i4", 9 * n).reshape(n, 9) for k, n in (("bv", nv), ("tv", nv), ("bu", nu), ("tu", nu))}
    W = {k: rd("<f8", 3 * n).reshape(n, 3) for k, n in (("bv", nv), ("tv", nv), ("bu", nu), ("tu", nu))}
    return L, W


def geoms():
    ref = O.load_ref()
    out = {}
    cwd = os.getcwd()
    for name, (ngx, ngz, ny, dsty, ntx, ntz, r, ci, ck, stag) in GEOMS.items():
        td = tempfile.mkdtemp()
        os.chdir(td)
        try:
            with open("geom_circ_rough.txt", "w") as f:
                for lab, v in (("r_p1=", r), ("c_i_p1", ci), ("c_k_p1", ck), ("stag=", stag)):
                    f.write("%-7s%10.6f\n" % (lab[:7], v))
            ref.ref_geom(ngx, ngz, ny, dsty, ntx, ntz, (td + "/").encode())
        finally:
            os.chdir(cwd)
        fn = [f for f in os.listdir(td) if f.startswith("boundary_")][0]
        L, W = read_boundary(os.path.join(td, fn))
        for k in L:
            out["%s_%s_l" % (name, k)] = L[k]
            out["%s_%s_w" % (name, k)] = W[k]
        out["%s_params" % name] = np.array([ngx, ngz, ny, dsty, ntx, ntz, r, ci, ck, stag])
    np.savez_compressed(os.path.join(OUT, "iblists.npz"), **out)


if __name__ == "__main__":
    if O.load_ref() is None:
        sys.exit("oracle/_ref/libref.so is not built (make -C oracle -f Makefile.ref)")
    fft_kat()
    ygrids()
    bands()
    geoms()
    print("wrote", sorted(f for f in os.listdir(OUT) if f.endswith(".npz")))
