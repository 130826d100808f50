"""ctypes wrapper of the CPU FP64 oracle (oracle/libcfd_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, as the checker / the timed CPU baseline.  The
product (cfd_amd/) never imports it.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libcfd_oracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libref.so")

INT_FIELDS = "ngalx ngalz ny dsty dsty1 ppp ygrid geometry_type ntilex ntilez extras24 flag_ctpress ngalx2 ngalz2".split()
DBL_FIELDS = "Re alp bet dyq posth posth1 ytanh_beta CFL mpgx r c_i c_k stag".split()
FIELDS = ("u1", "u2", "u3", "p", "Nu1", "Nu2", "Nu3", "du1", "du2", "du3", "psi", "div")
GRID = {"u1": "u", "u2": "v", "u3": "u", "p": "p", "Nu1": "u", "Nu2": "v", "Nu3": "u",
        "du1": "u", "du2": "v", "du3": "u", "psi": "p", "div": "p"}
SCALARS = ("t", "dt", "dti", "mpgx", "dgx", "QxT", "Qx", "dtc")


class OrParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in INT_FIELDS] + [(n, ctypes.c_double) for n in DBL_FIELDS]


_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_lib = None


def build():
    """Compile the oracle (gcc) in place."""
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = ctypes.CDLL(LIB)
        lib.or_create.restype = ctypes.c_void_p
        lib.or_create.argtypes = [ctypes.POINTER(OrParams), ctypes.c_char_p, ctypes.c_int]
        for name in ("or_destroy", "or_dims", "or_get_ygrid", "or_get_metrics", "or_get_iblist", "or_get_dg",
                     "or_set_field", "or_get_field", "or_set_scalars", "or_get_scalars", "or_rhs0",
                     "or_nonlinear", "or_solve_u", "or_solve_v", "or_meanflow_ctU", "or_solve_p", "or_vcorr",
                     "or_divergence", "or_substep", "or_step", "or_flowrate", "or_meanflow_ctP", "or_stats_init", "or_stats",
                     "or_stats_dims", "or_stats_get", "or_stats_reset", "or_set_compat_lastplane",
                     "or_plane_partition", "or_band_info"):
            getattr(lib, name).argtypes = None
        lib.or_flowrate.restype = ctypes.c_double
        _lib = lib
    return _lib


def to_or_params(p):
    """From any object with the cfd_params attribute names (ngalx2, ngalz2
    default to 0: one resolution in every band)."""
    c = OrParams()
    for n in INT_FIELDS:
        setattr(c, n, int(getattr(p, n, 0)))
    for n in DBL_FIELDS:
        setattr(c, n, float(getattr(p, n)))
    return c


def plane_partition(N20, ngalx, ngalz, np_ranks):
    """proc_lims_planes (start.f90:1652-1866): {"u": [(first, last)], "v":
    [...]} per rank, planes without walls; None when np_ranks < 3."""
    lim = (ctypes.c_int * (4 * np_ranks))()
    if load().or_plane_partition((ctypes.c_int * 20)(*[int(v) for v in N20]), ctypes.c_int(ngalx),
                                 ctypes.c_int(ngalz), ctypes.c_int(np_ranks), lim):
        return None
    lim = list(lim)
    return {g: [(lim[(k * 2) * np_ranks + r], lim[(k * 2 + 1) * np_ranks + r]) for r in range(np_ranks)]
            for k, g in enumerate(("u", "v"))}


def set_threads(n):
    load().or_set_threads(ctypes.c_int(int(n)))


class Oracle:
    """CPU FP64 restatement of the substep, same operator names as Channel."""

    def __init__(self, params):
        self.lib = load()
        self.params = params
        self._c = to_or_params(params)
        err = ctypes.create_string_buffer(256)
        h = self.lib.or_create(ctypes.byref(self._c), err, 256)
        if not h:
            raise RuntimeError("oracle: " + err.value.decode())
        self.h = ctypes.c_void_p(h)
        d = (ctypes.c_int * 16)()
        self.lib.or_dims(self.h, d)
        d = list(d)
        self.N1, self.N2, self.Mx, self.Ncol = d[0:4]
        self.nyu, self.nyv, self.nyp = d[4:7]
        self.jlo_u, self.jlo_v, self.jlo_p = d[7:10]
        self._ny = {"u": self.nyu, "v": self.nyv, "p": self.nyp}

    def close(self):
        if getattr(self, "h", None):
            self.lib.or_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def shape(self, name):
        return (self.Ncol, self._ny[GRID[name]])

    def set_field(self, name, arr):
        arr = np.ascontiguousarray(arr, dtype=np.complex128)
        assert arr.shape == self.shape(name), (name, arr.shape, self.shape(name))
        self.lib.or_set_field(self.h, ctypes.c_int(FIELDS.index(name)), arr.view(np.float64).ctypes.data_as(_D))

    def get_field(self, name):
        out = np.zeros(self.shape(name), dtype=np.complex128)
        self.lib.or_get_field(self.h, ctypes.c_int(FIELDS.index(name)), out.view(np.float64).ctypes.data_as(_D))
        return out

    def scalars(self):
        v = np.zeros(8)
        self.lib.or_get_scalars(self.h, v.ctypes.data_as(_D))
        return dict(zip(SCALARS, v.tolist()))

    def set_scalars(self, **kw):
        cur = self.scalars()
        cur.update({k: float(v) for k, v in kw.items()})
        v = np.array([cur[k] for k in SCALARS])
        self.lib.or_set_scalars(self.h, v.ctypes.data_as(_D))

    def ygrid(self):
        out = [np.zeros(self.nyu), np.zeros(self.nyv), np.zeros(self.nyu), np.zeros(self.nyv),
               np.zeros(3 * self.nyu), np.zeros(3 * self.nyv)]
        self.lib.or_get_ygrid(self.h, *[a.ctypes.data_as(_D) for a in out])
        return dict(yu=out[0], yv=out[1], dthdyu=out[2], dthdyv=out[3],
                    dyu2i=out[4].reshape(-1, 3), dyv2i=out[5].reshape(-1, 3))

    def metrics(self):
        gw, gwi, dtv = np.zeros(2), np.zeros(2), ctypes.c_double()
        self.lib.or_get_metrics(self.h, gw.ctypes.data_as(_D), gwi.ctypes.data_as(_D), ctypes.byref(dtv))
        return dict(gw=gw, gwi=gwi, dtv=dtv.value)

    def iblist(self, grid, wall):
        n = self.lib.or_get_iblist(self.h, ctypes.c_int(grid), ctypes.c_int(wall), None, None)
        l9 = np.zeros((n, 9), dtype=np.int32)
        w3 = np.zeros((n, 3))
        self.lib.or_get_iblist(self.h, ctypes.c_int(grid), ctypes.c_int(wall), l9.ctypes.data_as(_I),
                               w3.ctypes.data_as(_D))
        return l9, w3

    def dg(self):
        out = np.zeros((self.Ncol, self.nyp, 3))
        self.lib.or_get_dg(self.h, out.ctypes.data_as(_D))
        return out

    def flowrate(self):
        return self.lib.or_flowrate(self.h)

    def RHS0(self, kRK):
        self.lib.or_rhs0(self.h, ctypes.c_int(kRK))

    def nonlinear(self, kRK):
        self.lib.or_nonlinear(self.h, ctypes.c_int(kRK))

    def solveU(self, comp, kRK):
        self.lib.or_solve_u(self.h, ctypes.c_int(comp), ctypes.c_int(kRK))

    def solveV(self, kRK):
        self.lib.or_solve_v(self.h, ctypes.c_int(kRK))

    def meanflow_ctU(self, kRK):
        self.lib.or_meanflow_ctU(self.h, ctypes.c_int(kRK))

    def meanflow_ctP(self):
        self.lib.or_meanflow_ctP(self.h)

    def solveP(self, kRK):
        self.lib.or_solve_p(self.h, ctypes.c_int(kRK))

    def v_corr(self, kRK):
        self.lib.or_vcorr(self.h, ctypes.c_int(kRK))

    def divergence(self):
        self.lib.or_divergence(self.h)

    def substep(self, kRK):
        self.lib.or_substep(self.h, ctypes.c_int(kRK))

    def step(self, n=1):
        self.lib.or_step(self.h, ctypes.c_int(int(n)))

    # -- statistics (stats.f90, spectra.f90) -----------------------------------
    STATS = ("Um", "U2m", "Vm", "V2m", "Wm", "W2m", "Pm", "P2m", "UVm", "UWm", "VWm", "wxm", "wx2m")
    SPECTRA = ("spU", "spV", "spW", "spUV", "spP")
    _STAT_GRID = "uuvvuuppuuvvv"

    def set_compat_lastplane(self, N20, np_ranks):
        """The as-shipped last-plane nonlinear term of np_ranks ranks
        (ops_in_planes, FOU3D.f90:1553-1613); 0: per plane."""
        rc = self.lib.or_set_compat_lastplane(self.h, (ctypes.c_int * 20)(*[int(v) for v in N20]),
                                              ctypes.c_int(int(np_ranks)))
        if rc:
            raise ValueError("oracle last-plane replay: no partition (%d)" % rc)

    def stats_init(self, N20):
        rc = self.lib.or_stats_init(self.h, (ctypes.c_int * 20)(*[int(v) for v in N20]))
        if rc:
            raise ValueError("oracle stats: band table or tiles do not fit (%d)" % rc)

    def stats(self):
        if self.lib.or_stats(self.h):
            raise RuntimeError("oracle stats: not initialised")

    def stats_dims(self):
        d = (ctypes.c_int * 6)()
        self.lib.or_stats_dims(self.h, d)
        d = list(d)
        return dict(dnx=d[0], dnz=d[1], nyc={"u": d[2], "v": d[3], "p": d[4]}, istat=d[5])

    def stats_get(self):
        """Accumulators: profiles {name: (ny,)}, conditional sums {name:
        (bottom, top)} as (nyc, dnz, dnx) arrays, spectra {name: (Ncol, ny)}."""
        d = self.stats_dims()
        out = {"istat": d["istat"], "prof": {}, "cond": {}, "spec": {}}
        for q, name in enumerate(self.STATS):
            g = self._STAT_GRID[q]
            a = np.zeros(self._ny[g])
            self.lib.or_stats_get(self.h, ctypes.c_int(q), ctypes.c_int(0), a.ctypes.data_as(_D))
            out["prof"][name] = a
            cb = np.zeros((d["nyc"][g], d["dnz"], d["dnx"]))
            ct = np.zeros_like(cb)
            self.lib.or_stats_get(self.h, ctypes.c_int(q), ctypes.c_int(1), cb.ctypes.data_as(_D))
            self.lib.or_stats_get(self.h, ctypes.c_int(q), ctypes.c_int(2), ct.ctypes.data_as(_D))
            out["cond"][name + "C"] = (cb, ct)
        for q, (name, g) in enumerate(zip(self.SPECTRA, "uvuup")):
            a = np.zeros((self.Ncol, self._ny[g]))
            self.lib.or_stats_get(self.h, ctypes.c_int(q), ctypes.c_int(3), a.ctypes.data_as(_D))
            out["spec"][name] = a
        return out

    def stats_reset(self):
        self.lib.or_stats_reset(self.h)


# ---------------------------------------------------------------------------
# compiled reference routines (oracle/_ref/libref.so, built from the
# reference sources by oracle/Makefile.ref) -- optional pins
# ---------------------------------------------------------------------------
_ref = None


def load_ref():
    global _ref
    if _ref is None and os.path.exists(REF_LIB):
        _ref = ctypes.CDLL(REF_LIB)
    return _ref


def ref_cft(n, c, ise, isa, m, iopt):
    load_ref().ref_cft(ctypes.c_int(n), c.ctypes.data_as(_D), ctypes.c_int(ise), ctypes.c_int(isa), ctypes.c_int(m),
                       ctypes.c_int(iopt))


def ref_rft(n, c, isa, m, iopt):
    load_ref().ref_rft(ctypes.c_int(n), c.ctypes.data_as(_D), ctypes.c_int(isa), ctypes.c_int(m), ctypes.c_int(iopt))


def ref_plane_partition(N20, ngalx, ngalz, np_ranks):
    """The compiled reference's proc_lims_planes (start.f90) in the layout of
    plane_partition; np_ranks >= 3."""
    lim = (ctypes.c_int * (4 * np_ranks))()
    load_ref().ref_plane_partition((ctypes.c_int * 20)(*[int(v) for v in N20]), ctypes.c_int(ngalx),
                                   ctypes.c_int(ngalz), ctypes.c_int(np_ranks), lim)
    lim = list(lim)
    return {g: [(lim[(k * 2) * np_ranks + r], lim[(k * 2 + 1) * np_ranks + r]) for r in range(np_ranks)]
            for k, g in enumerate(("u", "v"))}


def or_cft(n, c, ise, isa, m, iopt):
    load().or_cft(ctypes.c_int(n), c.ctypes.data_as(_D), ctypes.c_int(ise), ctypes.c_int(isa), ctypes.c_int(m),
                  ctypes.c_int(iopt))


def or_rft(n, c, isa, m, iopt):
    load().or_rft(ctypes.c_int(n), c.ctypes.data_as(_D), ctypes.c_int(isa), ctypes.c_int(m), ctypes.c_int(iopt))
