"""Host-side mirror of the reference hot path on one GPU.

``Channel`` owns one ``cfd_handle`` and exposes the reference's hot-path
subroutines under their own names (littleharsh.f90:163-189):
``RHS0`` (RHS0_u1/u2/u3), ``nonlinear``, ``solveU``, ``solveV``,
``meanflow_ctU``, ``solveP``, ``v_corr``, plus ``substep`` (one pass of the
RK loop body) and ``step`` (whole time steps, the time loop restored).

Fields cross the boundary in the reference layout: complex128 arrays of shape
(Ncol, ny_grid), i.e. ``f(j, column)`` with y fastest, columns ordered k
outer / i inner (start.f90:1467-1481).
"""
import ctypes
import os

import numpy as np

from . import _lib
from .params import Params

FIELDS = ("u1", "u2", "u3", "p", "Nu1", "Nu2", "Nu3", "du1", "du2", "du3", "psi", "div")
GRID = {"u1": "u", "u2": "v", "u3": "u", "p": "p", "Nu1": "u", "Nu2": "v", "Nu3": "u",
        "du1": "u", "du2": "v", "du3": "u", "psi": "p", "div": "p"}
SCALARS = ("t", "dt", "dti", "mpgx", "dgx", "QxT", "Qx", "dtc")
KCLASSES = ("zinv", "xpass", "zfwd", "elem", "ysolve", "press", "ib")

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)


def _dp(a):
    return a.ctypes.data_as(_D)


def build_ygrid(params):
    """Host-only y grid (cfd_build_ygrid): dict of yu, yv, dthdyu, dthdyv,
    dyu2i (nyu,3), dyv2i (nyv,3)."""
    lib = _lib.load()
    p = params.to_c()
    nyu, nyv = params.nyu, params.nyv
    out = [np.zeros(nyu), np.zeros(nyv), np.zeros(nyu), np.zeros(nyv), np.zeros(3 * nyu), np.zeros(3 * nyv)]
    _lib.check(lib.cfd_build_ygrid(ctypes.byref(p), *[_dp(a) for a in out]))
    return dict(yu=out[0], yv=out[1], dthdyu=out[2], dthdyv=out[3],
                dyu2i=out[4].reshape(nyu, 3), dyv2i=out[5].reshape(nyv, 3))


def build_iblist(params, grid, wall):
    """Host-only IB list (cfd_build_iblist): (int32 (n,9), float64 (n,3))."""
    lib = _lib.load()
    p = params.to_c()
    n = ctypes.c_int(0)
    _lib.check(lib.cfd_build_iblist(ctypes.byref(p), grid, wall, None, None, ctypes.byref(n)))
    l9 = np.zeros((n.value, 9), dtype=np.int32)
    w3 = np.zeros((n.value, 3))
    _lib.check(lib.cfd_build_iblist(ctypes.byref(p), grid, wall, l9.ctypes.data_as(_I), _dp(w3), ctypes.byref(n)))
    return l9, w3


TUNING_KEYS = ("fft", "seg_rows", "solve_threads")
TUNING_DEFAULT = dict(fft=0, seg_rows=13, solve_threads=256)


def _tuning_c(t):
    bad = set(t) - set(TUNING_KEYS)
    if bad:
        raise ValueError("unknown tuning keys %s" % sorted(bad))
    d = dict(TUNING_DEFAULT, **t)
    return (ctypes.c_int * 3)(*(int(d[k]) for k in TUNING_KEYS))


class Channel:
    """One rough channel resident on one GPU."""

    def __init__(self, params=None, tuning=None, **kw):
        """tuning: None (defaults) or a dict with keys fft (0 four-step, 1
        Stockham), seg_rows (5, 9, 13, 21) and solve_threads (64, 128, 256);
        missing keys take the defaults 0, 13, 256."""
        self.params = (params or Params()).replace(**kw) if kw else (params or Params())
        self._lib = _lib.load()
        self._c = self.params.to_c()
        h = ctypes.c_void_p()
        if tuning is None:
            rc = self._lib.cfd_create(ctypes.byref(self._c), ctypes.byref(h))
        else:
            rc = self._lib.cfd_create_tuned(ctypes.byref(self._c), _tuning_c(tuning), ctypes.byref(h))
        _lib.check(rc, None)
        self._h = h
        d = (ctypes.c_int * 16)()
        self._call("cfd_dims", d)
        d = list(d)
        self.N1, self.N2, self.Mx, self.Ncol = d[0:4]
        self.nyu, self.nyv, self.nyp = d[4:7]
        self.jlo_u, self.jlo_v, self.jlo_p = d[7:10]
        self.nib_u, self.nib_v = d[14], d[15]
        self._ny = {"u": self.nyu, "v": self.nyv, "p": self.nyp}

    # -- plumbing ---------------------------------------------------------
    def _call(self, name, *args):
        if self._h is None:
            raise _lib.CfdError("channel is closed")
        rc = getattr(self._lib, name)(self._h, *args)
        _lib.check(rc, self._h)

    def close(self):
        if getattr(self, "_h", None) is not None:
            self._lib.cfd_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- state --------------------------------------------------------------
    def shape(self, name):
        return (self.Ncol, self._ny[GRID[name]])

    def set_field(self, name, arr):
        arr = np.ascontiguousarray(arr, dtype=np.complex128)
        if arr.shape != self.shape(name):
            raise ValueError("%s: expected shape %s, got %s" % (name, self.shape(name), arr.shape))
        self._call("cfd_set_field", FIELDS.index(name), arr.view(np.float64).ctypes.data_as(_D))

    def get_field(self, name):
        out = np.zeros(self.shape(name), dtype=np.complex128)
        self._call("cfd_get_field", FIELDS.index(name), out.view(np.float64).ctypes.data_as(_D))
        return out

    def scalars(self):
        v = np.zeros(8)
        self._call("cfd_get_scalars", _dp(v))
        return dict(zip(SCALARS, v.tolist()))

    def set_scalars(self, **kw):
        cur = self.scalars()
        for k, val in kw.items():
            if k not in cur:
                raise KeyError(k)
            cur[k] = float(val)
        v = np.array([cur[k] for k in SCALARS])
        self._call("cfd_set_scalars", _dp(v))

    def ygrid(self):
        out = [np.zeros(self.nyu), np.zeros(self.nyv), np.zeros(self.nyu), np.zeros(self.nyv),
               np.zeros(3 * self.nyu), np.zeros(3 * self.nyv)]
        self._call("cfd_get_ygrid", *[_dp(a) for a in out])
        return dict(yu=out[0], yv=out[1], dthdyu=out[2], dthdyv=out[3],
                    dyu2i=out[4].reshape(-1, 3), dyv2i=out[5].reshape(-1, 3))

    # -- reference operators ---------------------------------------------------
    def RHS0(self, kRK):
        self._call("cfd_rhs0", kRK)

    def nonlinear(self, kRK):
        self._call("cfd_nonlinear", kRK)

    def solveU(self, comp, kRK):
        self._call("cfd_solve_u", comp, kRK)

    def solveV(self, kRK):
        self._call("cfd_solve_v", kRK)

    def meanflow_ctU(self, kRK):
        self._call("cfd_meanflow_ctU", kRK)

    def solveP(self, kRK):
        self._call("cfd_solve_p", kRK)

    def v_corr(self, kRK):
        self._call("cfd_vcorr", kRK)

    def substep(self, kRK):
        self._call("cfd_substep", kRK)

    def step(self, nsteps=1):
        self._call("cfd_step", int(nsteps))

    def sync(self):
        self._call("cfd_sync")

    def device_bytes(self):
        """Device memory held by this solver (bytes)."""
        v = ctypes.c_double(0.0)
        self._call("cfd_device_bytes", ctypes.byref(v))
        return v.value

    # -- field files ------------------------------------------------------------
    def record_out(self, dirout, iter=0):
        """Write u1, u2, u3, p as the reference's record_out does (file names
        <f>_<N1>x<N2>x<ny>_t<int(10 t)>.dat in dirout); returns the u1 path."""
        self._call("cfd_record_out", os.fsencode(dirout), int(iter))
        t = self.scalars()["t"]
        return os.path.join(dirout, "u1_%04dx%04dx%04d_t%05d.dat" % (self.N1, self.N2, self.params.ny,
                                                                     int(10.0 * t)))

    def read_in(self, dirin, fnameimb, flag_init=2):
        """read_in + the flag_init branch of start: reads dirin/u1<tail> ...
        with tail = fnameimb[2:]; returns iter0."""
        it = ctypes.c_int(0)
        self._call("cfd_read_in", os.fsencode(dirin), os.fsencode(fnameimb), int(flag_init), ctypes.byref(it))
        return it.value

    def four_to_phys(self, name):
        """Physical values of a field, (planes, Ngalz, Ngalx): four_to_phys's
        zero padding (without the kz = -N2/2 mode) and unscaled inverse
        transforms."""
        out = np.zeros((self._ny[GRID[name]], self.params.ngalz, self.params.ngalx))
        self._call("cfd_four_to_phys", FIELDS.index(name), out.ctypes.data_as(_D))
        return out

    def four_to_phys_planes(self, name, plane0, nplanes, dev_ptr):
        """Planes [plane0, plane0 + nplanes) of four_to_phys into the device
        buffer at dev_ptr (nplanes x Ngalz x Ngalx doubles on this channel's
        device), queued on the channel's stream without any device
        allocation; call sync() before reading the buffer."""
        self._call("cfd_four_to_phys_planes", FIELDS.index(name), int(plane0), int(nplanes),
                   ctypes.c_void_p(int(dev_ptr)))

    # -- statistics and spectra -------------------------------------------------
    def stats(self):
        """Add one sample of the spectra and statistics of the current state
        (spectra.f90, stats.f90)."""
        self._call("cfd_stats")

    def write_stats(self, dirout, iter=0):
        """write_spect + write_stats with the current t and mpgx, then zero
        the accumulators; returns the stats file path."""
        self._call("cfd_write_stats", os.fsencode(dirout), int(iter))
        return dirout + "stats_" + self._stats_tag(self.scalars()["t"])

    def stats_request(self, sample=True, write_dir=None, iter=0):
        """Sample and/or write at the reference's places inside the next
        substep with kRK = 1 (the nonlinear term's flagst / flagwr)."""
        flags = (1 if sample else 0) | (2 if write_dir is not None else 0)
        self._call("cfd_stats_request", flags, os.fsencode(write_dir) if write_dir is not None else None, int(iter))

    def stats_dims(self):
        """dnx, dnz, wall-band plane counts of the u, v, p grids, istat."""
        d = (ctypes.c_int * 6)()
        self._call("cfd_stats_dims", d)
        d = list(d)
        return dict(dnx=d[0], dnz=d[1], nyc={"u": d[2], "v": d[3], "p": d[4]}, istat=d[5])

    def _stats_tag(self, t):
        return "%04dx%04dx%04d_%05d.dat" % (self.N1, self.N2, self.params.ny, int(10.0 * t))

    # -- learned closure ---------------------------------------------------------
    def closure_attach(self, net, every=1, clip=False):
        """Add the network's correction in every nonlinear term, evaluated at
        kRK = 1 of every `every`-th step and reused in between (None detaches);
        clip: zero the components that would feed energy into the resolved
        field (u c < 0)."""
        self._call("cfd_closure_attach", net.handle if net is not None else None, int(every))
        self._call("cfd_closure_clip", 1 if clip else 0)

    def closure_phys(self, which):
        """Physical planes of the last closure sample / application: 0..2 the
        velocities, 3..5 the corrections; (planes, Ngalz, Ngalx)."""
        ny = self.nyv if which % 3 == 1 else self.nyu
        out = np.zeros((ny, self.params.ngalz, self.params.ngalx))
        self._call("cfd_closure_phys", int(which), out.ctypes.data_as(_D))
        return out

    # -- timing ----------------------------------------------------------------
    HIST = ("t", "dtv", "dtc", "dt", "err", "Qx", "QxT", "mpgx", "dgx", "Umax", "Uslip")

    def hist(self):
        """The history-file record of the current state (record_out,
        FOU3D.f90:2504-2534) as a dict; sets the Qx scalar."""
        v = np.zeros(len(self.HIST))
        self._call("cfd_hist", _dp(v))
        return dict(zip(self.HIST, v.tolist()))

    def set_mode_filter(self, alpha, order=8, eta_c=0.5):
        """Modal filter after every RK step: sigma = exp(-alpha ((eta -
        eta_c) / (1 - eta_c))^order) for eta = max(|i|/(N1/2), |kz|/(N2/2))
        above eta_c; alpha = 0 switches it off."""
        self._call("cfd_set_mode_filter", float(alpha), int(order), float(eta_c))

    def set_compat_lastplane(self, np_ranks):
        """The reference's as-shipped nonlinear term of a run on np_ranks
        ranks (every plane of a rank takes the x/z terms of the rank's last
        plane, FOU3D.f90:1553-1613); 0 restores the per-plane term."""
        self._call("cfd_set_compat_lastplane", int(np_ranks))

    def meanflow_ctP(self):
        self._call("cfd_meanflow_ctP")

    def tuning(self):
        t = (ctypes.c_int * 3)()
        self._call("cfd_get_tuning", t)
        return dict(zip(TUNING_KEYS, list(t)))

    def set_tuning(self, **t):
        """Change seg_rows / solve_threads (the FFT engine is fixed)."""
        self._call("cfd_set_tuning", _tuning_c(dict(self.tuning(), **t)))

    def profile(self, enable=True):
        self._call("cfd_profile", 1 if enable else 0)

    def profile_read(self):
        ms = np.zeros(len(KCLASSES))
        cnt = np.zeros(len(KCLASSES), dtype=np.int32)
        self._call("cfd_profile_read", _dp(ms), cnt.ctypes.data_as(_I))
        return {k: (float(m), int(c)) for k, m, c in zip(KCLASSES, ms, cnt)}
