"""The oracle with a coarser central band (multi-band resolution,
start.f90:89-111, 1420-1557), CPU.

Pins, each independent of the oracle's own arithmetic:

* band limits and column sets: the band table's y limits equal the product's
  cfd_band_table (itself pinned against the compiled reference); the short
  columns are the wall bands' modes outside the central band's N1c x N2c.
* solves: each band of a short column is a closed system with boundary rows
  at the band interface (linear extrapolation to zero at the interface face,
  gridweighting(1,2) and (3,1), start.f90:494-500): numpy dense solves of the
  u, v and pressure systems per band.
* projection: after solveP + v_corr every band's divergence is round-off and
  a short column's v vanishes on its interface faces.
* nonlinear term: a numpy restatement that transforms each plane on its
  band's grid (irfft2/rfft2 of the wall bands' or the central band's size,
  the central planes with the long columns only) and zeroes a short column's
  fluxes on central planes (they are never filled, FOU3D.f90:1662-1758).
* a Stokes mode in a short column decays by the closed-form RK3/CN factor
  of its band's eigenvalue; the Orr-Sommerfeld mode (a long column) grows on
  the multi-band grid as on the uniform grid of the central band's size.
"""
import ctypes

import numpy as np
import pytest

from cfd_amd import Params, ic
from oracle.oracle import Oracle, set_threads

from .helpers import random_field, rel_l2
from .test_oracle_ops import helmholtz_matrix, wavenumbers
from .test_oracle_substep_kats import RK_A, _os_run

MB = Params(ngalx=24, ngalz=48, ny=33, dsty=3, ygrid=1, ytanh_beta=2.0, geometry_type=3, ntilex=4, ntilez=8, r=1.2,
            c_i=3.5, c_k=3.5, stag=0.5, extras24=0, Re=2000.0, alp=1.0, bet=4.0, posth=0.06, CFL=0.5, ngalx2=12,
            ngalz2=24)
MB_SMOOTH = MB.replace(geometry_type=0, dsty=0, posth=0.0)


@pytest.fixture(autouse=True, scope="module")
def _one_thread():
    set_threads(1)
    yield
    import os
    set_threads(len(os.sched_getaffinity(0)))


def band_info(o):
    b = (ctypes.c_int * 11)()
    sh = np.zeros(o.params.Ncol, np.uint8)
    o.lib.or_band_info(o.h, b, sh.ctypes.data_as(ctypes.c_void_p))
    b = list(b)
    return dict(mb=b[0], N1c=b[1], N2c=b[2], N3=b[3:7], N4=b[7:11], short=sh.astype(bool))


def segments(p, info, col, grid):
    """Rows (j) of each band system of a column: [lo, hi] inclusive."""
    lo = {"u": -p.dsty, "v": -p.dsty, "p": -p.dsty + 1}[grid]
    hi = {"u": p.ny + p.dsty, "v": p.ny - 1 + p.dsty, "p": p.ny - 1 + p.dsty}[grid]
    if not info["short"][col]:
        return [(lo, hi)]
    N3, N4 = info["N3"], info["N4"]
    t1 = {"u": N4[1] + 1, "v": N3[1] + 1, "p": N4[1]}[grid]
    b3 = {"u": N4[2], "v": N3[2], "p": N4[2] + 1}[grid]
    return [(lo, t1), (b3, hi)]


def mb_state(p, seed):
    """Random fields with a short column's rows between its bands at zero."""
    o = Oracle(p)
    info = band_info(o)
    rng = np.random.default_rng(seed)
    f = {}
    for n in ("u1", "u2", "u3", "p", "Nu1", "Nu2", "Nu3", "du1", "du2", "du3"):
        a = random_field(p, n, rng, amp=1.0 if n[0] == "u" else 0.3)
        grid = {"u2": "v", "Nu2": "v", "du2": "v", "p": "p"}.get(n, "u")
        jlo = {"u": -p.dsty, "v": -p.dsty, "p": -p.dsty + 1}[grid]
        for col in np.nonzero(info["short"])[0]:
            keep = np.zeros(a.shape[1], bool)
            for lo, hi in segments(p, info, col, grid):
                keep[lo - jlo:hi - jlo + 1] = True
            a[col, ~keep] = 0.0
        if n == "u2":  # no flow through the walls or a short column's interface faces
            for col in range(p.Ncol):
                for lo, hi in segments(p, info, col, "v"):
                    a[col, lo - jlo] = a[col, hi - jlo] = 0.0
        f[n] = a
    f["u1"][0, :] += 1.0
    for n, a in f.items():
        o.set_field(n, a)
    o.set_scalars(t=0.0, dt=3e-3, dti=1 / 3e-3, mpgx=-2e-3, QxT=o.flowrate() + 0.02)
    return o, f, info


def test_band_table_and_columns_match_the_product():
    from cfd_amd import _lib
    o = Oracle(MB)
    info = band_info(o)
    assert info["mb"] == 1 and info["N1c"] == 8 and info["N2c"] == 16
    N = (ctypes.c_int * 20)()
    assert _lib.load().cfd_band_table(ctypes.byref(MB.to_c()), N) == 0
    N = list(N)
    assert info["N3"] == [N[4 * b + 2] for b in range(4)]
    assert info["N4"] == [N[4 * b + 3] for b in range(4)]
    assert np.array_equal(info["short"], MB.short_columns())
    # a central band of several planes between the wall bands
    assert info["N3"][1] + 2 < info["N3"][2]


def _gw_interface(g, info, jlo_u, jlo_v):
    """u-grid boundary weights at the band interfaces: linear extrapolation
    to zero at the interface faces yv(N(3,1)+1) and yv(N(3,2))."""
    yu, yv = g["yu"], g["yv"]
    N3, N4 = info["N3"], info["N4"]
    U = lambda j: yu[j - jlo_u]  # noqa: E731
    V = lambda j: yv[j - jlo_v]  # noqa: E731
    gw12 = (U(N4[1] + 1) - V(N3[1] + 1)) / (V(N3[1] + 1) - U(N4[1]))
    gw31 = -(V(N3[2]) - U(N4[2])) / (V(N3[2]) - U(N4[2] + 1))
    return gw12, gw31


@pytest.mark.parametrize("comp", [1, 2, 3])
def test_multiband_helmholtz_vs_dense(comp):
    p = MB_SMOOTH
    o, f, info = mb_state(p, 11 + comp)
    g = o.ygrid()
    m = o.metrics()
    kx, kz = wavenumbers(p)
    gw12, gw31 = _gw_interface(g, info, -p.dsty, -p.dsty)
    if comp == 2:
        o.solveV(2)
        got, du0, u0, d2, grid = o.get_field("u2"), f["du2"], f["u2"], g["dyv2i"], "v"
    else:
        o.solveU(comp, 2)
        got, du0, u0, d2, grid = o.get_field("u%d" % comp), f["du%d" % comp], f["u%d" % comp], g["dyu2i"], "u"
    beta = RK_A[1] * 3e-3 / p.Re
    jlo = -p.dsty
    checked = 0
    for col in range(p.Ncol):
        k2 = -(kx[col % p.Mx] ** 2) - kz[col // p.Mx] ** 2  # k2F = -k^2 (def_k, start.f90:965-986)
        for lo, hi in segments(p, info, col, grid):
            a, b = lo - jlo, hi - jlo + 1
            if grid == "u":
                g0 = m["gw"][0] if lo == jlo else gw31
                g1 = m["gw"][1] if hi == p.ny + p.dsty else gw12
            else:
                g0 = g1 = 0.0
            A = helmholtz_matrix(d2[a:b], k2, beta, g0, g1)
            r = u0[col, a:b] + 3e-3 * du0[col, a:b]
            r[0] = r[-1] = 0.0
            want = np.linalg.solve(A, r)
            assert np.abs(got[col, a:b] - want).max() <= 1e-12 * max(1.0, np.abs(want).max()), (col, lo, hi)
            checked += 1
    assert checked > p.Ncol  # short columns contributed two systems


def test_multiband_pressure_vs_dense_and_projection():
    p = MB_SMOOTH
    o, f, info = mb_state(p, 21)
    g = o.ygrid()
    kx, kz = wavenumbers(p)
    o.solveP(1)
    psi, div = o.get_field("psi"), o.get_field("div")
    dtirk = (1 / 3e-3) / (2 * RK_A[0])
    du, dv = g["dthdyu"], g["dthdyv"]
    jp, ju, jv = -p.dsty + 1, -p.dsty, -p.dsty
    for col in range(0, p.Ncol, 7):
        k2 = kx[col % p.Mx] ** 2 + kz[col // p.Mx] ** 2
        for lo, hi in segments(p, info, col, "p"):
            n = hi - lo + 1
            A = np.zeros((n, n))
            for q in range(n):
                j = lo + q
                fl = dv[j - 1 - jv] if q > 0 else 0.0   # flux through the face below (dropped at the band end)
                fu_ = dv[j - jv] if q < n - 1 else 0.0  # and above
                A[q, q] = -du[j - ju] * (fl + fu_) - k2
                if q > 0:
                    A[q, q - 1] = du[j - ju] * dv[j - 1 - jv]
                if q < n - 1:
                    A[q, q + 1] = du[j - ju] * dv[j - jv]
            r = dtirk * div[col, lo - jp:hi - jp + 1]
            if col == 0:
                A[0, :] = 0.0
                A[0, 0] = 1.0 / (g["yu"][2 - ju] - g["yu"][1 - ju]) ** 2
                r[0] = 0.0
            want = np.linalg.solve(A, r)
            got = psi[col, lo - jp:hi - jp + 1]
            assert np.abs(got - want).max() <= 1e-10 * max(1.0, np.abs(want).max()), (col, lo, hi)
    o.v_corr(1)
    u1, u2 = o.get_field("u1"), o.get_field("u2")
    div = o.get_field("div")
    scale = np.abs(u1).max()
    for col in range(p.Ncol):
        for lo, hi in segments(p, info, col, "p"):
            assert np.abs(div[col, lo - jp:hi - jp + 1]).max() <= 1e-11 * scale, col
        if info["short"][col]:
            for lo, hi in segments(p, info, col, "v"):
                assert u2[col, lo - jv] == 0 and u2[col, hi - jv] == 0


def nl_numpy_mb(p, g, m, info, u1, u2, u3):
    """The multi-band nonlinear term with numpy FFTs: the x/z derivative parts
    and the y fluxes of every plane on its band's grid, then the y derivatives
    per band system of every column."""
    yu, yv = g["yu"], g["yv"]
    nyu, nyv = p.nyu, p.nyv
    jlo = -p.dsty
    short = info["short"]
    N3, N4 = info["N3"], info["N4"]
    gw = m["gwi"]
    # interpolations per band system
    u1i = np.zeros((p.Ncol, nyv), complex)
    u3i = np.zeros_like(u1i)
    u2i = np.zeros((p.Ncol, nyu), complex)
    for col in range(p.Ncol):
        for (lo, hi) in segments(p, info, col, "v"):
            j = np.arange(lo, hi + 1) - jlo
            w = (yv[j] - yu[j]) / (yu[j + 1] - yu[j])
            u1i[col, j] = w * u1[col, j + 1] + (1 - w) * u1[col, j]
            u3i[col, j] = w * u3[col, j + 1] + (1 - w) * u3[col, j]
        for (lo, hi) in segments(p, info, col, "u"):
            j = np.arange(lo + 1, hi) - jlo
            u2i[col, j] = ((yu[j] - yv[j - 1]) * u2[col, j] + (yv[j] - yu[j]) * u2[col, j - 1]) / (yv[j] - yv[j - 1])
    kx, kz = wavenumbers(p)
    ikx = 1j * np.tile(kx, p.N2)
    ikz = 1j * np.repeat(kz, p.Mx)
    kk = np.arange(p.N2)
    fine = dict(ngx=p.ngalx, ngz=p.ngalz, N1=p.N1, N2=p.N2, cols=np.ones(p.Ncol, bool))
    coarse = dict(ngx=p.ngalx2, ngz=p.ngalz2, N1=p.N1c, N2=p.N2c, cols=~short)

    def view(grid, j):
        cen = (N4[1] < j <= N4[2]) if grid == "u" else (N3[1] < j <= N3[2])
        return coarse if cen else fine

    def phys(v, f, q):
        slot = np.where(kk < p.N2 // 2, kk, kk - p.N2 + v["ngz"])
        S = np.zeros((v["ngz"], v["ngx"] // 2 + 1), complex)
        a = np.where(v["cols"], f[:, q], 0).reshape(p.N2, p.Mx)
        for k in range(p.N2):
            for i in range(p.Mx):
                if a[k, i] != 0:
                    S[slot[k], i] = a[k, i]
        return np.fft.irfft2(S, s=(v["ngz"], v["ngx"])) * (v["ngx"] * v["ngz"])

    def four(v, a):
        slot = np.where(kk < p.N2 // 2, kk, kk - p.N2 + v["ngz"])
        F = np.fft.rfft2(a) / (v["ngx"] * v["ngz"])
        out = np.zeros((p.N2, p.Mx), complex)
        for k in range(p.N2):
            for i in range(p.Mx):
                col = k * p.Mx + i
                if v["cols"][col]:
                    out[k, i] = F[slot[k], i]
        return out.reshape(-1)

    def odd(v):
        slot = np.where(kk < p.N2 // 2, kk, kk - p.N2 + v["ngz"])
        return np.repeat(slot == v["ngz"] - v["N2"] // 2, p.Mx)

    Nu1 = np.zeros((p.Ncol, nyu), complex)
    Nu3 = np.zeros_like(Nu1)
    VV = np.zeros_like(Nu1)
    Nu2 = np.zeros((p.Ncol, nyv), complex)
    UV = np.zeros_like(Nu2)
    VW = np.zeros_like(Nu2)
    for q in range(1, nyu - 1):
        v = view("u", q + jlo)
        a, b, c = phys(v, u1, q), phys(v, u2i, q), phys(v, u3, q)
        uu, uw, vv, ww = four(v, a * a), four(v, a * c), four(v, b * b), four(v, c * c)
        VV[:, q] = vv
        Nu1[:, q] = np.where(odd(v), 0, ikx * uu + ikz * uw)
        Nu3[:, q] = np.where(odd(v), 0, ikx * uw + ikz * ww)
    for q in range(nyv):
        v = view("v", q + jlo)
        a, b, c = phys(v, u1i, q), phys(v, u2, q), phys(v, u3i, q)
        uv, vw = four(v, a * b), four(v, c * b)
        UV[:, q], VW[:, q] = uv, vw
        if 0 < q < nyv - 1:
            Nu2[:, q] = np.where(odd(v), 0, ikx * uv + ikz * vw)
    dthu, dthv = g["dthdyu"], g["dthdyv"]
    for col in range(p.Ncol):
        for (lo, hi) in segments(p, info, col, "u"):
            j = np.arange(lo + 1, hi) - jlo
            Nu1[col, j] += (UV[col, j] - UV[col, j - 1]) * dthu[j]
            Nu3[col, j] += (VW[col, j] - VW[col, j - 1]) * dthu[j]
        for (lo, hi) in segments(p, info, col, "v"):
            j = np.arange(lo + 1, hi) - jlo
            Nu2[col, j] += (VV[col, j + 1] - VV[col, j]) * dthv[j]
    return Nu1, Nu2, Nu3


def test_multiband_nonlinear_vs_numpy():
    p = MB_SMOOTH
    o, f, info = mb_state(p, 31)
    g, m = o.ygrid(), o.metrics()
    o.nonlinear(2)
    Nu1, Nu2, Nu3 = nl_numpy_mb(p, g, m, info, f["u1"], f["u2"], f["u3"])
    jlo = -p.dsty
    for name, want, grid in (("Nu1", Nu1, "u"), ("Nu2", Nu2, "v"), ("Nu3", Nu3, "u")):
        got = o.get_field(name)
        err, ref = 0.0, 0.0
        for col in range(p.Ncol):
            for lo, hi in segments(p, info, col, grid):
                a, b = lo - jlo + 1, hi - jlo
                err += np.sum(np.abs(got[col, a:b] - want[col, a:b]) ** 2)
                ref += np.sum(np.abs(want[col, a:b]) ** 2)
        assert np.sqrt(err / ref) < 1e-12, (name, np.sqrt(err / ref))


def test_multiband_stokes_short_column_closed_form():
    """u = v(y) cos(beta z) in a short column (kz beyond the central band's
    modes), one eigenvector of each wall band's viscous operator (the band-1
    system ends at the interface ghost row with gridweighting(1,2), the
    band-3 system starts with gridweighting(3,1)): no nonlinear term, no
    pressure, and each band decays by its own closed-form RK3/CN factor."""
    p = MB_SMOOTH.replace(flag_ctpress=1, mpgx=0.0, Re=400.0)
    o = Oracle(p)
    info = band_info(o)
    g, m = o.ygrid(), o.metrics()
    jlo = -p.dsty
    gw12, gw31 = _gw_interface(g, info, jlo, jlo)
    kq = p.N2c // 2 + 1                  # first z mode outside the central band
    assert kq < p.N2 // 2
    c1, c2 = kq * p.Mx, (p.N2 - kq) * p.Mx
    assert info["short"][c1] and info["short"][c2]
    beta = kq * p.bet
    col = np.zeros(p.nyu)
    lams = []
    for (lo, hi), (ga, gb) in zip(segments(p, info, c1, "u"), ((m["gw"][0], gw12), (gw31, m["gw"][1]))):
        a, b = lo - jlo, hi - jlo
        d2 = g["dyu2i"][a:b + 1]
        n = b - a - 1
        L = np.zeros((n, n))
        for r in range(n):
            q = r + 1
            L[r, r] = d2[q, 1] - beta * beta
            if r > 0:
                L[r, r - 1] = d2[q, 0]
            else:
                L[r, r] += -ga * d2[q, 0]
            if r < n - 1:
                L[r, r + 1] = d2[q, 2]
            else:
                L[r, r] += -gb * d2[q, 2]
        w, V = np.linalg.eig(L)
        kk = np.argmax(w.real)
        vint = V[:, kk].real / np.abs(V[:, kk].real).max()
        col[a + 1:b] = vint
        col[a], col[b] = -ga * vint[0], -gb * vint[-1]
        lams.append(w[kk].real)
    f = {n: np.zeros(o.shape(n), complex) for n in ("u1", "u2", "u3", "p")}
    f["u1"][c1] = f["u1"][c2] = 0.5 * col
    for n in f:
        o.set_field(n, f[n])
    o.set_scalars(t=0.0, dt=0.0, mpgx=0.0, QxT=0.0)
    amp = [1.0, 1.0]
    for _ in range(5):
        o.step(1)
        dt = o.scalars()["dt"]
        for a_k in RK_A:
            for b in range(2):
                z = a_k * dt * lams[b] / p.Re
                amp[b] *= (1.0 + z) / (1.0 - z)
    got = o.get_field("u1")
    want = np.zeros_like(got)
    for b, (lo, hi) in enumerate(segments(p, info, c1, "u")):
        want[c1, lo - jlo:hi - jlo + 1] = 0.5 * amp[b] * col[lo - jlo:hi - jlo + 1]
    want[c2] = want[c1]
    assert rel_l2(got, want) <= 1e-12, rel_l2(got, want)
    assert abs(lams[0] - lams[1]) <= 1e-9 * abs(lams[0])  # the tanh grid is symmetric
    for n in ("u2", "u3", "p"):
        assert np.abs(o.get_field(n)).max() <= 1e-13


def test_multiband_orr_sommerfeld_matches_uniform_central_grid():
    """The Orr-Sommerfeld mode (alpha = 1, a long column) on a grid whose wall
    bands are twice as fine in x as its central band grows and travels as on
    the uniform grid of the central band's size: the products of the base
    flow and the mode are exact on both grids, so the rates agree to the
    mode's amplitude (1e-6), and both are Orszag's to the y-discretisation
    error.  Both run with dt = 0.2 fixed (the CFL step would differ: the
    viscous limit dtv takes the wall bands' largest wavenumber)."""
    gm, om = _os_run(49, 10.0, 20.0, ngalx=12, ngalx2=6, dt=0.2)
    gu, ou = _os_run(49, 10.0, 20.0, ngalx=6, dt=0.2)
    assert abs(gm - gu) <= 1e-8 and abs(om - ou) <= 1e-8, (gm, gu, om, ou)
    assert abs(gm - 0.00373967) <= 1.5e-3 and abs(om + 0.23752649) <= 4e-3
