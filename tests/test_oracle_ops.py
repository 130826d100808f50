"""Pins of the oracle's spectral operators (CPU).

Each check is independent of the oracle's own arithmetic: dense linear
algebra (numpy.linalg.solve) for the tridiagonal family, numpy's FFT for the
nonlinear term and the IB transforms, closed forms on polynomial profiles for
the Laplacian and gradients, and invariants of the projection (discrete
divergence-free), the flow-rate correction (Q = QxT) and the laminar channel
(non-mean modes stay zero).
"""
import numpy as np
import pytest

from cfd_amd import Params
from oracle.oracle import Oracle

from .helpers import SHIPPED, SMOOTH, random_field, rel_l2

RK_A = (4 / 15, 1 / 15, 1 / 6)
RK_G = (8 / 15, 5 / 12, 3 / 4)
RK_C = (0.0, -17 / 60, -5 / 12)
RK_D = (8 / 15, 2 / 15, 1 / 3)

SMALL_IB = Params(ngalx=48, ngalz=36, ny=21, dsty=3, ygrid=1, geometry_type=3, ntilex=2, ntilez=2, r=3.1,
                  c_i=12.0, c_k=9.0, stag=0.5, extras24=0, Re=1500.0, posth=0.07)


def wavenumbers(p):
    kx = p.alp * np.arange(p.Mx)
    k = np.arange(p.N2)
    kz = np.where(k < p.N2 // 2, p.bet * k, -p.bet * (p.N2 - k))
    return kx, kz


def state(p, seed, walls_zero=True):
    o = Oracle(p)
    rng = np.random.default_rng(seed)
    f = {}
    for n in ("u1", "u2", "u3", "p", "Nu1", "Nu2", "Nu3", "du1", "du2", "du3"):
        f[n] = random_field(p, n, rng, amp=1.0 if n[0] == "u" else 0.3)
    if walls_zero:
        f["u2"][:, 0] = f["u2"][:, -1] = 0.0
    f["u1"][0, :] += 1.0
    for n, a in f.items():
        o.set_field(n, a)
    o.set_scalars(t=0.0, dt=3e-3, dti=1 / 3e-3, mpgx=-2e-3, QxT=o.flowrate() + 0.02)
    return o, f


# --------------------------------------------------------------------------
# tridiagonal solves vs dense solves
# --------------------------------------------------------------------------
def helmholtz_matrix(d2, k2, beta, g0, g1):
    n = d2.shape[0]
    A = np.zeros((n, n))
    A[0, 0], A[0, 1] = 1.0, g0
    A[-1, -1], A[-1, -2] = 1.0, g1
    for q in range(1, n - 1):
        A[q, q - 1] = -beta * d2[q, 0]
        A[q, q] = 1.0 - beta * (d2[q, 1] + k2)
        A[q, q + 1] = -beta * d2[q, 2]
    return A


@pytest.mark.parametrize("comp", [1, 2, 3])
def test_solve_vs_dense(comp):
    p = SMOOTH
    o, f = state(p, 3 + comp)
    kRK = 2
    dt = o.scalars()["dt"]
    g = o.ygrid()
    m = o.metrics()
    kx, kz = wavenumbers(p)
    name = "u%d" % comp
    if comp == 2:
        o.solveV(kRK)
        d2, g0, g1 = g["dyv2i"], 0.0, 0.0
    else:
        o.solveU(comp, kRK)
        d2, g0, g1 = g["dyu2i"], m["gw"][0], m["gw"][1]
    got = o.get_field(name)
    beta = RK_A[kRK - 1] * dt / p.Re
    for col in range(0, p.Ncol, 7):
        i, k = col % p.Mx, col // p.Mx
        r = f[name][col] + dt * f["d" + name][col]
        r[0] = r[-1] = 0.0
        A = helmholtz_matrix(d2, -(kx[i] ** 2) - kz[k] ** 2, beta, g0, g1)
        x = np.linalg.solve(A, r)
        assert np.abs(got[col] - x).max() <= 1e-12 * max(1.0, np.abs(x).max())


def test_pressure_solve_vs_dense_and_projection():
    p = SHIPPED
    o, f = state(p, 11)
    kRK = 1
    s = o.scalars()
    g = o.ygrid()
    kx, kz = wavenumbers(p)
    o.solveP(kRK)
    psi, div = o.get_field("psi"), o.get_field("div")
    du, dv = g["dthdyu"], g["dthdyv"]
    n = p.nyp
    for col in range(0, p.Ncol, 5):
        i, k = col % p.Mx, col // p.Mx
        # D_y G_y - k^2 on the p grid with no flux through the walls
        M = np.zeros((n, n))
        for q in range(n):
            j = q + 1  # u-grid index of p row q
            up = du[j] * dv[j] if q < n - 1 else 0.0
            lo = du[j] * dv[j - 1] if q > 0 else 0.0
            M[q, q] = -up - lo - kx[i] ** 2 - kz[k] ** 2
            if q < n - 1:
                M[q, q + 1] = up
            if q > 0:
                M[q, q - 1] = lo
        rhs = div[col] * s["dti"] / (RK_A[kRK - 1] * 2)
        if col == 0:
            M[0, :] = 0.0
            M[0, 0] = 1.0
            rhs[0] = 0.0
        x = np.linalg.solve(M, rhs)
        assert np.abs(psi[col] - x).max() <= 1e-10 * max(1e-12, np.abs(x).max())
    o.v_corr(kRK)
    d = o.get_field("div")
    assert np.abs(d).max() <= 1e-12 * np.abs(o.get_field("u1")).max()


def test_meanflow_holds_flow_rate():
    p = SHIPPED
    o, f = state(p, 12)
    QxT = o.scalars()["QxT"]
    o.meanflow_ctU(2)
    assert abs(o.flowrate() - QxT) <= 1e-13 * abs(QxT)
    u = o.get_field("u1")
    assert np.array_equal(u[1:], f["u1"][1:])  # only the (0,0) mode changes
    du = (u[0] - f["u1"][0]).real
    assert np.all(np.imag(u[0]) == np.imag(f["u1"][0]))
    nn = p.ny - 2
    # the correction is g * u11, u11 the discrete (I - beta D2) solve of -dRK dt
    # on the tip-to-tip rows with the wall ghost rows (flowrate_corr)
    s = o.scalars()
    g = o.ygrid()
    m = o.metrics()
    j0 = p.dsty  # array index of j = 0
    d2 = g["dyu2i"][j0:j0 + nn + 3]
    A = helmholtz_matrix(d2, 0.0, RK_A[1] * s["dt"] / p.Re, m["gw"][0], m["gw"][1])
    b = np.full(nn + 3, -RK_D[1] * s["dt"])
    b[0] = b[-1] = 0.0
    u11 = np.linalg.solve(A, b)
    assert np.allclose(du[j0:j0 + nn + 3], s["dgx"] * u11, rtol=1e-11, atol=1e-15)
    assert np.all(du[:j0] == 0) and np.all(du[j0 + nn + 3:] == 0)


# --------------------------------------------------------------------------
# RHS0: closed forms
# --------------------------------------------------------------------------
@pytest.mark.parametrize("kRK", [1, 2, 3])
def test_rhs0_closed_form(kRK):
    p = SMOOTH.replace(dsty=2, posth=0.05)
    o = Oracle(p)
    g = o.ygrid()
    kx, kz = wavenumbers(p)
    yu, yv = g["yu"], g["yv"]
    rng = np.random.default_rng(kRK)
    a, b, c = rng.standard_normal((3, p.Ncol)) + 1j * rng.standard_normal((3, p.Ncol))
    U = a[:, None] + b[:, None] * yu + c[:, None] * yu ** 2
    V = a[:, None] + b[:, None] * yv + c[:, None] * yv ** 2
    P = random_field(p, "p", rng)
    N = {n: random_field(p, n, rng) for n in ("Nu1", "Nu2", "Nu3")}
    o.set_field("u1", U)
    o.set_field("u3", 2 * U)
    o.set_field("u2", V)
    o.set_field("p", P)
    for n, v in N.items():
        o.set_field(n, v)
    o.set_scalars(mpgx=-3e-3)
    o.RHS0(kRK)
    A, C, D = RK_A[kRK - 1], RK_C[kRK - 1], RK_D[kRK - 1]
    k2 = -(kx[None, :] ** 2 + kz[:, None] ** 2).reshape(-1, 1)
    lapU = 2 * c[:, None] + k2 * U
    lapV = 2 * c[:, None] + k2 * V
    ikx = 1j * np.tile(kx, p.N2)[:, None]
    ikz = 1j * np.repeat(kz, p.Mx)[:, None]
    want1 = (A / p.Re * lapU - C * N["Nu1"])[:, 1:-1] - D * ikx * P
    want1[0] += -D * (-3e-3)
    want3 = (A / p.Re * 2 * lapU - C * N["Nu3"])[:, 1:-1] - D * ikz * P
    want2 = (A / p.Re * lapV - C * N["Nu2"])[:, 1:-1] - D * g["dthdyv"][1:-1] * (P[:, 1:] - P[:, :-1])
    for name, want in (("du1", want1), ("du2", want2), ("du3", want3)):
        got = o.get_field(name)[:, 1:-1]
        assert rel_l2(got, want) < 1e-9, name


# --------------------------------------------------------------------------
# nonlinear term vs numpy FFT
# --------------------------------------------------------------------------
def nl_numpy(p, g, m, u1, u2, u3):
    yu, yv = g["yu"], g["yv"]
    nyu, nyv = p.nyu, p.nyv
    u1i = ((yv - yu[:-1]) * u1[:, 1:] + (yu[1:] - yv) * u1[:, :-1]) / (yu[1:] - yu[:-1])
    u3i = ((yv - yu[:-1]) * u3[:, 1:] + (yu[1:] - yv) * u3[:, :-1]) / (yu[1:] - yu[:-1])
    u2i = np.zeros((p.Ncol, nyu), complex)
    j = np.arange(1, nyu - 1)
    u2i[:, j] = ((yu[j] - yv[j - 1]) * u2[:, j] + (yv[j] - yu[j]) * u2[:, j - 1]) / (yv[j] - yv[j - 1])
    u2i[:, 0] = m["gwi"][0] * (u2[:, 0] - u2i[:, 1]) + u2i[:, 1]
    u2i[:, -1] = m["gwi"][1] * (u2[:, -1] - u2i[:, -2]) + u2i[:, -2]
    ngx, ngz, N2, Mx = p.ngalx, p.ngalz, p.N2, p.Mx
    kk = np.arange(N2)
    slot = np.where(kk < N2 // 2, kk, kk - N2 + ngz)

    def phys(f, q):
        S = np.zeros((ngz, ngx // 2 + 1), complex)
        S[slot, :Mx] = f[:, q].reshape(N2, Mx)
        return np.fft.irfft2(S, s=(ngz, ngx)) * (ngx * ngz)

    def four(a):
        F = np.fft.rfft2(a) / (ngx * ngz)
        return F[slot, :Mx].reshape(-1)

    kx, kz = wavenumbers(p)
    ikx = 1j * np.tile(kx, N2)
    ikz = 1j * np.repeat(kz, Mx)
    odd = np.repeat(slot == ngz - N2 // 2, Mx)
    Nu1 = np.zeros((p.Ncol, nyu), complex)
    Nu3 = np.zeros_like(Nu1)
    VV = np.zeros_like(Nu1)
    Nu2 = np.zeros((p.Ncol, nyv), complex)
    UV = np.zeros_like(Nu2)
    VW = np.zeros_like(Nu2)
    m1 = m3 = 0.0
    m2 = 1e-6
    for q in range(nyu):
        a, b, c = phys(u1, q), phys(u2i, q), phys(u3, q)
        m1, m3 = max(m1, np.abs(a).max()), max(m3, np.abs(c).max())
        uu, uw, vv, ww = four(a * a), four(a * c), four(b * b), four(c * c)
        VV[:, q] = vv
        Nu1[:, q] = np.where(odd, 0, ikx * uu + ikz * uw)
        Nu3[:, q] = np.where(odd, 0, ikx * uw + ikz * ww)
    for q in range(nyv):
        a, b, c = phys(u1i, q), phys(u2, q), phys(u3i, q)
        if 0 < q < nyv - 1:
            m2 = max(m2, (np.abs(b) * g["dthdyv"][q]).max())
        uv, vw = four(a * b), four(c * b)
        UV[:, q], VW[:, q] = uv, vw
        Nu2[:, q] = np.where(odd, 0, ikx * uv + ikz * vw)
    Nu1[:, 1:-1] += (UV[:, 1:] - UV[:, :-1]) * g["dthdyu"][1:-1]
    Nu3[:, 1:-1] += (VW[:, 1:] - VW[:, :-1]) * g["dthdyu"][1:-1]
    Nu2[:, 1:-1] += (VV[:, 2:-1] - VV[:, 1:-2]) * g["dthdyv"][1:-1]
    dtc = min(1 / (p.alp * (p.N1 // 2) * m1), 1 / (p.bet * (p.N2 // 2) * m3), 1 / m2)
    return Nu1, Nu2, Nu3, dtc


@pytest.mark.parametrize("case", ["smooth", "ib"])
def test_nonlinear_vs_numpy(case):
    p = SMOOTH if case == "smooth" else SMALL_IB
    o, f = state(p, 21)
    g = o.ygrid()
    m = o.metrics()
    du0 = {n: f[n] for n in ("du1", "du2", "du3")}
    o.nonlinear(1)
    Nu1, Nu2, Nu3, dtc = nl_numpy(p, g, m, f["u1"], f["u2"], f["u3"])
    for name, want in (("Nu1", Nu1), ("Nu2", Nu2), ("Nu3", Nu3)):
        got = o.get_field(name)[:, 1:-1]
        assert rel_l2(got, want[:, 1:-1]) < 1e-12, name
    for n, N in (("du1", Nu1), ("du2", Nu2), ("du3", Nu3)):
        want = du0[n][:, 1:-1] - RK_G[0] * N[:, 1:-1]
        assert rel_l2(o.get_field(n)[:, 1:-1], want) < 1e-12, n
    s = o.scalars()
    assert abs(s["dtc"] - dtc) <= 1e-13 * dtc
    assert s["dt"] == min(2.5 * m["dtv"], p.CFL * s["dtc"])


def test_ib_forcing_vs_numpy():
    """solveU with IB planes = predictor, physical-space forcing through numpy
    FFTs, dense solve."""
    p = SMALL_IB
    o, f = state(p, 31)
    kRK = 3
    dt = o.scalars()["dt"]
    o.solveU(1, kRK)
    got = o.get_field("u1")
    r = f["u1"] + dt * f["du1"]
    r[:, 0] = r[:, -1] = 0.0
    ngx, ngz, N2, Mx = p.ngalx, p.ngalz, p.N2, p.Mx
    kk = np.arange(N2)
    slot = np.where(kk < N2 // 2, kk, kk - N2 + ngz)

    def phys(a):
        S = np.zeros((ngz, ngx // 2 + 1), complex)
        S[slot, :Mx] = a.reshape(N2, Mx)
        return np.fft.irfft2(S, s=(ngz, ngx)) * (ngx * ngz)

    def four(a):
        return (np.fft.rfft2(a) / (ngx * ngz))[slot, :Mx].reshape(-1)

    lists = [o.iblist(0, 0), o.iblist(0, 1)]
    jlo = -p.dsty
    for L, W in lists:
        for j in np.unique(L[:, 2]):
            q = j - jlo
            R, U = phys(r[:, q]), phys(f["u1"][:, q])
            for e, w in zip(L[L[:, 2] == j], W[L[:, 2] == j]):
                i, k = e[0] - 1, e[1] - 1
                vf = w[1] * U[e[4] - 1, e[3] - 1] + w[2] * U[e[7] - 1, e[6] - 1]
                R[k, i] = w[0] * (vf - U[k, i] + R[k, i])
            r[:, q] = four(R)
    g = o.ygrid()
    m = o.metrics()
    kx, kz = wavenumbers(p)
    beta = RK_A[kRK - 1] * dt / p.Re
    err = 0.0
    for col in range(0, p.Ncol, 3):
        i, k = col % p.Mx, col // p.Mx
        A = helmholtz_matrix(g["dyu2i"], -(kx[i] ** 2) - kz[k] ** 2, beta, m["gw"][0], m["gw"][1])
        x = np.linalg.solve(A, r[col])
        err = max(err, np.abs(got[col] - x).max() / max(1.0, np.abs(x).max()))
    assert err < 1e-11


# --------------------------------------------------------------------------
# whole substep invariants
# --------------------------------------------------------------------------
def test_laminar_channel_stays_laminar():
    p = SMOOTH
    o = Oracle(p)
    g = o.ygrid()
    u1 = np.zeros((p.Ncol, p.nyu), complex)
    u1[0] = np.where(np.abs(g["yu"]) < 1, 1.5 * (1 - g["yu"] ** 2), 0.0)
    o.set_field("u1", u1)
    Q = o.flowrate()
    o.set_scalars(t=0.0, dt=0.0, mpgx=0.0, QxT=Q)
    o.step(5)
    a = o.get_field("u1")
    # non-mean modes and v, w stay at FFT round-off of the mean flow
    assert np.abs(a[1:]).max() < 1e-15 and np.abs(o.get_field("u2")).max() < 1e-15
    assert np.abs(o.get_field("u3")).max() < 1e-15
    assert abs(o.flowrate() - Q) < 1e-14 * Q
    s = o.scalars()
    assert s["t"] > 0 and np.isfinite(s["mpgx"])
