"""Seeded initial conditions for the rough channel.

Laminar tip-to-tip Poiseuille profile u = 1.5 (1 - y^2) (bulk velocity 1,
the scale of the bulk Reynolds number of input.in:10) on the (0,0) mode,
zero inside the canopies, plus a random perturbation: Gaussian complex
coefficients for the modes 1 <= |(i, kz)| <= kmax drawn from a Philox stream,
times the wall-vanishing envelope (1 - y^2)^2, scaled to ``amp`` times the
centreline velocity.  ``project`` then removes the divergence with the
solver's own discrete projection (solveP + v_corr), so the field is
divergence-free in the operator the time stepping uses.
"""
import numpy as np


def mode_index(N2, kz):
    """Reference column k index (0-based) of z wavenumber index kz."""
    return kz if kz >= 0 else N2 + kz


def initial_fields(params, ygrid, seed=0, amp=0.1, kmax=8):
    p = params
    Mx, N2, Ncol = p.Mx, p.N2, p.Ncol
    yu, yv = ygrid["yu"], ygrid["yv"]
    rng = np.random.Generator(np.random.Philox(seed))
    Uc = 1.5
    env_u = np.where(np.abs(yu) < 1.0, (1.0 - yu ** 2) ** 2, 0.0)
    env_v = np.where(np.abs(yv) < 1.0, (1.0 - yv ** 2) ** 2, 0.0)
    u = [np.zeros((Ncol, p.nyu), complex), np.zeros((Ncol, p.nyv), complex), np.zeros((Ncol, p.nyu), complex)]
    modes = []
    for i in range(0, min(kmax, Mx - 1) + 1):
        for kz in range(-kmax, kmax + 1):
            if i == 0 and kz <= 0:
                continue  # i = 0: kz < 0 is the conjugate of kz > 0
            if not (1 <= i * i + kz * kz <= kmax * kmax):
                continue
            if kz <= -N2 // 2 or kz >= N2 // 2:
                continue
            modes.append((i, kz))
    coef = rng.standard_normal((len(modes), 3, 2))
    # physical RMS of the sum over modes (each mode and its conjugate)
    rms = np.sqrt(np.sum(coef ** 2))
    scale = amp * Uc / max(rms, 1e-300)
    for m, (i, kz) in enumerate(modes):
        col = mode_index(N2, kz) * Mx + i
        for c in range(3):
            a = scale * (coef[m, c, 0] + 1j * coef[m, c, 1])
            env = env_v if c == 1 else env_u
            u[c][col, :] += a * env
            if i == 0:
                colc = mode_index(N2, -kz) * Mx
                u[c][colc, :] += np.conj(a) * env
    u[0][0, :] = np.where(np.abs(yu) < 1.0, Uc * (1.0 - yu ** 2), 0.0)
    return dict(u1=u[0], u2=u[1], u3=u[2],
                p=np.zeros((Ncol, p.nyp), complex))


def flowrate(u1, ygrid, ny):
    """flowrateIm of the (0,0) u column, tip to tip (littleharsh.f90:411-428)."""
    yv = ygrid["yv"]
    dsty = (len(yv) - ny) // 2
    Q = 0.0
    for j in range(1, ny):
        Q = Q + (yv[j + dsty] - yv[j - 1 + dsty]) * u1[0, j + dsty].real
    return Q


def project(solver, fields):
    """Divergence-free projection with the solver's own solveP + v_corr.

    With dt = dti = 1 and kRK = 1 the correction subtracts
    (aRK+bRK) G psi with psi = (D G)^{-1} D u / (aRK+bRK), i.e. the exact
    discrete projection; the pressure it leaves behind is discarded.
    """
    for k in ("u1", "u2", "u3"):
        solver.set_field(k, fields[k])
    solver.set_field("p", np.zeros_like(fields["p"]))
    solver.set_scalars(dt=1.0, dti=1.0)
    solver.solveP(1)
    solver.v_corr(1)
    out = {k: solver.get_field(k) for k in ("u1", "u2", "u3")}
    out["p"] = np.zeros_like(fields["p"])
    solver.set_field("p", out["p"])
    return out
