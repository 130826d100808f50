"""Shared helpers of the test suite: parameter sets, random spectral fields,
comparison metrics."""
import numpy as np

from cfd_amd import Params

# the geometry the reference ships (geom_circ_rough.txt) on 24 x 24 tiles,
# with the 24-point extras of the linked geometry variant
SHIPPED = Params(ngalx=48, ngalz=48, ny=25, dsty=4, dsty1=0, ppp=5, ygrid=0, geometry_type=3, ntilex=2,
                 ntilez=2, extras24=1, flag_ctpress=0, Re=2800.0, alp=1.0, bet=2.0, dyq=10.0, posth=0.089,
                 r=5.43723, c_i=12.0, c_k=12.0, stag=0.5, CFL=0.5)

# tanh grid, small staggered posts, uneven x/z sizes with radix 3
SMALLPOSTS = Params(ngalx=72, ngalz=48, ny=33, dsty=3, ygrid=1, ytanh_beta=2.0, geometry_type=3, ntilex=6,
                    ntilez=4, extras24=0, flag_ctpress=0, Re=2800.0, alp=1.0, bet=2.0, posth=0.06, r=1.6,
                    c_i=6.0, c_k=6.0, stag=0.5, CFL=0.5)

# the lateral grid and posts of the 384 x 192 x 192 workload on a short
# channel (tanh grid, 16 x 8 staggered posts per wall)
WIDE = Params(ngalx=384, ngalz=192, ny=24, dsty=3, ygrid=1, ytanh_beta=2.0, geometry_type=3, ntilex=16,
              ntilez=8, extras24=0, flag_ctpress=0, Re=6887.0, alp=1.0, bet=2.0, posth=0.12, r=1.2378,
              c_i=12.0, c_k=12.0, stag=0.5, CFL=0.5)

# the lateral grid and post array of the 768 x 384 x 384 workload on a short
# channel (32 x 16 posts per wall)
WIDE768 = Params(ngalx=768, ngalz=384, ny=16, dsty=2, ygrid=1, ytanh_beta=2.0, geometry_type=3, ntilex=32,
                 ntilez=16, extras24=0, flag_ctpress=0, Re=10440.0, alp=1.0, bet=2.0, posth=0.1, r=1.1,
                 c_i=12.0, c_k=12.0, stag=0.5, CFL=0.5)

# the lateral grid and posts of the 192 x 192 x 96 closure grid on a short
# channel (four-step lengths z 96, x 192; posts of r = 0.62 cells)
COARSE = Params(ngalx=192, ngalz=96, ny=24, dsty=3, ygrid=1, ytanh_beta=2.0, geometry_type=3, ntilex=16,
                ntilez=8, extras24=0, flag_ctpress=0, Re=6887.0, alp=1.0, bet=2.0, posth=0.12, r=0.619,
                c_i=6.25, c_k=6.25, stag=0.5, CFL=0.5)

SMOOTH = Params(ngalx=36, ngalz=24, ny=25, dsty=0, ygrid=0, geometry_type=0, ntilex=2, ntilez=2, Re=2000.0,
                alp=1.0, bet=2.0, posth=0.0, CFL=0.5)

# smooth walls with nyu = 28 = 2 x (13 + 1): the segmented column solve
# places a separator on the top boundary row; 24-point transforms take the
# runtime-plan FFT kernels
EDGE = Params(ngalx=24, ngalz=24, ny=27, dsty=0, ygrid=0, geometry_type=0, ntilex=2, ntilez=2, Re=1500.0,
              alp=1.0, bet=2.0, posth=0.0, CFL=0.5)

FIELD_GRID = {"u1": "u", "u2": "v", "u3": "u", "p": "p", "Nu1": "u", "Nu2": "v", "Nu3": "u",
              "du1": "u", "du2": "v", "du3": "u", "psi": "p", "div": "p"}


def ny_of(p, name):
    return {"u": p.nyu, "v": p.nyv, "p": p.nyp}[FIELD_GRID[name]]


def random_field(p, name, rng, amp=1.0, decay=0.35, real_mean=False):
    """Random spectral field, amplitude decaying with |k|, Hermitian-consistent
    on the kx = 0 line (a real physical field), zero kz = -N2/2 oddball."""
    Mx, N2 = p.Mx, p.N2
    ny = ny_of(p, name)
    i = np.arange(Mx)[None, :]
    kz = np.arange(N2)
    kz = np.where(kz < N2 // 2, kz, kz - N2)[:, None]
    w = np.exp(-decay * np.sqrt(i ** 2 + kz ** 2))  # (N2, Mx)
    f = (rng.standard_normal((N2, Mx, ny)) + 1j * rng.standard_normal((N2, Mx, ny))) * w[..., None] * amp
    # kx = 0 line: f(-kz) = conj f(kz); kz = 0: real
    for k in range(1, N2 // 2):
        f[N2 - k, 0, :] = np.conj(f[k, 0, :])
    f[0, 0, :] = f[0, 0, :].real
    f[N2 // 2, :, :] = 0.0  # kz = -N2/2
    return f.reshape(N2 * Mx, ny)


def rel_l2(a, b):
    den = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (den if den > 0 else 1.0)


def max_rel(a, b):
    den = np.abs(b).max()
    return np.abs(a - b).max() / (den if den > 0 else 1.0)

# rough walls whose y bands are strictly increasing (4 x 12 posts per wall on
# a 48 x 33 x 72 tanh grid: bands -3, 14, 18, 35), so the reference's plane
# partition over np ranks exists
BANDED = Params(ngalx=48, ngalz=72, ny=33, dsty=3, ygrid=1, ytanh_beta=2.0, geometry_type=3, ntilex=4,
                ntilez=12, extras24=0, flag_ctpress=0, Re=2800.0, alp=1.0, bet=2.0, posth=0.06, r=1.0,
                c_i=6.0, c_k=3.0, stag=0.5, CFL=0.5)
