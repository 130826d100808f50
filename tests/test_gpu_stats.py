"""Statistics and spectra on the GPU (stats.f90, spectra.f90) against the
oracle, through the C ABI and the files the library writes.

* one sample from identical state: the spectra of u1, u2, u3 and p bit for
  bit (the same products and sums in the same order), spUV to 1e-14; plane
  sums and tile-conditional sums within 1e-12 of their natural scale
  (different FFTs on the two sides); headers, tables, grid records;
  accumulators zero after a write;
* on five configurations: smooth tanh grid, the 384 x 192 post array, the
  smallest smooth grid, the mixed-precision solver and the Stockham engine;
* a request served inside the next substep with kRK = 1: the sample is that
  of the start of the substep, the header carries the new t and the old
  mpgx, the name the old t; the substep itself is unchanged;
* two samples with a step between them against the oracle;
* band tables whose limits are out of order are refused;
* the 384 x 192 x 192 workload: Parseval between the stats and spectrum
  files and the wall-band identity of the conditional sums.
"""
import dataclasses
import os

import numpy as np
import pytest

from cfd_amd import Channel, fieldfile, statsfile
from cfd_amd._lib import CfdError
from oracle.oracle import Oracle

from .helpers import EDGE, SMALLPOSTS, WIDE, random_field

pytestmark = pytest.mark.gpu

SM = dataclasses.replace(SMALLPOSTS, geometry_type=0)  # smooth walls: h = 0.25, bands in order
WIDE40 = dataclasses.replace(WIDE, ny=40)
CASES = {"smooth_tanh": (SM, None), "wide40": (WIDE40, None), "edge": (EDGE, None),
         "wide40_mixed": (dataclasses.replace(WIDE40, precision=1), None), "wide40_stockham": (WIDE40, "stockham")}
NAMES = ("u1", "u2", "u3", "p")


def _channel(p, fft=None):
    if fft:
        os.environ["CFD_AMD_FFT"] = fft
    try:
        return Channel(p)
    finally:
        os.environ.pop("CFD_AMD_FFT", None)


def _pair(p, seed, fft=None, oracle=True):
    print("case stats", p.ngalx, p.ny, p.ngalz, "seed", seed, fft, flush=True)
    ch = _channel(p, fft)
    orc = Oracle(dataclasses.replace(p, precision=0)) if oracle else None
    rng = np.random.default_rng(seed)
    for n in NAMES:
        f = random_field(p, n, rng, amp=1.0 if n != "p" else 0.3)
        if n == "u1":
            f[0, :] += 1.0
        if n == "u2":
            f[:, 0] = 0.0  # no flow through the walls
            f[:, -1] = 0.0
        ch.set_field(n, f)
        if orc:
            orc.set_field(n, f)
    sc = dict(t=0.5, dt=2e-3, dti=500.0, mpgx=-1e-3, QxT=(orc.flowrate() if orc else 2.0) + 0.01)
    ch.set_scalars(**sc)
    if orc:
        orc.set_scalars(**sc)
        orc.stats_init(fieldfile.band_table(p))
    return ch, orc


def _fold(raw, p):
    """recvwrspec (spectra.f90:142-170) restated: (Ncol, ny) per-column sums
    -> buffSP (ny, N2/2+1, Mx): kz >= 0 at row kz, kz < 0 added at row -kz,
    kz = -N2/2 at row N2/2; kx > 0 doubled."""
    ny = raw.shape[1]
    r = raw.T.reshape(ny, p.N2, p.Mx)
    out = np.zeros((ny, p.N2 // 2 + 1, p.Mx))
    out[:, :p.N2 // 2] = r[:, :p.N2 // 2]
    for k in range(p.N2 // 2, p.N2):
        out[:, p.N2 - k] += r[:, k]
    out[:, :, 1:] *= 2.0
    return out


def _read_all(d, ch, t_name):
    tag = "%04dx%04dx%04d_%05d.dat" % (ch.N1, ch.N2, ch.params.ny, int(10.0 * t_name))
    dims = ch.stats_dims()
    h, prof, cond = statsfile.read_stats(os.path.join(d, "stats_" + tag), dims["dnx"], dims["dnz"])
    spec = {}
    for s in ("spU", "spV", "spW", "spUV", "spP"):
        hs, spec[s] = statsfile.read_spectrum(os.path.join(d, s + "_" + tag))
        assert hs["istat"] == h["istat"] and hs["iter"] == h["iter"] and hs["t"] == h["t"]
    return h, prof, cond, spec


SCALE = {"Um": "U2m", "Vm": "V2m", "Wm": "W2m", "Pm": "P2m", "wxm": "wx2m"}
PRODUCT = {"UVm": ("U2m", "V2m"), "UWm": ("U2m", "W2m"), "VWm": ("V2m", "W2m")}


def _scale(d, k, count):
    """Natural size of the rounding of sum k: sqrt(count * sum x^2) for a
    first moment (Cauchy-Schwarz), sqrt(sum x^2 sum y^2) for a product, the
    value itself for a second moment.  d maps names to arrays; count is the
    number of terms of one sum."""
    if k in SCALE:
        return np.sqrt(count * np.abs(d[SCALE[k]]).max())
    if k in PRODUCT:
        a, b = PRODUCT[k]
        return np.sqrt(np.abs(d[a]).max() * np.abs(d[b]).max())
    return np.abs(d[k]).max()


def _compare(p, prof, cond, spec, ref, npts):
    for s in ("spU", "spV", "spW", "spP"):
        assert (spec[s] == _fold(ref["spec"][s], p)).all(), s
    e = _fold(ref["spec"]["spUV"], p)
    assert np.abs(spec["spUV"] - e).max() <= 1e-14 * np.abs(e).max()
    R = ref["prof"]
    for k, v in prof.items():
        sc = _scale(R, k, npts)
        assert np.abs(v - R[k]).max() <= 1e-12 * sc, (k, np.abs(v - R[k]).max() / sc)
    # conditional sums: bottom + top, as write_stats adds them
    C = {k[:-1]: cb + ct for k, (cb, ct) in ref["cond"].items()}  # "UmC" -> "Um"
    for k, v in cond.items():
        e = C[k[:-1]]
        assert v.shape == e.shape, k
        sc = _scale(C, k[:-1], 2 * p.ntilex * p.ntilez)
        assert np.abs(v - e).max() <= 1e-12 * sc, (k, np.abs(v - e).max() / sc)


@pytest.mark.parametrize("case", list(CASES))
def test_one_sample_against_oracle(tmp_path, case):
    p, fft = CASES[case]
    ch, orc = _pair(p, 300, fft)
    try:
        ch.stats()
        orc.stats()
        d = str(tmp_path) + "/"
        ch.write_stats(d, iter=7)
        h, prof, cond, spec = _read_all(d, ch, 0.5)
        assert (h["t"], h["mpgx"], h["iter"], h["istat"], h["nband"]) == (0.5, -1e-3, 7, 1, 3)
        assert (h["Re"], h["alp"], h["bet"]) == (p.Re, p.alp, p.bet)
        N = fieldfile.band_table(p)
        ngal = list(N)
        for b in (1, 2, 3):
            ngal[4 * b:4 * b + 2] = [p.ngalx, p.ngalz]
        assert h["table"] == ngal
        g = ch.ygrid()
        assert (h["yu"] == g["yu"]).all() and (h["yv"] == g["yv"]).all()
        assert (h["dthdyu"] == g["dthdyu"]).all() and (h["dthdyv"] == g["dthdyv"]).all()
        assert h["dthetai_u"] == 1.0 and h["dthetai_v"] == 1.0
        hs, _ = statsfile.read_spectrum(os.path.join(d, "spU_%04dx%04dx%04d_00005.dat" % (ch.N1, ch.N2, p.ny)))
        assert hs["table"] == N
        _compare(p, prof, cond, spec, orc.stats_get(), p.ngalx * p.ngalz)
        # the write zeroed the accumulators
        ch.write_stats(d, iter=8)
        h2, prof2, cond2, spec2 = _read_all(d, ch, 0.5)
        assert h2["istat"] == 0 and all((v == 0).all() for v in prof2.values())
        assert all((v == 0).all() for v in spec2.values()) and all((v == 0).all() for v in cond2.values())
    finally:
        ch.close()
        orc.close()


def test_request_served_inside_the_substep(tmp_path):
    p = SM  # flag_ctpress = 0: the substep moves mpgx
    a, _ = _pair(p, 310, oracle=False)
    b, _ = _pair(p, 310, oracle=False)
    try:
        s0 = a.scalars()
        da, db = str(tmp_path / "a") + "/", str(tmp_path / "b") + "/"
        os.makedirs(da)
        os.makedirs(db)
        a.stats_request(sample=True, write_dir=da, iter=11)
        a.substep(1)
        s1 = a.scalars()
        assert s1["t"] > s0["t"] and s1["mpgx"] != s0["mpgx"]
        ha, pa, ca, sa = _read_all(da, a, s0["t"])
        assert ha["t"] == s1["t"] and ha["mpgx"] == s0["mpgx"] and ha["iter"] == 11 and ha["istat"] == 1
        # the same sample taken directly, then the same substep
        b.stats()
        b.write_stats(db, iter=11)
        hb, pb, cb, sb = _read_all(db, b, s0["t"])
        for k in pa:
            assert (pa[k] == pb[k]).all(), k
        for k in ca:
            assert (ca[k] == cb[k]).all(), k
        for k in sa:
            assert (sa[k] == sb[k]).all(), k
        b.substep(1)
        for n in ("u1", "u2", "u3", "p"):
            assert (a.get_field(n) == b.get_field(n)).all(), n
        # served once: the next kRK = 1 substep writes nothing new
        before = sorted(os.listdir(da))
        a.substep(2)
        a.substep(3)
        a.substep(1)
        assert sorted(os.listdir(da)) == before
    finally:
        a.close()
        b.close()


def test_two_samples_with_a_step_between(tmp_path):
    p = SM
    ch, orc = _pair(p, 320)
    try:
        ch.stats()
        orc.stats()
        ch.step(1)
        orc.step(1)
        ch.stats()
        orc.stats()
        d = str(tmp_path) + "/"
        t = ch.scalars()["t"]
        ch.write_stats(d, iter=3)
        h, prof, cond, spec = _read_all(d, ch, t)
        assert h["istat"] == 2
        ref = orc.stats_get()
        # one RK3 step apart, the two sides' fields differ at the 1e-12 level
        for s in ("spU", "spV", "spW", "spUV", "spP"):
            e = _fold(ref["spec"][s], p)
            assert np.abs(spec[s] - e).max() <= 1e-10 * np.abs(e).max(), s
        for k in ("Um", "U2m", "Vm", "V2m", "Wm", "W2m", "Pm", "P2m", "UVm", "wxm", "wx2m"):
            R = ref["prof"][k]
            sc = np.sqrt(p.ngalx * p.ngalz * ref["prof"][SCALE[k]]).max() if k in SCALE else np.abs(R).max()
            assert np.abs(prof[k] - R).max() <= 1e-10 * sc, k
    finally:
        ch.close()
        orc.close()


def test_unordered_band_table_is_refused(tmp_path):
    ch = Channel(SMALLPOSTS)  # tall posts on a short grid: N(4,1) > N(4,2)
    try:
        with pytest.raises(CfdError):
            ch.stats()
        with pytest.raises(CfdError):
            ch.stats_request(sample=True)
        with pytest.raises(CfdError):
            ch.write_stats(str(tmp_path) + "/")
    finally:
        ch.close()


def test_workload_parseval_and_wall_bands(tmp_path):
    """Full 384 x 192 x 192 workload: the stats file's second moments equal
    Ngx Ngz times the spectrum files' sums; the conditional sums of each wall
    distance add up to the plane sums of the two wall planes."""
    from cfd_amd import ic, workloads
    p, _ = workloads.get("rt395_384x192x192")
    print("case stats rt395_384x192x192", flush=True)
    ch = Channel(p)
    try:
        g = ch.ygrid()
        f = ic.project(ch, ic.initial_fields(p, g, seed=5, amp=0.1))
        for n in NAMES:
            ch.set_field(n, f[n])
        ch.set_scalars(t=2.0, dt=0.0, mpgx=0.0, QxT=ic.flowrate(f["u1"], g, p.ny))
        ch.step(1)
        ch.stats()
        ch.step(1)
        ch.stats()
        t = ch.scalars()["t"]
        d = str(tmp_path) + "/"
        ch.write_stats(d, iter=2)
        h, prof, cond, spec = _read_all(d, ch, t)
        assert h["istat"] == 2
        n = p.ngalx * p.ngalz
        for m2, s in (("U2m", "spU"), ("V2m", "spV"), ("W2m", "spW"), ("P2m", "spP"), ("UVm", "spUV")):
            lhs, rhs = prof[m2], n * spec[s].sum(axis=(1, 2))
            assert np.abs(lhs - rhs).max() <= 1e-12 * np.abs(prof["U2m"]).max(), m2
        N = fieldfile.band_table(p)
        nb = lambda r, b: N[4 * b + r - 1]  # noqa: E731
        # wall distance j1: bottom plane j = N(r,0) + j1 (profile index j1),
        # top plane j = N(r,3) + 1 - j1; v's first moment flips sign on top
        for k, sign, r in (("Um", 1.0, 4), ("Vm", -1.0, 3), ("Wm", 1.0, 4), ("U2m", 1.0, 4), ("V2m", 1.0, 3)):
            c = cond[k + "C"].sum(axis=(1, 2))
            top = nb(r, 3) + 1 - nb(r, 0)
            e = np.array([prof[k][j1] + sign * prof[k][top - j1] for j1 in range(c.shape[0])])
            assert np.abs(c - e).max() <= 1e-11 * _scale(prof, k, n), k
    finally:
        ch.close()
