"""Flow diagnostics on spectral fields in the reference column layout
(shape (Ncol, ny_g), column 0 = the kx = kz = 0 mode).

mean_profile  U(y) of the u1 centres (real part of mode (0,0)).
wall_shear    (1/Re) dU/dy at the two walls, from the first interior centre
              and its ghost: the ghosts are the linear extrapolation that
              vanishes at the wall face (gridweighting, start.f90:494-508), so
              the one-sided difference is the wall gradient of that profile.
compare       error of one state against a reference state: relative L2 per
              velocity component, mean profile, wall shear and the mean
              pressure gradient (the constant-flow-rate forcing, i.e. the
              total drag per unit volume).
"""
import numpy as np


def mean_profile(u1):
    return np.asarray(u1)[0, :].real.copy()


def wall_shear(u1, ygrid, Re):
    U = mean_profile(u1)
    yu = ygrid["yu"]
    bottom = (U[1] - U[0]) / (yu[1] - yu[0]) / Re
    top = -(U[-1] - U[-2]) / (yu[-1] - yu[-2]) / Re
    return bottom, top


def _rel(a, b):
    den = np.linalg.norm(b)
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / (den if den > 0 else 1.0))


def compare(test, ref, ygrid, Re):
    """test, ref: dicts with u1, u2, u3 (spectral) and mpgx."""
    out = {"rel_l2_" + k: _rel(test[k], ref[k]) for k in ("u1", "u2", "u3")}
    Ut, Ur = mean_profile(test["u1"]), mean_profile(ref["u1"])
    out["mean_profile_max_rel"] = float(np.abs(Ut - Ur).max() / np.abs(Ur).max())
    tb, tt = wall_shear(test["u1"], ygrid, Re)
    rb, rt = wall_shear(ref["u1"], ygrid, Re)
    out["wall_shear_rel"] = float(max(abs(tb - rb) / abs(rb), abs(tt - rt) / abs(rt)))
    out["mpgx_rel"] = float(abs(test["mpgx"] - ref["mpgx"]) / abs(ref["mpgx"]))
    return out
