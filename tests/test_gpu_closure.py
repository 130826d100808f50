"""The learned closure on the solver side (cfd_closure_*), on a small pair of
grids (128 x 33 x 64 fine, 64 x 33 x 32 coarse, tanh grid, dsty 3):

* a sample: the input planes equal numpy's inverse transforms of the
  box-filtered fine velocity, the target planes those of
  filt(N_fine(u)) - N_coarse(filt(u)) computed from two more solvers on the
  host's filter; the RMS values and the scales they set;
* inference: with the network attached, the nonlinear term changes by the
  forward transform of the network's correction (numpy rfft/fft of the
  applied physical planes, the coarse modes; the (0,0) column of u1 and u3
  less its plane-width weighted y mean), and that correction is the
  network's own output on the same BF16 input;
* training on samples of a short fine run lowers the held-out loss of a new
  sample below half that of predicting no correction;
* backscatter clipping zeroes exactly the components with u c < 0;
* the full-DNS fallback: the fine solver takes over from the last coarse
  checkpoint, zero-padded onto its modes, exactly as a fine run from that
  state; without a failure the coarse run is that of run_until.
"""
import dataclasses

import numpy as np
import pytest

from cfd_amd import Channel, Params, closure, cnn

from .helpers import random_field

pytestmark = pytest.mark.gpu

FINE = Params(ngalx=128, ngalz=64, ny=33, dsty=3, ygrid=1, ytanh_beta=2.0, geometry_type=0, posth=0.06,
              ntilex=4, ntilez=2, Re=2800.0, alp=1.0, bet=2.0, CFL=0.5, flag_ctpress=0)
COARSE = dataclasses.replace(FINE, ngalx=64, ngalz=32)
VEL = ("u1", "u2", "u3")
NL = ("Nu1", "Nu2", "Nu3")


def _fields(p, seed):
    rng = np.random.default_rng(seed)
    f = {}
    for n in ("u1", "u2", "u3", "p"):
        a = random_field(p, n, rng, amp=0.3)
        if n == "u1":
            a[0, :] += 1.0
        if n == "u2":
            a[:, 0] = 0.0
            a[:, -1] = 0.0
        f[n] = a
    return f


def _load(ch, f):
    for n, a in f.items():
        ch.set_field(n, a)
    ch.set_scalars(t=0.0, dt=2e-3, dti=500.0, mpgx=-1e-3, QxT=2.0)


def _slots(p):
    N2 = 2 * (p.ngalz // 3)
    k = np.arange(N2)
    return np.where(k < N2 // 2, k, k - N2 + p.ngalz)


def _phys(p, f):
    N1, N2 = 2 * (p.ngalx // 3), 2 * (p.ngalz // 3)
    Mx = N1 // 2 + 1
    planes = f.T.reshape(-1, N2, Mx)
    C = np.zeros((planes.shape[0], p.ngalz, p.ngalx // 2 + 1), dtype=complex)
    C[:, _slots(p), :Mx] = planes
    return np.fft.irfft(np.fft.ifft(C, axis=1) * p.ngalz, n=p.ngalx, axis=2) * p.ngalx


def _spec(p, P):
    """physical planes -> the solver's spectral modes (forward, 1/(Ngx Ngz))"""
    N1, N2 = 2 * (p.ngalx // 3), 2 * (p.ngalz // 3)
    Mx = N1 // 2 + 1
    F = np.fft.fft(np.fft.rfft(P, axis=2) / p.ngalx, axis=1) / p.ngalz
    return F[:, _slots(p), :Mx].reshape(P.shape[0], -1).T


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _reference_sample(f):
    """(filtered velocity, target) on the coarse modes, from two solvers"""
    fine = Channel(FINE)
    _load(fine, f)
    fine.nonlinear(2)
    nf = {n: fine.get_field(n) for n in NL}
    fine.close()
    uc = closure.filtered_start(COARSE, f, FINE)
    co = Channel(COARSE)
    _load(co, uc)
    co.nonlinear(2)
    nc = {n: co.get_field(n) for n in NL}
    co.close()
    filt = closure.filtered_start(COARSE, nf, FINE)
    tgt = {n: filt[n] - nc[n] for n in NL}
    for n in NL:
        tgt[n][:, 0] = 0.0
        tgt[n][:, -1] = 0.0
    return uc, tgt


def test_snapshot_matches_host_pipeline():
    print("case closure snapshot", FINE.ngalx, COARSE.ngalx, flush=True)
    f = _fields(FINE, 11)
    uc, tgt = _reference_sample(f)
    fine, co = Channel(FINE), Channel(COARSE)
    net = cnn.CNN(co.nyu, COARSE.ngalz, COARSE.ngalx)
    try:
        _load(fine, f)
        net.dataset(2)
        rms = closure.snapshot(fine, co, net, 0)
        for k, n in enumerate(VEL):
            got, ref = co.closure_phys(k), _phys(COARSE, uc[n])
            assert _rel(got, ref) <= 1e-12, (n, _rel(got, ref))
            assert abs(rms[k] - np.sqrt(np.mean(ref ** 2))) <= 1e-12 * rms[k]
        for k, n in enumerate(NL):
            got, ref = co.closure_phys(3 + k), _phys(COARSE, tgt[n])
            assert _rel(got, ref) <= 1e-10, (n, _rel(got, ref))
            assert abs(rms[3 + k] - np.sqrt(np.mean(ref ** 2))) <= 1e-10 * rms[3 + k]
        assert np.allclose(net.scales(), rms.astype(np.float32))
    finally:
        net.close()
        fine.close()
        co.close()


def test_attached_network_corrects_the_nonlinear_term():
    print("case closure inference", COARSE.ngalx, flush=True)
    f = closure.filtered_start(COARSE, _fields(FINE, 12), FINE)
    a, b = Channel(COARSE), Channel(COARSE)
    net = cnn.CNN(a.nyu, COARSE.ngalz, COARSE.ngalx)
    try:
        w = cnn.init_params(seed=2, last_scale=1.0)
        net.set_params(w)
        s = np.array([1.0, 0.3, 0.3, 0.05, 0.05, 0.05], np.float32)
        net.set_scales(s)
        _load(a, f)
        _load(b, f)
        a.nonlinear(2)
        b.closure_attach(net)
        b.nonlinear(2)
        b.closure_attach(None)
        # the applied correction is the network's output on the same input
        u = [b.closure_phys(k) for k in range(3)]
        ny, nz, nx = b.nyu, COARSE.ngalz, COARSE.ngalx
        x = np.zeros((ny, nz, nx, 4), np.float32)
        x[..., 0] = u[0].astype(np.float32) / s[0]
        x[:ny - 1, ..., 1] = u[1].astype(np.float32) / s[1]
        x[..., 2] = u[2].astype(np.float32) / s[2]
        net.set_input(x)
        y = net.forward()
        corr = [b.closure_phys(3 + k) for k in range(3)]
        exp0 = np.zeros_like(corr[0])
        exp0[1:-1] = y[1:-1, ..., 0] * s[3]
        exp1 = np.zeros_like(corr[1])
        exp1[1:-1] = y[1:ny - 2, ..., 1] * s[4]
        assert _rel(corr[0], exp0) <= 1e-6 and _rel(corr[1], exp1) <= 1e-6
        # the nonlinear term moved by the forward transform of the correction,
        # its (0,0) u1 and u3 columns less their width-weighted y mean (no net
        # force on the flow)
        yv = b.ygrid()["yv"]
        w = yv[1:b.nyu - 1] - yv[:b.nyu - 2]
        for k, n in enumerate(NL):
            d = b.get_field(n) - a.get_field(n)
            e = _spec(COARSE, corr[k])
            if k != 1:
                e[0, 1:-1] -= (w * e[0, 1:-1].real).sum() / w.sum()
                assert abs((w * d[0, 1:-1].real).sum()) <= 1e-10 * np.abs(e[0]).max() * w.sum()
            assert _rel(d[:, 1:-1], e[:, 1:-1]) <= 1e-10, (n, _rel(d[:, 1:-1], e[:, 1:-1]))
        # backscatter clipping, same state: a component with u c < 0 is zero,
        # every other value is the unclipped correction
        _load(b, f)
        b.closure_attach(net, clip=True)
        b.nonlinear(2)
        b.closure_attach(None)
        for k in range(3):
            assert np.array_equal(b.closure_phys(k), u[k])
            cc = b.closure_phys(3 + k)
            keep = u[k] * corr[k] >= 0
            assert np.array_equal(cc[keep], corr[k][keep]) and (cc[~keep] == 0).all()
            assert (~keep).any() and keep.any()
        # attached with every = 2: the network evaluates at kRK = 1 of steps
        # 0, 2, ... and the other substeps reuse the held correction, so the
        # kRK = 2 term moves by that same correction
        _load(a, f)
        _load(b, f)
        b.closure_attach(net, every=2)
        b.nonlinear(1)
        held = [b.closure_phys(3 + k) for k in range(3)]
        _load(b, f)
        b.set_field("u1", b.get_field("u1") * 1.01)  # a different state: no new evaluation
        a.set_field("u1", a.get_field("u1") * 1.01)
        a.nonlinear(2)
        b.nonlinear(2)
        b.closure_attach(None)
        for k, n in enumerate(NL):
            assert (b.closure_phys(3 + k) == held[k]).all()
            d = b.get_field(n) - a.get_field(n)
            e = _spec(COARSE, held[k])
            if k != 1:
                e[0, 1:-1] -= (w * e[0, 1:-1].real).sum() / w.sum()
            assert _rel(d[:, 1:-1], e[:, 1:-1]) <= 1e-10, (n, _rel(d[:, 1:-1], e[:, 1:-1]))
    finally:
        net.close()
        a.close()
        b.close()


def test_attached_network_on_the_mixed_path():
    """The mixed path (FP32 nonlinear-term storage) takes the same correction:
    the network reads the FP64 fields through the FP64 transforms, so the
    correction is the FP64 path's bit for bit, and the nonlinear term moves
    by it to FP32 rounding of the stored x/z terms."""
    # the 128 x 64 grid: the mixed path needs four-step lengths in x and z
    print("case closure inference mixed", FINE.ngalx, flush=True)
    pm = dataclasses.replace(FINE, precision=1)
    f = _fields(FINE, 13)
    chans = [Channel(FINE), Channel(FINE), Channel(pm), Channel(pm)]
    net = cnn.CNN(chans[0].nyu, FINE.ngalz, FINE.ngalx)
    try:
        net.set_params(cnn.init_params(seed=3, last_scale=1.0))
        net.set_scales(np.array([1.0, 0.3, 0.3, 0.05, 0.05, 0.05], np.float32))
        for ch in chans:
            _load(ch, f)
        a64, b64, a32, b32 = chans
        for a, b in ((a64, b64), (a32, b32)):
            a.nonlinear(2)
            b.closure_attach(net)
            b.nonlinear(2)
            b.closure_attach(None)
        for k in range(3):
            assert np.array_equal(b32.closure_phys(3 + k), b64.closure_phys(3 + k))
        for n in NL:
            d64 = b64.get_field(n) - a64.get_field(n)
            d32 = b32.get_field(n) - a32.get_field(n)
            scale = np.abs(a64.get_field(n)[:, 1:-1]).max()
            assert np.abs(d32[:, 1:-1] - d64[:, 1:-1]).max() <= 1e-6 * scale, n
            assert _rel(d32[:, 1:-1], d64[:, 1:-1]) <= 1e-3, (n, _rel(d32[:, 1:-1], d64[:, 1:-1]))
    finally:
        net.close()
        for ch in chans:
            ch.close()


def test_training_beats_no_correction_on_a_new_sample():
    print("case closure training", FINE.ngalx, COARSE.ngalx, flush=True)
    fine, co = Channel(FINE), Channel(COARSE)
    net = cnn.CNN(co.nyu, COARSE.ngalz, COARSE.ngalx)
    zero = cnn.CNN(co.nyu, COARSE.ngalz, COARSE.ngalx)
    try:
        from cfd_amd import ic
        closure.start_fields(fine, FINE, seed=0, amp=0.3)
        net.set_params(cnn.init_params(seed=1))
        zero.set_params(np.zeros(cnn.nparams(), np.float32))
        net.dataset(17)
        rng = np.random.default_rng(0)
        for k in range(48):
            fine.step(1)
            closure.snapshot(fine, co, net, k % 16)
            for _ in range(4):
                net.train_slot(int(rng.integers(0, min(k + 1, 16))), 2e-3)
        fine.step(3)
        closure.snapshot(fine, co, net, 16)
        zero.set_scales(net.scales())
        trained = net.eval_slot(16)
        zero.dataset(1)
        # the no-correction loss: mean square of the normalised target
        net2 = cnn.CNN(co.nyu, COARSE.ngalz, COARSE.ngalx)
        try:
            net2.set_params(np.zeros(cnn.nparams(), np.float32))
            net2.set_scales(net.scales())
            net2.dataset(1)
            closure.snapshot(fine, co, net2, 0)
            base = net2.eval_slot(0)
        finally:
            net2.close()
        print("held-out loss %.4g, no correction %.4g" % (trained, base), flush=True)
        assert trained < 0.5 * base, (trained, base)
        del ic
    finally:
        net.close()
        zero.close()
        fine.close()
        co.close()


def test_full_dns_fallback():
    """run_with_fallback: with a drag bound the coarse run cannot meet, the
    fine solver takes over from the last checkpoint (the start), zero-padded
    onto its modes, and runs to t_end exactly as a fine solver started from
    that state does; with a loose bound the coarse run finishes alone, step
    for step the run of run_until."""
    print("case closure fallback", FINE.ngalx, COARSE.ngalx, flush=True)
    from cfd_amd import ic
    fine, co = Channel(FINE), Channel(COARSE)
    f0, q0 = closure.start_fields(fine, FINE, seed=0, amp=0.3)
    fc = closure.filtered_start(COARSE, f0, FINE)

    def load_coarse(ch):
        for k in ("u1", "u2", "u3", "p"):
            ch.set_field(k, fc[k])
        ch.set_scalars(t=0.0, dt=0.0, mpgx=0.0, QxT=ic.flowrate(fc["u1"], ch.ygrid(), COARSE.ny))

    load_coarse(co)
    hist, switch, _ = closure.run_with_fallback(co, fine, t_end=0.05, max_steps=200, drag_cap=-1.0,
                                                checkpoint_every=5)
    assert switch == 0.0 and len(hist) and (hist[:, 2] == 1).all() and hist[-1, 0] >= 0.05
    # the same fine run from the prolongated start
    ref = Channel(FINE)
    fp = closure.prolong(FINE, fc, COARSE)
    for k in ("u1", "u2", "u3", "p"):
        ref.set_field(k, fp[k])
    ref.set_scalars(t=0.0, dt=0.0, dti=0.0, mpgx=0.0, QxT=ic.flowrate(fc["u1"], co.ygrid(), COARSE.ny))
    ref.step(len(hist))
    for k in ("u1", "u2", "u3"):
        assert np.array_equal(ref.get_field(k), fine.get_field(k)), k
    # loose bound: no switch, the coarse history of run_until
    load_coarse(co)
    h2, sw2, _ = closure.run_with_fallback(co, fine, t_end=0.05, max_steps=200, drag_cap=1e9)
    load_coarse(co)
    h3, _ = closure.run_until(co, 0.05, 200, 1e9)
    assert sw2 is None and (h2[:, 2] == 0).all() and np.array_equal(h2[:, :2], h3)
    # a lower bound above every drag of that run: the switch at the start
    load_coarse(co)
    h4, sw4, _ = closure.run_with_fallback(co, fine, t_end=0.01, max_steps=200, drag_cap=1e9,
                                           drag_lo=h3[:, 1].max() + 1.0)
    assert sw4 == 0.0 and (h4[:, 2] == 1).all()
    for ch in (fine, co, ref):
        ch.close()


def test_mode_filter_scales_each_mode():
    """cfd_set_mode_filter: after one RK step the filtered run's velocity is
    sigma(column) times the unfiltered run's, sigma = exp(-alpha ((eta -
    eta_c) / (1 - eta_c))^order) above eta_c with eta = max(|i|/(N1/2),
    |kz|/(N2/2)): bit for bit where sigma = 1, to 1e-12 relative elsewhere
    (the table's exp/pow round differently from numpy's, by up to a few
    1e-15 times the exponent); alpha = 0 switches it off; bad arguments are
    refused."""
    print("case mode filter", COARSE.ngalx, flush=True)
    f = closure.filtered_start(COARSE, _fields(FINE, 21), FINE)
    a, b = Channel(COARSE), Channel(COARSE)
    for ch in (a, b):
        _load(ch, f)
    alpha, order, eta_c = 36.0, 8, 0.3
    b.set_mode_filter(alpha, order, eta_c)
    a.step(1)
    b.step(1)
    N1, N2, Mx = COARSE.N1, COARSE.N2, COARSE.Mx
    k = np.arange(N2)
    kz = np.where(k < N2 // 2, k, N2 - k)
    eta = np.maximum(np.arange(Mx)[None, :] / (N1 // 2), kz[:, None] / (N2 // 2)).reshape(-1)
    sig = np.where(eta > eta_c, np.exp(-alpha * ((eta - eta_c) / (1 - eta_c)) ** order), 1.0)
    assert sig.min() < 1e-10 and (sig == 1.0).sum() > 0
    for n in VEL:
        ua, ub = a.get_field(n), b.get_field(n)
        assert np.array_equal(ub[sig == 1.0], ua[sig == 1.0]), n
        assert np.all(np.abs(ub - ua * sig[:, None]) <= 1e-12 * np.abs(ua) * sig[:, None]), n
        assert np.abs(ub[sig < 0.5]).max() < np.abs(ua[sig < 0.5]).max(), n
    for bad in ((-1.0, 8, 0.3), (1.0, 0, 0.3), (1.0, 8, 1.0)):
        with pytest.raises(RuntimeError):
            b.set_mode_filter(*bad)
    b.set_mode_filter(0.0)
    a.step(1)
    b2 = Channel(COARSE)
    for n in VEL + ("p",):
        b2.set_field(n, b.get_field(n))
    b2.set_scalars(**b.scalars())
    b.step(1)
    b2.step(1)
    for n in VEL:
        assert np.array_equal(b.get_field(n), b2.get_field(n)), n
    for ch in (a, b, b2):
        ch.close()
