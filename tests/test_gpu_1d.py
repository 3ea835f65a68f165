"""GPU parity of the 1-D HS2 path (rhs_HsF2 + breaking + rkf45 driver)
against the FP64 oracle.

Tolerances: FP64 device path relL2 <= 1e-11 on a right-hand side, B exact
to 1e-12, trajectories <= 1e-9 with identical step counts (the device takes
GSL's error norm over the same full state as the reference); FP32
transforms with FP64 state: right-hand side <= 2e-5, trajectory <= 1e-4.
"""
import numpy as np
import pytest
from conftest import relerr

from hawassi_c_amd import HW_PREC_FP32, HW_PREC_FP64
from hawassi_c_amd import workloads1d as W1
from hawassi_c_amd.model1d import Hawassi1D
from oracle import oracle as O
from test_oracle1d import crest_state

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def model(kind, n, **kw):
    return getattr(W1, kind)(n, t_end=4.0, **kw)


@pytest.mark.parametrize("kind,n,prec", [
    ("channel", 256, HW_PREC_FP64), ("channel", 2048, HW_PREC_FP64), ("slope", 256, HW_PREC_FP64),
    ("slope", 1024, HW_PREC_FP64), ("slope", 8192, HW_PREC_FP64),
    ("channel", 1024, HW_PREC_FP32), ("slope", 1024, HW_PREC_FP32), ("slope", 16384, HW_PREC_FP32),
])
def test_rhs1d_parity(kind, n, prec):
    say("rhs1d", kind, n, prec)
    m = model(kind, n)
    y = W1.random_state(m, 3, amp=0.05, kfrac=0.3)
    o = O.Oracle1D(m)
    d = Hawassi1D(m, prec)
    tol = 1e-11 if prec == HW_PREC_FP64 else 2e-5
    for t in (0.37, 1.3):
        ref = o.rhs(t, y)
        got = d.rhs(t, y)
        e = relerr(got, ref)
        say("  t", t, "relL2", e)
        assert e < tol


@pytest.mark.parametrize("prec", [HW_PREC_FP64, HW_PREC_FP32])
def test_breaking1d_sequence(prec):
    say("breaking1d", prec)
    m = model("slope", 1024)
    y, _ = crest_state(m, 150.0, 0.3, 3.0)
    o = O.Oracle1D(m)
    d = Hawassi1D(m, prec)
    for t in (1.0, 1.5, 1.8, 2.4):
        ref = o.rhs(t, y)
        got = d.rhs(t, y)
        Bd, fd = d.breaking_state()
        Bo = o.B()
        say("  t", t, "flag", o.flag(), fd, "B>0", int((Bo > 0).sum()), int((Bd > 0).sum()), "rel", relerr(got, ref))
        assert fd == o.flag()
        if prec == HW_PREC_FP64:
            assert np.abs(Bd - Bo).max() < 1e-12
            assert relerr(got, ref) < 1e-11
        else:
            assert np.abs(Bd - Bo).max() < 1e-4
            assert relerr(got, ref) < 2e-5
    assert int((Bo > 0).sum()) > 0


def test_batch_uneven_sizes():
    say("batch uneven")
    ms = [model("slope", 256), model("slope", 512), model("channel", 512), model("slope", 1024)]
    d = Hawassi1D(ms, HW_PREC_FP64)
    for r, m in enumerate(ms):
        y = W1.random_state(m, 10 + r, amp=0.05, kfrac=0.3)
        ref = O.Oracle1D(m).rhs(0.9, y)
        got = d.rhs(0.9, y, r)
        say("  r", r, m.n, relerr(got, ref))
        assert relerr(got, ref) < 1e-11


def test_batch_of_any_sizes_in_size_classes():
    """Sizes 64x apart in one batch, interleaved and with a three-depth
    realization among them: each size class (n, depth count) has its own
    kernel and stream.  Every realization's right-hand side equals the
    oracle's, and the batch integrated together ends bit for bit where each
    realization integrated alone ends (same kernels, same steps)."""
    say("batch size classes")
    ms = [model("slope", 256, ndepth=2), model("slope", 16384, ndepth=2), model("channel", 512),
          model("slope", 16384, ndepth=2), model("slope", 1024, ndepth=3), model("slope", 256, ndepth=2)]
    d = Hawassi1D(ms, HW_PREC_FP32)
    for r, m in enumerate(ms):
        y = W1.random_state(m, 20 + r, amp=0.05, kfrac=0.3)
        e = relerr(d.rhs(0.6, y, r), O.Oracle1D(m).rhs(0.6, y))
        say("  r", r, m.n, "relL2", e)
        assert e < 2e-5
    d = Hawassi1D(ms, HW_PREC_FP32, t0=0.0, h0=1e-3)
    for r, m in enumerate(ms):
        d.set_state(W1.random_state(m, 30 + r, amp=0.02, kfrac=0.2), r)
    d.integrate(0.05)
    for r, m in enumerate(ms):
        a = Hawassi1D([m], HW_PREC_FP32, t0=0.0, h0=1e-3)
        a.set_state(W1.random_state(m, 30 + r, amp=0.02, kfrac=0.2))
        a.integrate(0.05)
        same = np.array_equal(a.get_state(0), d.get_state(r))
        sa, sb = a.stats(0), d.stats(r)
        say("  r", r, "alone == batch", same, sa, sb)
        # (rhs_evals may differ: a batch with breaking anywhere runs rkf45's
        # dydt_out evaluation for every realization, without side effects
        # where breaking is off)
        assert same and (sa["steps_accepted"], sa["steps_rejected"]) == (sb["steps_accepted"], sb["steps_rejected"])


def _traj(ms, prec, t_out):
    d = Hawassi1D(ms, prec, t0=0.0, h0=1e-4)
    os_, ys, ts, hs = [], [], [], []
    for m in ms:
        os_.append(O.Oracle1D(m))
        ys.append(np.zeros(4 * m.n))
        ts.append(0.0)
        hs.append(1e-4)
    for t1 in t_out:
        d.integrate(t1)
        for r, m in enumerate(ms):
            rc, ts[r], hs[r], _ = os_[r].integrate(ts[r], t1, hs[r], ys[r])
            assert rc == 0
    return d, os_, ys, hs


def test_trajectory1d_fp64_matches_step_for_step():
    say("trajectory1d fp64")
    ms = [model("channel", 512), model("slope", 256)]
    d, os_, ys, hs = _traj(ms, HW_PREC_FP64, np.arange(0.1, 1.01, 0.1))
    for r, m in enumerate(ms):
        got = d.get_state(r)
        t, h = d.time(r)
        st = d.stats(r)
        e = relerr(got, ys[r])
        say("  r", r, "relL2", e, "steps", st, "oracle rhs", os_[r].rhs_count(), "h", h, hs[r])
        assert abs(t - 1.0) < 1e-12
        assert e < 1e-9
        assert st["rhs_evals"] == os_[r].rhs_count()
        assert abs(h - hs[r]) < 1e-9 * hs[r]


def test_trajectory1d_breaking_fp64():
    say("trajectory1d breaking")
    m = model("slope", 1024)
    y0, _ = crest_state(m, 150.0, 0.3, 3.0)
    o = O.Oracle1D(m)
    d = Hawassi1D(m, HW_PREC_FP64, t0=1.0, h0=1e-3)
    d.set_state(y0)
    y = y0.copy()
    t, h = 1.0, 1e-3
    for t1 in (1.02, 1.04, 1.06):
        d.integrate(t1)
        rc, t, h, _ = o.integrate(t, t1, h, y)
        assert rc == 0
    Bd, fd = d.breaking_state()
    e = relerr(d.get_state(), y)
    say("  relL2", e, "flag", o.flag(), fd, "B>0", int((o.B() > 0).sum()), int((Bd > 0).sum()))
    assert fd == o.flag() and np.abs(Bd - o.B()).max() < 1e-10
    assert e < 1e-9
    assert d.stats(0)["rhs_evals"] == o.rhs_count()


def test_trajectory1d_fp32():
    say("trajectory1d fp32")
    m = model("channel", 4096)
    d, os_, ys, _ = _traj([m], HW_PREC_FP32, np.arange(0.25, 1.01, 0.25))
    got = d.get_state(0)
    n = m.n
    eta_d = np.real(np.fft.ifft(got[:n] + 1j * got[n:2 * n]))
    eta_o = np.real(np.fft.ifft(ys[0][:n] + 1j * ys[0][n:2 * n]))
    e = relerr(eta_d, eta_o)
    say("  eta relL2", e)
    assert e < 1e-4


def test_round_limit_returns_step_error_and_resumes():
    """hw1d_set_round_limit bounds one integrate call's rounds (the first
    eager, the later ones replayed from the captured round graph, four
    between two reads of the active count): the call returns HW_ERR_STEP
    with the realizations where they got, and a later call without the limit
    takes them on to t1 and ends where an unlimited run ends."""
    say("round limit")
    ms = [model("slope", 512, ndepth=2), model("channel", 256)]
    a = Hawassi1D(ms, HW_PREC_FP64, t0=0.0, h0=1e-4)
    a.integrate(0.3)
    b = Hawassi1D(ms, HW_PREC_FP64, t0=0.0, h0=1e-4)
    b.set_round_limit(6)
    with pytest.raises(Exception, match="rounds"):
        b.integrate(0.3)
    t_mid = [b.time(r)[0] for r in range(2)]
    say("  after 6 rounds t =", t_mid)
    assert all(0.0 < t < 0.3 for t in t_mid)
    b.set_round_limit(0)
    b.integrate(0.3)
    for r in range(2):
        e = relerr(b.get_state(r), a.get_state(r))
        say("  r", r, "t", b.time(r)[0], "relL2 vs unlimited", e)
        assert b.time(r)[0] == 0.3 and e < 1e-10
