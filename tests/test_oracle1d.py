"""Pins the 1-D oracle (oracle/hawassi_oracle1d.c) and the 1-D setup.

* FFT against numpy and a brute-force DFT;
* flat HS2 linear dynamics: a small single Fourier mode oscillates at the
  exact dispersion frequency of its (freqspace) wavenumber;
* two-depth operator: with (gamma-, gamma+) = (1, 0) or (0, 1) and a
  multiplier symmetric under k -> -k it equals the flat operator on a
  Hermitian state; the identity fails if HSS drops a term;
* Up2IP: gamma = (1, 0) at D-, (0, 1) at D+, and the fit reproduces C^2 at
  the table depths (Cramer construction);
* breaking: a hand-built crest moving to +x is detected and then tracked
  with B following the reference's ramp; one moving to -x never marks
  nodes (the reference's scan loop runs only while ly < indEND).
"""
import math

import numpy as np
import pytest
from conftest import relerr

from hawassi_c_amd import setup1d as S
from hawassi_c_amd import workloads1d as W1
from oracle import oracle as O


def test_fft1d_matches_numpy_and_brute_force():
    g = np.random.default_rng(0)
    for n in (8, 64, 512):
        a = g.standard_normal(n) + 1j * g.standard_normal(n)
        assert relerr(O.fft1d(a, -1), np.fft.fft(a)) < 1e-14
        assert relerr(O.fft1d(a, +1), np.fft.ifft(a) * n) < 1e-14
    for n in (16, 15, 30, 50):
        a = g.standard_normal(n) + 1j * g.standard_normal(n)
        W = np.exp(-2j * np.pi * np.outer(np.arange(n), np.arange(n)) / n)
        assert relerr(O.fft1d(a, -1), W @ a) < 1e-14
    # lengths of 2, 3 and 5 factors (1dBBr001HarmBar has 2000 nodes)
    for n in (2000, 300, 4000, 3 * 5 * 5):
        a = g.standard_normal(n) + 1j * g.standard_normal(n)
        assert relerr(O.fft1d(a, -1), np.fft.fft(a)) < 2e-15
        assert relerr(O.fft1d(a, +1), np.fft.ifft(a) * n) < 2e-15


def test_freqspace1d_and_alias():
    k = S.freqspace1d(200.0, 8)
    d = k[1] - k[0]
    assert np.allclose(np.diff(k[:4]), d) and k[0] < 0 and abs(k[0]) < 1e-2
    # the spacing is (fmax - fmin)/(n-1), slightly below 2 pi / L
    assert d < 2 * math.pi / 200.0
    a = S.alias1d(S.freqspace1d(100.0, 64), 4)
    assert a[15] == 0.5 and set(np.unique(a)) <= {0.0, 0.5, 1.0}


def _bare(m):
    """the model with source, damping and the mask switched off"""
    m.source_on = False
    m.damp = np.zeros(m.n)
    m.aal = np.ones(m.n)
    m.chi = np.ones(m.n)
    return m


def test_single_mode_flat_dispersion():
    m = _bare(W1.channel(256, t_end=2.0))
    n = m.n
    j = 9
    amp = 1e-7
    eta_hat = np.zeros(n, complex)
    eta_hat[j] = eta_hat[n - j] = amp * n / 2
    y = np.concatenate([eta_hat.real, eta_hat.imag, np.zeros(n), np.zeros(n)])
    o = O.Oracle1D(m)
    t1 = 1.7
    rc, t, h, _ = o.integrate(0.0, t1, 1e-3, y, 1e-14, 1e-10)
    assert rc == 0 and t == t1
    for b in (j, n - j):
        w = math.sqrt(S.GRAV * abs(m.k[b]) * math.tanh(abs(m.k[b]) * 1.0))
        got = y[b] + 1j * y[n + b]
        assert abs(got - eta_hat[b] * math.cos(w * t1)) < 1e-7 * abs(eta_hat[b])


def _sym_multiplier(n, length, d):
    js = np.where(np.arange(n) < n // 2, np.arange(n), np.arange(n) - n)
    ks = 2 * math.pi * js / length
    return S.phase_velocity_v(ks, d) ** 2


@pytest.mark.parametrize("which", ["min", "plus"])
def test_two_depth_reduces_to_flat(which):
    mf = _bare(W1.channel(256, t_end=2.0))
    n = mf.n
    sym = _sym_multiplier(n, 200.0, 1.0)
    mf.csqr, mf.lop = sym, mf.k ** 2 * sym / S.GRAV
    one, zero = np.ones(n), np.zeros(n)
    ops = dict(gam_min=one if which == "min" else zero, gam_plus=zero if which == "min" else one,
               csq_min=sym if which == "min" else 0.5 * sym, csq_plus=0.5 * sym if which == "min" else sym)
    mb = S.Model1D(**{**mf.__dict__, "flat": False, "ops": ops})
    y = W1.random_state(mf, 5, amp=0.05, kfrac=0.3)
    df = O.Oracle1D(mf).rhs(0.3, y)
    db = O.Oracle1D(mb).rhs(0.3, y)
    assert relerr(db, df) < 1e-12
    # a dropped HSS term breaks the identity
    bad = dict(ops)
    bad["gam_min"], bad["gam_plus"] = 0.5 * bad["gam_min"], 0.5 * bad["gam_plus"]
    dbad = O.Oracle1D(S.Model1D(**{**mb.__dict__, "ops": bad})).rhs(0.3, y)
    assert relerr(dbad, df) > 1e-3


def test_up2ip_closed_forms():
    m = W1.slope(512, t_end=2.0, breaking=False, ndepth=2)
    o = m.ops
    d = m.depth
    nu = 2 * math.pi / 2.0
    at_min, at_max = np.argmin(d), np.argmax(d)
    assert abs(o["gam_min"][at_min] - 1) < 1e-9 and abs(o["gam_plus"][at_min]) < 1e-9
    assert abs(o["gam_min"][at_max]) < 1e-9 and abs(o["gam_plus"][at_max] - 1) < 1e-9
    # Cramer fit: gamma- C^2(D-) + gamma+ C^2(D+) = C^2(D) at the peak wavenumber of D
    for D in (0.35, 0.6, 0.85):
        kap = S.invers_omega(nu, D)
        dd = (o["dplus"] - o["dmin"]) / 100
        if abs((D - o["dmin"]) / dd - round((D - o["dmin"]) / dd)) > 1e-9:
            continue
        gm = np.interp(D, d[np.argsort(d)], o["gam_min"][np.argsort(d)])
        gp = np.interp(D, d[np.argsort(d)], o["gam_plus"][np.argsort(d)])
        lhs = gm * S.phase_velocity_i(kap, o["dmin"]) ** 2 + gp * S.phase_velocity_i(kap, o["dplus"]) ** 2
        assert abs(lhs - S.phase_velocity_i(kap, D) ** 2) < 1e-3


def crest_state(m, x0, a, U, w=1.5):
    """eta: a Gaussian crest of height a at x0; u: U times the same shape
    (peak exactly at the crest node)."""
    x = m.x
    i0 = int(np.argmin(np.abs(x - x0)))
    x0 = x[i0]
    eta = a * np.exp(-((x - x0) / w) ** 2)
    u = U * np.exp(-((x - x0) / (2 * w)) ** 2)
    return np.concatenate([np.fft.fft(eta).real, np.fft.fft(eta).imag, np.fft.fft(u).real, np.fft.fft(u).imag]), i0


def test_breaking_detect_and_track_right_moving_crest():
    m = W1.slope(1024, t_end=4.0)
    y, i0 = crest_state(m, 150.0, 0.3, 3.0)
    o = O.Oracle1D(m)
    o.rhs(1.0, y)
    assert o.flag() == 2                        # detected, kept for tracking
    assert np.all(o.B() == 0)                   # U* = U_I at detection: no node above it
    o.rhs(1.5, y)                               # half-way through T* = 1 s: U* = 0.6 U_I
    B = o.B()
    nz = np.nonzero(B)[0]
    assert len(nz) > 0 and nz.min() >= i0 - 1 and np.all(B <= 1) and np.all(B >= 0)
    # nodes right of the crest while |u| >= 0.2 U_I, B = |u|/U* - 1 clipped
    eta = np.real(np.fft.ifft(y[:m.n] + 1j * y[m.n:2 * m.n]))
    u = np.real(np.fft.ifft(y[2 * m.n:3 * m.n] + 1j * y[3 * m.n:]))
    ui = abs(u[int(np.argmax(eta))])
    us = ui + 0.5 * (0.2 * ui - ui)
    j = int(np.argmax(eta))
    exp = np.zeros(m.n)
    while j < m.breaking["xbend"] and abs(u[j]) >= 0.2 * ui:
        exp[j] = min(1.0, max(0.0, abs(u[j]) / us - 1))
        j += 1
    assert np.abs(B - exp).max() < 1e-12


def test_breaking_left_moving_crest_never_marks_nodes():
    m = W1.slope(1024, t_end=4.0)
    y, _ = crest_state(m, 150.0, 0.3, -3.0)
    o = O.Oracle1D(m)
    o.rhs(1.0, y)
    assert o.flag() == 0 and np.all(o.B() == 0)
    o.rhs(1.5, y)
    assert o.flag() == 0 and np.all(o.B() == 0)


def test_integrate_reaches_target_and_counts():
    m = W1.channel(256, t_end=2.0)
    o = O.Oracle1D(m)
    y = np.zeros(4 * m.n)
    rc, t, h, st = o.integrate(0.0, 0.5, 1e-4, y)
    assert rc == 0 and t == 0.5 and st[0] > 0
    # 5 stage evaluations + dydt_out per attempt, dydt_in once per evolve_apply
    assert o.rhs_count() == 6 * (st[0] + st[1]) + st[0]


# ------------------------------------------------------- three depths ---
def test_three_depth_middle_weight_reduces_to_flat_and_to_two_depths():
    """Up3IP / HSS with a middle depth: gamma = (0, 1, 0) with C2_mid the
    flat multiplier is the flat operator; a zero middle weight is the
    two-depth operator bit for bit."""
    mf = _bare(W1.channel(256, t_end=2.0))
    n = mf.n
    sym = _sym_multiplier(n, 200.0, 1.0)
    mf.csqr, mf.lop = sym, mf.k ** 2 * sym / S.GRAV
    one, zero = np.ones(n), np.zeros(n)
    ops = dict(gam_min=zero, gam_plus=zero, gam_mid=one, csq_min=0.3 * sym, csq_plus=2.0 * sym, csq_mid=sym)
    mb = S.Model1D(**{**mf.__dict__, "flat": False, "ops": ops})
    y = W1.random_state(mf, 5, amp=0.05, kfrac=0.3)
    df = O.Oracle1D(mf).rhs(0.3, y)
    db = O.Oracle1D(mb).rhs(0.3, y)
    assert relerr(db, df) < 1e-12
    m3 = W1.slope(512, t_end=2.0, breaking=False)
    o3 = S.up3ip(m3.depth, m3.k, 2 * math.pi / 2.0)
    m3.ops = dict(o3)
    y = W1.random_state(m3, 6, amp=0.05, kfrac=0.3)
    r3 = O.Oracle1D(m3).rhs(0.3, y)
    m3.ops["gam_mid"] = np.zeros(m3.n)
    r3z = O.Oracle1D(m3).rhs(0.3, y)
    m3.ops = {k: v for k, v in o3.items() if k not in ("gam_mid", "csq_mid")}
    r2 = O.Oracle1D(m3).rhs(0.3, y)
    assert np.array_equal(r3z, r2)
    assert relerr(r3, r2) > 1e-6


def test_up3ip_solves_the_interpolation_system():
    m = W1.slope(512, t_end=2.0, breaking=False)
    nu = 2 * math.pi / 2.0
    o = S.up3ip(m.depth, m.k, nu)
    Ds = (o["dmin"], o["dmid"], o["dplus"])
    for i in (0, 33, 100):
        D = o["table_depth"][i]
        kap = S.invers_omega(nu, D)
        M = np.array([[math.sqrt(S.GRAV * d) for d in Ds], [S.omega_i(kap, d) for d in Ds],
                      [S.phase_velocity_i(kap, d) ** 2 for d in Ds]])
        rhs = np.array([math.sqrt(S.GRAV * D), S.omega_i(kap, D), S.phase_velocity_i(kap, D) ** 2])
        assert np.allclose(o["table_gam"][:, i], np.linalg.solve(M, rhs), rtol=1e-9, atol=1e-9)
    assert np.allclose(o["table_gam"][:, 0], [1, 0, 0], atol=1e-12)
    assert np.allclose(o["table_gam"][:, 100], [0, 0, 1], atol=1e-12)


# ------------------------------------------------------- HS1 and HS3 ---
def test_hs1_single_mode_any_amplitude():
    """rhs_HsF1 is linear: a single mode of any amplitude keeps the exact
    dispersion frequency."""
    m = _bare(W1.channel(256, t_end=2.0))
    m.order = 1
    n = m.n
    j = 9
    eta_hat = np.zeros(n, complex)
    eta_hat[j] = eta_hat[n - j] = 0.3 * n / 2
    y = np.concatenate([eta_hat.real, eta_hat.imag, np.zeros(n), np.zeros(n)])
    rc, t, h, _ = O.Oracle1D(m).integrate(0.0, 1.7, 1e-3, y, 1e-14, 1e-10)
    assert rc == 0
    for b in (j, n - j):
        w = math.sqrt(S.GRAV * abs(m.k[b]) * math.tanh(abs(m.k[b]) * 1.0))
        assert abs(y[b] + 1j * y[n + b] - eta_hat[b] * math.cos(w * 1.7)) < 1e-7 * abs(eta_hat[b])


@pytest.mark.parametrize("kind", ["channel", "slope"])
def test_hs1_is_affine_and_hs3_increment_is_cubic(kind):
    m = getattr(W1, kind)(512, t_end=2.0, **({"breaking": False} if kind == "slope" else {}))
    m.source_on = False
    y = W1.random_state(m, 4, amp=0.05, kfrac=0.3)
    m.order = 1
    r1 = [O.Oracle1D(m).rhs(0.3, s * y) for s in (1.0, 2.0)]
    assert relerr(r1[1], 2.0 * r1[0]) < 1e-12          # linear, source off
    d = []
    for s in (1.0, 2.0):
        m.order = 3
        r3 = O.Oracle1D(m).rhs(0.3, s * y)
        m.order = 2
        r2 = O.Oracle1D(m).rhs(0.3, s * y)
        d.append(r3 - r2)
    assert np.abs(d[0]).max() > 1e-10
    assert relerr(d[1], 8.0 * d[0]) < 1e-8


def test_hs3_two_depth_with_unit_weights_equals_flat():
    mf = _bare(W1.channel(256, t_end=2.0))
    mf.order = 3
    n = mf.n
    sym = _sym_multiplier(n, 200.0, 1.0)
    mf.csqr, mf.lop = sym, mf.k ** 2 * sym / S.GRAV
    one, zero = np.ones(n), np.zeros(n)
    ops = dict(gam_min=one, gam_plus=zero, csq_min=sym, csq_plus=0.5 * sym)
    mb = S.Model1D(**{**mf.__dict__, "flat": False, "ops": ops})
    y = W1.random_state(mf, 5, amp=0.2, kfrac=0.3)
    assert relerr(O.Oracle1D(mb).rhs(0.3, y), O.Oracle1D(mf).rhs(0.3, y)) < 1e-12
    mf.order = 2
    assert relerr(O.Oracle1D(mb).rhs(0.3, y), O.Oracle1D(mf).rhs(0.3, y)) > 1e-8
