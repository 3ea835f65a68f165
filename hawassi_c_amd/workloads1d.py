"""Seeded synthetic 1-D workloads.

* ``channel(n)`` -- periodic domain [0, 200] m, uniform depth 1 m, damping
  zones of 20 m at both ends, Point influx at x = 50 m of a seeded
  random-phase JONSWAP record (Hs 0.05 m, Tp 1.5 s, gamma 3.3), flat HS2,
  no breaking, RKF45 eps_rel 1e-6, eps_abs 1e-7.  n = 4096 is the full size.
* ``slope(n, seed)`` -- [0, 400] m, depth 1.0 m up to x = 100 m, a linear
  slope to 0.2 m at x = 360 m, 0.2 m beyond; two-reference-depth operator
  (D- = 0.2, D+ = 1.0); Point influx at x = 50 m of a seeded JONSWAP record
  (Hs 0.10 m, Tp 2.0 s, gamma 3.3), unidirectional (Uni+); kinematic
  breaking (KBC 0.85, TC 0.2, Tchar 0.5); ode_solver_Ham's tolerances
  (eps_abs 1e-5, eps_rel 1e-4).
* ``slope_batch(nreal, seed)`` -- nreal realizations of ``slope`` with
  seeded phases and n drawn from {4096, 8192, 16384}.

The influx record is generated here (the reference's JONSWAP generator seeds
from the clock) and fed through the reference's Point-influx preparation
(modified_influx); see setup1d.
"""
from __future__ import annotations

import math

import numpy as np

from . import setup1d as S
from .workloads import jonswap, rng


def jonswap_record(t: np.ndarray, hs: float, tp: float, gamma: float, seed: int, ncomp: int = 256,
                   ramp_periods: float = 2.0) -> np.ndarray:
    """Random-phase JONSWAP elevation record with a (1 - cos) start ramp."""
    g = rng(seed)
    wp = 2 * math.pi / tp
    w = np.linspace(0.5 * wp, 3.0 * wp, ncomp)
    dw = w[1] - w[0]
    amp = np.sqrt(2 * jonswap(w, hs, tp, gamma) * dw)
    ph = g.uniform(0, 2 * math.pi, ncomp)
    out = (amp[None, :] * np.cos(w[None, :] * t[:, None] + ph[None, :])).sum(axis=1)
    tr = ramp_periods * tp
    return out * np.where(t < tr, 0.5 * (1 - np.cos(math.pi * np.clip(t / tr, 0, 1))), 1.0)


def _build(x, dx, depth, flat, hs, tp, xinflux, dampl, dampr, t_end, dt_sig, seed, breaking,
           propagation="Uni+", adjcoef=1.0, kbc=0.85, tc=0.2, tchar=0.5, eps_abs=1e-5, eps_rel=1e-4,
           name="1d", ndepth=3):
    n = len(x)
    xstart, xend = float(x[0]), float(x[-1])
    length = xend - xstart
    k = S.freqspace1d(length, n)
    idx_inf = S.closest(x, xinflux)
    xinflux = float(x[idx_inf])
    depthinf = float(depth[idx_inf])
    t = np.arange(0.0, t_end + 2 * dt_sig, dt_sig)
    eta_in = jonswap_record(t, hs, tp, 3.3, seed)
    hs_in = 4 * math.sqrt(float(np.var(eta_in)))
    nu = 2 * math.pi / tp
    kp = S.invers_omega(nu, depthinf)
    lam = 2 * math.pi / kp
    dampchar = S.dampzone(x, xstart + dampl, xend - dampr, dampl, dampr)
    cp_l = S.omega_i(kp, float(depth[0])) / kp
    cp_r = S.omega_i(kp, float(depth[-1])) / kp
    dampcoef = np.where(np.arange(n) < n // 2, 7 * cp_l / (0.9 * dampl), 7 * cp_r / (0.9 * dampr))
    aal = S.alias1d(k, 4)
    chi = S.chi_adjust(x, xinflux, adjcoef, lam, propagation)
    gx, gs = S.point_gamx(k, xinflux, xstart, dx, depthinf)
    sig, sig_skew = S.point_influx_signal(eta_in, t, depthinf)
    m = S.Model1D(n=n, x=x, dx=dx, k=k, aal=aal, chi=chi, damp=dampcoef * dampchar, depth=depth, flat=flat,
                  propagation=propagation, gamx=gx, gamx_skew=gs, t_sig=t, sig=sig, sig_skew=sig_skew,
                  eps_abs=eps_abs, eps_rel=eps_rel,
                  meta=dict(name=name, tp=tp, hs=hs, hs_in=hs_in, kp=kp, lam=lam, xinflux=xinflux,
                            depthinf=depthinf, seed=seed))
    if flat:
        m.csqr, m.lop = S.flat_operators(k, depthinf)
    elif ndepth == 3:
        # the reference's 1-D program always takes three reference depths for
        # a varying bottom (mid[0] = 3, 1d/main.cc:35-36; set_operator_bathy,
        # 1d/functions.c:2186-2193)
        m.ops = S.up3ip(depth, k, nu)
    else:
        m.ops = S.up2ip(depth, k, nu)
    if breaking:
        w = S.breaking_window(x, xinflux, adjcoef, lam, propagation)
        m.breaking = dict(KBC=kbc, TC=tc, Tstar=tchar * tp, MaxEtaInit=0.8 * hs_in / 2, delb=1.2, **w)
    return m


def channel(n: int = 4096, seed: int = 1, t_end: float = 60.0, dt_sig: float = 0.05) -> S.Model1D:
    x, dx = S.domain_n(0.0, 200.0, n)
    depth = np.ones(n)
    return _build(x, dx, depth, True, 0.05, 1.5, 50.0, 20.0, 20.0, t_end, dt_sig, seed, breaking=False,
                  eps_abs=1e-7, eps_rel=1e-6, name=f"channel{n}")


def slope_depth(x: np.ndarray) -> np.ndarray:
    return np.clip(1.0 - 0.8 * (x - 100.0) / 260.0, 0.2, 1.0)


def slope(n: int = 4096, seed: int = 2, t_end: float = 120.0, dt_sig: float = 0.05, breaking: bool = True,
          kbc: float = 0.85, ndepth: int = 3) -> S.Model1D:
    x, dx = S.domain_n(0.0, 400.0, n)
    return _build(x, dx, slope_depth(x), False, 0.10, 2.0, 50.0, 20.0, 20.0, t_end, dt_sig, seed,
                  breaking=breaking, kbc=kbc, name=f"slope{n}", ndepth=ndepth)


def slope_batch(nreal: int = 256, seed: int = 7, t_end: float = 120.0, sizes=(4096, 8192, 16384),
                breaking: bool = True, ndepth: int = 3) -> list[S.Model1D]:
    g = rng(seed)
    ns = g.choice(np.asarray(sizes), nreal)
    return [slope(int(n), seed=seed * 1000 + r, t_end=t_end, breaking=breaking, ndepth=ndepth)
            for r, n in enumerate(ns)]


def random_state(m: S.Model1D, seed: int, amp: float = 0.05, kfrac: float = 0.2) -> np.ndarray:
    """A seeded smooth random state in the reference's 4-plane layout."""
    g = rng(seed)
    n = m.n
    mask = np.abs(m.k) < kfrac * np.abs(m.k).max()
    planes = []
    for f, a in ((0, amp), (1, 3 * amp)):
        spec = (g.standard_normal(n) + 1j * g.standard_normal(n)) * mask
        phys = np.real(np.fft.ifft(spec))
        phys *= a / (np.abs(phys).max() + 1e-300)
        s = np.fft.fft(phys)
        planes += [s.real, s.imag]
    return np.concatenate(planes)
