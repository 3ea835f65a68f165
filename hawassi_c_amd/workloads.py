"""Seeded synthetic 2-D workloads.

* ``harbour(n)``       -- variable-depth harbour: 1.5 km x 1.5 km, depth ramping
  smoothly 20 m -> 4 m with seeded Gaussian bumps (amplitude 1 m, length
  scale 50 m), two-reference-depth operators, a breakwater (reflection 0.5)
  with a 100 m gap, directional JONSWAP line influx (Hs 2 m, Tp 8 s, s 15).
  n = 2048 is the full-size case; smaller n keep the same physics on a
  coarser grid for tests.
* ``short_crested(n)`` -- flat 30 m, 2 km x 2 km, absorbing strips of 200 m,
  directional JONSWAP line influx (Hs 3 m, Tp 10 s, s 10) at x = 250 m.
* ``shoal(n)`` + ``shoal_sea(model, seed)`` -- the learned-parareal case: a
  periodic 100 m x 100 m basin, depth 1.0 m over a Gaussian shoal rising to
  0.4 m, two-reference-depth operators, no influx, absorbing strips or
  adjustment zone, breaking off (so the dynamics is autonomous and keeps
  u = grad phi); the initial condition is a seeded directional JONSWAP sea
  (Hs 0.08 m, Tp 2 s, cos^2s spreading s = 10) in linear wave theory.
  ``shoal(n, shoal_depth=depth0)`` is the flat basin (flat-bottom operators).

Influx: the line signal eta_line(y_j, t) is a seeded random-phase sum of
directional JONSWAP components, used as the source signal with the
group-velocity influx kernel of the reference's area generation
(2d/functions.c:3052-3068) as spatial shape, bidirectional source.  The
spectral inversion of fun_area_generation (2d/functions.c:3175-3229) is not
applied, so wave heights are indicative only; the operator being evaluated
is the reference's.
"""
from __future__ import annotations

import math

import numpy as np

from . import setup2d as S


def splitmix64(seed: int):
    """SplitMix64 stream -> numpy Generator seed material."""
    x = seed & 0xFFFFFFFFFFFFFFFF
    while True:
        x = (x + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        yield z ^ (z >> 31)


def rng(seed: int) -> np.random.Generator:
    g = splitmix64(seed)
    return np.random.default_rng([next(g) for _ in range(4)])


def jonswap(w: np.ndarray, hs: float, tp: float, gamma: float) -> np.ndarray:
    """JONSWAP variance density (normalised afterwards to the target Hs)."""
    wp = 2 * math.pi / tp
    sig = np.where(w <= wp, 0.07, 0.09)
    s = w ** -5 * np.exp(-1.25 * (wp / w) ** 4) * gamma ** np.exp(-((w - wp) ** 2) / (2 * sig ** 2 * wp ** 2))
    m0 = np.trapezoid(s, w)
    return s * (hs / 4.0) ** 2 / m0


def directional_line_signal(y: np.ndarray, t: np.ndarray, hs: float, tp: float, gamma: float, s_spread: float,
                            depth: float, seed: int, ncomp: int = 256, theta0: float = 0.0,
                            ramp_periods: float = 2.0) -> np.ndarray:
    """eta(y_j, t) along a vertical line: random-phase directional sum, with a
    (1 - cos) start ramp over ramp_periods * Tp.  Returns [ny, nt]."""
    g = rng(seed)
    wp = 2 * math.pi / tp
    w = np.linspace(0.5 * wp, 3.0 * wp, ncomp)
    dw = w[1] - w[0]
    amp = np.sqrt(2 * jonswap(w, hs, tp, gamma) * dw)
    # cos^(2s) spreading by rejection sampling
    th = np.empty(ncomp)
    n = 0
    while n < ncomp:
        c = g.uniform(-math.pi / 2, math.pi / 2, 4 * ncomp)
        keep = c[g.uniform(0, 1, c.size) < np.cos(c) ** (2 * s_spread)]
        m = min(keep.size, ncomp - n)
        th[n:n + m] = keep[:m]
        n += m
    th += theta0
    phase = g.uniform(0, 2 * math.pi, ncomp)
    k = np.array([S.invers_omega(wi, depth) for wi in w])
    ky = k * np.sin(th)
    out = np.zeros((y.size, t.size))
    for m in range(ncomp):
        out += amp[m] * np.cos(w[m] * t[None, :] - ky[m] * y[:, None] + phase[m])
    tr = ramp_periods * tp
    ramp = np.where(t < tr, 0.5 * (1 - np.cos(math.pi * np.clip(t / tr, 0, 1))), 1.0)
    return out * ramp[None, :]


def _source_shapes(x, y, kx, line_x, meandepth, Lb, Lt):
    ng, sg = S.area_source_shapes_vertical(x, y, kx, line_x, y, meandepth, Lb, Lt)
    dx, dy = x[1] - x[0], y[1] - y[0]
    fact = (2 * math.pi) ** 2 / (dx * dy)
    # remove the reference's (2 pi)^2/(dx dy) factor: the signal here is the
    # line elevation itself (see module docstring)
    return ng / fact, sg / fact


def harbour(n: int = 2048, seed: int = 4, t_end: float = 60.0, dt_sig: float = 0.1, length: float = 1500.0,
            damp: float = 100.0, nadj: float = 1.0, eps_rel: float = 1e-2, eps_abs: float = 1e-3,
            with_wall: bool = True, breaking: bool = True, kbc: float = 0.85, ny: int | None = None) -> S.Model2D:
    """The harbour on an n x n grid (n x ny when ny is given, the y side
    then ny/n of the x side's length)."""
    m = ny or n
    x = S.interval_to_vector(0.0, length, n)
    y = S.interval_to_vector(0.0, length * m / n, m)
    dx, dy = x[1] - x[0], y[1] - y[0]
    kx = S.freqspace(x[-1] - x[0], n)
    ky = S.freqspace(y[-1] - y[0], m)
    g = rng(seed)
    X, Y = np.meshgrid(x, y)
    depth = 12.0 + 8.0 * np.cos(math.pi * X / length)
    nb = 24
    cx, cy = g.uniform(0, length, nb), g.uniform(0, length, nb)
    sgn = g.choice([-1.0, 1.0], nb)
    for i in range(nb):
        depth += sgn[i] * np.exp(-((X - cx[i]) ** 2 + (Y - cy[i]) ** 2) / (2 * 50.0 ** 2))
    # walls: a breakwater 12 m thick at x = 1000 m spanning y in [200, 1300]
    # with a 100 m gap centred at y = 750 m (wall.dat convention: 1 on the wall)
    wall_ind = np.zeros((m, n))
    if with_wall:
        on = (np.abs(X - 1000.0) <= 6.0) & (Y >= 200.0) & (Y <= 1300.0) & ~((Y > 700.0) & (Y < 800.0))
        wall_ind[on] = 1.0
    wallchar, cfsaw = S.wallchar_user(wall_ind, 0.5)
    cfsa = S.cfsa2d(x, y, damp, damp, damp, damp)
    if with_wall:
        cfsa = np.minimum(cfsa, cfsaw)
    fbdy = S.fbdy2d(x, y, damp, damp, damp, damp, cfsa, wall=with_wall)
    aal = S.aal2d(4, kx, ky)
    tp, hs = 8.0, 2.0
    line_x = 200.0
    ix = int(round((line_x - x[0]) / dx))
    meandepth = float(depth[:, ix].mean())
    wp = 2 * math.pi / tp
    kp = S.invers_omega(wp, meandepth)
    lam = 2 * math.pi / kp
    cpeak = float(S.exact_cp1d(kp, meandepth, 0.0))
    ops = S.disp_interpolation_2p(depth, kx, ky, wp)
    chi = S.chi_adjust_line(x, y, line_x, nadj, lam, vertical=True)
    t_sig = np.arange(0.0, t_end + 2 * dt_sig, dt_sig)
    sig = directional_line_signal(y, t_sig, hs, tp, 3.3, 15.0, meandepth, seed + 1000)
    ng, sg = _source_shapes(x, y, kx, line_x, meandepth, damp, damp)
    brk = S.breaking_params(hs, tp, kp, cfsa, chi, KBC=kbc, TC=0.2, Tchar=0.5) if breaking else None
    return S.Model2D(nx=n, ny=m, x=x, y=y, kx=kx, ky=ky, flat=False, cpeak=cpeak, aal=aal, fbdy=fbdy,
                     wallchar=wallchar, chi=chi, depth=depth, wall=with_wall, ops=ops, source_on=True,
                     vertical=True, propagation=2, new_gam=ng, skew_gam=sg, t_sig=t_sig, sig=sig,
                     skew_sig=np.zeros_like(sig), eps_abs=eps_abs, eps_rel=eps_rel, breaking=brk,
                     meta=dict(name=f"harbour{n}", tp=tp, hs=hs, kp=kp, lam=lam, line_x=line_x))


def short_crested(n: int = 1024, seed: int = 3, t_end: float = 60.0, dt_sig: float = 0.1, length: float = 2000.0,
                  damp: float = 200.0, depth0: float = 30.0, nadj: float = 1.0, eps_rel: float = 1e-2,
                  eps_abs: float = 1e-3, ny: int | None = None, wall_refl: float | None = None) -> S.Model2D:
    """The flat short-crested sea on an n x n grid (n x ny when ny is given).
    wall_refl: a breakwater (wall.dat convention) 3/5 of the way across x,
    spanning the middle 3/4 of y with a gap at its centre, with that
    reflection coefficient: the flat-bottom wall branches of
    fun_delphiH1hat (2d/fun_odesolver.c:2100-2239)."""
    m = ny or n
    x = S.interval_to_vector(0.0, length, n)
    y = S.interval_to_vector(0.0, length * m / n, m)
    dx, dy = x[1] - x[0], y[1] - y[0]
    kx = S.freqspace(x[-1] - x[0], n)
    ky = S.freqspace(y[-1] - y[0], m)
    KX, KY = kx[None, :], ky[:, None]
    depth = np.full((m, n), depth0)
    cfsa = S.cfsa2d(x, y, damp, damp, damp, damp)
    wall_on = wall_refl is not None
    wallchar = np.ones((m, n))
    gwx = gwy = None
    if wall_on:
        X, Y = np.meshgrid(x, y)
        ly = y[-1]
        on = ((np.abs(X - 0.6 * length) <= 2.5 * dx) & (Y >= 0.125 * ly) & (Y <= 0.875 * ly) &
              ~(np.abs(Y - 0.5 * ly) < 0.06 * ly))
        wall_ind = np.where(on, 1.0, 0.0)
        wallchar, cfsaw = S.wallchar_user(wall_ind, float(wall_refl))
        cfsa = np.minimum(cfsa, cfsaw)
        gwx, gwy = S.gradient_clipped(wallchar, dx, dy)   # fun_gradient of wallchar (2d/functions.c:741)
    fbdy = S.fbdy2d(x, y, damp, damp, damp, damp, cfsa, wall=wall_on)
    aal = S.aal2d(4, kx, ky)
    tp, hs = 10.0, 3.0
    wp = 2 * math.pi / tp
    kp = S.invers_omega(wp, depth0)
    lam = 2 * math.pi / kp
    cpeak = float(S.exact_cp1d(kp, depth0, 0.0))
    om2d = S.exact_disp2d(KX, KY, depth0)
    l2d = om2d * om2d / S.GRAV
    c2d = S.exact_cp2d(KX, KY, depth0)
    line_x = 250.0
    chi = S.chi_adjust_line(x, y, line_x, nadj, lam, vertical=True)
    t_sig = np.arange(0.0, t_end + 2 * dt_sig, dt_sig)
    sig = directional_line_signal(y, t_sig, hs, tp, 3.3, 10.0, depth0, seed + 1000)
    ng, sg = _source_shapes(x, y, kx, line_x, depth0, damp, damp)
    return S.Model2D(nx=n, ny=m, x=x, y=y, kx=kx, ky=ky, flat=True, cpeak=cpeak, aal=aal, fbdy=fbdy,
                     wallchar=wallchar, chi=chi, depth=depth, wall=wall_on,
                     refl_one=wall_on and float(wall_refl) == 1.0, grad_wall_x=gwx, grad_wall_y=gwy, c2d=c2d, l2d=l2d,
                     source_on=True, vertical=True, propagation=2, new_gam=ng, skew_gam=sg, t_sig=t_sig, sig=sig,
                     skew_sig=np.zeros_like(sig), eps_abs=eps_abs, eps_rel=eps_rel,
                     meta=dict(name=f"short_crested{n}", tp=tp, hs=hs, kp=kp, lam=lam, line_x=line_x))


def shoal(n: int = 512, length: float = 100.0, depth0: float = 1.0, shoal_depth: float = 0.4,
          shoal_width: float = 15.0, tp: float = 2.0, eps_rel: float = 1e-2, eps_abs: float = 1e-3) -> S.Model2D:
    """Periodic basin with a Gaussian shoal; see the module docstring."""
    x = S.interval_to_vector(0.0, length, n)
    y = S.interval_to_vector(0.0, length, n)
    kx = S.freqspace(x[-1] - x[0], n)
    ky = S.freqspace(y[-1] - y[0], n)
    X, Y = np.meshgrid(x, y)
    r2 = (X - 0.5 * length) ** 2 + (Y - 0.5 * length) ** 2
    depth = depth0 - (depth0 - shoal_depth) * np.exp(-r2 / (2 * shoal_width ** 2))
    one = np.ones((n, n))
    wp = 2 * math.pi / tp
    kp = S.invers_omega(wp, depth0)
    if shoal_depth == depth0:
        # no shoal: the flat-bottom operators
        om2d = S.exact_disp2d(kx[None, :], ky[:, None], depth0)
        return S.Model2D(nx=n, ny=n, x=x, y=y, kx=kx, ky=ky, flat=True, cpeak=float(S.exact_cp1d(kp, depth0, 0.0)),
                         aal=S.aal2d(4, kx, ky), fbdy=0.0 * one, wallchar=one, chi=one.copy(),
                         depth=np.full((n, n), depth0), wall=False, c2d=S.exact_cp2d(kx[None, :], ky[:, None], depth0),
                         l2d=om2d * om2d / S.GRAV, source_on=False, eps_abs=eps_abs, eps_rel=eps_rel, breaking=None,
                         meta=dict(name=f"flatbasin{n}", tp=tp, kp=kp, lam=2 * math.pi / kp, length=length))
    ops = S.disp_interpolation_2p(depth, kx, ky, wp)
    return S.Model2D(nx=n, ny=n, x=x, y=y, kx=kx, ky=ky, flat=False, cpeak=float(S.exact_cp1d(kp, depth0, 0.0)),
                     aal=S.aal2d(4, kx, ky), fbdy=0.0 * one, wallchar=one, chi=one.copy(), depth=depth, wall=False,
                     ops=ops, source_on=False, eps_abs=eps_abs, eps_rel=eps_rel, breaking=None,
                     meta=dict(name=f"shoal{n}", tp=tp, kp=kp, lam=2 * math.pi / kp, length=length))


def shoal_sea(model: S.Model2D, seed: int, hs: float = 0.08, tp: float = 2.0, gamma: float = 3.3,
              s_spread: float = 10.0, kmax_modes: int = 30, theta0: float = 0.0):
    """Seeded directional JONSWAP initial sea on the basin's Fourier modes
    (|k| below kmax_modes fundamental wavenumbers, mean direction theta0
    from +x, cos^(2s) spreading), linear theory at the far-field depth: eta
    and the surface potential phi = (g / omega) eta shifted by a quarter
    period.  Returns (eta, phi) [ny, nx]."""
    n = model.nx
    L = model.x[-1] - model.x[0]
    d0 = float(model.depth.max())
    g = rng(seed)
    k1 = 2 * math.pi / L
    kxv = np.fft.fftfreq(n) * n * k1
    kyv = np.fft.fftfreq(n) * n * k1
    KX, KY = kxv[None, :], kyv[:, None]
    K = np.hypot(KX, KY)
    th = np.angle(np.exp(1j * (np.arctan2(KY, KX) - theta0)))
    om = np.sqrt(S.GRAV * K * np.tanh(K * d0))
    wp = 2 * math.pi / tp
    with np.errstate(divide="ignore", invalid="ignore"):
        sig = np.where(om <= wp, 0.07, 0.09)
        spec = np.where(K > 0, om ** -5 * np.exp(-1.25 * (wp / om) ** 4)
                        * gamma ** np.exp(-((om - wp) ** 2) / (2 * sig ** 2 * wp ** 2)), 0.0)
    spread = np.where(np.abs(th) < math.pi / 2, np.cos(th) ** (2 * s_spread), 0.0)
    keep = (K > 0) & (K < kmax_modes * k1)
    amp = np.sqrt(np.where(keep, spec * spread, 0.0))
    a = amp * np.exp(1j * g.uniform(0, 2 * math.pi, (n, n)))
    eta = np.real(np.fft.ifft2(a))
    phi = np.real(np.fft.ifft2(-1j * np.where(keep, S.GRAV / np.where(om > 0, om, 1.0), 0.0) * a))
    sc = hs / 4.0 / max(float(eta.std()), 1e-300)
    return eta * sc, phi * sc


def random_state(model: S.Model2D, seed: int, amp: float = 0.05, kfrac: float = 0.35) -> np.ndarray:
    """A seeded smooth random state (full 6-plane layout): random fields whose
    spectra are limited to |k| < kfrac * k_max, forward transformed."""
    g = rng(seed)
    ny, nx = model.ny, model.nx
    KK = np.sqrt(model.kx[None, :] ** 2 + model.ky[:, None] ** 2)
    mask = KK < kfrac * KK.max()
    planes = []
    for f in range(3):
        spec = (g.standard_normal((ny, nx)) + 1j * g.standard_normal((ny, nx))) * mask
        phys = np.real(np.fft.ifft2(spec))
        phys *= (amp if f == 0 else 3 * amp) / (np.abs(phys).max() + 1e-300)
        s = np.fft.fft2(phys)
        planes += [s.real.ravel(), s.imag.ravel()]
    return np.concatenate(planes)
