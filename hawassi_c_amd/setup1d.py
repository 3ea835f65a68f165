"""Host-side model setup for the 1-D HS2 path (FP64, numpy).

Restates, in this project's own code, the reference's 1-D set-up routines
whose outputs ``rhs_HsF2`` reads every call (the globals of 1d/header.h:
k, aal, dampcoef*dampchar, ChiAdj, Csqr/L or the two-depth operator, gamX,
the influx signal splines and the breaking parameters).  Each function cites
the reference routine it follows; everything runs once per realization.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .setup2d import GRAV, _spline_eval_vec, spline_coeffs, spline_eval  # noqa: F401  (same natural spline)

PROPAGATION = {"Uni+": 0, "Uni-": 1, "Bi": 2, "None": 3}


# ---------------------------------------------------------------- grid --
def domain(xstart: float, xend: float, dx: float):
    """read_input domain set-up (1d/read_input.c:95-106): Nx = length/dx + 1,
    made even, then dx = length/(Nx-1)."""
    length = xend - xstart
    nx = int(length / dx) + 1
    if nx % 2:
        nx -= 1
    dx = length / (nx - 1)
    return xstart + dx * np.arange(nx), dx


def domain_n(xstart: float, xend: float, nx: int):
    """The same grid for a given (even) number of points."""
    dx = (xend - xstart) / (nx - 1)
    return xstart + dx * np.arange(nx), dx


def freqspace1d(length: float, n: int) -> np.ndarray:
    """freqspace (1d/functions.c:1173-1197): FFT-ordered, spacing
    (fmax-fmin)/(n-1); unlike 2-D, k[0] is left as computed (not 0)."""
    delta = 2 * math.pi / length
    fmin = -0.5 * 2 * math.pi / (length / (n - 1))
    fmax = -fmin - delta if (n // 2) * 2 == n else -fmin
    d = (fmax - fmin) / (n - 1)
    k = np.empty(n)
    h = n // 2
    k[:h] = fmin + (h + np.arange(h)) * d
    k[h:] = fmin + np.arange(n - h) * d
    return k


def closest(a: np.ndarray, ref: float) -> int:
    """closest (1d/functions.c:1696-1707): first index of the smallest
    |a_i - ref| below 10 * (a[n/2] - a[n/2-1]); 0 if none."""
    n = len(a)
    m = 10 * (a[n // 2] - a[n // 2 - 1])
    idx = 0
    for i in range(n):
        if abs(a[i] - ref) < m:
            m = abs(a[i] - ref)
            idx = i
    return idx


# ---------------------------------------------------------- dispersion --
def omega_i(k: float, d: float) -> float:
    """omega_i (1d/functions.c:1397-1413), exact dispersion."""
    return 0.0 if k == 0 else (-1.0 if k < 0 else 1.0) * math.sqrt(GRAV * k * math.tanh(d * k))


def omega_v(k: np.ndarray, d) -> np.ndarray:
    k = np.asarray(k, float)
    s = np.where(k < 0, -1.0, 1.0)
    with np.errstate(invalid="ignore"):
        om = s * np.sqrt(GRAV * k * np.tanh(d * k))
    return np.where(k == 0, 0.0, om)


def phase_velocity_v(k: np.ndarray, d) -> np.ndarray:
    """phase_velocity_i (1d/functions.c:1415-1428), vectorised."""
    k = np.asarray(k, float)
    kk = np.where(k == 0, 1.0, k)
    s = np.where(k < 0, -1.0, 1.0)
    cp = s * np.sqrt(GRAV * kk * np.tanh(d * kk)) / kk
    return np.where(k == 0, np.sqrt(GRAV * np.asarray(d, float)) * np.ones_like(k), cp)


def phase_velocity_i(k: float, d: float) -> float:
    if k == 0:
        return math.sqrt(GRAV * d)
    return (-1.0 if k < 0 else 1.0) * math.sqrt(GRAV * k * math.tanh(d * k)) / k


def group_velocity_i(k: float, d: float) -> float:
    """group_velocity_i (1d/functions.c:1361-1377)."""
    if k == 0:
        return math.sqrt(GRAV * d)
    t = math.tanh(d * k)
    return (-1.0 if k < 0 else 1.0) * math.sqrt(GRAV) / 2 / math.sqrt(k * t) * (t + k * (1 - t * t) * d)


def invers_omega(w: float, d: float) -> float:
    """invers_omega (1d/functions.c:984-1020): Newton from k = 0 on
    w - omega(k) with derivative -Ug(k), stopped when |k_n - k_{n-1}| <
    1e-3 |k_n| (gsl_root_test_delta(x, x0, 0, 1e-3)), at most 1000 steps;
    0 if it does not stop."""
    x = 0.0
    for _ in range(1000):
        f = w - omega_i(x, d)
        df = -group_velocity_i(x, d)
        x0, x = x, x - f / df
        if abs(x - x0) < 1e-3 * abs(x):
            return x
    return 0.0


# --------------------------------------------------------------- masks --
def dampzone(x: np.ndarray, bleft: float, bright: float, dampl: float, dampr: float) -> np.ndarray:
    """dampzone (1d/functions.c:1331-1345): 1 - hhR*hhL with cosine tapers."""
    hR0 = np.where(bright - x < 0, -1.0, 1.0)
    hR1 = 1 + np.where(bright + dampr - x < 0, -1.0, 1.0)
    hR2 = (1 + np.cos(math.pi * (x - bright) / dampr)) / 2
    hhR = np.maximum(hR0, np.minimum(hR1, hR2))
    hL0 = 1 + np.where(x - bleft + dampl < 0, -1.0, 1.0)
    hL1 = np.where(x - bleft < 0, -1.0, 1.0)
    hL2 = (1 - np.cos(math.pi * (x - bleft + dampl) / dampl)) / 2
    hhL = np.minimum(hL0, np.maximum(hL1, hL2))
    return 1 - hhR * hhL


def alias1d(k: np.ndarray, frac: int) -> np.ndarray:
    """alias (1d/functions.c:1773-1784): 1 inside |k| < k[idk], 0 outside,
    0.5 at idk = int(N/frac) - 1 itself."""
    n = len(k)
    idk = int(n / frac) - 1
    k0 = k[idk]
    d = k0 - np.abs(k)
    with np.errstate(invalid="ignore", divide="ignore"):
        al = (d / np.abs(d) + 1) / 2
    al[idk] = 0.5
    return al


def cf(x: np.ndarray, xi: float, dl: float) -> np.ndarray:
    """cf (1d/x_space.c:1-16): 0 left of xi, cosine ramp to 1 over dl."""
    out = np.zeros_like(x)
    with np.errstate(invalid="ignore", divide="ignore"):
        # fmax, as C: a NaN argument (0/0 at x = xi + dl) yields the other
        r = np.fmax((x - (xi + dl)) / np.abs(x - (xi + dl)), (1 - np.cos((x - xi) * math.pi / dl)) / 2)
    out[x > xi] = r[x > xi]
    out[x == xi] = 0.5 * r[x == xi]
    return out


def chi_adjust(x: np.ndarray, xinflux: float, adjcoef: float, lam: float, propagation: str) -> np.ndarray:
    """ChiAdj of x_space (1d/x_space.c:135-162)."""
    L = adjcoef * lam
    if propagation == "Uni+":
        return cf(x, xinflux, L)
    if propagation == "Uni-":
        return 1 - cf(x, xinflux - L, L)
    if propagation == "Bi":
        return 1 - cf(x, xinflux - L, L) + cf(x, xinflux, L)
    return np.ones_like(x)


# ------------------------------------------------------------ operators --
def up2ip(depth: np.ndarray, k: np.ndarray, nu: float, n: int = 101):
    """Up2IP (1d/functions.c:2195-2245): gamma_-(D), gamma_+(D) from a
    Cramer fit of C^2 and omega at the peak frequency on a 101-point depth
    table between min and max depth, natural spline in D; spectral
    multipliers Csq_min/plus = Cp(k, D_-/+)^2."""
    dmin, dplus = float(depth.min()), float(depth.max())
    dd = (dplus - dmin) / (n - 1)
    D = dmin + np.arange(n) * dd
    gm = np.empty(n)
    gp = np.empty(n)
    for i in range(n):
        kap = invers_omega(nu, D[i])
        cpm = phase_velocity_i(kap, dmin) ** 2
        cpp = phase_velocity_i(kap, dplus) ** 2
        cpd = phase_velocity_i(kap, D[i]) ** 2
        om_m, om_p, om_d = omega_i(kap, dmin), omega_i(kap, dplus), omega_i(kap, D[i])
        det = cpm * om_p - cpp * om_m
        gm[i] = (om_p * cpd - cpp * om_d) / det
        gp[i] = (-om_m * cpd + cpm * om_d) / det
    cgm, cgp = spline_coeffs(D, gm), spline_coeffs(D, gp)
    if depth.size <= 4096:
        gam_min = np.array([spline_eval(D, gm, cgm, float(d)) for d in depth])
        gam_plus = np.array([spline_eval(D, gp, cgp, float(d)) for d in depth])
    else:
        gam_min = _spline_eval_vec(D, gm, cgm, depth)
        gam_plus = _spline_eval_vec(D, gp, cgp, depth)
    return dict(gam_min=gam_min, gam_plus=gam_plus, csq_min=phase_velocity_v(k, dmin) ** 2,
                csq_plus=phase_velocity_v(k, dplus) ** 2, dmin=dmin, dplus=dplus)


def up3ip(depth: np.ndarray, k: np.ndarray, nu: float, mid: str = "default", n: int = 101):
    """Up3IP (1d/functions.c:2247-2340): three reference depths D-, Dmid (the
    midpoint, or the value of mid_depth_ref when it is not "default":
    read_kinematics sets mid[1] = atof(mid_depth_ref) before define_operator,
    1d/functions.c:3560-3562, 1d/main.cc:203-204), D+.  On a 101-point depth table the
    weights solve sum_j g_j X(D_j) = X(D) for X = sqrt(g D), omega and Cp^2
    at kappa(nu) by Cramer's rule, splined in D; Csq_* = Cp(k, D_*)^2."""
    dmin, dplus = float(depth.min()), float(depth.max())
    dmid = (dplus + dmin) / 2 if mid == "default" else float(mid)
    dd = (dplus - dmin) / (n - 1)
    D = dmin + np.arange(n) * dd
    G = np.empty((3, n))
    c0 = [math.sqrt(GRAV * dmin), math.sqrt(GRAV * dmid), math.sqrt(GRAV * dplus)]
    for i in range(n):
        kap = invers_omega(nu, D[i])
        cp = [phase_velocity_i(kap, d) ** 2 for d in (dmin, dmid, dplus)]
        om = [omega_i(kap, d) for d in (dmin, dmid, dplus)]
        c0d, omd, cpd = math.sqrt(GRAV * D[i]), omega_i(kap, D[i]), phase_velocity_i(kap, D[i]) ** 2
        A = om[1] * cp[2] - om[2] * cp[1]
        B = -(om[0] * cp[2]) + om[2] * cp[0]
        C = om[0] * cp[1] - om[1] * cp[0]
        Dd = -(c0[1] * cp[2]) + c0[2] * cp[1]
        E = c0[0] * cp[2] - c0[2] * cp[0]
        F = -(c0[0] * cp[1]) + c0[1] * cp[0]
        Gg = c0[1] * om[2] - c0[2] * om[1]
        H = -(c0[0] * om[2]) + c0[2] * om[0]
        J = c0[0] * om[1] - c0[1] * om[0]
        det = c0[0] * A + c0[1] * B + c0[2] * C
        G[0, i] = (A * c0d + Dd * omd + Gg * cpd) / det
        G[1, i] = (B * c0d + E * omd + H * cpd) / det
        G[2, i] = (C * c0d + F * omd + J * cpd) / det
    gam = [_spline_eval_vec(D, G[j], spline_coeffs(D, G[j]), depth) for j in range(3)]
    return dict(gam_min=gam[0], gam_mid=gam[1], gam_plus=gam[2], csq_min=phase_velocity_v(k, dmin) ** 2,
                csq_mid=phase_velocity_v(k, dmid) ** 2, csq_plus=phase_velocity_v(k, dplus) ** 2,
                dmin=dmin, dmid=dmid, dplus=dplus, table_depth=D, table_gam=G)


def flat_operators(k: np.ndarray, depthinf: float):
    """fun_csqr / operator_L (1d/functions.c:1430-1436, 1796-1800)."""
    csqr = phase_velocity_v(k, depthinf) ** 2
    return csqr, k * k * csqr / GRAV


# --------------------------------------------------------------- influx --
def cumtrapz(a: np.ndarray, dt: float) -> np.ndarray:
    """cumtrapz (1d/functions.c:1840-1847)."""
    out = np.zeros_like(a)
    out[1:] = np.cumsum(0.5 * dt * (a[:-1] + a[1:]))
    return out


def point_influx_signal(eta_in: np.ndarray, t: np.ndarray, depthinf: float):
    """modified_influx for Point influx (1d/functions.c:1968-2030): the line
    elevation filtered by the group velocity in frequency space, and the same
    for its time integral.  ww = freqspace over the record length
    (1d/functions.c:1265)."""
    n = len(t)
    dt = t[1] - t[0]
    skew = cumtrapz(eta_in, dt)
    ww = freqspace1d(t[-1] - t[0], n)
    kom = np.array([0.0 if w == 0 else invers_omega(w, depthinf) for w in ww])
    gv = np.array([group_velocity_i(kk, depthinf) for kk in kom])
    sig = np.real(np.fft.ifft(gv * np.fft.fft(eta_in)))
    sig_skew = np.real(np.fft.ifft(gv * np.fft.fft(skew)))
    return sig, sig_skew


def point_gamx(k: np.ndarray, xinflux: float, xstart: float, dx: float, depthinf: float):
    """Point generation kernel of x_space (1d/x_space.c:46-54, 101-104):
    gamX = exp(-i k (Xinflux - Xstart)) / dx, gamX_skew = i omega(k) gamX."""
    gx = np.exp(-1j * k * (xinflux - xstart)) / dx
    gs = 1j * omega_v(k, depthinf) * gx
    return gx, gs


# ------------------------------------------------------------- breaking --
def breaking_window(x: np.ndarray, xinflux: float, adjcoef: float, lam: float, propagation: str):
    """parambreak's search windows (1d/breaking.c:13-33).  For "Bi" the
    reference sets ILL/IFF but not lenIF, which stays 0, so its criterion
    never finds a peak (1d/breaking.c:67-72); that is kept."""
    n = len(x)
    w = dict(xbstart=0, xbend=n - 1, xbendL=0, xbstartR=0, lenIF=0, ILL=0, IFF=0)
    if propagation == "Uni+":
        w["xbstart"] = closest(x, xinflux + 0.1 * adjcoef * lam)
        w["lenIF"] = w["xbend"] - w["xbstart"]
    elif propagation == "Uni-":
        w["xbend"] = closest(x, xinflux - 0.1 * adjcoef * lam)
        w["lenIF"] = w["xbend"]
    elif propagation == "Bi":
        w["xbendL"] = closest(x, xinflux - 0.1 * adjcoef * lam)
        w["xbstartR"] = closest(x, xinflux + 0.1 * adjcoef * lam)
        w["ILL"] = w["xbendL"]
        w["IFF"] = w["xbend"] - w["xbstartR"]
    else:
        w["lenIF"] = w["xbend"]
    return w


@dataclass
class Model1D:
    n: int
    x: np.ndarray
    dx: float
    k: np.ndarray
    aal: np.ndarray
    chi: np.ndarray
    damp: np.ndarray            # dampcoef * dampchar
    depth: np.ndarray
    flat: bool
    csqr: np.ndarray | None = None
    lop: np.ndarray | None = None
    ops: dict | None = None
    source_on: bool = True
    propagation: str = "Uni+"
    gamx: np.ndarray | None = None
    gamx_skew: np.ndarray | None = None
    t_sig: np.ndarray | None = None
    sig: np.ndarray | None = None
    sig_skew: np.ndarray | None = None
    breaking: dict | None = None
    eps_abs: float = 1e-5
    eps_rel: float = 1e-4
    order: int = 2                      # 1: rhs_HsF1, 2: rhs_HsF2, 3: rhs_HsF3
    wall: dict | None = None            # idx, refl, t0, gamw_skew (define_wall)
    meta: dict = field(default_factory=dict)
