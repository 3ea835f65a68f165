"""Host-side model setup for the 2-D AB2 path (FP64, numpy).

Restates, in this project's own code, the set-up functions of the reference
whose outputs the right-hand side reads every call (the reference's globals
``dominfo``, ``Oprt``, ``ChiAdj``, masks).  Each function cites the reference
routine it follows (paths relative to the reference repository).  Everything
here runs once per run, on the host; the per-step work is in the HIP library.

Array convention: fields are ``[ny, nx]`` row-major (row = y index, column =
x index), the reference's ``i*Nx+j`` layout.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

GRAV = 9.80665  # 2d/var_initialization.h:246


# --------------------------------------------------------------- grids --
def cal_npoint(xi: float, xf: float, dx: float) -> int:
    """calNpoint (2d/functions1.c:178-181)."""
    return int(round((xf - xi) * 1000 / (dx * 1000) + 1))


def interval_to_vector(a: float, b: float, n: int) -> np.ndarray:
    """fun_intervaltovector (2d/functions1.c:204-210)."""
    dxh = (b - a) / (n - 1)
    return a + np.arange(n) * dxh


def freqspace(length: float, n: int) -> np.ndarray:
    """freqspace (2d/functions.c:1810-1832): FFT-ordered wavenumbers with the
    reference's spacing (fmax-fmin)/(n-1) and k[0] forced to 1e-10."""
    delta = 2 * math.pi / length
    fmin = -0.5 * 2 * math.pi / (length / (n - 1))
    fmax = -fmin - delta if (n // 2) * 2 == n else -fmin
    d = (fmax - fmin) / (n - 1)
    k = np.empty(n)
    for i in range(n // 2):
        k[i] = fmin + (n // 2 + i) * d
    for ni, i in enumerate(range(n // 2, n)):
        k[i] = fmin + ni * d
    k[0] = 1e-10
    return k


def sign(v):
    """fun_sign (2d/functions1.c:170-172): -1 for v<0 else +1."""
    return np.where(np.asarray(v) < 0, -1.0, 1.0)


def heav(v):
    """fun_heav (2d/functions1.c:174-176)."""
    return np.where(np.asarray(v) < 0, 0.0, 1.0)


def aal2d(cutfrac: int, kx: np.ndarray, ky: np.ndarray) -> np.ndarray:
    """fun_aal2D (2d/functions.c:1834-1879): radial anti-alias mask."""
    nx, ny = len(kx), len(ky)
    idx = max(int(math.floor(nx / cutfrac)) - 1, 0)
    idy = max(int(math.floor(ny / cutfrac)) - 1, 0)
    kxmax, kymax = abs(kx[idx]), abs(ky[idy])
    KK = np.sqrt(kx[None, :] ** 2 + ky[:, None] ** 2)
    if idx > 0 and idy > 0:
        kmax = math.sqrt(kxmax ** 2 + kymax ** 2)
        return np.where(np.abs(KK) > kmax, 0.0, 1.0)
    alx = np.where(np.abs(kx) > kxmax, 0.0, 1.0)
    aly = np.where(np.abs(ky) > kymax, 0.0, 1.0)
    return aly[:, None] * alx[None, :]


# ---------------------------------------------------------- dispersion --
def exact_disp1d(k, D, U=0.0):
    """fun_exact_disp1d_val (2d/functions.c:64-68)."""
    k = np.asarray(k, dtype=float)
    return sign(k) * np.sqrt(GRAV * k * np.tanh(k * D)) + k * U


def exact_cp1d(k, D, U=0.0):
    """fun_exact_Cp1d_val (2d/functions.c:70-80); Cp(0) = sqrt(g D) + U."""
    k = np.asarray(k, dtype=float)
    with np.errstate(invalid="ignore", divide="ignore"):
        cp = (sign(k) * np.sqrt(GRAV * k * np.tanh(k * D)) + k * U) / k
    return np.where(k == 0, math.sqrt(GRAV * D) + U, cp)


def exact_cp2d(kx, ky, D, ux=0.0, uy=0.0):
    """fun_exact_Cp_val (2d/functions.c:101-112)."""
    nk = np.sqrt(kx ** 2 + ky ** 2)
    with np.errstate(invalid="ignore", divide="ignore"):
        cp = np.sqrt(GRAV * np.tanh(nk * D) / nk) + (kx * ux + ky * uy) / nk
    return np.where(nk == 0, math.sqrt(GRAV * D), cp)


def exact_disp2d(kx, ky, D, thetadir=0.0, ux=0.0, uy=0.0):
    """fun_exact_disp_val (2d/functions.c:114-120): directional sign."""
    s = sign(kx * math.cos(thetadir) + ky * math.sin(thetadir))
    nk = np.sqrt(kx ** 2 + ky ** 2)
    return s * np.sqrt(GRAV * nk * np.tanh(nk * D)) + kx * ux + ky * uy


def omega_i(k, d):
    """omega_i (2d/functions1.c:945-956)."""
    return 0.0 if k == 0 else (-1 if k < 0 else 1) * math.sqrt(GRAV * k * math.tanh(d * k))


def group_velocity_i(k, d):
    """group_velocity_i (2d/functions1.c:958-970)."""
    if k == 0:
        return math.sqrt(GRAV * d)
    t = math.tanh(d * k)
    return (-1 if k < 0 else 1) * math.sqrt(GRAV) / 2 / math.sqrt(k * t) * (t + k * (1 - t * t) * d)


def invers_omega(w: float, depth: float) -> float:
    """invers_omega (2d/functions1.c:997-1031): Newton from k=0 on
    w - omega(k), stopped when |k_n - k_{n-1}| < 1e-9 |k_n| (gsl_root_test_delta
    with epsabs 0, epsrel 1e-9), at most 1000 iterations."""
    x = 0.0
    for _ in range(1000):
        f = w - omega_i(x, depth)
        df = -group_velocity_i(x, depth)
        x0, x = x, x - f / df
        if abs(x - x0) < 1e-9 * abs(x):
            return x
    return 0.0


# ----------------------------------------------------- natural spline --
def spline_coeffs(x: np.ndarray, y: np.ndarray) -> np.ndarray:
    """Natural cubic spline half second derivatives c_i (gsl_interp_cspline
    convention: c_0 = c_{n-1} = 0), computed by the Thomas algorithm.  Works
    along the last axis of y."""
    x = np.asarray(x, float)
    y = np.asarray(y, float)
    n = x.shape[0]
    c = np.zeros_like(y)
    if n < 3:
        return c
    h = np.diff(x)
    m = n - 2
    diag = 2.0 * (h[:-1] + h[1:])
    off = h[1:].copy()
    dy = np.diff(y, axis=-1)
    rhs = 3.0 * (dy[..., 1:] / h[1:] - dy[..., :-1] / h[:-1])
    diag = np.broadcast_to(diag, rhs.shape).copy()
    for i in range(1, m):
        w = off[i - 1] / diag[..., i - 1]
        diag[..., i] -= w * off[i - 1]
        rhs[..., i] -= w * rhs[..., i - 1]
    c[..., m] = rhs[..., m - 1] / diag[..., m - 1]
    for i in range(m - 2, -1, -1):
        c[..., i + 1] = (rhs[..., i] - off[i] * c[..., i + 2]) / diag[..., i]
    return c


def spline_index(x: np.ndarray, t: float) -> int:
    """Interval search of gsl_interp_bsearch."""
    ilo, ihi = 0, len(x) - 1
    while ihi > ilo + 1:
        i = (ihi + ilo) // 2
        if x[i] > t:
            ihi = i
        else:
            ilo = i
    return ilo


def spline_eval(x, y, c, t):
    """Evaluate the natural spline (gsl cspline eval) at scalar t, along the
    last axis of y/c."""
    i = spline_index(x, t)
    h = x[i + 1] - x[i]
    dx = t - x[i]
    dyv = y[..., i + 1] - y[..., i]
    ci, cip = c[..., i], c[..., i + 1]
    b = dyv / h - h * (cip + 2.0 * ci) / 3.0
    d = (cip - ci) / (3.0 * h)
    return y[..., i] + dx * (b + dx * (ci + dx * d))


# ------------------------------------------- two-depth interpolation --
def inv_om_spline(Om: np.ndarray, kk: np.ndarray, nu: float) -> float:
    """fun_invOm (2d/functions.c:1764-1778): natural spline k(Om) at nu."""
    c = spline_coeffs(Om, kk)
    return float(spline_eval(Om, kk, c, nu))


def disp_interpolation_2p(depth: np.ndarray, kx: np.ndarray, ky: np.ndarray, nu: float,
                          ux: float = 0.0, uy: float = 0.0, mdir: float = 0.0) -> dict:
    """funOprt_DispInterpolation_2p (2d/functions.c:1534-1625).

    depth: positive still-water depth [ny, nx] (the reference's -bathy).
    Fits gamma_-(D), gamma_+(D) so that gamma_- C(D-) + gamma_+ C(D+) = C(D)
    and gamma_- w^2(D-) + gamma_+ w^2(D+) = w^2(D) at the wavenumber whose
    frequency is nu, on a 101-point depth table, then splines them in D.
    Returns the interpolation weights and the min/max-depth multipliers."""
    U = math.sqrt(ux * ux + uy * uy)
    KK = np.sqrt(kx[None, :] ** 2 + ky[:, None] ** 2)
    Dmin, Dmax = float(depth.min()), float(depth.max())
    N = 101
    dD = (Dmax - Dmin) / (N - 1)
    dkk = float(KK.max()) / (N - 1)
    DD = Dmin + np.arange(N) * dD
    kk = np.arange(N) * dkk
    DD[N - 1] = DD[N - 1] + 0.1 * dD
    gm = np.empty(N)
    gp = np.empty(N)
    for i in range(N):
        Om = exact_disp1d(kk, DD[i], U)
        kap = inv_om_spline(Om, kk, nu)
        cpD = float(exact_cp1d(kap, DD[i], U))
        cpm = float(exact_cp1d(kap, Dmin, U))
        cpp = float(exact_cp1d(kap, Dmax, U))
        o2D = float(exact_disp1d(kap, DD[i], U)) ** 2
        o2m = float(exact_disp1d(kap, Dmin, U)) ** 2
        o2p = float(exact_disp1d(kap, Dmax, U)) ** 2
        det = cpm * o2p - cpp * o2m
        gm[i] = (o2p * cpD - cpp * o2D) / det
        gp[i] = (-o2m * cpD + cpm * o2D) / det
    cgm = spline_coeffs(DD, gm)
    cgp = spline_coeffs(DD, gp)
    flat_d = depth.ravel()
    gam_m = np.array([spline_eval(DD, gm, cgm, d) for d in flat_d]).reshape(depth.shape) \
        if depth.size <= 4096 else _spline_eval_vec(DD, gm, cgm, flat_d).reshape(depth.shape)
    gam_p = np.array([spline_eval(DD, gp, cgp, d) for d in flat_d]).reshape(depth.shape) \
        if depth.size <= 4096 else _spline_eval_vec(DD, gp, cgp, flat_d).reshape(depth.shape)
    KX, KY = kx[None, :], ky[:, None]
    om_m = exact_disp2d(KX, KY, Dmin, mdir, ux, uy)
    om_p = exact_disp2d(KX, KY, Dmax, mdir, ux, uy)
    c_m = exact_cp2d(KX, KY, Dmin, ux, uy)
    c_p = exact_cp2d(KX, KY, Dmax, ux, uy)
    return dict(gam_min=gam_m, gam_plus=gam_p, om_min=om_m, om_plus=om_p,
                om2_min=om_m ** 2, om2_plus=om_p ** 2, l2d_min=om_m ** 2 / GRAV,
                l2d_plus=om_p ** 2 / GRAV, c2d_min=c_m, c2d_plus=c_p,
                csq_min=c_m ** 2, csq_plus=c_p ** 2, d_min=Dmin, d_max=Dmax,
                table_depth=DD, table_gam_min=gm, table_gam_plus=gp)


def disp_interpolation_3p(depth: np.ndarray, kx: np.ndarray, ky: np.ndarray, nu: float, mid: str = "default",
                          ux: float = 0.0, uy: float = 0.0, mdir: float = 0.0) -> dict:
    """funOprt_DispInterpolation_3p (2d/functions.c:1627-1762): three
    reference depths D-, Dmid (the midpoint, or the depth_ref value), D+.
    On a 100-point depth table the weights solve
        sum_j gamma_j C0(D_j) = C0(D),  sum_j gamma_j w^2(D_j) = w^2(D),
        sum_j gamma_j Cp(D_j) = Cp(D)
    at the wavenumber whose frequency is nu (C0 = sqrt(g D)), by Cramer's
    rule, and are splined in D.  Returns the 2p keys plus the *_mid ones."""
    U = math.sqrt(ux * ux + uy * uy)
    KK = np.sqrt(kx[None, :] ** 2 + ky[:, None] ** 2)
    Dmin, Dmax = float(depth.min()), float(depth.max())
    kkmax = float(KK.max())
    if kkmax < nu:
        kkmax = 2 * nu
    Dmid = (Dmax + Dmin) / 2 if mid == "default" else float(mid)
    N = 100
    dD = (Dmax - Dmin) / (N - 1)
    dkk = kkmax / (N - 1)
    DD = Dmin + np.arange(N) * dD
    kk = np.arange(N) * dkk
    DD[N - 1] = DD[N - 1] + 0.1 * dD
    G = np.empty((3, N))
    c0 = [math.sqrt(GRAV * Dmin), math.sqrt(GRAV * Dmid), math.sqrt(GRAV * Dmax)]
    for i in range(N):
        Om = exact_disp1d(kk, DD[i], U)
        kap = inv_om_spline(Om, kk, nu)
        cp = [float(exact_cp1d(kap, d, U)) for d in (Dmin, Dmid, Dmax)]
        o2 = [float(exact_disp1d(kap, d, U)) ** 2 for d in (Dmin, Dmid, Dmax)]
        cpD = float(exact_cp1d(kap, DD[i], U))
        o2D = float(exact_disp1d(kap, DD[i], U)) ** 2
        c0D = math.sqrt(GRAV * DD[i])
        A = o2[1] * cp[2] - o2[2] * cp[1]
        B = -(o2[0] * cp[2]) + o2[2] * cp[0]
        Cc = o2[0] * cp[1] - o2[1] * cp[0]
        D_ = -(c0[1] * cp[2]) + c0[2] * cp[1]
        E = c0[0] * cp[2] - c0[2] * cp[0]
        F = -(c0[0] * cp[1]) + c0[1] * cp[0]
        G_ = c0[1] * o2[2] - c0[2] * o2[1]
        H = -(c0[0] * o2[2]) + c0[2] * o2[0]
        I_ = c0[0] * o2[1] - c0[1] * o2[0]
        det = c0[0] * A + c0[1] * B + c0[2] * Cc
        G[0, i] = (A * c0D + D_ * o2D + G_ * cpD) / det
        G[1, i] = (B * c0D + E * o2D + H * cpD) / det
        G[2, i] = (Cc * c0D + F * o2D + I_ * cpD) / det
    flat_d = depth.ravel()
    gam = [_spline_eval_vec(DD, G[j], spline_coeffs(DD, G[j]), flat_d).reshape(depth.shape) for j in range(3)]
    KX, KY = kx[None, :], ky[:, None]
    out = dict(d_min=Dmin, d_mid=Dmid, d_max=Dmax, table_depth=DD, table_gam=G)
    for j, (name, d) in enumerate((("min", Dmin), ("mid", Dmid), ("plus", Dmax))):
        om = exact_disp2d(KX, KY, d, mdir, ux, uy)
        out["gam_" + name] = gam[j]
        out["om_" + name] = om
        out["om2_" + name] = om ** 2
        out["l2d_" + name] = om ** 2 / GRAV
        out["c2d_" + name] = exact_cp2d(KX, KY, d, ux, uy)
        out["csq_" + name] = exact_cp1d(KK, d, U) ** 2
    return out


def _spline_eval_vec(x, y, c, t):
    """Vectorised gsl cspline eval at many points (same interval rule)."""
    t = np.asarray(t, float)
    idx = np.searchsorted(x, t, side="right") - 1
    idx = np.clip(idx, 0, len(x) - 2)
    h = x[idx + 1] - x[idx]
    dx = t - x[idx]
    dy = y[idx + 1] - y[idx]
    ci, cip = c[idx], c[idx + 1]
    b = dy / h - h * (cip + 2.0 * ci) / 3.0
    d = (cip - ci) / (3.0 * h)
    return y[idx] + dx * (b + dx * (ci + dx * d))


# ------------------------------------------------------------ damping --
def fun_closest(a: np.ndarray, ref: float) -> int:
    """fun_closest (2d/functions1.c:786-799): first index of the minimum."""
    return int(np.argmin(np.abs(ref - a)))


def cfsa2d(x, y, Lleft, Lright, Lbottom, Ltop) -> np.ndarray:
    """fun_cfSA2D (2d/functions.c:2084-2132): 1 inside, cos-tapered to 0 in
    the absorbing strips."""
    nx, ny = len(x), len(y)
    ixl = fun_closest(x, x[0] + Lleft)
    ixr = fun_closest(x, x[nx - 1] - Lright)
    iyb = fun_closest(y, y[0] + Lbottom)
    iyt = fun_closest(y, y[ny - 1] - Ltop)
    hxl = np.where(x - x[0] < 0, 0.0, 1.0)
    hxr = np.where(x - x[ixr] < 0, 0.0, 1.0)
    hyb = np.where(y - y[0] < 0, 0.0, 1.0)
    hyt = np.where(y - y[iyt] < 0, 0.0, 1.0)
    with np.errstate(divide="ignore", invalid="ignore"):
        A = sign(x - x[ixl]); B = (1 - np.cos((x - x[0]) * math.pi / Lleft)) / 2
        cfxl = hxl * np.maximum(A, B) if Lleft != 0 else np.ones(nx)
        C = sign(x - x[nx - 1]); D = (1 - np.cos((x - x[ixr]) * math.pi / Lright)) / 2
        cfxr = hxr * np.maximum(C, D) if Lright != 0 else np.zeros(nx)
        E = sign(y - y[iyb]); F = (1 - np.cos((y - y[0]) * math.pi / Lbottom)) / 2
        cfyb = hyb * np.maximum(E, F) if Lbottom != 0 else np.ones(ny)
        G = sign(y - y[ny - 1]); H = (1 - np.cos((y - y[iyt]) * math.pi / Ltop)) / 2
        cfyt = hyt * np.maximum(G, H) if Ltop != 0 else np.zeros(ny)
    # the reference's ternary max returns B when A>B is false, including NaN
    cfx = cfxl - cfxr
    cfy = cfyb - cfyt
    return cfy[:, None] * cfx[None, :]


def fbdy2d(x, y, Lleft, Lright, Lbottom, Ltop, cfsa: np.ndarray, wall: bool, model: str = "HS2") -> np.ndarray:
    """fun_fbdy2D (2d/functions.c:1958-2082): damping characteristic
    (1-cfSA)/(0.9 L), with the strip length L carried across loop iterations
    exactly as the reference's running variable does, and x10 with walls for
    models other than HS1."""
    nx, ny = len(x), len(y)
    ixl = fun_closest(x, x[0] + Lleft)
    ixr = fun_closest(x, x[nx - 1] - Lright)
    iyb = fun_closest(y, y[0] + Lbottom)
    iyt = fun_closest(y, y[ny - 1] - Ltop)
    # The reference keeps L in a variable that survives loop iterations: at
    # j = 0 it is always set to Lleft, so the points between the strips carry
    # Lleft (or 7 if Lleft is 0) unless a bottom/top strip sets them.
    def z7(v):
        return 7.0 if v == 0 else v
    jj = np.arange(nx)[None, :]
    ii = np.arange(ny)[:, None]
    mid = (jj > ixl) & (jj < ixr)
    Lgrid = np.full((ny, nx), z7(Lleft))
    Lgrid = np.where(jj >= ixr, z7(Lright), Lgrid)
    Lgrid = np.where(mid & (ii < iyb), z7(Lbottom), Lgrid)
    Lgrid = np.where(mid & (ii > iyt), z7(Ltop), Lgrid)
    out = (1 - cfsa) / (0.9 * Lgrid)
    if wall and model != "HS1":
        out = 10 * out
    return out


# --------------------------------------------------------------- walls --
def cutoff(R: float, model: str = "HS2") -> float:
    """fun_cutoff (2d/functions.c:225-258), keeping the reference's operator
    precedence: -b + sqrt(b^2 - 4a(c-R)) / (2a)."""
    if model == "HS1":
        a, b, c = 0.8086, 0.2413, 0.0044
    else:
        a, b, c = 0.7289, 0.3229, -0.00058
    if R == 1:
        return 1.0
    if R == 0:
        return 0.0
    return -b + math.sqrt(b * b - 4 * a * (c - R)) / (2 * a)


def wallchar_user(wall_ind: np.ndarray, refl_coef: float, model: str = "HS2"):
    """fun_wallchar, user_defined branch (2d/functions.c:264-283).
    wall_ind: the wall.dat values (1 on the wall, 0 elsewhere).
    Returns (wallchar, cfSAwall)."""
    co = cutoff(refl_coef, model)
    cfsaw = 1 - wall_ind
    cfsaw = np.where(cfsaw < 1, 0.0, cfsaw)
    return 1 - wall_ind * co, cfsaw


def wallchar_rect(x, y, walls, refl_coef: float, model: str = "HS2", wallmethod: str = "energy"):
    """fun_wallchar, rectangle branch (2d/functions.c:285-345): walls given in
    input.txt as (X0, Y0, X1, Y1) corners (read at 2d/functions.c:184-186).

    * Each wall covers the points with X0' <= x < X1' and Y0' <= y < Y1',
      the fun_Heav2d profiles (2d/functions1.c:336-367) with bottom = X0 - x[0]
      and top = x[-1] - X1 (a bound of exactly 0 is dropped: X0 on the first
      grid point starts the wall there, X1 on the last point includes it);
      X0' = x[0] + (X0 - x[0]) and X1' = x[-1] - (x[-1] - X1) as the
      reference rounds them.
    * wallchar = 1 - cutoff(R) on every wall point and 1 elsewhere, unless
      wallmethod is "influxing" (wallchar = 1 everywhere, 2d/functions.c:313).
    * cfSAwall = 0 on the index box round((X0 - x[0])/dx) .. round((X1 -
      x[0])/dx) (inclusive, C round) of the FIRST wall only and 1 elsewhere
      (2d/functions.c:322-337): the reference damps inside wall 0 alone.
    The wall index lists (idxWall, dirWall, 2d/functions.c:346-389) feed only
    fun_swall, which nothing calls; they are not built.
    Returns (wallchar, cfSAwall), [ny, nx]."""
    x = np.asarray(x, float)
    y = np.asarray(y, float)
    nx, ny = len(x), len(y)
    heav = np.ones((ny, nx))
    for (x0, y0, x1, y1) in walls:
        hx = _heav2d(x, x0 - x[0], x[nx - 1] - x1)
        hy = _heav2d(y, y0 - y[0], y[ny - 1] - y1)
        heav = heav * (1 - hy[:, None] * hx[None, :])
    wallchar = np.ones((ny, nx))
    if wallmethod != "influxing":
        wallchar = np.where(heav == 0, 1 - cutoff(refl_coef, model), 1.0)
    cfsaw = np.ones((ny, nx))
    if len(walls):
        from .io2d import c_round
        dx, dy = x[1] - x[0], y[1] - y[0]
        x0, y0, x1, y1 = walls[0]
        ix0, ix1 = c_round((x0 - x[0]) / dx), c_round((x1 - x[0]) / dx)
        iy0, iy1 = c_round((y0 - y[0]) / dy), c_round((y1 - y[0]) / dy)
        # (a box reaching past the grid is cut at its edges; the reference
        # would write outside its arrays there)
        cfsaw[max(iy0, 0):max(iy1 + 1, 0), max(ix0, 0):max(ix1 + 1, 0)] = 0.0
    return wallchar, cfsaw


def read_wall_dat(path, nx: int, ny: int) -> np.ndarray:
    """wall.dat of a user_defined wall (2d/functions.c:268-283): Ny rows of
    Nx values read with fscanf("%lf "), 1 on the wall and 0 elsewhere.
    The reference's loop tests !fscanf(...), so
    * past the end of the file (fscanf returns EOF, -1) the loop goes on and
      every remaining point keeps wallchar's initial value 1
      (2d/functions.c:731), i.e. is read as wall;
    * a value that does not parse (fscanf returns 0) ends its row, and every
      later row at its first point (the stream stays on that token), so
      those points keep the no-wall values."""
    from .io2d import CStream
    out = np.zeros((ny, nx))
    cs = CStream(open(path, "rb").read())
    stuck = False
    for i in range(ny):
        for j in range(nx):
            cs.skip_ws()
            if cs.eof():
                out[i, j] = 1.0
                continue
            if stuck:
                break
            v = cs.scan("%lf ")
            if not v:
                stuck = True
                break
            out[i, j] = v[0]
    return out


def gradient_clipped(f: np.ndarray, dx: float, dy: float):
    """fun_gradient (2d/functions.c:392-461): periodic forward difference
    clipped to <=0 plus periodic backward difference clipped to >=0."""
    fx = (np.roll(f, -1, axis=1) - f) / dx
    fy = (np.roll(f, -1, axis=0) - f) / dy
    bx = (f - np.roll(f, 1, axis=1)) / dx
    by = (f - np.roll(f, 1, axis=0)) / dy
    fx = np.where(fx > 0, 0.0, fx)
    fy = np.where(fy > 0, 0.0, fy)
    bx = np.where(bx < 0, 0.0, bx)
    by = np.where(by < 0, 0.0, by)
    return fx + bx, fy + by


# ------------------------------------------------------- adjustment χ --
def cfv_profile(xi: np.ndarray, Xi: float, dl: float) -> np.ndarray:
    """cfv/cfh ramp (2d/functions.c:787-799)."""
    return heav(xi - Xi) * np.maximum(sign(xi - (Xi + dl)), (1 - np.cos((xi - Xi) * math.pi / dl)) / 2.0)


def chi_adjust_line(x, y, line_x: float, nadj: float, lam: float, vertical: bool = True) -> np.ndarray:
    """ChiAdj for an influx line spanning the whole domain side
    (2d/functions.c:917-929 vertical / 975-986 horizontal):
    chi = 1 - cfv(X - Nadj*lam) + cfv(X), a smooth dip to 0 around the line."""
    nx, ny = len(x), len(y)
    if nadj == 0.0:
        return np.ones((ny, nx))
    dl = nadj * lam
    if vertical:
        ix = int(round((line_x - x[0]) / (x[1] - x[0])))
        X = x[ix]
        prof = 1 - cfv_profile(x, X - dl, dl) + cfv_profile(x, X, dl)
        return np.broadcast_to(prof[None, :], (ny, nx)).copy()
    iy = int(round((line_x - y[0]) / (y[1] - y[0])))
    Y = y[iy]
    prof = 1 - cfv_profile(y, Y - dl, dl) + cfv_profile(y, Y, dl)
    return np.broadcast_to(prof[:, None], (ny, nx)).copy()


# ----------------------------------------------------- source shapes --
def exact_ug(k: np.ndarray, d: float, U: float = 0.0) -> np.ndarray:
    """fun_exact_Ug_val (2d/functions.c:82-92)."""
    out = np.empty_like(k, dtype=float)
    for i, ki in enumerate(k):
        if ki == 0:
            out[i] = math.sqrt(GRAV * d) + U
        else:
            t = math.tanh(d * ki)
            out[i] = (-1 if ki < 0 else 1) * math.sqrt(GRAV) / 2 / math.sqrt(ki * t) * (t + ki * (1 - t * t) * d) + U
    return out


def area_source_shapes_vertical(x, y, kx, line_x: float, line_y: np.ndarray, meandepth: float,
                                Lbottom: float, Ltop: float, U: float = 0.0):
    """Spatial influx shapes newGam/SkewGam for a vertical influx line
    (2d/functions.c:3018-3096): per line row, the inverse transform of the
    group velocity Ug(kx) shifted to the line position, its x-derivative
    (spectral), scaled by (2 pi)^2/(dx dy) and masked in y by the absorbing
    strips and the line extent.  Returns ([ny,nx], [ny,nx])."""
    nx, ny = len(x), len(y)
    dx, dy = x[1] - x[0], y[1] - y[0]
    fact = (2 * math.pi) ** 2 / (dx * dy)
    ug_hat = exact_ug(kx, meandepth, U)
    # ifft_1d_real_real: make_1d_sym then real(ifft)/N
    sym = _make_1d_sym(ug_hat)
    ugx = np.real(np.fft.ifft(sym))
    dist = line_x - x[0]
    ugx = np.roll(ugx, int(round(dist / dx)))
    temp = np.fft.fft(ugx)
    temp2 = 1j * kx * temp
    dgam = np.real(np.fft.ifft(_make_1dcom_sym(temp2)))
    heav1 = _heav2d(y, Lbottom, Ltop)
    heav2 = _heav2d(y, line_y[0] - y[0], y[ny - 1] - line_y[-1])
    G = np.zeros((ny, nx))
    DG = np.zeros((ny, nx))
    for yl in line_y:
        ind = int(round((yl - y[0]) / dy))
        G[ind, :] = ugx
        DG[ind, :] = dgam
    mask = (heav1 * heav2)[:, None]
    return fact * G * mask, fact * DG * mask


def _heav2d(Y, bottom, top):
    """fun_Heav2d profile (2d/functions1.c:336-367) along one axis."""
    n = len(Y)
    hb = np.ones(n) if bottom == 0 else np.where(Y - (Y[0] + bottom) < 0, 0.0, 1.0)
    ht = np.zeros(n) if top == 0 else np.where(Y - (Y[n - 1] - top) < 0, 0.0, 1.0)
    return hb * (1 - ht)


def _make_1d_sym(a):
    """make_1d_sym (2d/functions1.c:370-377): mirror the first half."""
    a = np.array(a, dtype=float)
    n = len(a)
    for j in range(1, n // 2):
        a[n - j] = a[j]
    return a


def _make_1dcom_sym(a):
    """make_1dcom_sym (2d/functions1.c:379-387): Hermitian mirror."""
    a = np.array(a, dtype=complex)
    n = len(a)
    a[0] = a[0].real
    a[n // 2] = a[n // 2].real
    for j in range(1, n // 2):
        a[n - j] = np.conj(a[j])
    return a


# ------------------------------------------------------------ breaking --
def breaking_params(hs: float, tp: float, kp: float, cfsa: np.ndarray, chi: np.ndarray,
                    KBC: float = 0.85, TC: float = 0.2, Tchar: float = 0.5, delb: float = 1.5) -> dict:
    """parambreak (2d/breaking.c:1-21) with pb.delb = 1.5 (2d/functions.c:204):
    onset threshold MaxEtaInit = 0.9 Hs/2, T* = Tchar Tp, search radius
    lambda_p/4, and the checked region pb.Char = 1 where cfSA = 1 and
    ChiAdj = 1 (outside absorbing strips, walls and the adjustment zone)."""
    char = np.where((cfsa < 1) | (chi < 1), 0.0, cfsa)
    return dict(KBC=KBC, TC=TC, Tchar=Tchar, delb=delb, MaxEtaInit=hs / 2.0 * 0.9, Tstar=Tchar * tp,
                Rsearch=2 * math.pi / kp / 4.0, char=char)


# ------------------------------------------------------- model bundle --
@dataclass
class Model2D:
    """Everything the 2-D RHS reads (the reference's globals), FP64."""
    nx: int
    ny: int
    x: np.ndarray
    y: np.ndarray
    kx: np.ndarray
    ky: np.ndarray
    flat: bool
    cpeak: float
    aal: np.ndarray
    fbdy: np.ndarray
    wallchar: np.ndarray
    chi: np.ndarray
    depth: np.ndarray
    wall: bool = False
    refl_one: bool = False
    c2d: np.ndarray | None = None
    l2d: np.ndarray | None = None
    grad_wall_x: np.ndarray | None = None
    grad_wall_y: np.ndarray | None = None
    ops: dict = field(default_factory=dict)
    source_on: bool = False
    vertical: bool = True
    propagation: int = 2      # 0 Uni+, 1 Uni-, 2 bidirectional
    new_gam: np.ndarray | None = None
    skew_gam: np.ndarray | None = None
    t_sig: np.ndarray | None = None
    sig: np.ndarray | None = None       # [nline, nt]
    skew_sig: np.ndarray | None = None  # [nline, nt]
    eps_abs: float = 1e-3
    eps_rel: float = 1e-2
    breaking: dict | None = None        # see breaking_params()
    order: int = 2                      # 1: HS1 (rhs_AB1_2D*_uv, linear), 2: HS2 (rhs_AB2_2D*_uv)
    meta: dict = field(default_factory=dict)

    @property
    def n(self):
        return self.nx * self.ny
