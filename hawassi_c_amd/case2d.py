"""A reference 2-D case directory (input.txt, bath.dat, influx.dat) turned
into the Model2D the library integrates: the setup the reference's main runs
before its time loop (2d/main.c:103-150), for what the device path supports
(grids of 2, 3 and 5 factors, walls from wall.dat or input.txt rectangles,
no runup).

* grid, wavenumbers, bathymetry, anti-alias mask, damping: fun_param_init
  (2d/functions.c:667-874);
* the influx line signal: fun_generation_influx (2d/functions.c:2634-2895),
  harmonic (harmonic_influx, :2200-2227), jonswap (fun_jonswap and
  jonswap_influx, :2226-2446, seeded) or user_defined (influx.dat), with
  the ramp of fun_ramp2dori / fun_tapered2dori (:2522-2595);
* the sea characteristics (Hs, peak frequency, kp, wavelength) and the
  nonlinear-adjustment zone ChiAdj: fun_calcseacharacteristic (:876-990);
* the area source: fun_area_generation (:2991-3276) -- the spatial kernels
  newGam / SkewGam and the signals Signal / SkewSignal that fun_source
  splines in time (the library builds the same natural-spline tables);
* the operators: fun_operator_setup (:1462-1533) with
  funOprt_DispInterpolation_2p or _3p (depth_ref 2 or 3) for a varying bottom.
"""
from __future__ import annotations

import math
import os

import numpy as np

from . import io2d
from . import setup2d as S


def _sgn(v):
    return np.where(np.asarray(v) < 0, -1.0, 1.0)


def _heav(v):
    return np.where(np.asarray(v) < 0, 0.0, 1.0)


def tapered_ramp(t, xy, xmin, xmax, LrampT, LrampXY, dxy, dt, ramp, taper):
    """fun_tapered2d / fun_tapered2dori (2d/functions.c:2455-2495,
    2522-2570): [Nt][n] = cfT(t) cfX(xy), cos ramps of length LrampT in time
    (both ends) and LrampXY along the line (both ends)."""
    t = np.asarray(t, float)
    xy = np.asarray(xy, float)
    c = io2d.c_round
    if taper == 0:
        cfX = np.ones(len(xy))
    else:
        idl = c((xmin + LrampXY - xy[0]) / dxy)
        idr = c((xmax - LrampXY - xy[0]) / dxy)
        cfXl = _heav(xy - xmin) * np.maximum(_sgn(xy - xy[idl]), 0.5 * (1.0 - np.cos((xy - xmin) * math.pi / LrampXY)))
        cfXr = _heav(xy - xy[idr]) * np.maximum(_sgn(xy - xmax), 0.5 * (1.0 - np.cos((xy - xy[idr]) * math.pi / LrampXY)))
        cfX = cfXl - cfXr
    if ramp == 0:
        cfT = np.ones(len(t))
    else:
        idb = c(LrampT / dt)
        idt = c((t[-1] - LrampT - t[0]) / dt)
        cfTb = _heav(t - t[0]) * np.maximum(_sgn(t - t[idb]), 0.5 * (1 - np.cos((t - t[0]) * math.pi / LrampT)))
        cfTt = _heav(t - t[idt]) * np.maximum(_sgn(t - t[-1]), 0.5 * (1 - np.cos((t - t[idt]) * math.pi / LrampT)))
        cfT = cfTb - cfTt
    return cfT[:, None] * cfX[None, :]


def ug2d(k, kt, depth):
    """funOprt_Ug2d with ej = (1, 0) (2d/functions.c:13-30): [Nt][N] group
    velocity at K = |(k, kt)|, signed by k."""
    K = np.sqrt(k[None, :] ** 2 + kt[:, None] ** 2)
    with np.errstate(invalid="ignore", divide="ignore"):
        th = np.tanh(depth * K)
        ug = _sgn(k[None, :] + 0.0 * kt[:, None]) * math.sqrt(S.GRAV) / 2.0 / np.sqrt(K * th) * (th + K * (1 - th * th) * depth)
    return np.where(K == 0, math.sqrt(S.GRAV * depth), ug)


def real_2d_sym(a):
    """real_2d_sym (2d/functions1.c:562-581) on [Nt][N]: mirror row 0 and row
    Nt/2 in N, column 0 in Nt, and the right half from the point mirror."""
    a = np.array(a, float)
    nt, n = a.shape
    for j in range(1, n // 2):
        a[0, n - j] = a[0, j]
        a[nt // 2, n - j] = a[nt // 2, j]
    for i in range(1, nt // 2):
        a[nt - i, 0] = a[i, 0]
    cols = np.arange(n // 2 + 1, n)
    src = n // 2 - np.arange(1, len(cols) + 1)
    for i in range(1, nt):
        a[i, cols] = a[nt - i, src]
    return a


def complex_2d_sym(a):
    """complex_2d_sym (2d/functions1.c:461-490) on [Nt][N]: the Hermitian
    projection the reference applies before a complex-to-real inverse."""
    a = np.array(a, complex)
    nt, n = a.shape
    for (i, j) in ((0, 0), (0, n // 2), (nt // 2, 0), (nt // 2, n // 2)):
        a[i, j] = a[i, j].real
    for j in range(1, n // 2):
        a[0, n - j] = np.conj(a[0, j])
        a[nt // 2, n - j] = np.conj(a[nt // 2, j])
    for i in range(1, nt // 2):
        a[nt - i, 0] = np.conj(a[i, 0])
    cols = np.arange(n // 2 + 1, n)
    src = n // 2 - np.arange(1, len(cols) + 1)
    for i in range(1, nt):
        a[i, cols] = np.conj(a[nt - i, src])
    return a


def influx_wavenumbers(ww, meandepth, flat: bool, U: float = 0.0):
    """k(omega) of the influx frequencies (2d/functions.c:2800-2821): Newton
    per frequency on a flat bottom, else a natural spline of omega(k) over
    10000 wavenumbers in [-1.1 kmax, 1.1 kmax] inverted at each omega."""
    wmax = float(np.max(ww))
    kmax = S.invers_omega(wmax, meandepth)
    if flat:
        return np.array([S.invers_omega(w, meandepth) for w in ww]), kmax
    nt = 10000
    dk = (2.2 * kmax) / (nt - 1)
    kt = -1.1 * kmax + np.arange(nt) * dk
    wt = S.exact_disp1d(kt, meandepth, U)
    c = S.spline_coeffs(wt, kt)
    kw = S._spline_eval_vec(wt, kt, c, np.asarray(ww, float))
    return np.where(np.asarray(ww) == 0, 0.0, kw), kmax


def area_signals(eta, k1, K1, meandepth, dxy, ramp, U: float = 0.0):
    """The signal half of fun_area_generation (2d/functions.c:3174-3226):
    the influx record eta [Nt][Nxy] divided in (t, line) Fourier space by the
    transform of the group-velocity kernel and reshaped so that the area
    source reproduces it; returns Signal, SkewSignal [Nxy][Nt] (ramped)."""
    KK = np.sqrt(k1[None, :] ** 2 + K1[:, None] ** 2)
    ug1 = ug2d(k1, K1, meandepth)
    # ifftw_2d_r2c then fftw_2d_c2c: the round trip of the symmetrised kernel
    gam_hat = np.fft.fft2(np.fft.ifft2(real_2d_sym(ug1)))
    ug_hat = ug2d(k1, K1, meandepth)
    ffact = dxy / (2 * math.pi)
    Ks = _sgn(K1[:, None] + k1[None, :]) * KK
    WW1 = S.exact_disp1d(Ks, meandepth, U)
    op_c = WW1 / K1[:, None]
    ins_hat = np.fft.fft2(eta)
    ff_hat = ffact * (ins_hat / gam_hat) / (2 * math.pi) * np.abs(K1)[:, None] / KK * ug_hat
    skew_hat = op_c * (ff_hat.imag - 1j * ff_hat.real) / WW1
    ff = np.real(np.fft.ifft2(complex_2d_sym(ff_hat)))
    skew = np.real(np.fft.ifft2(complex_2d_sym(skew_hat)))
    return (ff * ramp).T.copy(), (skew * ramp).T.copy()


def area_shapes(x, y, kx, ky, line_pts, meandepth, damp, vertical: bool, U: float = 0.0):
    """The spatial half of fun_area_generation (2d/functions.c:3018-3172):
    newGam, SkewGam [ny][nx]."""
    Lleft, Lright, Lbottom, Ltop = damp
    if vertical:
        lx = line_pts[0][0]
        return S.area_source_shapes_vertical(x, y, kx, lx, np.array([p[1] for p in line_pts]), meandepth,
                                             Lbottom, Ltop, U)
    # horizontal: the same construction with x and y exchanged
    ly = line_pts[0][1]
    ng, sg = S.area_source_shapes_vertical(y, x, ky, ly, np.array([p[0] for p in line_pts]), meandepth,
                                           Lleft, Lright, U)
    return ng.T.copy(), sg.T.copy()


def _bilinear(xd, yd, zd, x, y):
    """gsl_interp2d_bilinear at the points (x[j], y[i]) inside the data grid,
    -9999 elsewhere (fun_interp2, 2d/functions1.c:1102-1128)."""
    out = np.full((len(y), len(x)), -9999.0)
    jx = np.where((x >= xd[0]) & (x <= xd[-1]))[0]
    iy = np.where((y >= yd[0]) & (y <= yd[-1]))[0]
    if len(jx) == 0 or len(iy) == 0:
        return out
    xi = np.clip(np.searchsorted(xd, x[jx], side="right") - 1, 0, len(xd) - 2)
    yi = np.clip(np.searchsorted(yd, y[iy], side="right") - 1, 0, len(yd) - 2)
    t = (x[jx] - xd[xi]) / (xd[xi + 1] - xd[xi])
    u = (y[iy] - yd[yi]) / (yd[yi + 1] - yd[yi])
    T, U_ = t[None, :], u[:, None]
    z00 = zd[yi[:, None], xi[None, :]]
    z10 = zd[yi[:, None], xi[None, :] + 1]
    z01 = zd[yi[:, None] + 1, xi[None, :]]
    z11 = zd[yi[:, None] + 1, xi[None, :] + 1]
    out[np.ix_(iy, jx)] = (1 - T) * (1 - U_) * z00 + T * (1 - U_) * z10 + T * U_ * z11 + (1 - T) * U_ * z01
    return out


def chi_adjust_partial(x, y, line_pts, dista):
    """ChiAdj for an influx line shorter than the domain side
    (2d/functions.c:930-970): 1 - cfSA of the line's bounding box widened by
    Nadj lambda, interpolated bilinearly onto the grid; 1 outside the box,
    values below 1e-3 set to 0."""
    dx, dy = x[1] - x[0], y[1] - y[0]
    x1, x2 = line_pts[0][0] - dista, line_pts[-1][0] + dista
    y1, y2 = line_pts[0][1] - dista, line_pts[-1][1] + dista
    nxt = int((x2 - x1) / dx) + 1
    nyt = int((y2 - y1) / dy) + 1
    xt = x1 + np.arange(nxt) * dx
    yt = y1 + np.arange(nyt) * dy
    cf = S.cfsa2d(xt, yt, dista, dista, dista, dista)
    chi = _bilinear(xt, yt, cf, np.asarray(x, float), np.asarray(y, float))
    out = np.where(chi == -9999, 1.0, 1 - chi)
    return np.where((chi != -9999) & (np.abs(out) < 0.001), 0.0, out)


# ------------------------------------------------------------ JONSWAP --
def jonswap_spectrum(omg, omgp: float, gamJS: float, Hs: float) -> np.ndarray:
    """fun_jonswap (2d/functions.c:2226-2307) on the first len(omg)//2
    frequencies: the JONSWAP spectrum alpha g^2 (2 pi)^-4 f^-5
    exp(-5/4 (f/fp)^-4) gamma^exp(-(f/fp - 1)^2 / (2 sigma^2)), sigma 0.07 up
    to omega_p and 0.09 above, zero at the first frequency, scaled so that
    its trapezoid integral (functions1.c:236-246: uniform spacing omg[1] -
    omg[0]) is Hs^2/16.  The reference reads SPcorrection.dat into a
    corrected copy but returns the uncorrected spectrum, so no correction
    applies."""
    omg = np.asarray(omg, float)
    n = len(omg) // 2
    w = omg[:n]
    f = w / (2 * math.pi)
    fp = omgp / (2 * math.pi)
    sig = np.where(w <= omgp, 0.07, 0.09)
    js = np.zeros(n)
    with np.errstate(divide="ignore", over="ignore", invalid="ignore"):
        g = np.exp(-0.5 * ((f[1:] / fp - 1) / sig[1:]) ** 2)
        js[1:] = (0.0081 * S.GRAV ** 2 * (2 * math.pi) ** -4 * f[1:] ** -5
                  * np.exp(-(5.0 / 4.0) * (f[1:] / fp) ** -4) * gamJS ** g)
    m0 = (js.sum() - 0.5 * (js[0] + js[-1])) * (w[1] - w[0])
    return (Hs * Hs / 16.0) * js / m0


def jonswap_directions(nhalf: int, s_word: str, u: np.ndarray) -> np.ndarray:
    """The directional spreading of jonswap_influx (2d/functions.c:2318-2390):
    one direction per frequency, 0 for long-crested waves (s = "inf"); else
    inverse-CDF draws from cos^(2s)(theta) on theta = -pi/2 + j pi/(nhalf-1):
    the cumulative sum normalised to 1 plus exp(-12)/pi theta (strictly
    increasing), inverted by a natural spline (theta as a function of the
    CDF) at the uniform draws u."""
    if s_word == "inf":
        return np.zeros(nhalf)
    s = io2d._atof(s_word)
    theta = -math.pi / 2 + np.arange(nhalf) * (math.pi / (nhalf - 1))
    pdf = (math.gamma(s + 1.0) / (math.sqrt(math.pi) * math.gamma(s + 0.5))) * np.cos(theta) ** (2 * s)
    cdf = np.cumsum(pdf)
    cdf = cdf / cdf[-1] + math.exp(-12.0) / math.pi * theta
    c = S.spline_coeffs(cdf, theta)
    return S._spline_eval_vec(cdf, theta, c, np.asarray(u, float))


def jonswap_line_signal(t, ww, kw, Tp, gamJS, Hs, s_word, mdir, line_xy, line0, vertical: bool,
                        seed: int = 0) -> np.ndarray:
    """jonswap_influx (2d/functions.c:2310-2446): the line record
    eta(t_i, p) = sum_j a_j cos(-w_j t_i + phase_j + k_j sin(theta_j + mdir)
    (y_p - y_0)) (vertical line; cos and x for a horizontal one), a_j =
    sqrt(2 S_j dw), over the first Nt/2 influx frequencies ww and their
    wavenumbers kw.  [Nt][nline].

    The reference draws the directions and the phases after srand(time(0))
    each (unreproducible); here one seeded uniform sequence u serves both, as
    the two srand(time(0)) calls within one second give the reference one
    sequence twice: directions from u (short-crested), phases 2 pi u."""
    t = np.asarray(t, float)
    nt = len(t)
    nhalf = nt // 2
    spec = jonswap_spectrum(ww, 2 * math.pi / Tp, gamJS, Hs)
    dw = ww[1] - ww[0]
    a = np.sqrt(2 * spec * dw)
    u = np.random.default_rng(seed).random(nhalf)
    th = jonswap_directions(nhalf, s_word, u)
    ph = 2 * math.pi * u
    w = np.asarray(ww[:nhalf], float)
    k = np.asarray(kw[:nhalf], float)
    rel = np.asarray(line_xy, float) - line0
    proj = np.sin(th + mdir) if vertical else np.cos(th + mdir)
    # sum_j a_j cos(psi_jp - w_j t) = Re sum_j (a_j e^{i psi_jp}) e^{-i w_j t}
    E = a[None, :] * np.exp(1j * (ph[None, :] + (k * proj)[None, :] * rel[:, None]))
    out = np.empty((nt, len(rel)))
    for i0 in range(0, nt, 512):
        tt = t[i0:i0 + 512]
        out[i0:i0 + 512] = np.real(np.exp(-1j * np.outer(tt, w)) @ E.T)
    return out


def read_influx(path, nxy: int, nt: int):
    """influx.dat (2d/functions.c:2643-2683): columns = spaces on the first
    line (read_col, at most Nxy), then the line x, the line y and Nt rows of
    time and elevation."""
    raw = open(path, "rb").read()
    first = raw.split(b"\n", 1)[0]
    col = min(first.count(b" "), nxy)
    cs = io2d.CStream(raw)
    rd = lambda: (cs.scan("%lf ") or [0.0])[0]
    rd()
    xs = [rd() for _ in range(col)]
    rd()
    ys = [rd() for _ in range(col)]
    t = np.zeros(nt)
    eta = np.zeros((nt, col))
    for j in range(nt):
        t[j] = rd()
        for i in range(col):
            eta[j, i] = rd()
    return np.array(xs), np.array(ys), t, eta


def model_from_input(case_dir: str, eps_abs: float = 1e-3, eps_rel: float = 1e-2, jonswap_seed: int = 0) -> S.Model2D:
    """Builds the Model2D of a reference case directory.  jonswap_seed seeds
    the random directions and phases of a JONSWAP influx (the reference
    seeds them from the clock)."""
    p = io2d.read_input(os.path.join(case_dir, "input.txt"))
    if p.wavename not in ("harmonic", "user_defined", "jonswap"):
        raise NotImplementedError(f"influx '{p.wavename}' is not supported (harmonic, user_defined, jonswap)")
    if p.order == 0:
        raise ValueError(f"unknown dynamics '{p.model}'")
    # ---- fun_param_init
    nx = S.cal_npoint(p.x_start, p.x_end, p.dx)
    ny = S.cal_npoint(p.y_start, p.y_end, p.dy)
    for n in (nx, ny):
        m = n
        for f in (2, 3, 5):
            while m % f == 0:
                m //= f
        if n < 2 or m != 1:
            raise NotImplementedError(f"grid {nx}x{ny}: the device transforms take sides of 2, 3 and 5 factors")
    x = S.interval_to_vector(p.x_start, p.x_end, nx)
    y = S.interval_to_vector(p.y_start, p.y_end, ny)
    dx, dy = x[1] - x[0], y[1] - y[0]
    kx = S.freqspace(x[-1] - x[0], nx)
    ky = S.freqspace(y[-1] - y[0], ny)
    bathy, runup = io2d.read_bathy(os.path.join(case_dir, "bath.dat"), nx, ny, p.bathy_id, p.depth)
    if runup:
        raise NotImplementedError("runup (dry points) is not supported")
    flat = p.bathy_id == "flat"
    aal = S.aal2d(p.cutfrac, kx, ky)
    nt = S.cal_npoint(p.t_start, p.t_end, p.dt)
    if nt % 2:
        nt -= 1
    t = p.t_start + np.arange(nt) * p.dt
    dt = t[2] - t[1]
    cfsa = S.cfsa2d(x, y, *p.damp)
    # ---- fun_wallchar (2d/functions.c:260-390): any wall word but "no" has a
    # wall, wall.dat for "user_defined", else atoi(word) rectangles
    wall = p.wall != "no"
    wallchar = np.ones((ny, nx))
    gwx = gwy = None
    if wall:
        if p.wall == "user_defined":
            wallchar, cfsaw = S.wallchar_user(S.read_wall_dat(os.path.join(case_dir, "wall.dat"), nx, ny),
                                              p.refl_coef, p.model)
        else:
            wallchar, cfsaw = S.wallchar_rect(x, y, p.walls, p.refl_coef, p.model, p.wallmethod)
        # fun_fbdy2D: cfSA = min(cfSA, cfSAwall) (2d/functions.c:1964-1970)
        cfsa = np.minimum(cfsa, cfsaw)
        # fun_gradient of wallchar (2d/functions.c:741), read by the flat
        # bottom's reflecting-wall branch of fun_delphiH1hat
        gwx, gwy = S.gradient_clipped(wallchar, dx, dy)
    fbdy = S.fbdy2d(x, y, *p.damp, cfsa, wall=wall, model=p.model)
    vertical = p.orientation == "vertical"
    nxy = ny if vertical else nx
    # ---- fun_generation_influx
    Tp_in = p.Tp
    if p.wavename == "user_defined":
        lx, ly, t_file, eta_l = read_influx(os.path.join(case_dir, "influx.dat"), nxy, nt)
        t = t_file
        ix = np.array([io2d.c_round((v - x[0]) / dx) for v in lx])
        iy = np.array([io2d.c_round((v - y[0]) / dy) for v in ly])
        bath_inf = bathy[iy, ix]
        line_pts = list(zip(x[ix], y[iy]))
        lxy = y[iy] if vertical else x[ix]
        span = (lxy[0], lxy[-1])
        LrampT = p.ramp * Tp_in
        Linf = span[1] - span[0]
        r_ori = tapered_ramp(t, lxy, span[0], span[1], LrampT, (p.taper / 2.0 if vertical else p.taper) * Linf,
                             dy if vertical else dx, dt, p.ramp, p.taper) \
            if (p.ramp != 0.0 or p.taper != 0.0) else np.ones((nt, len(lxy)))
        eta = np.zeros((nt, nxy))
        idx = iy if vertical else ix
        eta[:, idx] = eta_l * r_ori
        meandepth = -float(np.mean(bath_inf))
    else:
        c = io2d.c_round
        ix0, iy0 = c((p.line_influx[0] - x[0]) / dx), c((p.line_influx[1] - y[0]) / dy)
        ix1, iy1 = c((p.line_influx[2] - x[0]) / dx), c((p.line_influx[3] - y[0]) / dy)
        nl = (iy1 - iy0) + 1 if vertical else (ix1 - ix0) + 1
        if vertical:
            mg = (y[iy1] - y[iy0]) / (x[ix1] - x[ix0]) if ix1 != ix0 else 0.0
            it = [c(((x[ix0] + j * mg / dy) - x[0]) / dx) for j in range(nl)]
            line_pts = [(x[it[j]], y[iy0 + j]) for j in range(nl)]
            bath_inf = np.array([bathy[iy0 + j, it[j]] for j in range(nl)])
        else:
            mg = (y[iy1] - y[iy0]) / (x[ix1] - x[ix0]) if iy1 != iy0 else 0.0
            it = [c(((y[iy0] + j * mg * dx) - y[0]) / dy) for j in range(nl)]
            line_pts = [(x[ix0 + j], y[it[j]]) for j in range(nl)]
            bath_inf = np.array([bathy[it[j], ix0 + j] for j in range(nl)])
        meandepth = -float(np.mean(bath_inf))
        lxy = np.array([q[1] if vertical else q[0] for q in line_pts])
        Linf = lxy[-1] - lxy[0]
        r_ori = tapered_ramp(t, lxy, lxy[0], lxy[-1], p.ramp * Tp_in, (p.taper / 2.0 if vertical else p.taper) * Linf,
                             dy if vertical else dx, dt, p.ramp, p.taper) \
            if (p.ramp != 0.0 or p.taper != 0.0) else np.ones((nt, nl))
        eta = np.zeros((nt, nxy))
        ii = np.arange(nl)
        if p.wavename == "harmonic":
            # harmonic_influx: the line coordinate is the grid's own (dominfo.y[ii])
            W0 = 2 * math.pi / Tp_in
            K0 = S.invers_omega(W0, meandepth)
            if ix0 == ix1:
                eta[:, ii + iy0] = r_ori * p.Hs / 2 * np.cos(K0 * math.sin(p.mdir) * y[ii][None, :] - W0 * t[:, None])
            else:
                eta[:, ii + ix0] = r_ori * p.Hs / 2 * np.cos(K0 * math.cos(p.mdir) * x[ii][None, :] - W0 * t[:, None])
    ww = S.freqspace(t[-1] - t[0], nt)
    kw, _ = influx_wavenumbers(ww, meandepth, flat, p.U)
    if p.wavename == "jonswap":
        # jonswap_influx on the line's points (wave_influx.x / .y), relative
        # to the line's first grid point, ramped like the harmonic record
        lx = np.array([q[0] for q in line_pts])
        ly = np.array([q[1] for q in line_pts])
        sig = jonswap_line_signal(t, ww, kw, Tp_in, p.gamJS, p.Hs, p.s, p.mdir, ly if vertical else lx,
                                  y[iy0] if vertical else x[ix0], vertical, seed=jonswap_seed)
        if vertical:
            eta[:, np.arange(nl) + iy0] = sig * r_ori
        else:
            eta[:, np.arange(nl) + ix0] = sig * r_ori
    # ---- fun_calcseacharacteristic
    ins_hat = np.fft.fft2(eta)
    Hs = p.Hs
    if p.wavename == "user_defined":
        Hs = 4.0 * math.sqrt(float(np.sum((eta - eta.mean()) ** 2)) / (eta.size + 1))
    a = np.abs(ins_hat)
    flat_i = int(np.argmax(a))
    id_t = flat_i // a.shape[1] if a.ravel()[flat_i] > a[0, 0] else -1
    wp = abs(ww[id_t])
    Tp = 2 * math.pi / wp
    kp = S.invers_omega(wp, meandepth)
    lam = 2 * math.pi / kp
    if p.model != "HS1" and p.Nadj != 0.0:
        lxy_all = np.array([q[1] if vertical else q[0] for q in line_pts])
        full = (lxy_all[0] == (y[0] if vertical else x[0])) and (lxy_all[-1] == (y[-1] if vertical else x[-1]))
        if full:
            chi = S.chi_adjust_line(x, y, line_pts[0][0] if vertical else line_pts[0][1], p.Nadj, lam,
                                    vertical=vertical)
        else:
            chi = chi_adjust_partial(x, y, line_pts, p.Nadj * lam)
    else:
        chi = np.ones((ny, nx))
    # ---- fun_ramp2d (over the whole side) and fun_area_generation
    side = y if vertical else x
    if p.ramp == 0.0 and p.taper == 0.0:
        r2d = np.ones((nt, nxy))
    else:
        Lside = side[-1] - side[0]
        r2d = tapered_ramp(t, side, side[0], side[-1], p.ramp * Tp, (p.taper / 2.0 if vertical else p.taper) * Lside,
                           dy if vertical else dx, dt, p.ramp, p.taper)
    k1 = ky[:nxy] if vertical else kx[:nxy]
    sig, skew = area_signals(eta, k1, kw, meandepth, dy if vertical else dx, r2d, p.U)
    ng, sg = area_shapes(x, y, kx, ky, line_pts, meandepth, p.damp, vertical, p.U)
    # ---- fun_operator_setup
    KX, KY = kx[None, :], ky[:, None]
    kw_ = {}
    if flat:
        cpeak = float(S.exact_cp1d(kp, p.depth, p.U))
        om = S.exact_disp2d(KX, KY, p.depth, p.mdir, p.current[0], p.current[1])
        kw_ = dict(c2d=S.exact_cp2d(KX, KY, p.depth, p.current[0], p.current[1]), l2d=om * om / S.GRAV)
        ops = {}
    else:
        cpeak = float(S.exact_cp1d(kp, meandepth, p.U))
        if p.ninterp == 2:
            ops = S.disp_interpolation_2p(-bathy, kx, ky, wp, p.current[0], p.current[1], p.mdir)
        elif p.ninterp == 3:
            ops = S.disp_interpolation_3p(-bathy, kx, ky, wp, p.mid, p.current[0], p.current[1], p.mdir)
        else:
            raise ValueError(f"depth_ref = {p.ninterp}: the reference accepts 2 or 3")
    brk = None
    if p.breaking == "yes":
        brk = S.breaking_params(Hs, Tp, kp, cfsa, chi, KBC=p.KBC, TC=p.TC, Tchar=p.Tchar, delb=p.delb)
    prop = {"Uni+": 0, "Uni-": 1}.get(p.propagation, 2)
    return S.Model2D(nx=nx, ny=ny, x=x, y=y, kx=kx, ky=ky, flat=flat, cpeak=cpeak, aal=aal, fbdy=fbdy,
                     wallchar=wallchar, chi=chi, depth=-bathy, wall=wall, refl_one=wall and p.refl_coef == 1.0,
                     grad_wall_x=gwx, grad_wall_y=gwy, ops=ops, source_on=True,
                     vertical=vertical, propagation=prop, new_gam=ng, skew_gam=sg, t_sig=t, sig=sig,
                     skew_sig=skew, eps_abs=eps_abs, eps_rel=eps_rel, breaking=brk, order=p.order,
                     meta=dict(name=os.path.basename(os.path.normpath(case_dir)), tp=Tp, hs=Hs, kp=kp, lam=lam,
                               wp=wp, meandepth=meandepth, line=line_pts, t_end=p.t_end, params=p),
                     **kw_)
