"""A 1-D reference case directory (input.txt and its data files) turned into
a ``Model1D``, and its run with the reference's outputs.

``case_from_dir(case_dir)`` follows the reference program's set-up order
(1d/main.cc:40-205):

1. ``read_input`` (io1d) -- read_input_file and influxlib's reads;
2. the influx record: user_defined influx.dat (read_data), harmonic
   (harmonic_influx, 1d/influx_lib.c:110-133), zero (1d/influx_lib.c:17-31);
   the JONSWAP generator draws its phases from ``srand(time(0))``
   (1d/influx_lib.c:195), so a JONSWAP case is taken as a user_defined
   influx.dat record;
3. the grid (1d/influx_lib.c:92-106): Nx = int(length/dx + 1) made even,
   dx = length/(Nx-1), Xinflux moved to the nearest node;
4. bath.dat (define_bath, 1d/functions.c:804-830), gauges.dat;
5. set_influx (1d/functions.c:1320-1329): nu_peak at the largest spectral
   amplitude of the record over the first half of the FFT (the last index
   among equal maxima); kp, Tp_peak, lambda_p, Hs_in = 4 sqrt(var),
   MaxEtaInit = 0.8 Hs_in/2;
6. modified_influx (1d/functions.c:1968-2041): for user_defined records the
   start and end ramp max(sign(t - t0 - nTramp Tp_peak), (1 - cos((t - t0)
   pi/(nTramp Tp_peak)))/2) times its reverse; insig_skew = cumtrapz(insig);
   Point: both filtered by Ug(kappa(omega)); AreaShort: the Ug(kappa(omega))
   weights normalised to sqrt(g D_inf)/Ug(0);
7. x_space (1d/x_space.c:19-163): damping zone and coefficient, k,
   the generation kernel gamX (Point: exp(-i k (Xinflux - Xstart))/dx;
   Area: FFT of the real part of IFFT(Ug) rolled to the influx node and
   normalised so that gamX[0] = sqrt(g D_inf)/dx; AreaShort: the same,
   windowed by a Gaussian of +-2 lambda_p, with the record re-filtered by
   Ug/|gamX| in frequency), gamX_skew = i omega(k) gamX, ChiAdj;
8. initial.dat (1d/main.cc:89-123, set_initial_ode 1d/functions.c:314-358);
9. parambreak (1d/breaking.c:1-33); define_operator (flat: Csqr, L;
   varying: Up3IP with the mid depth given by mid_depth_ref,
   1d/functions.c:2828-2860, 2186-2193, 3560-3562); alias with cutfracwn.

``run_case(case, out_dir)`` is ode_solver_Ham's loop (1d/functions.c:528-782):
integrate to every sample time of the record and write the gauge file and
the processed influx.  A wall (define_wall / wall_setup) adds its skew source
kernel and ChiAdj factor; the wall node's output is scaled by 1 + refl_coef.
With `kinematic = yes` it also writes the kinematics files
(Hawassi_Elev/Pressure/Vel/Acc_XX.dat, `kin_setup`, hw1d_kinematics).
Not supported (they raise): runup (a bathymetry above the still water level),
coupling, friction, and kinematics output together with breaking or a wall.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np

from . import io1d
from . import setup1d as S
from .setup2d import spline_coeffs, _spline_eval_vec

GRAV = S.GRAV


@dataclass
class Case1D:
    params: io1d.Input1D
    model: S.Model1D
    x: np.ndarray
    t_out: np.ndarray               # ttrace: every sample of the record
    gauges: np.ndarray
    y0: np.ndarray                  # initial state, 4 planes
    insig: np.ndarray               # the processed influx record
    derived: dict = field(default_factory=dict)


def harmonic_record(T0: float, Tend: float, Tp: float, dt: float, ampli: float):
    """harmonic_influx (1d/influx_lib.c:110-133): samples T0 + i dt while the
    previous sample is <= Tend, ramped over 4 Tp."""
    ts, es = [], []
    t, i = T0, 0
    while t <= Tend:
        t = T0 + i * dt
        r1 = -1.0 if (t - 4 * Tp) < 0 else 1.0
        r2 = (1 - math.cos(t * math.pi / 4 / Tp)) / 2
        ts.append(t)
        es.append(ampli * max(r1, r2) * math.sin(t * 2 * math.pi / Tp))
        i += 1
    return np.asarray(ts), np.asarray(es)


def peak_frequency(eta: np.ndarray, t: np.ndarray) -> float:
    """set_influx / nsqrspec / search_max (1d/functions.c:1256-1280,
    1320-1329): |ww| at the largest |FFT(eta)| among the first n/2 entries,
    the last of equal maxima winning."""
    n = len(eta)
    a = np.abs(np.fft.fft(eta))
    sp = np.sqrt(a * a / np.sum(a * a))
    ww = S.freqspace1d(t[-1] - t[0], n)
    idx, best = 0, sp[0]
    for i in range(n // 2):
        if sp[i] >= best:
            best, idx = sp[i], i
    return abs(float(ww[idx]))


def influx_ramp(t: np.ndarray, ntramp: float, tp_peak: float) -> np.ndarray:
    """modified_influx's ramp times end ramp (1d/functions.c:1973-1995).  At
    the sample where t - t0 equals nTramp Tp_peak the reference's sign term is
    0/0 and the cosine term is taken.  With no ramp (nTramp = 0) the
    reference's cosine term is cos(inf) = NaN for every sample and the record
    becomes NaN; that is refused here."""
    if ntramp == 0:
        raise ValueError("ramp_influx = no with a user_defined record: the reference's ramp is NaN (x/0)")
    tt = t - t[0]
    a = tt - ntramp * tp_peak
    with np.errstate(invalid="ignore", divide="ignore"):
        b1 = a / np.abs(a)
        b2 = (1 - np.cos(tt * math.pi / ntramp / tp_peak)) / 2
    ramp = np.where(b1 > b2, b1, b2)
    return ramp * ramp[::-1]


def area_gamx(k, x, xinflux, xstart, dx, depthinf, lam, short: bool, shift=None):
    """The Area / AreaShort kernel of x_space (1d/x_space.c:55-82): the real
    part of IFFT(Ug) rolled by int((Xinflux - Xstart)/dx) nodes (or `shift`),
    transformed and normalised to gamX[0] = sqrt(g D_inf)/dx (scaling_influx,
    1d/functions.c:1747-1759).  define_wall builds its kernel the same way,
    rolled by the wall's node index (1d/functions.c:19-30)."""
    n = len(k)
    ug = np.array([S.group_velocity_i(float(kk), depthinf) for kk in k])
    sp = np.real(np.fft.ifft(ug))
    sp = np.roll(sp, int((xinflux - xstart) / dx) if shift is None else int(shift))
    if short:
        lb = 2 * (0.4 * lam) ** 2
        i1, i2 = S.closest(x, xinflux - 2 * lam), S.closest(x, xinflux + 2 * lam)
        win = np.zeros(n)
        win[i1:i2 + 1] = np.exp(-(x[i1:i2 + 1] - xinflux) ** 2 / lb)
        sp = sp * win
    gx = np.fft.fft(sp) * (1.0 / dx)
    scale = gx[0].real * dx
    gx = gx * (math.sqrt(GRAV * depthinf) / scale)
    return gx


def _spline(xk, yk, t):
    c = spline_coeffs(xk, yk)
    return _spline_eval_vec(xk, yk, c, t)


def case_from_dir(case_dir: str) -> Case1D:
    p = io1d.read_input(os.path.join(case_dir, "input.txt"))
    for name, bad in (("coupling", "yes"), ("friction", "yes")):
        if getattr(p, name) == bad:
            raise NotImplementedError(f"1-D case option {name} = {bad} is not supported")
    if p.order == 0:
        raise NotImplementedError(f"dynamics {p.dynmodel!r} is not supported (HS1, HS2, HS3)")
    # 2. influx record
    if p.wavename == "user_defined" or p.wavename == "jonswap":
        f = io1d.case_file(case_dir, "influx.dat")
        if f is None:
            raise FileNotFoundError("influx.dat: a JONSWAP case needs its record as a user_defined influx.dat "
                                    "(the reference generator seeds its phases from the clock)")
        ttrace, eta = io1d.read_signal(f)
    elif p.wavename == "harmonic":
        ttrace, eta = harmonic_record(p.T0, p.Tend, p.Tp, p.dt, 0.5 * p.Hs)
    elif p.wavename == "zero":
        nt = int((p.Tend - p.T0) / p.dt + 1)
        ttrace, eta = p.T0 + p.dt * np.arange(nt), np.zeros(nt)
    else:
        raise NotImplementedError(f"wavename {p.wavename!r}")
    n_t = len(ttrace)
    dt = float(ttrace[1] - ttrace[0])
    # 3. grid
    length = p.Xend - p.Xstart
    nx = int(length / p.dx + 1)
    if nx % 2:
        nx -= 1
    dx = length / (nx - 1)
    x = p.Xstart + dx * np.arange(nx)
    idx_inf = S.closest(x, p.Xinflux)
    xinflux = float(x[idx_inf])
    # 4. bathymetry and gauges
    if p.bath == "user_defined":
        depth = io1d.read_column(os.path.join(case_dir, "bath.dat"), nx)
        if (depth < 0).any():
            raise NotImplementedError("runup (a bathymetry above the still water level) is not supported")
        depthinf = float(depth[idx_inf])
    elif p.bath == "flat":
        depth = np.full(nx, p.depthflat)
        depthinf = p.depthflat
    else:
        raise NotImplementedError(f"bathymetry {p.bath!r}")
    gauges = io1d.read_column(os.path.join(case_dir, "gauges.dat"), p.ngauge) if p.ngauge > 0 else np.zeros(0)
    # 5.-6. influx spectrum and processing
    d = dict(nu_peak=0.0, kp=0.0, Tp_peak=0.0, lambda_p=0.0, Hs_in=0.0, MaxEtaInit=0.0)
    insig = eta.copy()
    insig_skew = np.zeros(n_t)
    kom = gvom = None
    if p.wavename != "zero":
        nu = peak_frequency(eta, ttrace)
        kp = S.invers_omega(nu, depthinf)
        tp_peak = 2 * math.pi / nu
        if p.wavename == "jonswap":
            nu = 2 * math.pi / p.Tp
            kp = S.invers_omega(nu, depthinf)
        hs_in = 4 * math.sqrt(float(np.mean((eta - np.mean(eta)) ** 2)))
        d.update(nu_peak=nu, kp=kp, Tp_peak=tp_peak, lambda_p=2 * math.pi / kp, Hs_in=hs_in,
                 MaxEtaInit=hs_in / 2 * 0.8)
        if p.wavename == "user_defined":
            insig = influx_ramp(ttrace, p.nTramp, tp_peak) * insig
        insig_skew = S.cumtrapz(insig, dt)
        ww = S.freqspace1d(ttrace[-1] - ttrace[0], n_t)
        kom = np.array([0.0 if w == 0 else S.invers_omega(float(w), depthinf) for w in ww])
        gvom = np.array([S.group_velocity_i(float(kk), depthinf) for kk in kom])
        if p.influxing == "Point":
            insig = np.real(np.fft.ifft(gvom * np.fft.fft(insig)))
            insig_skew = np.real(np.fft.ifft(gvom * np.fft.fft(insig_skew)))
        elif p.influxing == "AreaShort":
            insig_hat = np.fft.fft(insig)
            gvom = gvom * math.sqrt(GRAV * depthinf) / gvom[0]
    # 7. x_space
    kp, lam = d["kp"], d["lambda_p"]
    dampchar = S.dampzone(x, p.Xstart + p.dampL, p.Xend - p.dampR, p.dampL, p.dampR)
    if p.wavename != "zero":
        cp_l, cp_r = S.omega_i(kp, float(depth[0])) / kp, S.omega_i(kp, float(depth[-1])) / kp
    else:
        cp_l, cp_r = math.sqrt(GRAV * depth[0]), math.sqrt(GRAV * depth[-1])
    dampcoef = np.where(np.arange(nx) < nx // 2, 7 * cp_l / (0.9 * p.dampL), 7 * cp_r / (0.9 * p.dampR))
    k = S.freqspace1d(length, nx)
    gx = gs = None
    if p.wavename != "zero":
        if p.influxing == "Point":
            gx = np.exp(-1j * k * (xinflux - p.Xstart)) / dx
        else:
            gx = area_gamx(k, x, xinflux, p.Xstart, dx, depthinf, lam, p.influxing == "AreaShort")
            if p.influxing == "AreaShort":
                # the record re-filtered by Ug(kappa) / |gamX(kappa)| / dx
                h = nx // 2
                absg = np.abs(np.concatenate([gx[h:], gx[:h]]))
                ktemp = np.concatenate([k[h:], k[:h]])
                ev = _spline(ktemp, absg, kom)
                gxom = np.where(np.abs(kom) > abs(ktemp[-1]), absg[-1], ev)
                insig = np.real(np.fft.ifft(insig_hat * gvom / gxom / dx))
                insig_skew = S.cumtrapz(insig, dt)
        gs = 1j * S.omega_v(k, depthinf) * gx
    if p.order == 1 or p.wavename == "zero":
        chi = np.ones(nx)
    else:
        chi = S.chi_adjust(x, xinflux, p.adjcoef, lam, p.propagation)
    # 8. initial state
    y0 = np.zeros(4 * nx)
    if p.initial == "user_defined":
        w0, u0 = io1d.read_initial(os.path.join(case_dir, "initial.dat"), nx)
        a = np.abs(np.fft.fft(w0))
        i_max = 0
        best = a[0]
        for i in range(nx // 2):
            if a[i] >= best:
                best, i_max = a[i], i
        kp0 = abs(float(k[i_max])) if i_max != 0 else abs(float(k[1]))
        d["nu_peak"] = S.omega_i(kp0, float(np.mean(depth)))
        d["kp"] = kp0
        hs_here = 2 * float(np.max(w0))
        if hs_here > d["Hs_in"]:
            d["Hs_in"] = hs_here
        d["MaxEtaInit"] = d["Hs_in"] / 2 * 0.8
        e_hat = np.fft.fft(w0 * (1 - dampchar))
        u_hat = np.fft.fft(u0 * (1 - dampchar))
        y0 = np.concatenate([e_hat.real, e_hat.imag, u_hat.real, u_hat.imag])
    # 9. operators, anti-aliasing, breaking
    flat = p.bath == "flat"
    m = S.Model1D(n=nx, x=x, dx=dx, k=k, aal=S.alias1d(k, p.cutfracwn), chi=chi, damp=dampcoef * dampchar,
                  depth=depth, flat=flat, source_on=p.wavename != "zero", propagation=p.propagation,
                  gamx=gx, gamx_skew=gs, t_sig=ttrace, sig=insig, sig_skew=insig_skew, order=p.order,
                  meta=dict(name=os.path.basename(os.path.normpath(case_dir)), xinflux=xinflux,
                            depthinf=depthinf, **d))
    if flat:
        m.csqr, m.lop = S.flat_operators(k, depthinf)
    else:
        m.ops = S.up3ip(depth, k, d["nu_peak"], mid=p.mid_temp if p.mid_temp else "default")
    if p.breaking == "yes" and p.order != 1:
        tstar = 5 * math.sqrt(depthinf * GRAV) if p.initial != "zero" else p.parTchar * d["Tp_peak"]
        w = S.breaking_window(x, xinflux, p.adjcoef, lam, p.propagation)
        m.breaking = dict(KBC=p.parKBC, TC=p.parTC, Tstar=tstar, MaxEtaInit=d["MaxEtaInit"], delb=1.2, **w)
    if p.wall == "yes":
        m.wall = wall_setup(p, m, x, k, dx, depthinf, xinflux, lam, float(ttrace[0]))
    return Case1D(params=p, model=m, x=x, t_out=ttrace, gauges=gauges, y0=y0, insig=insig, derived=d)


def wall_setup(p, m, x, k, dx, depthinf, xinflux, lam, t0) -> dict:
    """define_wall (1d/functions.c:1-47) and wall_setup (:49-66): the wall
    node, the skew source kernel i omega(k) gamW (negated when the wall lies
    on the influx's left, wallInfdir = 1), and ChiAdj multiplied by
    1 - cf(x, Xwall - adjcoef lambda_p) + cf(x, Xwall) (m.chi in place)."""
    if lam == 0:
        raise NotImplementedError("a wall without an influx record (lambda_p = 0) is not supported")
    idx = S.closest(x, p.Xwall)
    gw = area_gamx(k, x, xinflux, x[0], dx, depthinf, lam, short=False, shift=idx)
    gws = 1j * S.omega_v(k, depthinf) * gw
    if not p.Xwall > xinflux:
        gws = -gws
    dl = p.adjcoef * lam
    m.chi = m.chi * (1 - S.cf(x, p.Xwall - dl, dl) + S.cf(x, p.Xwall, dl))
    return dict(idx=idx, refl=p.refl_coef, t0=t0, gamw_skew=gws)


def run_case(case: Case1D, out_dir: str, precision=None, progress=None, fields: bool = True) -> dict:
    """ode_solver_Ham's loop (1d/functions.c:537-552, 604-666, 700-744): from
    t = ttrace[0] with h = 1e-4, integrate to each sample time and write
    Hawassi_eta_gauge.txt and (fields=True) the partitioned Constants_XX.dat
    and Hawassi_eta/u/phi_XX.dat; influx.dat gets the processed record
    (read_influx_file).  Returns the gauge series and the solver counters."""
    from . import _lib
    from .model1d import Hawassi1D
    os.makedirs(out_dir, exist_ok=True)
    if case.params.wavename != "zero":
        io1d.write_influx(os.path.join(out_dir, "influx.dat"), case.t_out, case.insig)
    prec = _lib.HW_PREC_FP64 if precision is None else precision
    hw = Hawassi1D([case.model], precision=prec, t0=float(case.t_out[0]), h0=1e-4)
    hw.set_state(case.y0)
    gw = io1d.GaugeWriter1D(os.path.join(out_dir, "Hawassi_eta_gauge.txt"), case.gauges, case.x) \
        if len(case.gauges) else None
    m = case.model
    fw = io1d.FieldWriter1D(out_dir, case.x, case.t_out, m.depth, m.k, case.derived["nu_peak"],
                            case.params.npartition) if fields else None
    n = m.n
    # phi^ = (-i / k) u^, 0 where k = 0 (1d/functions.c:576-590)
    with np.errstate(divide="ignore"):
        a = np.where(m.k == 0, 0.0, -1.0 / np.where(m.k == 0, 1.0, m.k))
    vals = np.zeros((len(case.t_out), len(case.gauges)))
    kin = kin_setup(case) if case.params.kinematic == "yes" else None
    kw = hk = None
    if kin is not None:
        kw = io1d.KinWriter1D(out_dir, case.t_out, case.params.npartition, case.x[kin.ix0:kin.ix0 + kin.nxout],
                              kin.z)
        # d/dt at the output time from a context of its own (no breaking or
        # wall state to disturb there, see kin_setup)
        hk = Hawassi1D([case.model], precision=prec, t0=float(case.t_out[0]), h0=1e-4)
    try:
        for i, t in enumerate(case.t_out):
            if i > 0:
                hw.integrate(float(t))
            y = hw.get_state()
            eta = np.real(np.fft.ifft(y[:n] + 1j * y[n:2 * n]))
            if m.wall is not None:
                # the wall node's output (1d/functions.c:598-601)
                eta[m.wall["idx"]] *= 1 + m.wall["refl"]
            if fw is not None:
                uh = y[2 * n:3 * n] + 1j * y[3 * n:]
                u = np.real(np.fft.ifft(uh))
                phi = np.real(np.fft.ifft(-(a * uh.imag) + 1j * (a * uh.real)))
                fw.write(eta, u, phi)
            if gw is not None:
                vals[i] = gw.write(float(t), eta)
            if kw is not None:
                # kinematic_modul(t, etaf, u_hatf, dtetaf, dtphi_hatf) with
                # dy = f(t, y) (1d/functions.c:568-595, 691) and tprev = t, as
                # the reference's loop leaves it (1d/functions.c:552)
                dy = hk.rhs(float(t), y)
                uh = y[2 * n:3 * n] + 1j * y[3 * n:]
                deta = np.real(np.fft.ifft(dy[:n] + 1j * dy[n:2 * n]))
                dvh = dy[2 * n:3 * n] + 1j * dy[3 * n:]
                dph = -(a * dvh.imag) + 1j * (a * dvh.real)
                with np.errstate(divide="ignore", invalid="ignore"):
                    res = kin.run(eta, uh, deta, dph, float(t), float(t))
                kw.write(eta[kin.ix0:kin.ix0 + kin.nxout], res)
            if progress is not None and i % 500 == 0:
                progress(i, float(t))
        stats = hw.stats(0)
    finally:
        if gw is not None:
            gw.close()
        if fw is not None:
            fw.close()
        if kw is not None:
            kw.close()
        if hk is not None:
            hk.close()
        hw.close()
    return dict(t=case.t_out, gauges=vals, stats=stats)


def kin_setup(case: Case1D):
    """The kinematics setup of a case with `kinematic = yes`: levels over
    [zsurf1, zsurf2] from read_kinematics (1d/functions.c:3555-3582: the
    expert min/max_zsurf, or -/+2.5 Hs_in, -/+0.6 Hs_in with an initial
    elevation) and the window [xinterv1, xinterv2].  The d/dt the module
    needs comes from an extra right-hand side at the output time, which
    breaking and the wall source would not take free of side effects, so
    those cases are refused."""
    from .io2d import _atof
    from .kinematics1d import Kinematics1D
    p, m = case.params, case.model
    if m.breaking is not None or m.wall is not None:
        raise NotImplementedError("kinematics output with breaking or a wall is not covered")
    hs = case.derived["Hs_in"]
    user = p.initial == "user_defined"
    z1 = (-0.6 if user else -2.5) * hs if p.min_z == "default" else _atof(p.min_z)
    z2 = (0.6 if user else 2.5) * hs if p.max_z == "default" else _atof(p.max_z)
    return Kinematics1D(case.x, m.depth, m.k, case.derived["nu_peak"], z1, z2, p.nsurf, p.ndeep, p.xinterv1,
                        p.xinterv2)
