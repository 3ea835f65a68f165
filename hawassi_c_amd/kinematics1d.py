"""Internal flow kinematics of the 1-D model at output times.

The reference computes them in `kinematic_modul` (1d/kinematic_modul.c:1-247)
after each output time of `ode_solver_Ham` when `kinematic = yes`
(1d/functions.c:655-698), on levels and a window set up in
`modified_influx` (1d/functions.c:3603-3667).  This module builds that setup
on the host (the output levels, the two sinhcosh2IP weight tables) and runs
the computation through `hw1d_kinematics` on the GPU.

Setup restated here:

* `kin_levels` -- the output levels (1d/functions.c:3617-3650): `nsurf`
  equidistant levels over [min_zsurf, max_zsurf], then `ndeep` deeper levels
  whose spacing keeps the linear-wave flux between them at Flux / ndeep
  (upper Riemann sum of sinh(kp (D - z)) / sinh(kp D) at the peak
  wavenumber kp of the deepest point), all sorted ascending;
* `sinhcosh2ip` -- the two-depth weights (1d/functions.c:1889-1930):
  gamma_-(D), gamma_+(D) solve gamma_- cosh(kap Dmin) + gamma_+ cosh(kap Dplus)
  = cosh(kap D) and the same with sinh, kap = invers_omega(nu_peak, D)
  (kap = k[2] at D = 0), by Cramer's rule;
* `Kinematics1D` -- the tables on H in [min_zsurf + min depth, max_zsurf +
  max depth] and on DZ in [0, max_zsurf + max depth] (NH + 1 = 101 points,
  1d/kinematic_modul.c:49-82), the window x[indxIP1 .. indxIP2] by
  `closest` (1d/functions.c:3607-3613).

The runup form (a bathymetry with a dry node, 1d/kinematic_modul.c:19-46)
is not covered.
"""
from __future__ import annotations

import ctypes as ct

import numpy as np

from . import _lib
from .setup1d import GRAV, closest, invers_omega

NH = 100                      # table intervals (1d/kinematic_modul.c:12)


def kin_levels(zsurf1: float, zsurf2: float, nsurf: int, ndeep: int, nu_peak: float, maxd: float) -> np.ndarray:
    """Output levels zoutIP (1d/functions.c:3617-3650), ascending."""
    dzsurf = (zsurf2 - zsurf1) / (nsurf - 1)
    kpeak = invers_omega(nu_peak, maxd)
    zsurf = zsurf1 + np.arange(nsurf) * dzsurf
    flux = 1.0 / (kpeak * np.sinh(kpeak * maxd)) * (np.cosh(kpeak * (maxd + zsurf1)) - 1.0)
    tol = flux / ndeep
    zdeep = np.empty(ndeep)
    zdeep[0] = -1.0 * (zsurf1 - dzsurf)
    for i in range(1, ndeep):
        zdeep[i] = zdeep[i - 1] + (tol / (np.sinh(kpeak * (maxd - zdeep[i - 1])) / np.sinh(kpeak * maxd)))
    return np.sort(np.concatenate([zsurf, -zdeep]), kind="stable")


def sinhcosh2ip(D: np.ndarray, dmin: float, dplus: float, nu_peak: float, k2: float):
    """gamma_min, gamma_plus of sinhcosh2IP (1d/functions.c:1889-1930)."""
    D = np.asarray(D, float)
    kap = np.array([k2 if d == 0 else invers_omega(nu_peak, float(d)) for d in D])
    cmin, cplus, cd = np.cosh(kap * dmin), np.cosh(kap * dplus), np.cosh(kap * D)
    smin, splus, sd = np.sinh(kap * dmin), np.sinh(kap * dplus), np.sinh(kap * D)
    det = cmin * splus - cplus * smin
    gmin = (splus * cd - cplus * sd) / det
    gplus = (-smin * cd + cmin * sd) / det
    return gmin, gplus


class Kinematics1D:
    """Setup of kinematic_modul for one 1-D model and its GPU evaluation.

    x, depth (positive down), k (FFT order, `freqspace1d`) of the model;
    nu_peak the peak frequency; zsurf1 / zsurf2 = min_zsurf / max_zsurf,
    nsurf / ndeep the level counts, xinterv1 / xinterv2 the window
    (&IF KINEMATIC of input.txt)."""

    def __init__(self, x, depth, k, nu_peak: float, zsurf1: float, zsurf2: float, nsurf: int, ndeep: int,
                 xinterv1: float, xinterv2: float, grav: float = GRAV):
        self.x = np.ascontiguousarray(x, float)
        self.depth = np.ascontiguousarray(depth, float)
        self.k = np.ascontiguousarray(k, float)
        self.n = self.x.size
        if np.any(self.depth == 0.0):
            raise NotImplementedError("kinematics on a bathymetry with a dry node (the runup form) is not covered")
        self.grav = float(grav)
        self.nu_peak = float(nu_peak)
        maxd, mind = float(self.depth.max()), float(self.depth.min())
        self.flat = int(mind == maxd)
        self.hplus = zsurf2 + maxd
        self.hmin = zsurf1 + mind
        dh = (self.hplus - self.hmin) / NH
        self.htab = self.hmin + np.arange(NH + 1) * dh
        self.gam_hmin, self.gam_hplus = sinhcosh2ip(self.htab, self.hmin, self.hplus, self.nu_peak, self.k[2])
        self.dztab = 0.0 + np.arange(NH + 1) * ((self.hplus - 0.0) / NH)
        self.gam_dzmin, self.gam_dzplus = sinhcosh2ip(self.dztab, 0.0, self.hplus, self.nu_peak, self.k[2])
        self.ix0 = closest(self.x, xinterv1)
        self.nxout = closest(self.x, xinterv2) - self.ix0 + 1
        self.z = kin_levels(zsurf1, zsurf2, nsurf, ndeep, self.nu_peak, maxd)
        self.dx = float(self.x[1] - self.x[0])

    def config(self):
        """The hw_kin1d_config of this setup (its arrays are kept alive by self)."""
        p = lambda a: a.ctypes.data_as(ct.POINTER(ct.c_double))  # noqa: E731
        self._keep = [np.ascontiguousarray(a, float) for a in
                      (self.depth, self.k, self.z, self.htab, self.gam_hmin, self.gam_hplus, self.dztab,
                       self.gam_dzmin, self.gam_dzplus)]
        d, k, z, ht, ghm, ghp, dzt, gdm, gdp = self._keep
        return _lib.HwKin1DConfig(self.n, self.ix0, self.nxout, z.size, self.flat, ht.size, self.dx, self.grav,
                                  self.hmin, self.hplus, p(d), p(k), p(z), p(ht), p(ghm), p(ghp), p(dzt), p(gdm),
                                  p(gdp))

    def run(self, eta, u_hat, deta, dphi_hat, t: float, tprev: float, velx=None, velz=None) -> dict:
        """kinematic_modul at time t: eta, deta physical [n]; u_hat, dphi_hat
        complex [n] (unnormalised forward DFT).  Returns the [nxout][nz]
        fields dtPhi, dxPhi, dzPhi, P, acc_x, acc_z."""
        lib = _lib.load()
        cfg = self.config()
        dp = lambda a: a.ctypes.data_as(ct.POINTER(ct.c_double))  # noqa: E731
        eta = np.ascontiguousarray(eta, float)
        deta = np.ascontiguousarray(deta, float)
        uh = np.ascontiguousarray(np.asarray(u_hat, complex).view(float))
        dph = np.ascontiguousarray(np.asarray(dphi_hat, complex).view(float))
        m = self.nxout * self.z.size
        vx = None if velx is None else np.ascontiguousarray(velx, float).reshape(m)
        vz = None if velz is None else np.ascontiguousarray(velz, float).reshape(m)
        out = np.empty(6 * m)
        rc = lib.hw1d_kinematics(ct.byref(cfg), dp(eta), dp(uh), dp(deta), dp(dph), float(t), float(tprev),
                                 None if vx is None else dp(vx), None if vz is None else dp(vz), dp(out))
        _lib.check(rc)
        out = out.reshape(6, self.nxout, self.z.size)
        return dict(dtPhi=out[0], dxPhi=out[1], dzPhi=out[2], P=out[3], acc_x=out[4], acc_z=out[5])
